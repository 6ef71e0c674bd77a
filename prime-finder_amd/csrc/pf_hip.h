// pf_hip.h -- shared HIP helpers for the gfx950 kernels of this library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

// Evaluate a HIP call; on failure print "<where>: HIP error <name> (<call>)"
// and jump to the enclosing function's `fail:` label.
#define PF_TRY(where, call)                                                    \
    do {                                                                       \
        hipError_t pf_e_ = (call);                                             \
        if (pf_e_ != hipSuccess) {                                             \
            fprintf(stderr, "%s: HIP error %s (%s)\n", where,                  \
                    hipGetErrorName(pf_e_), #call);                            \
            goto fail;                                                         \
        }                                                                      \
    } while (0)

// 32-bit modular arithmetic (host and device).
#include "pf_mont.h"
