/*
 * pf_runtime.h -- device-memory, stream and event helpers exported by the
 * library so that C and Python callers can keep inputs resident in HBM and
 * time kernels on the stream they are launched on, without linking HIP
 * themselves.  No counterpart in the reference (its callers use the CUDA
 * runtime directly, e.g. tests/bench_gpu.cu:38-66).
 *
 * All functions act on the calling thread's current device and return
 * 0 on success, -1 on a HIP error (pointer-returning ones return NULL).
 */
#pragma once
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int         pf_device_count(void);
const char *pf_device_name(void);          /* static buffer; "" on error */
int         pf_device_arch(char *buf, size_t cap); /* e.g. "gfx950" */
int         pf_device_cu_count(void);      /* -1 on error */

void *pf_malloc(size_t bytes);
int   pf_free(void *dptr);
int   pf_memcpy_htod(void *dst, const void *src, size_t bytes);
int   pf_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int   pf_memcpy_dtod(void *dst, const void *src, size_t bytes);
int   pf_memset(void *dst, int value, size_t bytes);
int   pf_sync(void);
/* Streaming device-to-device copy kernel on `stream` (bytes and both
 * pointers 16-byte aligned); also the calibration point for HBM counters. */
int   pf_device_copy(void *dst, const void *src, size_t bytes, void *stream);

void *pf_stream_create(void);              /* default priority */
int   pf_stream_destroy(void *stream);
int   pf_stream_sync(void *stream);

void *pf_event_create(void);
int   pf_event_destroy(void *ev);
int   pf_event_record(void *ev, void *stream);
int   pf_event_sync(void *ev);
float pf_event_elapsed_ms(void *start, void *stop);   /* -1 on error */

#ifdef __cplusplus
}
#endif
