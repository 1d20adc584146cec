/* voss.h — error codes, error strings and memory release of the voss C ABI.
 *
 * Drop-in for the reference header library/include/voss/voss.h:9-25: the
 * same codes, values and entry points, so existing ctypes / C callers bind
 * unchanged.  Implemented by voss_amd/csrc/capi.cpp.
 */
#ifndef VOSS_H
#define VOSS_H

#ifdef __cplusplus
extern "C" {
#endif

#define VOSS_OK                 0
#define VOSS_ERR_INVALID_N      1
#define VOSS_ERR_OUT_OF_RANGE   2
#define VOSS_ERR_NO_CUDA        3 /* no GPU device (name kept for ABI compatibility) */
#define VOSS_ERR_OUT_OF_MEMORY  4
#define VOSS_ERR_CUDA           5 /* device runtime error (HIP on this build) */
#define VOSS_ERR_INVALID_ARG    6
#define VOSS_ERR_INTERNAL       7

/* Fixed short text for a code (reference: library/src/core/error.cpp:20-32). */
const char* voss_strerror(int code);
/* Message of the last failure on the calling thread (error.cpp:10-18). */
const char* voss_get_last_error(void);
/* Release an array returned by the library; NULL is accepted (alloc.cpp:8-10). */
void voss_free(void* ptr);

#ifdef __cplusplus
}
#endif
#endif /* VOSS_H */
