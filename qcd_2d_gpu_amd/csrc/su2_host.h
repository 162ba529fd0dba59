/*
 * su2_host.h - internal declarations shared by the C++/HIP translation units
 * and the C host driver of libsu2mc (not part of the public ABI: hidden in
 * the shared library).
 */
#ifndef SU2_HOST_H
#define SU2_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* Records a printf-style message for su2_last_error() of the calling thread
 * and returns code. */
__attribute__((visibility("hidden"))) int su2_set_error(int code, const char *fmt, ...);

#ifdef __cplusplus
}
#endif

#endif /* SU2_HOST_H */
