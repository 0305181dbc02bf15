// Internal helpers shared by the library's translation units.
#ifndef LBM_INTERNAL_H
#define LBM_INTERNAL_H

#ifdef __cplusplus
extern "C" {
#endif

// Records a printf-style message as the last error and returns 1.
int lbm_fail(const char *fmt, ...);

#ifdef __cplusplus
}
#endif
#endif
