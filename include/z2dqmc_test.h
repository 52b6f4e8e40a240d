/*
 * z2dqmc_test.h -- kernel test hooks of libz2dqmc (no counterpart in the
 * reference): single-matrix runs of the FP64 dense kernels used by the
 * stabilization, so tests can check each kernel against LAPACK semantics.
 */
#ifndef Z2DQMC_TEST_H
#define Z2DQMC_TEST_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
/* C = op(A) op(B); ta/tb: 0 = N, 1 = T; triA/triB: use the upper triangle only */
int z2_test_gemm(int32_t n, const double *A, int32_t ta, const double *B, int32_t tb,
                 int32_t triA, int32_t triB, double *C);
/* QDRP_decompose (src/stabilization.f90:42-68) in place on A; Q (optional) = explicit Q */
int z2_test_qrp(int32_t n, double *A, double *D, int32_t *ipvt, double *tau, double *Q);
/* the blocked QR with columns pre-ordered by norm (same outputs as z2_test_qrp,
 * ipvt the pre-ordering) and its explicit Q */
int z2_test_qr_blocked(int32_t n, double *A, double *D, int32_t *ipvt, double *tau, double *Q);
/* reps launches of a plain copy kernel (k_copy) of n doubles: reads and writes
 * 8 n bytes each, the reference traffic for scaling hardware byte counters */
int z2_test_copy(int64_t n, int32_t reps);
/* X <- R^{-1} X (right = 0, ZTRSM L,U,N,N) or X R^{-T} (right = 1, ZTRSM R,U,C,N) */
int z2_test_trsm(int32_t n, const double *R, double *X, int32_t right);
#ifdef __cplusplus
}
#endif
#endif
