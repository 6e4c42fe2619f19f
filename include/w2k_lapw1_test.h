/*
 * w2k_lapw1_test.h -- internal kernels of libw2klapw1.so exposed for tests.
 * No counterpart in the reference; used by tests/ to check building blocks
 * (the FP64 MFMA GEMM) in isolation, and by bench.py to calibrate the
 * profiler's byte counters with a copy kernel.  Host pointers; the call allocates,
 * uploads, runs on the device and copies C back.
 */
#ifndef W2K_LAPW1_TEST_H
#define W2K_LAPW1_TEST_H
#ifdef __cplusplus
extern "C" {
#endif

/* C = alpha op(A) op(B) + beta C on device 0; tri/atri/btri as in the
 * library's internal GEMM (0 dense, 1 lower/upper, ...). */
int w2k_test_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                   const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri);

/* The same dense product with a device workspace of wsn doubles, which lets
 * a product with few output tiles and a long K split K across the chip. */
int w2k_test_dgemm_ws(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                      const double *B, int ldb, double beta, double *C, int ldc, long wsn);

/* reps launches of a plain device-to-device copy of n doubles (n even) on
 * device 0, for calibrating the HBM byte counters of a profiler; the mean
 * time of one copy in *ms_per_copy. */
int w2k_test_copy(long n, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
