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

/* The same product on a given tile kernel (variant 0: 64x64 tiles, 1:
 * 128x128 tiles), the two choices of the library's GEMM tile autotuner,
 * which must give bit-identical results. */
int w2k_test_dgemm_tiles(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                         const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri,
                         int variant);

/* The same product with a device workspace of wsn doubles, which lets a
 * product with few output tiles and a long K, or a product with a
 * triangular operand or output and a long K, split K across the chip
 * (tri / atri / btri as in w2k_test_dgemm). */
int w2k_test_dgemm_ws(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                      const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri,
                      long wsn);

/* Split-precision C (+)= A B^T (A: m x k, B: n x k, column-major FP64) with
 * `slices` INT8 digits per element (3..7), the kernel behind
 * w2k_set_split_precision; tri 0 full, 1 lower; accumulate 1 adds to C. */
int w2k_test_split_gemm(int m, int n, int k, const double *A, int lda, const double *B, int ldb, double *C, int ldc,
                        int tri, int accumulate, int slices);

/* reps launches of a plain device-to-device copy of n doubles (n even) on
 * device 0, for calibrating the HBM byte counters of a profiler; the mean
 * time of one copy in *ms_per_copy. */
int w2k_test_copy(long n, int reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
