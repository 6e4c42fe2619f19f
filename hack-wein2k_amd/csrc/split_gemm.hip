// Split-precision product C (+)= A B^T on the INT8 matrix cores, with FP64
// inputs and output (the HNS rank-K update in split precision).
//
// Every row r of A (and of B) is scaled by a power of two 2^E_r with
// max_k |X(r,k)| < 2^E_r and cut into S signed 7-bit digits:
//   X(r,k) = 2^E_r * sum_{p<S} d_p(r,k) 2^(-7(p+1)) + residual,
//   |residual| < 2^(E_r - 7S),  d_p in [-127, 127]
// (the digits are exact: scaling by 2^7 and removing the integer part are
// exact in FP64).  Then
//   (A B^T)(i,j) = 2^(E_i + F_j - 14) * sum_d 2^(-7d) * acc_d(i,j),
//   acc_d = sum_k sum_{p+q=d} dA_p(i,k) dB_q(j,k),
// keeping the digit pairs with p + q <= S - 1.  Each acc_d is an exact INT32
// sum (v_mfma_i32_16x16x64_i8; flushed to FP64 every 4096 k, so no overflow
// for any K), and the recombination multiplies by exact powers of two, so the
// only departures from the exact product are the dropped digit pairs and
// the residual: |error(i,j)| <= K * (S + 2) * 2^(E_i + F_j - 7S).
//
// This is the extended form of the HNS update with no reference counterpart;
// it is pinned against the FP64 product (tests/test_gpu_split.py).
#include <algorithm>
#include <cstdint>

#include "w2k_internal.h"

namespace w2k {

typedef int v4i __attribute__((ext_vector_type(4)));

constexpr int SG_T = 64;          // output tile (rows and columns) and k slab
constexpr int SG_ROWB = 80;       // LDS bytes per digit row (64 + 16 padding: conflict-free b128 reads)
constexpr int SG_FLUSH = 4096;    // k between INT32 -> FP64 flushes
constexpr int SG_EMIN = -1100;    // exponent of an all-zero row

__global__ void k_fill_int(int *__restrict__ p, int cnt, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = v;
}

// row exponents: expo[r] = max over k of the frexp exponent of X(r,k)
// (expo must hold SG_EMIN on entry).  Block: 64 rows x 4 column phases;
// grid.y splits K.
__global__ __launch_bounds__(256) void k_row_expo(const double *__restrict__ X, int rows, int ld, int K, int kchunk,
                                                  int *__restrict__ expo)
{
  const int t = threadIdx.x;
  const int r = blockIdx.x * 64 + (t & 63);
  const int ph = t >> 6;
  const int k0 = blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
  __shared__ int red[4][64];
  int e = SG_EMIN;
  if (r < rows) {
    const double *x = X + r;
    int k = k0 + ph;
    for (; k + 12 < k1; k += 16) {
      const double a = fabs(x[(size_t)ld * k]), b = fabs(x[(size_t)ld * (k + 4)]);
      const double c = fabs(x[(size_t)ld * (k + 8)]), d = fabs(x[(size_t)ld * (k + 12)]);
      const double m = fmax(fmax(a, b), fmax(c, d));
      if (m > 0.0) e = max(e, ilogb(m) + 1);
    }
    for (; k < k1; k += 4) {
      const double a = fabs(x[(size_t)ld * k]);
      if (a > 0.0) e = max(e, ilogb(a) + 1);
    }
  }
  red[ph][t & 63] = e;
  __syncthreads();
  if (t < 64 && r < rows) {
    const int m = max(max(red[0][t], red[1][t]), max(red[2][t], red[3][t]));
    if (m > SG_EMIN) atomicMax(expo + r, m);
  }
}

// digits: D[p][r][k] (int8, rows padded to rows_pad, k to kpad, zero there).
// Block: 64 rows x 64 k, read column-major (coalesced over r), transposed
// through LDS, written k-contiguous.
template <int S>
__global__ __launch_bounds__(256) void k_digits(const double *__restrict__ X, int rows, int ld, int K,
                                                const int *__restrict__ expo, int8_t *__restrict__ D, int rows_pad,
                                                int kpad)
{
  __shared__ int8_t sd[S][64][64 + 4];
  const int t = threadIdx.x;
  const int r0 = blockIdx.x * 64, k0 = blockIdx.y * 64;
  const int lr = t & 63;
  const int r = r0 + lr;
  const int e = (r < rows) ? expo[r] : SG_EMIN;
  for (int kk = t >> 6; kk < 64; kk += 4) {
    const int k = k0 + kk;
    double y = 0.0;
    if (r < rows && k < K && e > SG_EMIN) y = ldexp(X[(size_t)r + (size_t)ld * k], -e);
#pragma unroll
    for (int p = 0; p < S; p++) {
      const double tt = y * 128.0;
      const double dd = trunc(tt);
      sd[p][lr][kk] = (int8_t)(int)dd;
      y = tt - dd;
    }
  }
  __syncthreads();
  // 64 rows x 64 bytes per digit plane: 16 bytes per thread per plane
  const int wr = t >> 2, wc = (t & 3) * 16;
#pragma unroll
  for (int p = 0; p < S; p++) {
    int v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int8_t *s = &sd[p][wr][wc + 4 * u];
      v[u] = (int)((uint32_t)(uint8_t)s[0] | ((uint32_t)(uint8_t)s[1] << 8) | ((uint32_t)(uint8_t)s[2] << 16) |
                   ((uint32_t)(uint8_t)s[3] << 24));
    }
    v4i *dst = (v4i *)(D + ((size_t)p * rows_pad + r0 + wr) * kpad + k0 + wc);
    *dst = v4i{v[0], v[1], v[2], v[3]};
  }
}

struct SplitArgs {
  int m, n, kpad;
  int rows_pad_a, rows_pad_b;
  const int8_t *DA, *DB;  // digit planes
  const int *ea, *eb;     // row exponents
  double *C;
  int ldc;
  int tri;                // 0 full, 1 lower (i >= j)
  int accumulate;         // 1: C += product, 0: C = product
};

// One 64 x 64 output tile per block; 4 waves in 2 x 2, each 32 x 32 as 2 x 2
// MFMA blocks of 16 x 16; one INT32 accumulator set per digit order d.
template <int S>
__global__ __launch_bounds__(256) void k_split_gemm(SplitArgs g)
{
  __shared__ __align__(16) int8_t As[S * SG_T * SG_ROWB];  // [S][64][SG_ROWB]
  __shared__ __align__(16) int8_t Bs[S * SG_T * SG_ROWB];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int i0 = blockIdx.x * SG_T, j0 = blockIdx.y * SG_T;
  if (g.tri == 1 && i0 + SG_T - 1 < j0) return;
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
  v4i acc[S][2][2];
  double acc64[2][2][4];
#pragma unroll
  for (int d = 0; d < S; d++)
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) acc[d][a][b] = v4i{0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc64[a][b][r] = 0.0;
  // global -> LDS: each thread moves 16 bytes per digit plane of A and of B
  const int lr = t >> 2, lc = (t & 3) * 16;
  const int8_t *ga = g.DA + (size_t)(i0 + lr) * g.kpad + lc;
  const int8_t *gb = g.DB + (size_t)(j0 + lr) * g.kpad + lc;
  const size_t pa = (size_t)g.rows_pad_a * g.kpad, pb = (size_t)g.rows_pad_b * g.kpad;
  int since = 0;
  for (int k0 = 0; k0 < g.kpad; k0 += SG_T) {
    v4i va[S], vb[S];
#pragma unroll
    for (int p = 0; p < S; p++) {
      va[p] = *(const v4i *)(ga + pa * p + k0);
      vb[p] = *(const v4i *)(gb + pb * p + k0);
    }
    __syncthreads();  // previous slab's fragment reads are done
#pragma unroll
    for (int p = 0; p < S; p++) {
      *(v4i *)(As + (p * SG_T + lr) * SG_ROWB + lc) = va[p];
      *(v4i *)(Bs + (p * SG_T + lr) * SG_ROWB + lc) = vb[p];
    }
    __syncthreads();
    // A fragments of all digit planes; B fragments one plane at a time
    v4i fa[S][2];
    const int fr = lane & 15, fk = (lane >> 4) * 16;
#pragma unroll
    for (int p = 0; p < S; p++)
#pragma unroll
      for (int a = 0; a < 2; a++) fa[p][a] = *(const v4i *)(As + (p * SG_T + wr + 16 * a + fr) * SG_ROWB + fk);
#pragma unroll
    for (int q = 0; q < S; q++) {
      v4i fb[2];
#pragma unroll
      for (int b = 0; b < 2; b++) fb[b] = *(const v4i *)(Bs + (q * SG_T + wc + 16 * b + fr) * SG_ROWB + fk);
#pragma unroll
      for (int p = 0; p + q < S; p++)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++)
            acc[p + q][a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[p][a], fb[b], acc[p + q][a][b], 0, 0, 0);
    }
    since += SG_T;
    if (since >= SG_FLUSH || k0 + SG_T >= g.kpad) {
      since = 0;
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            double s = 0.0;
#pragma unroll
            for (int d = S - 1; d >= 0; d--) s = s * 0.0078125 + (double)acc[d][a][b][r];
            acc64[a][b][r] += s;
          }
#pragma unroll
      for (int d = 0; d < S; d++)
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int b = 0; b < 2; b++) acc[d][a][b] = v4i{0, 0, 0, 0};
    }
  }
  // D layout of the 16 x 16 MFMA: lane l, item r -> row 4 (l / 16) + r, column l % 16
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = j0 + wc + 16 * b + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wr + 16 * a + 4 * (lane >> 4) + r;
        if (i >= g.m || j >= g.n) continue;
        if (g.tri == 1 && i < j) continue;
        const int ex = g.ea[i], fx = g.eb[j];
        const double v = (ex > SG_EMIN && fx > SG_EMIN) ? ldexp(acc64[a][b][r], ex + fx - 14) : 0.0;
        double *cp = g.C + (size_t)i + (size_t)g.ldc * j;
        *cp = g.accumulate ? *cp + v : v;
      }
    }
}

static size_t pad64(size_t x) { return (x + 63) / 64 * 64; }

// workspace bytes for split_gemm_nt
size_t split_gemm_ws_bytes(int m, int n, int k, int slices)
{
  const size_t kp = pad64((size_t)k);
  return (size_t)slices * (pad64(m) + pad64(n)) * kp + sizeof(int) * (pad64(m) + pad64(n)) + 256;
}

template <int S>
static int split_gemm_t(hipStream_t s, int m, int n, int k, const double *A, int lda, const double *B, int ldb,
                        double *C, int ldc, int tri, int accumulate, void *ws)
{
  const size_t kp = pad64((size_t)k), mp = pad64(m), np = pad64(n);
  int8_t *DA = (int8_t *)ws;
  int8_t *DB = DA + (size_t)S * mp * kp;
  int *ea = (int *)(DB + (size_t)S * np * kp);
  int *eb = ea + mp;
  // exponents start at the all-zero value (a fill kernel, not a memset)
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((mp + np + 255) / 256)), dim3(256), 0, s, ea, (int)(mp + np),
                     SG_EMIN);
  W2K_STAGE(s, "k_fill_int");
  const int kchunk = 512;
  const unsigned ks = (unsigned)((k + kchunk - 1) / kchunk);
  hipLaunchKernelGGL(k_row_expo, dim3((unsigned)(mp / 64), ks), dim3(256), 0, s, A, m, lda, k, kchunk, ea);
  W2K_STAGE(s, "k_row_expo");
  hipLaunchKernelGGL(k_row_expo, dim3((unsigned)(np / 64), ks), dim3(256), 0, s, B, n, ldb, k, kchunk, eb);
  W2K_STAGE(s, "k_row_expo");
  hipLaunchKernelGGL(k_digits<S>, dim3((unsigned)(mp / 64), (unsigned)(kp / 64)), dim3(256), 0, s, A, m, lda, k, ea,
                     DA, (int)mp, (int)kp);
  W2K_STAGE(s, "k_digits");
  hipLaunchKernelGGL(k_digits<S>, dim3((unsigned)(np / 64), (unsigned)(kp / 64)), dim3(256), 0, s, B, n, ldb, k, eb,
                     DB, (int)np, (int)kp);
  W2K_STAGE(s, "k_digits");
  SplitArgs g;
  g.m = m; g.n = n; g.kpad = (int)kp;
  g.rows_pad_a = (int)mp; g.rows_pad_b = (int)np;
  g.DA = DA; g.DB = DB; g.ea = ea; g.eb = eb;
  g.C = C; g.ldc = ldc; g.tri = tri; g.accumulate = accumulate;
  hipLaunchKernelGGL(k_split_gemm<S>, dim3((unsigned)(mp / 64), (unsigned)(np / 64)), dim3(256), 0, s, g);
  W2K_STAGE(s, "k_split_gemm");
  return 0;
}

int split_gemm_nt(hipStream_t s, int m, int n, int k, const double *A, int lda, const double *B, int ldb, double *C,
                  int ldc, int tri, int accumulate, int slices, void *ws)
{
  if (m <= 0 || n <= 0) return 0;
  if (k <= 0 || lda < m || ldb < n || ldc < m || (tri != 0 && tri != 1) || !ws) return W2K_EINVAL;
  switch (slices) {
    case 3: return split_gemm_t<3>(s, m, n, k, A, lda, B, ldb, C, ldc, tri, accumulate, ws);
    case 4: return split_gemm_t<4>(s, m, n, k, A, lda, B, ldb, C, ldc, tri, accumulate, ws);
    case 5: return split_gemm_t<5>(s, m, n, k, A, lda, B, ldb, C, ldc, tri, accumulate, ws);
    case 6: return split_gemm_t<6>(s, m, n, k, A, lda, B, ldb, C, ldc, tri, accumulate, ws);
    case 7: return split_gemm_t<7>(s, m, n, k, A, lda, B, ldb, C, ldc, tri, accumulate, ws);
    default: return W2K_EINVAL;
  }
}

}  // namespace w2k
