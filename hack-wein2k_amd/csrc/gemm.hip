// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
//   C = alpha * op(A) * op(B) + beta * C        (column-major, Fortran layout)
//
// with optional structure: the output restricted to its lower or upper
// triangle (tiles outside skipped, diagonal tiles masked per element), and
// op(A) / op(B) triangular (k range clipped per tile, elements outside the
// triangle read as zero, so the caller's storage there may hold anything).
//
// Tile 64x64x16, 256 threads = 4 wave64 in a 2x2 arrangement, each wave a
// 32x32 block of four 16x16 MFMA accumulators.  Operands staged through LDS
// k-major so that the MFMA operand reads (lane l -> row l%16, k l/16) are
// bank-conflict free; LDS double-buffered so the next k-slab's global loads
// overlap the current slab's MFMAs.
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>

#include "w2k_internal.h"

namespace w2k {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int PAD = 2;  // LDS row padding (doubles)

struct GemmArgs {
  int m, n, k;
  int kslice;        // split-K: K range of one z-slice (0: no split)
  double *ws;        // split-K: m x n partial products per slice
  double alpha, beta;
  const double *A;
  int lda;
  const double *B;
  int ldb;
  double *C;
  int ldc;
  int ta, tb;        // 1 = transposed
  int tri;           // output: 0 full, 1 lower, 2 upper
  int atri, btri;    // op(A): 1 upper / 2 lower; op(B): 1 upper / 2 lower
};

// element (i, kk) of op(A), zero outside the matrix or outside its triangle
__device__ __forceinline__ double load_opA(const GemmArgs &g, int i, int kk)
{
  if (i >= g.m || kk >= g.k) return 0.0;
  if (g.atri == 1 && kk < i) return 0.0;
  if (g.atri == 2 && kk > i) return 0.0;
  return g.ta ? g.A[(size_t)kk + (size_t)g.lda * i] : g.A[(size_t)i + (size_t)g.lda * kk];
}
__device__ __forceinline__ double load_opB(const GemmArgs &g, int kk, int j)
{
  if (j >= g.n || kk >= g.k) return 0.0;
  if (g.btri == 1 && kk > j) return 0.0;
  if (g.btri == 2 && kk < j) return 0.0;
  return g.tb ? g.B[(size_t)j + (size_t)g.ldb * kk] : g.B[(size_t)kk + (size_t)g.ldb * j];
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_dgemm(GemmArgs g)
{
  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][BN + PAD];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * BM, j0 = blockIdx.y * BN;
  // output-triangle tile skip
  if (g.tri == 1 && i0 + BM - 1 < j0) return;
  if (g.tri == 2 && j0 + BN - 1 < i0) return;
  // k range from operand structure
  int kbeg = 0, kend = g.k;
  if (g.atri == 1) kbeg = max(kbeg, i0);
  if (g.atri == 2) kend = min(kend, i0 + BM);
  if (g.btri == 1) kend = min(kend, j0 + BN);
  if (g.btri == 2) kbeg = max(kbeg, j0);
  if (SPLIT) {
    kbeg = max(kbeg, (int)blockIdx.z * g.kslice);
    kend = min(kend, ((int)blockIdx.z + 1) * g.kslice);
  }
  kbeg = (kbeg / BK) * BK;

  const int lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;

  d4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  // with a K split, elements past this slice's end read as zero
  GemmArgs gl = g;
  if (SPLIT) gl.k = min(g.k, kend);
  // register staging for the global loads of one slab: 4 A + 4 B per thread
  double ra[4], rb[4];
  auto gload = [&](int k0) {
    if (!g.ta) {  // 64 rows contiguous per k column
      for (int r = 0; r < 4; r++) ra[r] = load_opA(gl, i0 + (tid & 63), k0 + (tid >> 6) + 4 * r);
    } else {      // 16 k contiguous per row
      for (int r = 0; r < 4; r++) ra[r] = load_opA(gl, i0 + (tid >> 4) + 16 * r, k0 + (tid & 15));
    }
    if (g.tb) {   // 64 columns contiguous per k row
      for (int r = 0; r < 4; r++) rb[r] = load_opB(gl, k0 + (tid >> 6) + 4 * r, j0 + (tid & 63));
    } else {      // 16 k contiguous per column
      for (int r = 0; r < 4; r++) rb[r] = load_opB(gl, k0 + (tid & 15), j0 + (tid >> 4) + 16 * r);
    }
  };
  auto sstore = [&](int buf) {
    if (!g.ta) {
      for (int r = 0; r < 4; r++) As[buf][(tid >> 6) + 4 * r][tid & 63] = ra[r];
    } else {
      for (int r = 0; r < 4; r++) As[buf][tid & 15][(tid >> 4) + 16 * r] = ra[r];
    }
    if (g.tb) {
      for (int r = 0; r < 4; r++) Bs[buf][(tid >> 6) + 4 * r][tid & 63] = rb[r];
    } else {
      for (int r = 0; r < 4; r++) Bs[buf][tid & 15][(tid >> 4) + 16 * r] = rb[r];
    }
  };

  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = (k0 + BK < kend);
      if (more) gload(k0 + BK);
#pragma unroll
      for (int kq = 0; kq < BK; kq += 4) {
        double av[2], bv[2];
        for (int a = 0; a < 2; a++) av[a] = As[buf][kq + lk][wi + 16 * a + li];
        for (int b = 0; b < 2; b++) bv[b] = Bs[buf][kq + lk][wj + 16 * b + li];
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
      }
      if (more) {
        sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
  }
  // epilogue: lane holds rows lane/16 + 4r (r = 0..3), column lane%16 of each
  // 16x16 block (v_mfma_f64_16x16x4_f64 result layout, checked on gfx950)
  for (int a = 0; a < 2; a++) {
    for (int b = 0; b < 2; b++) {
      const int j = j0 + wj + 16 * b + li;
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wi + 16 * a + lk + 4 * r;
        if (i >= g.m || j >= g.n) continue;
        if (g.tri == 1 && i < j) continue;
        if (g.tri == 2 && i > j) continue;
        if (SPLIT) {
          g.ws[((size_t)blockIdx.z * g.n + j) * g.m + i] = acc[a][b][r];
          continue;
        }
        double *cp = g.C + (size_t)i + (size_t)g.ldc * j;
        double v = g.alpha * acc[a][b][r];
        if (g.beta != 0.0) v += g.beta * *cp;
        *cp = v;
      }
    }
  }
}

// Large products: tile 128x128x16, 256 threads = 4 wave64 in a 2x2
// arrangement, each wave a 64x64 block of sixteen 16x16 MFMA accumulators
// (eight operand reads from LDS per sixteen MFMAs).  The transposes are
// template parameters, and every global load is unconditional (indices
// clamped into the matrix, the value masked afterwards) so that a slab's
// loads are issued together; LDS double-buffered as above.
constexpr int LM = 128, LN = 128, LK = 16;

template <int TA, int TB>
__global__ __launch_bounds__(256, 2) void k_dgemm_big(GemmArgs g)
{
  __shared__ double As[2][LK][LM + PAD];
  __shared__ double Bs[2][LK][LN + PAD];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * LM, j0 = blockIdx.y * LN;
  if (g.tri == 1 && i0 + LM - 1 < j0) return;
  if (g.tri == 2 && j0 + LN - 1 < i0) return;
  int kbeg = 0, kend = g.k;
  if (g.atri == 1) kbeg = max(kbeg, i0);
  if (g.atri == 2) kend = min(kend, i0 + LM);
  if (g.btri == 1) kend = min(kend, j0 + LN);
  if (g.btri == 2) kbeg = max(kbeg, j0);
  kbeg = (kbeg / LK) * LK;

  const int lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 1) * 64, wj = (wave >> 1) * 64;
  const int li = lane & 15, lk = lane >> 4;
  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  // this thread's eight elements of each operand per slab: op(A)(ia, ka + step),
  // op(B)(kb + step, jb) with the step along k (no transpose) or the row / column
  const int ia = TA ? i0 + (tid >> 4) : i0 + (tid & 127);
  const int ka = TA ? (tid & 15) : (tid >> 7);
  const int jb = TB ? j0 + (tid & 127) : j0 + (tid >> 4);
  const int kb = TB ? (tid >> 7) : (tid & 15);
  double ra[8], rb[8];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int i = TA ? ia + 16 * r : ia;
      const int kk = TA ? k0 + ka : k0 + ka + 2 * r;
      bool ok = i < g.m && kk < kend;
      if (g.atri == 1) ok = ok && kk >= i;
      if (g.atri == 2) ok = ok && kk <= i;
      const int ic = min(i, g.m - 1), kc = min(kk, g.k - 1);
      const double v = TA ? g.A[(size_t)kc + (size_t)g.lda * ic] : g.A[(size_t)ic + (size_t)g.lda * kc];
      ra[r] = ok ? v : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int j = TB ? jb : jb + 16 * r;
      const int kk = TB ? k0 + kb + 2 * r : k0 + kb;
      bool ok = j < g.n && kk < kend;
      if (g.btri == 1) ok = ok && kk <= j;
      if (g.btri == 2) ok = ok && kk >= j;
      const int jc = min(j, g.n - 1), kc = min(kk, g.k - 1);
      const double v = TB ? g.B[(size_t)jc + (size_t)g.ldb * kc] : g.B[(size_t)kc + (size_t)g.ldb * jc];
      rb[r] = ok ? v : 0.0;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 8; r++) {
      if (TA) As[buf][ka][(tid >> 4) + 16 * r] = ra[r];
      else As[buf][ka + 2 * r][tid & 127] = ra[r];
      if (TB) Bs[buf][kb + 2 * r][tid & 127] = rb[r];
      else Bs[buf][kb][(tid >> 4) + 16 * r] = rb[r];
    }
  };

  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += LK) {
      const bool more = (k0 + LK < kend);
      if (more) gload(k0 + LK);
#pragma unroll
      for (int kq = 0; kq < LK; kq += 4) {
        double av[4], bv[4];
#pragma unroll
        for (int a = 0; a < 4; a++) av[a] = As[buf][kq + lk][wi + 16 * a + li];
#pragma unroll
        for (int b = 0; b < 4; b++) bv[b] = Bs[buf][kq + lk][wj + 16 * b + li];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++)
            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
      }
      if (more) {
        sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
  }
  // epilogue: as in k_dgemm (lane holds rows lane/16 + 4r, column lane%16)
#pragma unroll
  for (int a = 0; a < 4; a++) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int j = j0 + wj + 16 * b + li;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wi + 16 * a + lk + 4 * r;
        if (i >= g.m || j >= g.n) continue;
        if (g.tri == 1 && i < j) continue;
        if (g.tri == 2 && i > j) continue;
        double *cp = g.C + (size_t)i + (size_t)g.ldc * j;
        double v = g.alpha * acc[a][b][r];
        if (g.beta != 0.0) v += g.beta * *cp;
        *cp = v;
      }
    }
  }
}

// C = alpha * sum_z ws[z] + beta * C, fixed summation order; with a
// triangular output (tri 1 lower, 2 upper) only that triangle is written
__global__ void k_splitk_reduce(int m, int n, int nz, const double *__restrict__ ws, double alpha, double beta,
                                double *__restrict__ C, int ldc, int tri)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  if ((tri == 1 && i < j) || (tri == 2 && i > j)) return;
  double v = 0.0;
  for (int z = 0; z < nz; z++) v += ws[((size_t)z * n + j) * m + i];
  double *cp = C + (size_t)i + (size_t)ldc * j;
  v *= alpha;
  if (beta != 0.0) v += beta * *cp;
  *cp = v;
}

// ------------------------------------------------------------------ complex
// Planar complex product C = alpha op(A) op(B) + beta C (alpha, beta real)
// in one kernel: each work-group forms a 64x64 tile of both Re C and Im C
// from the same staged operands (four real MFMA products per k-step, the
// imaginary parts of op(A) and op(B) signed for conjugation at staging), so
// C is read and written once instead of four times as by four real GEMMs.
// op: transpose TA/TB (template), conjugation sa/sb = -1 (runtime); output
// restricted to a triangle (tri 1 lower, 2 upper) as in k_dgemm.
struct ZGemmArgs {
  int m, n, k;
  double alpha, beta;
  const double *Ar, *Ai;
  int lda;
  const double *Br, *Bi;
  int ldb;
  double *Cr, *Ci;
  int ldc;
  double sa, sb;  // sign of the imaginary parts of op(A), op(B)
  int tri;
};

template <int TA, int TB>
__global__ __launch_bounds__(256, 2) void k_zgemm(ZGemmArgs g)
{
  __shared__ double Ar_s[2][BK][BM + PAD], Ai_s[2][BK][BM + PAD];
  __shared__ double Br_s[2][BK][BN + PAD], Bi_s[2][BK][BN + PAD];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * BM, j0 = blockIdx.y * BN;
  if (g.tri == 1 && i0 + BM - 1 < j0) return;
  if (g.tri == 2 && j0 + BN - 1 < i0) return;
  const int lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // C (beta != 0) is read before the K loop, from clamped addresses, so that
  // its loads overlap the products instead of following them
  double pcr[2][2][4], pci[2][2][4];
  if (g.beta != 0.0) {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int j = min(j0 + wj + 16 * b + li, g.n - 1);
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = min(i0 + wi + 16 * a + lk + 4 * r, g.m - 1);
          const size_t o = (size_t)i + (size_t)g.ldc * j;
          pcr[a][b][r] = g.Cr[o];
          pci[a][b][r] = g.Ci[o];
        }
      }
  }
  d4 cr[2][2], ci[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      cr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      ci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  // this thread's four elements of each operand per slab (rows or k steps as
  // in k_dgemm), loaded unconditionally from clamped addresses and masked
  double rar[4], rai[4], rbr[4], rbi[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = TA ? i0 + (tid >> 4) + 16 * r : i0 + (tid & 63);
      const int kk = TA ? k0 + (tid & 15) : k0 + (tid >> 6) + 4 * r;
      const bool ok = i < g.m && kk < g.k;
      const int ic = min(i, g.m - 1), kc = min(kk, g.k - 1);
      const size_t o = TA ? (size_t)kc + (size_t)g.lda * ic : (size_t)ic + (size_t)g.lda * kc;
      const double vr = g.Ar[o], vi = g.Ai[o];
      rar[r] = ok ? vr : 0.0;
      rai[r] = ok ? g.sa * vi : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int j = TB ? j0 + (tid & 63) : j0 + (tid >> 4) + 16 * r;
      const int kk = TB ? k0 + (tid >> 6) + 4 * r : k0 + (tid & 15);
      const bool ok = j < g.n && kk < g.k;
      const int jc = min(j, g.n - 1), kc = min(kk, g.k - 1);
      const size_t o = TB ? (size_t)jc + (size_t)g.ldb * kc : (size_t)kc + (size_t)g.ldb * jc;
      const double vr = g.Br[o], vi = g.Bi[o];
      rbr[r] = ok ? vr : 0.0;
      rbi[r] = ok ? g.sb * vi : 0.0;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ka = TA ? (tid & 15) : (tid >> 6) + 4 * r, ia = TA ? (tid >> 4) + 16 * r : (tid & 63);
      Ar_s[buf][ka][ia] = rar[r];
      Ai_s[buf][ka][ia] = rai[r];
      const int kb = TB ? (tid >> 6) + 4 * r : (tid & 15), jb = TB ? (tid & 63) : (tid >> 4) + 16 * r;
      Br_s[buf][kb][jb] = rbr[r];
      Bi_s[buf][kb][jb] = rbi[r];
    }
  };
  if (g.k > 0) {
    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < g.k; k0 += BK) {
      const bool more = (k0 + BK < g.k);
      if (more) gload(k0 + BK);
#pragma unroll
      for (int kq = 0; kq < BK; kq += 4) {
        double ar[2], ai[2], br[2], bi[2], nbi[2];
        for (int a = 0; a < 2; a++) {
          ar[a] = Ar_s[buf][kq + lk][wi + 16 * a + li];
          ai[a] = Ai_s[buf][kq + lk][wi + 16 * a + li];
        }
        for (int b = 0; b < 2; b++) {
          br[b] = Br_s[buf][kq + lk][wj + 16 * b + li];
          bi[b] = Bi_s[buf][kq + lk][wj + 16 * b + li];
          nbi[b] = -bi[b];
        }
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++) {
            cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
            cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], nbi[b], cr[a][b], 0, 0, 0);
            ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
            ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
          }
      }
      if (more) {
        sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
  }
  for (int a = 0; a < 2; a++) {
    for (int b = 0; b < 2; b++) {
      const int j = j0 + wj + 16 * b + li;
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wi + 16 * a + lk + 4 * r;
        if (i >= g.m || j >= g.n) continue;
        if (g.tri == 1 && i < j) continue;
        if (g.tri == 2 && i > j) continue;
        const size_t o = (size_t)i + (size_t)g.ldc * j;
        double vr = g.alpha * cr[a][b][r], vi = g.alpha * ci[a][b][r];
        if (g.beta != 0.0) {
          vr += g.beta * pcr[a][b][r];
          vi += g.beta * pci[a][b][r];
        }
        g.Cr[o] = vr;
        g.Ci[o] = vi;
      }
    }
  }
}

// The fused kernel over one slice of K per blockIdx.z (slice z: k in
// [z kslice, (z + 1) kslice)), its raw sums into wsr/wsi[(z n + j) m + i];
// k_zsplitk_reduce adds the slices and applies alpha, beta and C.  A kernel
// of its own: folding the K range into k_zgemm changed that kernel's code
// and cost the tridiagonalisation's rank-64 updates 24 ms per 100-atom
// k-point.
template <int TA, int TB>
__global__ __launch_bounds__(256, 2) void k_zgemm_split(ZGemmArgs g, int kslice, double *__restrict__ wsr, double *__restrict__ wsi)
{
  __shared__ double Ar_s[2][BK][BM + PAD], Ai_s[2][BK][BM + PAD];
  __shared__ double Br_s[2][BK][BN + PAD], Bi_s[2][BK][BN + PAD];
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * BM, j0 = blockIdx.y * BN;
  if (g.tri == 1 && i0 + BM - 1 < j0) return;
  if (g.tri == 2 && j0 + BN - 1 < i0) return;
  const int lane = tid & 63, wave = tid >> 6;
  const int wi = (wave & 1) * 32, wj = (wave >> 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int kbeg = blockIdx.z * kslice, kend = min(g.k, kbeg + kslice);
  d4 cr[2][2], ci[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) {
      cr[a][b] = d4{0.0, 0.0, 0.0, 0.0};
      ci[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  // this thread's four elements of each operand per slab (rows or k steps as
  // in k_dgemm), loaded unconditionally from clamped addresses and masked
  double rar[4], rai[4], rbr[4], rbi[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = TA ? i0 + (tid >> 4) + 16 * r : i0 + (tid & 63);
      const int kk = TA ? k0 + (tid & 15) : k0 + (tid >> 6) + 4 * r;
      const bool ok = i < g.m && kk < kend;
      const int ic = min(i, g.m - 1), kc = min(kk, g.k - 1);
      const size_t o = TA ? (size_t)kc + (size_t)g.lda * ic : (size_t)ic + (size_t)g.lda * kc;
      const double vr = g.Ar[o], vi = g.Ai[o];
      rar[r] = ok ? vr : 0.0;
      rai[r] = ok ? g.sa * vi : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int j = TB ? j0 + (tid & 63) : j0 + (tid >> 4) + 16 * r;
      const int kk = TB ? k0 + (tid >> 6) + 4 * r : k0 + (tid & 15);
      const bool ok = j < g.n && kk < kend;
      const int jc = min(j, g.n - 1), kc = min(kk, g.k - 1);
      const size_t o = TB ? (size_t)jc + (size_t)g.ldb * kc : (size_t)kc + (size_t)g.ldb * jc;
      const double vr = g.Br[o], vi = g.Bi[o];
      rbr[r] = ok ? vr : 0.0;
      rbi[r] = ok ? g.sb * vi : 0.0;
    }
  };
  auto sstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int ka = TA ? (tid & 15) : (tid >> 6) + 4 * r, ia = TA ? (tid >> 4) + 16 * r : (tid & 63);
      Ar_s[buf][ka][ia] = rar[r];
      Ai_s[buf][ka][ia] = rai[r];
      const int kb = TB ? (tid >> 6) + 4 * r : (tid & 15), jb = TB ? (tid & 63) : (tid >> 4) + 16 * r;
      Br_s[buf][kb][jb] = rbr[r];
      Bi_s[buf][kb][jb] = rbi[r];
    }
  };
  if (kbeg < kend) {
    gload(kbeg);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = kbeg; k0 < kend; k0 += BK) {
      const bool more = (k0 + BK < kend);
      if (more) gload(k0 + BK);
#pragma unroll
      for (int kq = 0; kq < BK; kq += 4) {
        double ar[2], ai[2], br[2], bi[2], nbi[2];
        for (int a = 0; a < 2; a++) {
          ar[a] = Ar_s[buf][kq + lk][wi + 16 * a + li];
          ai[a] = Ai_s[buf][kq + lk][wi + 16 * a + li];
        }
        for (int b = 0; b < 2; b++) {
          br[b] = Br_s[buf][kq + lk][wj + 16 * b + li];
          bi[b] = Bi_s[buf][kq + lk][wj + 16 * b + li];
          nbi[b] = -bi[b];
        }
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++) {
            cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], br[b], cr[a][b], 0, 0, 0);
            cr[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], nbi[b], cr[a][b], 0, 0, 0);
            ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[a], bi[b], ci[a][b], 0, 0, 0);
            ci[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[a], br[b], ci[a][b], 0, 0, 0);
          }
      }
      if (more) {
        sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
      }
    }
  }
  // raw sums of this K slice (no triangle: the split needs a full output)
  for (int a = 0; a < 2; a++) {
    for (int b = 0; b < 2; b++) {
      const int j = j0 + wj + 16 * b + li;
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wi + 16 * a + lk + 4 * r;
        if (i >= g.m || j >= g.n) continue;
        const size_t o = ((size_t)blockIdx.z * g.n + j) * g.m + i;
        wsr[o] = cr[a][b][r];
        wsi[o] = ci[a][b][r];
      }
    }
  }
}

// C = alpha sum_z ws[z] + beta C for both planes, fixed summation order
__global__ void k_zsplitk_reduce(int m, int n, int nz, const double *__restrict__ wsr, const double *__restrict__ wsi,
                                 double alpha, double beta, double *Cr, double *Ci, int ldc)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
  if (i >= m) return;
  double vr = 0.0, vi = 0.0;
  for (int z = 0; z < nz; z++) {
    const size_t o = ((size_t)z * n + j) * m + i;
    vr += wsr[o];
    vi += wsi[o];
  }
  const size_t o = (size_t)i + (size_t)ldc * j;
  vr *= alpha;
  vi *= alpha;
  if (beta != 0.0) {
    vr += beta * Cr[o];
    vi += beta * Ci[o];
  }
  Cr[o] = vr;
  Ci[o] = vi;
}

int zgemm_fused(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *Ar, const double *Ai,
                int lda, const double *Br, const double *Bi, int ldb, double beta, double *Cr, double *Ci, int ldc,
                int tri, double *ws, size_t wsn)
{
  if (m <= 0 || n <= 0) return 0;
  ZGemmArgs g;
  g.m = m; g.n = n; g.k = k > 0 ? k : 0; g.alpha = alpha; g.beta = beta;
  g.Ar = Ar; g.Ai = Ai; g.lda = lda; g.Br = Br; g.Bi = Bi; g.ldb = ldb; g.Cr = Cr; g.Ci = Ci; g.ldc = ldc;
  g.sa = (ta == 'C' || ta == 'R') ? -1.0 : 1.0;
  g.sb = (tb == 'C' || tb == 'R') ? -1.0 : 1.0;
  g.tri = tri;
  const bool ttA = (ta == 'T' || ta == 'C'), ttB = (tb == 'T' || tb == 'C');
  dim3 grid((m + BM - 1) / BM, (n + BN - 1) / BN);
  // few output tiles and a long K (the rule of dgemm's first split): K in
  // slices, partial sums in ws (two planes of m x n per slice)
  constexpr int ksl = W2K_ZSPLIT_KMIN / 2;
  const long tiles = (long)grid.x * grid.y;
  if (ws && tri == 0 && tiles < 256 && k >= W2K_ZSPLIT_KMIN) {
    int nz = (int)std::min<long>((k + ksl - 1) / ksl, std::max<long>(1, 1024 / tiles));
    nz = (int)std::min<size_t>((size_t)nz, wsn / ((size_t)2 * m * n));
    if (nz >= 2) {
      const int kslice = ((k + nz - 1) / nz + BK - 1) / BK * BK;
      nz = (k + kslice - 1) / kslice;
      double *wsr = ws, *wsi = ws + (size_t)nz * m * n;
      grid.z = nz;
      if (!ttA && !ttB) hipLaunchKernelGGL((k_zgemm_split<0, 0>), grid, dim3(256), 0, s, g, kslice, wsr, wsi);
      else if (ttA && !ttB) hipLaunchKernelGGL((k_zgemm_split<1, 0>), grid, dim3(256), 0, s, g, kslice, wsr, wsi);
      else if (!ttA && ttB) hipLaunchKernelGGL((k_zgemm_split<0, 1>), grid, dim3(256), 0, s, g, kslice, wsr, wsi);
      else hipLaunchKernelGGL((k_zgemm_split<1, 1>), grid, dim3(256), 0, s, g, kslice, wsr, wsi);
      W2K_STAGE(s, "k_zgemm_split");
      hipLaunchKernelGGL(k_zsplitk_reduce, dim3((m + 255) / 256, n), dim3(256), 0, s, m, n, nz, wsr, wsi, alpha, beta,
                         Cr, Ci, ldc);
      W2K_STAGE(s, "k_zsplitk_reduce");
      return 0;
    }
  }
  if (!ttA && !ttB) hipLaunchKernelGGL((k_zgemm<0, 0>), grid, dim3(256), 0, s, g);
  else if (ttA && !ttB) hipLaunchKernelGGL((k_zgemm<1, 0>), grid, dim3(256), 0, s, g);
  else if (!ttA && ttB) hipLaunchKernelGGL((k_zgemm<0, 1>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_zgemm<1, 1>), grid, dim3(256), 0, s, g);
  W2K_STAGE(s, "k_zgemm");
  return 0;
}

// ------------------------------------------------------------------ tile choice
// Products that both tile kernels can serve (at least 128 x 128 outputs and
// K >= 64) get their kernel from a per-device table keyed by the shape class
// (transposes, structure flags, floor(log2) of m, n and k): the first call of
// a class runs the heuristic's kernel between two events, the next call of
// the class the other kernel, and once both durations are in (read with
// hipEventQuery, never waited for) the faster one is kept.  The two kernels
// accumulate every product in the same order (MFMA steps of four along k,
// zero products only added for elements outside a triangle), so their
// results are bit-identical and the choice never changes a result.
// W2K_GEMM_TUNE=0 keeps the heuristic (128x128 tiles for outputs of at least
// 256 x 256 with K >= 256).
namespace {
struct TuneKey {
  int dev, ta, tb, tri, atri, btri, lm, ln, lk;
  bool operator<(const TuneKey &o) const
  {
    const int a[9] = {dev, ta, tb, tri, atri, btri, lm, ln, lk};
    const int b[9] = {o.dev, o.ta, o.tb, o.tri, o.atri, o.btri, o.lm, o.ln, o.lk};
    return std::lexicographical_compare(a, a + 9, b, b + 9);
  }
};
struct TuneEntry {
  int choice = -1;
  float ms[2] = {-1.f, -1.f};
  int pending[2] = {0, 0};
  hipEvent_t ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
};
std::mutex tune_mu;
std::map<TuneKey, TuneEntry> tune_table;
int ilog2(int x) { return 31 - __builtin_clz((unsigned)x); }
bool tuning_on()
{
  static const bool on = !(getenv("W2K_GEMM_TUNE") && atoi(getenv("W2K_GEMM_TUNE")) == 0);
  return on;
}
}  // namespace

// variant 0: 64x64 tiles, 1: 128x128 tiles
static int launch_tiles(hipStream_t s, const GemmArgs &g, int variant)
{
  if (variant == 1) {
    dim3 gb((g.m + LM - 1) / LM, (g.n + LN - 1) / LN);
    if (!g.ta && !g.tb) hipLaunchKernelGGL((k_dgemm_big<0, 0>), gb, dim3(256), 0, s, g);
    else if (g.ta && !g.tb) hipLaunchKernelGGL((k_dgemm_big<1, 0>), gb, dim3(256), 0, s, g);
    else if (!g.ta && g.tb) hipLaunchKernelGGL((k_dgemm_big<0, 1>), gb, dim3(256), 0, s, g);
    else hipLaunchKernelGGL((k_dgemm_big<1, 1>), gb, dim3(256), 0, s, g);
    W2K_STAGE(s, "k_dgemm_big");
    return 0;
  }
  dim3 grid((g.m + BM - 1) / BM, (g.n + BN - 1) / BN);
  hipLaunchKernelGGL(k_dgemm<false>, grid, dim3(256), 0, s, g);
  W2K_STAGE(s, "k_dgemm");
  return 0;
}

static GemmArgs gemm_args(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                          const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri)
{
  GemmArgs g;
  g.m = m; g.n = n; g.k = k > 0 ? k : 0; g.alpha = alpha; g.beta = beta;
  g.A = A; g.lda = lda; g.B = B; g.ldb = ldb; g.C = C; g.ldc = ldc;
  g.ta = (ta == 'T' || ta == 't'); g.tb = (tb == 'T' || tb == 't');
  g.tri = tri; g.atri = atri; g.btri = btri;
  g.kslice = 0; g.ws = nullptr;
  return g;
}

int dgemm_tiles(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri, int variant)
{
  if (m <= 0 || n <= 0) return 0;
  const GemmArgs g = gemm_args(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, atri, btri);
  return launch_tiles(s, g, (variant == 1 && g.k > 0) ? 1 : 0);
}

int dgemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
          const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri, double *ws,
          size_t wsn)
{
  if (m <= 0 || n <= 0) return 0;
  GemmArgs g = gemm_args(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, atri, btri);
  // few output tiles and a long K: split K over z-slices so that the whole
  // chip works, then reduce the slices in a fixed order.  Products with a
  // triangular operand or output and up to 4096 tiles are split into slices
  // of about 1024 as well: their K range grows across the tiles, and
  // unsplit, the tiles with the full K set the duration.  The rule depends
  // on the shape and the workspace only, so a product is always summed the
  // same way.
  dim3 grid((m + BM - 1) / BM, (n + BN - 1) / BN);
  const long tiles = (long)grid.x * grid.y;
  const bool structured = tri || atri || btri;
  int nz = 0;
  if (ws && tri == 0 && tiles < 256 && k >= 2 * 256)
    nz = (int)std::min<long>((k + 255) / 256, std::max<long>(1, 1024 / tiles));
  else if (ws && structured && tiles >= 256 && tiles <= 4096 && k >= 1024)
    nz = std::min((k + 1023) / 1024, 4);
  if (nz >= 2) {
    nz = (int)std::min<size_t>((size_t)nz, wsn / ((size_t)m * n));
    if (nz >= 2) {
      g.kslice = ((k + nz - 1) / nz + BK - 1) / BK * BK;
      nz = (k + g.kslice - 1) / g.kslice;
      g.ws = ws;
      grid.z = nz;
      hipLaunchKernelGGL(k_dgemm<true>, grid, dim3(256), 0, s, g);
      W2K_STAGE(s, "k_dgemm<split>");
      hipLaunchKernelGGL(k_splitk_reduce, dim3((m + 255) / 256, n), dim3(256), 0, s, m, n, nz, ws, alpha, beta, C,
                         ldc, tri);
      W2K_STAGE(s, "k_splitk_reduce");
      return 0;
    }
  }
  int variant = (m >= 256 && n >= 256 && k >= 256) ? 1 : 0;
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  // (not inside the allocation-free entry point, which creates no events)
  if (m >= 128 && n >= 128 && k >= 64 && tuning_on() && !g_arena && hipStreamIsCapturing(s, &cap) == hipSuccess &&
      cap == hipStreamCaptureStatusNone) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    const TuneKey key{dev, g.ta, g.tb, tri, atri, btri, ilog2(m), ilog2(n), ilog2(k)};
    std::unique_lock<std::mutex> lk(tune_mu);
    TuneEntry &e = tune_table[key];
    for (int v = 0; v < 2; v++) {
      if (e.pending[v] && hipEventQuery(e.ev[v][1]) == hipSuccess) {
        float ms = -1.f;
        if (hipEventElapsedTime(&ms, e.ev[v][0], e.ev[v][1]) == hipSuccess) e.ms[v] = ms;
        e.pending[v] = 0;
      }
    }
    (void)hipGetLastError();  // a query of an unfinished event is not an error here
    if (e.choice < 0 && e.ms[0] >= 0.f && e.ms[1] >= 0.f) e.choice = e.ms[1] < e.ms[0] ? 1 : 0;
    if (e.choice >= 0) {
      variant = e.choice;
    } else {
      int v = -1;
      for (int u : {variant, 1 - variant})
        if (v < 0 && e.ms[u] < 0.f && !e.pending[u]) v = u;
      if (v >= 0) {
        for (int q = 0; q < 2; q++)
          if (!e.ev[v][q]) W2K_HIP(hipEventCreate(&e.ev[v][q]));
        W2K_HIP(hipEventRecord(e.ev[v][0], s));
        int rc;
        if ((rc = launch_tiles(s, g, v))) return rc;
        W2K_HIP(hipEventRecord(e.ev[v][1], s));
        e.pending[v] = 1;
        return 0;
      }
    }
  }
  return launch_tiles(s, g, g.k > 0 ? variant : 0);
}

}  // namespace w2k
