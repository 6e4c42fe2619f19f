// SECLR4 of the complex build (lapw1c) on the GPU: seclr4.F:155-358 with the
// Z-routines ZPOTRF, ZHCGST (zhcgst.f:116-163), ZHETRD2/ZHEEVX2
// (zheevx2.f:305-367), DSTEBZ2 and the back-transforms.
//
// Conventions of the reference kept exactly (and restated by the oracle):
//   - HS is conjugated on entry (seclr4.F:159-166): the lower triangle then
//     holds H, the strict upper triangle conj(S) and HSDIAG diag(S);
//   - S = U^H U (ZPOTRF 'U') on that upper triangle;
//   - ZHCGST reduces with V = conj(U): C = V^-H H V^-1;
//   - the Hermitian tridiagonalisation uses ZLARFG reflectors (beta real,
//     tau complex), so the tridiagonal matrix is real and the bisection and
//     inverse iteration are those of the real build;
//   - Z = V^-1 Q Zt.
// Complex matrices are planar on the device (real parts, then imaginary parts
// one matrix further), and every complex product is four real MFMA GEMMs.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <functional>
#include <vector>

#include <hip/hip_ext.h>

#include "seclr4_shared.h"

namespace w2k {

// C = alpha op(A) op(B) + beta C, planar complex; op: 'N', 'T' (transpose),
// 'C' (conjugate transpose), 'R' (conjugate, not transposed); alpha, beta real
static int zgemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *Ar,
                 const double *Ai, int lda, const double *Br, const double *Bi, int ldb, double beta, double *Cr,
                 double *Ci, int ldc, int tri = 0, int atri = 0, int btri = 0, double *ws = nullptr, size_t wsn = 0)
{
  // products with dense operands and no split K: one fused complex kernel
  // (C read and written once); short K (rank updates) is where it pays
  if (atri == 0 && btri == 0 && ws == nullptr && k <= 256)
    return zgemm_fused(s, ta, tb, m, n, k, alpha, Ar, Ai, lda, Br, Bi, ldb, beta, Cr, Ci, ldc, tri);
  // dense products with a long K onto few output tiles (the back-transform's
  // V^H Z and Gram matrices): the fused kernel with split K reads each
  // operand plane once and feeds 16 MFMAs from 8 LDS reads per step, against
  // 4 from 4 in the real 64 x 64 kernel
  const long ztiles = (long)((m + 63) / 64) * ((n + 63) / 64);
  if (atri == 0 && btri == 0 && tri == 0 && ws != nullptr && ztiles < 256 && k >= W2K_ZSPLIT_KMIN &&
      wsn >= (size_t)4 * m * n)
    return zgemm_fused(s, ta, tb, m, n, k, alpha, Ar, Ai, lda, Br, Bi, ldb, beta, Cr, Ci, ldc, tri, ws, wsn);
  const double sa = (ta == 'C' || ta == 'R') ? -1.0 : 1.0, sb = (tb == 'C' || tb == 'R') ? -1.0 : 1.0;
  const char ra = (ta == 'T' || ta == 'C') ? 'T' : 'N', rb = (tb == 'T' || tb == 'C') ? 'T' : 'N';
  int rc;
  // Cr = alpha (Ar Br - sa sb Ai Bi) + beta Cr ; Ci = alpha (sb Ar Bi + sa Ai Br) + beta Ci
  if ((rc = dgemm(s, ra, rb, m, n, k, alpha, Ar, lda, Br, ldb, beta, Cr, ldc, tri, atri, btri, ws, wsn))) return rc;
  if ((rc = dgemm(s, ra, rb, m, n, k, -alpha * sa * sb, Ai, lda, Bi, ldb, 1.0, Cr, ldc, tri, atri, btri, ws, wsn)))
    return rc;
  if ((rc = dgemm(s, ra, rb, m, n, k, alpha * sb, Ar, lda, Bi, ldb, beta, Ci, ldc, tri, atri, btri, ws, wsn)))
    return rc;
  if ((rc = dgemm(s, ra, rb, m, n, k, alpha * sa, Ai, lda, Br, ldb, 1.0, Ci, ldc, tri, atri, btri, ws, wsn)))
    return rc;
  return 0;
}

// D (rows x cols, ld rows) = X + s Y (ld ldx)
__global__ void k_zsum(int rows, int cols, const double *__restrict__ X, const double *__restrict__ Y, int ldx,
                       double s, double *__restrict__ D)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= rows || j >= cols) return;
  const size_t o = (size_t)i + (size_t)ldx * j;
  D[(size_t)i + (size_t)rows * j] = X[o] + s * Y[o];
}

// three-product combination on the output (triangle tri): mode 0 after T =
// A_r B_r: Cr = a T + b Cr, Ci -= a T; mode 1 after T = Im(op A) Im(op B):
// Cr -= a T, Ci -= a T
__global__ void k_z3m_comb(int m, int n, const double *__restrict__ T, double a, double b, double *__restrict__ Cr,
                           double *__restrict__ Ci, int ldc, int tri, int mode)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= m || j >= n) return;
  if ((tri == 1 && i < j) || (tri == 2 && i > j)) return;
  const double t = a * T[(size_t)i + (size_t)m * j];
  const size_t o = (size_t)i + (size_t)ldc * j;
  if (mode == 0) {
    Cr[o] = (b != 0.0 ? b * Cr[o] : 0.0) + t;
    Ci[o] -= t;
  } else {
    Cr[o] -= t;
    Ci[o] -= t;
  }
}

// K up to which a complex product below the three-product size runs on the
// fused kernel (triangular operands read as dense): 1024 measured 2 ms
// faster than 512 on the 100-atom Cholesky, 4096 no faster than 1024
constexpr int ZFUSED_KMAX = 1024;

// zgemm for the large products: three real products instead of four
// (Ci = (Ar + Im op A)(Br + Im op B) - Ar Br - Im op A Im op B), which takes a
// quarter of the arithmetic off at the cost of three elementwise passes and
// an error of the imaginary part bounded by |A| |B| rather than by the
// products of the parts.  Smaller products, and calls inside the
// allocation-free entry point, go to the fused kernel (K <= ZFUSED_KMAX) or zgemm.
static int zgemm3(Ctx &c, hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *Ar,
                  const double *Ai, int lda, const double *Br, const double *Bi, int ldb, double beta, double *Cr,
                  double *Ci, int ldc, int tri = 0, int atri = 0, int btri = 0, double *ws = nullptr, size_t wsn = 0)
{
  const bool big = (double)m * n * k >= 1073741824.0 && k > 256;
  if (!big || g_arena || m <= 0 || n <= 0 || k <= 0) {
    // the triangular operands of these products (U, U^-1, V^-1) hold zeros in
    // their other triangle, so the fused kernel may read them as dense: one
    // launch instead of four for the many small products of the recursions
    if (k <= ZFUSED_KMAX)
      return zgemm_fused(s, ta, tb, m, n, k, alpha, Ar, Ai, lda, Br, Bi, ldb, beta, Cr, Ci, ldc, tri);
    return zgemm(s, ta, tb, m, n, k, alpha, Ar, Ai, lda, Br, Bi, ldb, beta, Cr, Ci, ldc, tri, atri, btri, ws, wsn);
  }
  const double sa = (ta == 'C' || ta == 'R') ? -1.0 : 1.0, sb = (tb == 'C' || tb == 'R') ? -1.0 : 1.0;
  const char ra = (ta == 'T' || ta == 'C') ? 'T' : 'N', rb = (tb == 'T' || tb == 'C') ? 'T' : 'N';
  // stored extents of A and B
  const int arows = ra == 'T' ? k : m, acols = ra == 'T' ? m : k;
  const int brows = rb == 'T' ? n : k, bcols = rb == 'T' ? k : n;
  const size_t na = (size_t)arows * acols, nbb = (size_t)brows * bcols, nt = (size_t)m * n;
  int rc;
  if ((rc = c.w_3m.alloc(na + nbb + nt))) return rc;
  double *SA = c.w_3m.p, *SB = SA + na, *T = SB + nbb;
  hipLaunchKernelGGL(k_zsum, dim3((arows + 255) / 256, acols), dim3(256), 0, s, arows, acols, Ar, Ai, lda, sa, SA);
  W2K_STAGE(s, "k_zsum");
  hipLaunchKernelGGL(k_zsum, dim3((brows + 255) / 256, bcols), dim3(256), 0, s, brows, bcols, Br, Bi, ldb, sb, SB);
  W2K_STAGE(s, "k_zsum");
  const dim3 gc((m + 255) / 256, n);
  if ((rc = dgemm(s, ra, rb, m, n, k, alpha, SA, arows, SB, brows, beta, Ci, ldc, tri, atri, btri, ws, wsn))) return rc;
  if ((rc = dgemm(s, ra, rb, m, n, k, 1.0, Ar, lda, Br, ldb, 0.0, T, m, tri, atri, btri, ws, wsn))) return rc;
  hipLaunchKernelGGL(k_z3m_comb, gc, dim3(256), 0, s, m, n, T, alpha, beta, Cr, Ci, ldc, tri, 0);
  W2K_STAGE(s, "k_z3m_comb");
  if ((rc = dgemm(s, ra, rb, m, n, k, sa * sb, Ai, lda, Bi, ldb, 0.0, T, m, tri, atri, btri, ws, wsn))) return rc;
  hipLaunchKernelGGL(k_z3m_comb, gc, dim3(256), 0, s, m, n, T, alpha, beta, Cr, Ci, ldc, tri, 1);
  W2K_STAGE(s, "k_z3m_comb");
  return 0;
}

__device__ __forceinline__ void cmad(double ar, double ai, double br, double bi, double &cr, double &ci)
{
  cr += ar * br - ai * bi;
  ci += ar * bi + ai * br;
}

// ------------------------------------------------------------------ set-up
// U (upper, diagonal from HSDIAG) = conj of the stored upper triangle;
// H full Hermitian from the stored lower triangle (conjugated back)
__global__ void k_zprep(int n, const double *__restrict__ hs, const double *__restrict__ hsdiag, double *Ur, double *Ui,
                        double *Hr, double *Hi)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const size_t o = (size_t)i + (size_t)n * j;
  if (i < j) {
    Ur[o] = hs[2 * o];
    Ui[o] = -hs[2 * o + 1];
    const size_t t = (size_t)j + (size_t)n * i;
    Hr[o] = hs[2 * t];
    Hi[o] = hs[2 * t + 1];
  } else if (i == j) {
    Ur[o] = hsdiag[2 * i];
    Ui[o] = -hsdiag[2 * i + 1];
    Hr[o] = hs[2 * o];
    Hi[o] = 0.0;
  } else {
    Ur[o] = 0.0;
    Ui[o] = 0.0;
    Hr[o] = hs[2 * o];
    Hi[o] = -hs[2 * o + 1];
  }
}

// unblocked complex Cholesky S = U^H U of the kb x kb diagonal block at k0
__global__ __launch_bounds__(256) void k_zpotf2(int n, int k0, int kb, double *__restrict__ Ur,
                                                double *__restrict__ Ui, int *__restrict__ info)
{
  __shared__ double ar[NB][NB + 1], ai[NB][NB + 1];
  __shared__ double piv;
  const int t = threadIdx.x;
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    const size_t o = (size_t)(k0 + r) + (size_t)n * (k0 + c);
    ar[r][c] = (r <= c) ? Ur[o] : 0.0;
    ai[r][c] = (r <= c) ? Ui[o] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; j++) {
    if (t == 0) {
      double ajj = ar[j][j];
      if (!(ajj > 0.0)) {
        if (*info == 0) *info = k0 + j + 1;
        ajj = 1.0;
      }
      piv = sqrt(ajj);
      ar[j][j] = piv;
      ai[j][j] = 0.0;
    }
    __syncthreads();
    for (int c = j + 1 + t; c < kb; c += blockDim.x) {
      ar[j][c] /= piv;
      ai[j][c] /= piv;
    }
    __syncthreads();
    // a[r][c] -= conj(a[j][r]) a[j][c], j < r <= c
    const int m = kb - j - 1;
    for (int idx = t; idx < m * m; idx += blockDim.x) {
      const int r = j + 1 + idx % m, c = j + 1 + idx / m;
      if (r <= c) {
        const double xr = ar[j][r], xi = -ai[j][r], yr = ar[j][c], yi = ai[j][c];
        ar[r][c] -= xr * yr - xi * yi;
        ai[r][c] -= xr * yi + xi * yr;
      }
    }
    __syncthreads();
  }
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    if (r <= c) {
      const size_t o = (size_t)(k0 + r) + (size_t)n * (k0 + c);
      Ur[o] = ar[r][c];
      Ui[o] = ai[r][c];
    }
  }
}

// R = S^-1 for a 64 x 64 upper triangular complex S in LDS (zero below the
// diagonal; R zero on entry), by all 256 threads of the block: the 16 x 16
// diagonal blocks by column back-substitution (threads 0..63, at most 16
// steps each), then two doubling steps X12 = -X11 (S12 X22) over the block's
// threads, through the product buffer P.  The longest dependent chain is
// about 140 complex multiply-adds instead of 2080 for back-substitution
// over all 64 rows.
__device__ void ztri_inv64(double (*sr)[NB + 1], double (*si)[NB + 1], double (*rr)[NB + 1], double (*ri)[NB + 1],
                           double (*pr)[NB / 2 + 1], double (*pi)[NB / 2 + 1], int t)
{
  static_assert(NB == 64, "64 x 64 blocks");
  if (t < NB) {
    const int b0 = (t / 16) * 16, c = t;
    for (int r = c; r >= b0; r--) {
      double xr = (r == c) ? 1.0 : 0.0, xi = 0.0;
      for (int q = r + 1; q <= c; q++) {
        xr -= sr[r][q] * rr[q][c] - si[r][q] * ri[q][c];
        xi -= sr[r][q] * ri[q][c] + si[r][q] * rr[q][c];
      }
      const double dr = sr[r][r], di = si[r][r], den = dr * dr + di * di;
      rr[r][c] = (xr * dr + xi * di) / den;
      ri[r][c] = (xi * dr - xr * di) / den;
    }
  }
  __syncthreads();
  for (int h = 16; h < NB; h *= 2) {
    const int nout = (NB / (2 * h)) * h * h;  // outputs of all pairs of this step
    for (int o = t; o < nout; o += blockDim.x) {
      const int pp = o / (h * h), rc = o % (h * h), r = rc % h, c = rc / h, b = 2 * h * pp;
      double ar = 0.0, ai = 0.0;  // P(r, c) = sum_{q <= c} S(b + r, b + h + q) X22(q, c)
      for (int q = 0; q <= c; q++) {
        const double s_r = sr[b + r][b + h + q], s_i = si[b + r][b + h + q];
        const double x_r = rr[b + h + q][b + h + c], x_i = ri[b + h + q][b + h + c];
        ar += s_r * x_r - s_i * x_i;
        ai += s_r * x_i + s_i * x_r;
      }
      pr[h * pp + r][c] = ar;
      pi[h * pp + r][c] = ai;
    }
    __syncthreads();
    for (int o = t; o < nout; o += blockDim.x) {
      const int pp = o / (h * h), rc = o % (h * h), r = rc % h, c = rc / h, b = 2 * h * pp;
      double ar = 0.0, ai = 0.0;  // X12(r, c) = -sum_{q >= r} X11(r, q) P(q, c)
      for (int q = r; q < h; q++) {
        const double x_r = rr[b + r][b + q], x_i = ri[b + r][b + q];
        const double p_r = pr[h * pp + q][c], p_i = pi[h * pp + q][c];
        ar += x_r * p_r - x_i * p_i;
        ai += x_r * p_i + x_i * p_r;
      }
      rr[b + r][b + h + c] = -ar;
      ri[b + r][b + h + c] = -ai;
    }
    __syncthreads();
  }
}

// inverse of the kb x kb upper block at (j0, j0) of U (cj = 0) or of conj(U)
// (cj = 1) into W (zero below the diagonal); a block shorter than 64 is
// padded with the identity
__global__ __launch_bounds__(256) void k_ztrtri_diag(int n, int j0, int kb, const double *__restrict__ Ur,
                                                     const double *__restrict__ Ui, int cj, double *__restrict__ Wr,
                                                     double *__restrict__ Wi)
{
  __shared__ double sr[NB][NB + 1], si[NB][NB + 1], rr[NB][NB + 1], ri[NB][NB + 1];
  __shared__ double pr[NB / 2][NB / 2 + 1], pi[NB / 2][NB / 2 + 1];
  const int t = threadIdx.x;
  const double sg = cj ? -1.0 : 1.0;
  for (int idx = t; idx < NB * NB; idx += blockDim.x) {
    const int r = idx % NB, c = idx / NB;
    double vr = (r == c) ? 1.0 : 0.0, vi = 0.0;
    if (r < kb && c < kb) {
      const size_t o = (size_t)(j0 + r) + (size_t)n * (j0 + c);
      vr = (r <= c) ? Ur[o] : 0.0;
      vi = (r <= c) ? sg * Ui[o] : 0.0;
    }
    sr[r][c] = vr;
    si[r][c] = vi;
    rr[r][c] = 0.0;
    ri[r][c] = 0.0;
  }
  __syncthreads();
  ztri_inv64(sr, si, rr, ri, pr, pi, t);
  for (int idx = t; idx < kb * kb; idx += blockDim.x) {
    const int r = idx % kb, c = idx / kb;
    const size_t o = (size_t)(j0 + r) + (size_t)n * (j0 + c);
    Wr[o] = (r <= c) ? rr[r][c] : 0.0;
    Wi[o] = (r <= c) ? ri[r][c] : 0.0;
  }
}

__global__ void k_zfill(size_t cnt, double *p, double v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = v;
}

__global__ void k_negate(size_t cnt, double *p)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = -p[i];
}

// Hermitian completion: C(j,i) = conj C(i,j) for i > j, real diagonal
__global__ void k_zmirror_lower(int n, double *Cr, double *Ci)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || i < j) return;
  if (i == j) {
    Ci[(size_t)i + (size_t)n * i] = 0.0;
    return;
  }
  const size_t lo = (size_t)i + (size_t)n * j, up = (size_t)j + (size_t)n * i;
  Cr[up] = Cr[lo];
  Ci[up] = -Ci[lo];
}

// ------------------------------------------------------------------ tridiagonalisation
// ZLATRD panels (lower), as the real build (seclr4.hip) with complex
// arithmetic: V(:,k) at panel column k, W(:,k) at nb+k, V again at 2nb+k;
// the SYMV adds A_IJ v_J to tile I and A_IJ^H v_I to tile J.
struct ZTriWork {
  int n, p0, nb;
  double *ar, *ai;          // n x n Hermitian (both triangles)
  double *Pr, *Pi;          // panel buffer n x 3 TNB
  double *d, *e;            // real tridiagonal
  double *taur, *taui;      // reflector scalars
  double *scr, *sci;        // 1 / (alpha - beta)
  double *part;             // SYMV partials: 256 per tile (zpart_off)
  double *dotr, *doti;      // tiles x 2 TNB partial V^H v, W^H v
  double *vy;               // partial Re(v^H y) per tile pair
  double *norm;             // per block partial |x|^2
  int *counter;
};

constexpr int ZCOL_ROWS = 64;  // rows per block of the column kernel (one per lane)

// ZLARFG of column i on (A(i+1,i), A(i+2:n,i)) from the squared-norm
// partials of the column kernel (one per block of its rows), evaluated by
// every SYMV wave with the same fixed-order arithmetic (so no grid-wide step
// separates the two kernels); the wave with store set writes E, TAU and the
// scale 1/(alpha - beta) of v.
__device__ __forceinline__ void ztri_reflector(const ZTriWork &g, int i, int lane, bool store, double &scr,
                                               double &sci)
{
  const int n = g.n;
  const int nnorm = (n - i + ZCOL_ROWS - 1) / ZCOL_ROWS;  // blocks of the column kernel
  double x[4];
#pragma unroll
  for (int u = 0; u < 4; u++) x[u] = g.norm[min(lane + 64 * u, nnorm - 1)];
  const double alr = g.ar[(size_t)n * i + i + 1], ali = g.ai[(size_t)n * i + i + 1];
  double xs = 0.0;
#pragma unroll
  for (int u = 0; u < 4; u++) xs += (lane + 64 * u < nnorm) ? x[u] : 0.0;
  for (int b = lane + 256; b < nnorm; b += 64) xs += g.norm[b];
  xs = wave_sum(xs);
  xs = __shfl(xs, 0, 64);
  const double xnorm = sqrt(xs);
  double taur = 0.0, taui = 0.0, beta = alr;
  scr = 0.0;
  sci = 0.0;
  if (!(xnorm == 0.0 && ali == 0.0)) {
    beta = -copysign(sqrt(alr * alr + ali * ali + xnorm * xnorm), alr);
    taur = (beta - alr) / beta;
    taui = -ali / beta;
    const double zr = alr - beta, zi = ali, den = zr * zr + zi * zi;
    scr = zr / den;
    sci = -zi / den;
  }
  if (store && lane == 0) {
    g.e[i] = beta;
    g.taur[i] = taur;
    g.taui[i] = taui;
    g.scr[i] = scr;
    g.sci[i] = sci;
  }
}

// y = A(r0:n, r0:n) v for Hermitian A from the tiles on and below the
// diagonal (one wave per 64 x 64 tile (I, J), J <= I): A_IJ v_J to the rows
// of tile I and, for J < I, A_IJ^H v_I to the rows of tile J.  A tile is
// taken in four quarters of 16 columns: the real and imaginary parts of a
// quarter-row fill 32 of the lane's registers, row sums stay in the lane and
// column sums use a wave-wide reduce-scatter (no LDS); v(C0 + j) reaches the
// row sums from lane j by readlane.  The first quarter's loads are issued
// before the reflector is formed and each further quarter's as soon as the
// previous one has been reduced, so a tile wave always has a quarter in
// flight.  The 4 nt waves after the tiles, four per row block, write V(:,k)
// and the partial dot products V^H v, W^H v of the block's rows.
//
// Each tile wave writes its partial sums into a block of its own, 256
// doubles at bid * 256: the row sums (Re, Im: 64 each), then the column sums
// (Re, Im).  Consecutive waves so write consecutive memory.  The writes are
// about 3% of the kernel's bytes but cost far more inside its read stream: on
// the order-12445 trailing matrix the kernel takes 222 us without them, 268
// us with the sums laid out as tiles x n vectors (every 512-byte segment in
// another place) and 251 us with this layout.
__device__ __forceinline__ double dbl_readlane(double v, int src)
{
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// columns of the tile taken per step of the SYMV's tile loop (8-column steps
// at four or five waves per SIMD measured slower on the 100-atom cell)
constexpr int ZQC = 16;
static_assert(ZQC == 16, "the column sums take 32 entries per step");

__global__ __launch_bounds__(64, 3) void k_ztri_symv(ZTriWork g, int i, int nt)
{
  const int n = g.n, r0 = i + 1, ii = i - g.p0, nb = g.nb;
  const int lane = threadIdx.x;
  const int nblk = nt * (nt + 1) / 2;
  const int bid = blockIdx.x;
  const double *air = g.ar + (size_t)n * i, *aii = g.ai + (size_t)n * i;
  // tile waves: the tile (I, J), J <= I, and its first 16 columns, loaded
  // before the reflector is formed so that their latency overlaps it
  const bool tile = bid < nblk;
  int I = 0, J = 0;
  if (tile) {
    I = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= bid) I++;
    while (I * (I + 1) / 2 > bid) I--;
    J = bid - I * (I + 1) / 2;
  }
  const int C0t = r0 + SYMV_T * J;
  const int mt = min(SYMV_T, n - C0t);
  const int Rlt = min(r0 + SYMV_T * I + lane, n - 1);
  const double *apr = g.ar + (size_t)Rlt + (size_t)n * C0t, *api = g.ai + (size_t)Rlt + (size_t)n * C0t;
  double q[2 * ZQC];  // q[j] = Re A(R, C0 + c0 + j), q[ZQC + j] = Im
  if (tile) {
#pragma unroll
    for (int j = 0; j < ZQC; j++) {
      const int jj = min(j, mt - 1);
      q[j] = apr[(size_t)n * jj];
      q[ZQC + j] = api[(size_t)n * jj];
    }
  }
  double sr, si;
  ztri_reflector(g, i, lane, bid == 0, sr, si);
  // v(x): 1 at r0, A(x, i) (sr + i si) below it, 0 outside [r0, n)
  auto vof = [&](int x, double &vr, double &vi) {
    const int xc = min(max(x, r0), n - 1);
    const double ar = air[xc], ai = aii[xc];
    vr = 0.0;
    vi = 0.0;
    if (x == r0) {
      vr = 1.0;
    } else if (x > r0 && x < n) {
      vr = ar * sr - ai * si;
      vi = ar * si + ai * sr;
    }
  };
  if (bid >= nblk) {
    // four waves per row block: V then W, each Re then Im of conj(X(R,k)) v(R)
    const int I = (bid - nblk) >> 2, pass = (bid - nblk) & 3;
    const int R = r0 + SYMV_T * I + lane;
    const bool live = R < n;
    const int Rs = live ? R : r0;
    double vrr, vri;
    vof(R, vrr, vri);
    if (pass == 0) {
      double *Vr = g.Pr + (size_t)n * ii, *Vi = g.Pi + (size_t)n * ii;
      double *V2r = g.Pr + (size_t)n * (2 * nb + ii), *V2i = g.Pi + (size_t)n * (2 * nb + ii);
      if (live) { Vr[R] = vrr; Vi[R] = vri; V2r[R] = vrr; V2i[R] = vri; }
      for (int q = I * SYMV_T + lane; q < r0; q += nt * SYMV_T) { Vr[q] = 0.0; Vi[q] = 0.0; V2r[q] = 0.0; V2i[q] = 0.0; }
    }
    if (ii > 0) {
      double *dr = g.dotr + (size_t)I * 2 * TNB, *di = g.doti + (size_t)I * 2 * TNB;
      const int half = pass >> 1, im = pass & 1;
      const double *Xr = g.Pr + (size_t)n * (half * nb), *Xi = g.Pi + (size_t)n * (half * nb);
      double t[TNB];  // 32 values per wave keep the branch's registers at the tile loop's level
#pragma unroll
      for (int k = 0; k < TNB; k++) {
        const double ur = Xr[(size_t)n * k + Rs], ui = Xi[(size_t)n * k + Rs];
        const bool use = live && k < ii;
        t[k] = use ? (im ? ur * vri - ui * vrr : ur * vrr + ui * vri) : 0.0;
      }
      const double sq = wave_transpose_sum32x(t, lane);  // lanes 2k, 2k + 1: column k
      if (!(lane & 1)) (im ? di : dr)[half * TNB + (lane >> 1)] = sq;
    }
    return;
  }
  const int R = r0 + SYMV_T * I + lane;
  const int C0 = r0 + SYMV_T * J;
  const bool live = R < n;
  const bool off = (I != J);
  const int m = min(SYMV_T, n - C0);
  double vcr, vci, vrr, vri;
  vof(C0 + lane, vcr, vci);  // v of the lane's column
  vof(R, vrr, vri);          // v of the lane's row
  double yr = 0.0, yi = 0.0, vyc = 0.0;
  // a full off-diagonal tile (all 64 rows and columns inside the matrix)
  // needs no masks; wave-uniform
  const bool full = off && m == SYMV_T && r0 + SYMV_T * (I + 1) <= n;
  // only the lower triangle of the trailing matrix is kept current: a
  // diagonal tile uses its lower part by rows and its strict lower part by
  // columns (A^H v), the column sums folded into the rows they belong to
#pragma unroll 1
  for (int c0 = 0; c0 < m; c0 += ZQC) {  // steps of ZQC columns (wave-uniform); q holds step c0
    if (!full) {
#pragma unroll
      for (int j = 0; j < ZQC; j++) {
        const bool ok = live && c0 + j < m && (off || c0 + j <= lane);
        q[j] = ok ? q[j] : 0.0;
        q[ZQC + j] = ok ? q[ZQC + j] : 0.0;
      }
    }
    {
      // four independent accumulator pairs (a chain of 8 multiply-adds
      // instead of 32 per quarter), added in a fixed order
      double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int j = 0; j < ZQC; j++)
        cmad(q[j], q[ZQC + j], dbl_readlane(vcr, c0 + j), dbl_readlane(vci, c0 + j), ar[j & 3], ai[j & 3]);
      yr += (ar[0] + ar[1]) + (ar[2] + ar[3]);
      yi += (ai[0] + ai[1]) + (ai[2] + ai[3]);
    }
#pragma unroll
    for (int j = 0; j < ZQC; j++) {  // conj(A(R,c)) v(R); the diagonal tile's diagonal only by rows
      const bool use = off || c0 + j < lane;
      const double ar = q[j], ai = q[ZQC + j];
      q[j] = use ? ar * vrr + ai * vri : 0.0;
      q[ZQC + j] = use ? ar * vri - ai * vrr : 0.0;
    }
    // lanes 2e and 2e + 1: entry e summed over the wave (Re of column e in
    // lanes 0..31, Im of column e - 16 in lanes 32..63)
    const double cs = wave_transpose_sum32x(q, lane);
    if (c0 + ZQC < m) {  // the next step's columns (q is free again)
#pragma unroll
      for (int j = 0; j < ZQC; j++) {
        const int jj = min(c0 + ZQC + j, m - 1);
        q[j] = apr[(size_t)n * jj];
        q[ZQC + j] = api[(size_t)n * jj];
      }
    }
    if (off) {
      const double co = xor32(cs, lane);  // the other part of the same column
      const int cl = c0 + ((lane >> 1) & (ZQC - 1));
      if (!(lane & 1) && cl < m) g.part[(size_t)bid * 256 + 128 + ((lane & 32) ? 64 : 0) + cl] = cs;
      const double vcl_r = __shfl(vcr, cl, 64), vcl_i = __shfl(vci, cl, 64);
      if (lane < 32 && !(lane & 1) && cl < m) vyc += vcl_r * cs + vcl_i * co;  // Re(conj(v_c) (A^H v)_c)
    } else {
      // column c0 + j of the tile is row C0 + c0 + j: lane c0 + j takes it
      const int src = lane - c0;
      const bool mine = src >= 0 && src < ZQC;
      const double fr = __shfl(cs, mine ? 2 * src : 0, 64), fi = __shfl(cs, mine ? 32 + 2 * src : 32, 64);
      if (mine) {
        yr += fr;
        yi += fi;
      }
    }
  }
  if (live) {
    g.part[(size_t)bid * 256 + lane] = yr;
    g.part[(size_t)bid * 256 + 64 + lane] = yi;
  }
  double vy = vrr * yr + vri * yi + vyc;  // Re(conj(v) y)
  vy = wave_sum(vy);
  if (lane == 0) g.vy[bid] = vy;
}

// Finish W(:, fin - p0) (fin >= 0) and update column col (col >= 0) with the
// panel so far (ZLATRD); write the squared-norm partial of its
// sub-subdiagonal part per block (ZLARFG itself is evaluated by the next
// SYMV's waves).  A block takes ZCOL_ROWS = 64 consecutive rows, one per
// lane, in each of its ZCOL_WAVES waves, so that every load of a SYMV partial
// or a panel column is one contiguous 512-byte row segment; the waves split
// the SYMV partials (chunk w, w + 8, ...) and the panel columns (four each)
// and their row sums meet in LDS.  Loads are issued in batches with clamped
// indices and masked values, so that a batch is in flight at once.
constexpr int ZCOL_WAVES = 8;
static_assert(ZCOL_ROWS == SYMV_T, "a column-kernel block reads the partials of one SYMV row block");

// offset of partial c (c < nt) of the rows of SYMV row block I, Re (Im at
// +64): the row sums of tile (I, c) for c <= I, else the column sums of tile
// (c, I); the lane's row at + lane
__device__ __forceinline__ size_t zpart_off(int I, int c)
{
  return c <= I ? (size_t)(I * (I + 1) / 2 + c) * 256 : (size_t)(c * (c + 1) / 2 + I) * 256 + 128;
}

__global__ __launch_bounds__(64 * ZCOL_WAVES) void k_ztri_col(ZTriWork g, int fin, int col, int ntile, int nch,
                                                              int nvy)
{
  constexpr int NT = 64 * ZCOL_WAVES;   // threads
  constexpr int KW = TNB / ZCOL_WAVES;  // panel columns per wave
  constexpr int YB = 16;                // SYMV-partial chunks per wave and batch
  constexpr int VB = 10;                // first batch of v^H y pairs per thread (nvy <= 2 NT VB)
  constexpr int DG = NT / 32;           // groups of the V^H v, W^H v sums (32 column pairs each)
  constexpr int DB = 8;                 // first batch of tiles per group (ntile <= DG DB)
  const int n = g.n, p0 = g.p0, nb = g.nb, t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int rbase = fin >= 0 ? fin + 1 : col;
  const int r = rbase + blockIdx.x * ZCOL_ROWS + lane;
  const bool live = r < n;
  const int rs = live ? r : rbase;
  const int ii = fin >= 0 ? fin - p0 : -1;
  const int jj = col - p0;
  __shared__ double abr[2 * TNB], abi[2 * TNB];  // a_k = V_k^H v, b_k = W_k^H v
  __shared__ double wcr[TNB], wci[TNB], vcr[TNB], vci[TNB];
  __shared__ double dred[2][DG][2 * TNB];
  __shared__ double red[6][ZCOL_WAVES][ZCOL_ROWS];  // per wave: y, z (W a + V b), a0 (re, im)
  __shared__ double sh[3][ZCOL_WAVES];
  const double *Vr = g.Pr, *Vi = g.Pi, *Wr = g.Pr + (size_t)n * nb, *Wi = g.Pi + (size_t)n * nb;
  double *acr = g.ar + (size_t)n * (col >= 0 ? col : 0), *aci = g.ai + (size_t)n * (col >= 0 ? col : 0);
  // ---- per-block scalars: column col of the panel, v^H y, V^H v and W^H v
  if (col >= 0 && t < jj) {
    vcr[t] = Vr[(size_t)n * t + col];
    vci[t] = Vi[(size_t)n * t + col];
    if (t < ii || fin < 0) {
      wcr[t] = Wr[(size_t)n * t + col];
      wci[t] = Wi[(size_t)n * t + col];
    }
  }
  if (fin >= 0) {
    // v^H y: every partial once, in pairs
    {
      const int vlast = (nvy - 1) & ~1;
      double2 xv[VB];
#pragma unroll
      for (int u = 0; u < VB; u++) xv[u] = *reinterpret_cast<const double2 *>(g.vy + min(2 * t + 2 * NT * u, vlast));
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int u = 0; u < VB; u++) {
        const int j = 2 * t + 2 * NT * u;
        s0 += (j < nvy) ? xv[u].x : 0.0;
        s1 += (j + 1 < nvy) ? xv[u].y : 0.0;
      }
      for (int j = 2 * t + 2 * NT * VB; j < nvy; j += 2 * NT) {
        s0 += g.vy[j];
        s1 += (j + 1 < nvy) ? g.vy[j + 1] : 0.0;
      }
      const double vs = wave_sum(s0 + s1);
      if (lane == 0) sh[0][w] = vs;
    }
    // row col's SYMV partials (for W(col, ii))
    if (col >= 0) {
      double pr = 0.0, pi = 0.0;
      for (int c = t; c < nch; c += NT) {  // row col is row 0 of row block 0
        const size_t o = zpart_off(0, c);
        pr += g.part[o];
        pi += g.part[o + 64];
      }
      pr = wave_sum(pr);
      pi = wave_sum(pi);
      if (lane == 0) {
        sh[1][w] = pr;
        sh[2][w] = pi;
      }
    }
    // V^H v, W^H v: columns 2dq, 2dq + 1 of tiles dp, dp + DG, ...
    {
      const int dq = t & 31, dp = t >> 5;
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      {
        double2 xd[DB], zd[DB];
#pragma unroll
        for (int u = 0; u < DB; u++) {
          const size_t o = (size_t)min(dp + DG * u, ntile - 1) * 2 * TNB + 2 * dq;
          xd[u] = *reinterpret_cast<const double2 *>(g.dotr + o);
          zd[u] = *reinterpret_cast<const double2 *>(g.doti + o);
        }
#pragma unroll
        for (int u = 0; u < DB; u++) {
          const bool k = dp + DG * u < ntile;
          s0 += k ? xd[u].x : 0.0;
          s1 += k ? xd[u].y : 0.0;
          s2 += k ? zd[u].x : 0.0;
          s3 += k ? zd[u].y : 0.0;
        }
      }
      for (int b = dp + DG * DB; b < ntile; b += DG) {
        const size_t o = (size_t)b * 2 * TNB + 2 * dq;
        const double2 x = *reinterpret_cast<const double2 *>(g.dotr + o);
        const double2 z = *reinterpret_cast<const double2 *>(g.doti + o);
        s0 += x.x;
        s1 += x.y;
        s2 += z.x;
        s3 += z.y;
      }
      dred[0][dp][2 * dq] = s0;
      dred[0][dp][2 * dq + 1] = s1;
      dred[1][dp][2 * dq] = s2;
      dred[1][dp][2 * dq + 1] = s3;
    }
  }
  // ---- this wave's share of the rows' sums
  double kvr[KW], kvi[KW], kwr[KW], kwi[KW];  // V(r, k), W(r, k), k = KW w + u
#pragma unroll
  for (int u = 0; u < KW; u++) {
    const size_t o = (size_t)n * (KW * w + u) + rs;
    kvr[u] = Vr[o];
    kvi[u] = Vi[o];
    kwr[u] = Wr[o];
    kwi[u] = Wi[o];
  }
  double yr = 0.0, yi = 0.0;
  if (fin >= 0) {
    for (int c0 = w; c0 < nch; c0 += ZCOL_WAVES * YB) {
      double xr[YB], xi[YB];
#pragma unroll
      for (int u = 0; u < YB; u++) {
        const size_t o = zpart_off(blockIdx.x, min(c0 + ZCOL_WAVES * u, nch - 1)) + lane;
        xr[u] = g.part[o];
        xi[u] = g.part[o + 64];
      }
#pragma unroll
      for (int u = 0; u < YB; u++) {  // rows past n: never written, selected away
        const bool k = live && c0 + ZCOL_WAVES * u < nch;
        yr += k ? xr[u] : 0.0;
        yi += k ? xi[u] : 0.0;
      }
    }
  }
  __syncthreads();
  if (fin >= 0 && t < 2 * TNB) {
    double qr = 0.0, qi = 0.0;
#pragma unroll
    for (int p = 0; p < DG; p++) {
      qr += dred[0][p][t];
      qi += dred[1][p][t];
    }
    const bool use = (t & (TNB - 1)) < ii;
    abr[t] = use ? qr : 0.0;
    abi[t] = use ? qi : 0.0;
  }
  __syncthreads();
  {
    double zr = 0.0, zi = 0.0, a0r = 0.0, a0i = 0.0;
#pragma unroll
    for (int u = 0; u < KW; u++) {
      const int k = KW * w + u;
      if (fin >= 0 && k < ii) {  // W(r,k) a_k + V(r,k) b_k
        zr += kwr[u] * abr[k] - kwi[u] * abi[k] + kvr[u] * abr[TNB + k] - kvi[u] * abi[TNB + k];
        zi += kwr[u] * abi[k] + kwi[u] * abr[k] + kvr[u] * abi[TNB + k] + kvi[u] * abr[TNB + k];
      }
      // V(r,k) conj(W(col,k)) + W(r,k) conj(V(col,k)), k < jj, k != ii; panel
      // columns k >= jj hold values of an earlier panel (or never written
      // memory): selected away, never multiplied by zero
      const bool use = col >= 0 && k < jj && k != ii;
      const double wcr_ = use ? wcr[min(k, TNB - 1)] : 0.0, wci_ = use ? wci[min(k, TNB - 1)] : 0.0;
      const double vcr_ = use ? vcr[min(k, TNB - 1)] : 0.0, vci_ = use ? vci[min(k, TNB - 1)] : 0.0;
      const double vr_ = use ? kvr[u] : 0.0, vi_ = use ? kvi[u] : 0.0;
      const double wr_ = use ? kwr[u] : 0.0, wi_ = use ? kwi[u] : 0.0;
      a0r += vr_ * wcr_ + vi_ * wci_ + wr_ * vcr_ + wi_ * vci_;
      a0i += vi_ * wcr_ - vr_ * wci_ + wi_ * vcr_ - wr_ * vci_;
    }
    red[0][w][lane] = yr;
    red[1][w][lane] = yi;
    red[2][w][lane] = zr;
    red[3][w][lane] = zi;
    red[4][w][lane] = a0r;
    red[5][w][lane] = a0i;
  }
  __syncthreads();
  // W(q, ii) = 0 above this kernel's rows
  if (fin >= 0) {
    double *Wir = g.Pr + (size_t)n * (nb + fin - p0), *Wii = g.Pi + (size_t)n * (nb + fin - p0);
    for (int q = blockIdx.x * NT + t; q < rbase; q += gridDim.x * NT) {
      Wir[q] = 0.0;
      Wii[q] = 0.0;
    }
  }
  if (w != 0) return;
  // ---- wave 0 finishes its 64 rows
  double sy[6];
#pragma unroll
  for (int a = 0; a < 6; a++) {
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < ZCOL_WAVES; p++) s += red[a][p][lane];
    sy[a] = s;
  }
  double wrr = 0.0, wri = 0.0, viir = 0.0, viii = 0.0, wciir = 0.0, wciii = 0.0;
  if (fin >= 0) {
    const double tr = g.taur[fin], ti = g.taui[fin];
    viir = Vr[(size_t)n * ii + rs];
    viii = Vi[(size_t)n * ii + rs];
    double vty = 0.0;
#pragma unroll
    for (int p = 0; p < ZCOL_WAVES; p++) vty += sh[0][p];
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < TNB; k++) ssum += abr[k] * abr[TNB + k] + abi[k] * abi[TNB + k];
    const double alpha2 = -0.5 * (tr * tr + ti * ti) * (vty - 2.0 * ssum);
    const double yr_ = sy[0] - sy[2], yi_ = sy[1] - sy[3];
    wrr = tr * yr_ - ti * yi_ + alpha2 * viir;
    wri = tr * yi_ + ti * yr_ + alpha2 * viii;
    double *Wir = g.Pr + (size_t)n * (nb + ii), *Wii = g.Pi + (size_t)n * (nb + ii);
    if (live) {
      Wir[r] = wrr;
      Wii[r] = wri;
    }
    if (col >= 0) {
      // row col = fin + 1 of the same formula (v(col) = 1), in every lane
      double cr = 0.0, ci = 0.0;
      for (int k = 0; k < ii; k++) {
        cr += wcr[k] * abr[k] - wci[k] * abi[k] + vcr[k] * abr[TNB + k] - vci[k] * abi[TNB + k];
        ci += wcr[k] * abi[k] + wci[k] * abr[k] + vcr[k] * abi[TNB + k] + vci[k] * abr[TNB + k];
      }
      double pr = 0.0, pi = 0.0;
#pragma unroll
      for (int p = 0; p < ZCOL_WAVES; p++) {
        pr += sh[1][p];
        pi += sh[2][p];
      }
      const double sr = pr - cr, si = pi - ci;
      wciir = tr * sr - ti * si + alpha2;
      wciii = tr * si + ti * sr;
    }
  }
  if (col < 0) return;
  double xr = acr[rs], xi = aci[rs];
  {
    double a0r = sy[4], a0i = sy[5];
    if (fin >= 0) {  // k = ii: this kernel's W(r, ii) and W(col, ii)
      const double vr_ = vcr[ii], vi_ = vci[ii];
      a0r += viir * wciir + viii * wciii + wrr * vr_ + wri * vi_;
      a0i += viii * wciir - viir * wciii + wri * vr_ - wrr * vi_;
    }
    xr -= a0r;
    xi -= a0i;
  }
  double nrm = 0.0;
  if (live && r >= col) {
    if (r == col) xi = 0.0;
    acr[r] = xr;
    aci[r] = xi;
    if (r == col) {
      g.d[col] = xr;
      if (col >= n - 1) {
        g.taur[col] = 0.0;
        g.taui[col] = 0.0;
        g.scr[col] = 0.0;
        g.sci[col] = 0.0;
      }
    }
    if (r >= col + 2) nrm = xr * xr + xi * xi;
  }
  nrm = wave_sum(nrm);
  if (lane == 0) g.norm[blockIdx.x] = nrm;
}

__global__ void k_ztri_store(ZTriWork g, int ncols)
{
  const int n = g.n;
  const int k = blockIdx.y;
  const int i = g.p0 + k;
  const int r = i + 2 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncols || r >= n) return;
  g.ar[(size_t)r + (size_t)n * i] = g.Pr[(size_t)n * k + r];
  g.ai[(size_t)r + (size_t)n * i] = g.Pi[(size_t)n * k + r];
}

// ------------------------------------------------------------------ back-transforms
__global__ void k_zvpanel(int n, int p0, int nb, const double *__restrict__ ar, const double *__restrict__ ai,
                          double *__restrict__ vr, double *__restrict__ vi)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (r >= n || k >= nb) return;
  const int i = p0 + k;
  double xr = 0.0, xi = 0.0;
  if (r == i + 1) xr = 1.0;
  else if (r > i + 1) { xr = ar[(size_t)r + (size_t)n * i]; xi = ai[(size_t)r + (size_t)n * i]; }
  vr[(size_t)r + (size_t)n * k] = xr;
  vi[(size_t)r + (size_t)n * k] = xi;
}

// T of I - V T V^H from G = V^H V (ZLARFT 'F','C'), whose recurrence
// T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i) says T = M^-1 with M upper
// triangular, M(i,i) = 1/tau_i and M(q,i) = G(q,i) above the diagonal.  A
// reflector with tau_i = 0 is the identity: its row and column of M are
// cleared (M(i,i) = 1) and T(i,i) is set to 0 afterwards.  nb <= 64.
__global__ __launch_bounds__(256) void k_zlarft(int nb, const double *__restrict__ gr, const double *__restrict__ gi,
                                                const double *__restrict__ taur, const double *__restrict__ taui,
                                                double *__restrict__ tmr, double *__restrict__ tmi)
{
  __shared__ double sr[NB][NB + 1], si[NB][NB + 1], rr[NB][NB + 1], ri[NB][NB + 1];
  __shared__ double pr[NB / 2][NB / 2 + 1], pi[NB / 2][NB / 2 + 1];
  __shared__ int zero[NB];
  const int t = threadIdx.x;
  if (t < NB) zero[t] = !(t < nb && (taur[t] != 0.0 || taui[t] != 0.0));
  __syncthreads();
  for (int idx = t; idx < NB * NB; idx += blockDim.x) {
    const int r = idx % NB, c = idx / NB;
    double vr = 0.0, vi = 0.0;
    if (r == c) {
      if (zero[r]) {
        vr = 1.0;
      } else {
        const double a = taur[r], b = taui[r], den = a * a + b * b;
        vr = a / den;
        vi = -b / den;
      }
    } else if (r < c && !zero[r] && !zero[c]) {
      vr = gr[r + nb * c];
      vi = gi[r + nb * c];
    }
    sr[r][c] = vr;
    si[r][c] = vi;
    rr[r][c] = 0.0;
    ri[r][c] = 0.0;
  }
  __syncthreads();
  ztri_inv64(sr, si, rr, ri, pr, pi, t);
  for (int idx = t; idx < nb * nb; idx += blockDim.x) {
    const int r = idx % nb, c = idx / nb;
    const bool keep = r <= c && !(r == c && zero[r]);
    tmr[r + nb * c] = keep ? rr[r][c] : 0.0;
    tmi[r + nb * c] = keep ? ri[r][c] : 0.0;
  }
}

__global__ void k_zgather_cols(int n, int m, const int *__restrict__ perm, const double *__restrict__ sr,
                               const double *__restrict__ si, double *__restrict__ dr, double *__restrict__ di)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= n || c >= m) return;
  dr[(size_t)r + (size_t)n * c] = sr[(size_t)r + (size_t)n * perm[c]];
  di[(size_t)r + (size_t)n * c] = si[(size_t)r + (size_t)n * perm[c]];
}

__global__ void k_interleave(size_t cnt, const double *__restrict__ re, const double *__restrict__ im,
                             double *__restrict__ out)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  out[2 * i] = re[i];
  out[2 * i + 1] = im[i];
}

// ------------------------------------------------------------------ driver
int seclr4_complex(Ctx &c, double el, double eu, int nume, int *ne_out, int *nbelw_out)
{
  const int n = c.n;
  hipStream_t s = c.stream;
  int rc;
  const size_t nn = (size_t)n * n;
  // planar complex workspaces: U, H then C (then reflectors), W = V^-1, X
  if ((rc = c.w_a.alloc(2 * nn)) || (rc = c.w_b.alloc(2 * nn)) || (rc = c.w_c.alloc(2 * nn))) return rc;
  if ((rc = c.w_x.alloc(2 * nn))) return rc;
  if ((rc = c.w_gk.alloc((size_t)1 << 22))) return rc;
  double *Ur = c.w_a.p, *Ui = Ur + nn, *Hr = c.w_b.p, *Hi = Hr + nn, *Wr = c.w_c.p, *Wi = Wr + nn;
  double *Xr = c.w_x.p, *Xi = Xr + nn;
  double *ws = c.w_gk.p;
  const size_t wsn = c.w_gk.n;
  Seclr4Bufs B;
  if ((rc = seclr4_buffers(c, n, 1, B))) return rc;

  W2K_HIP(hipEventRecord(c.ev[6], s));
  // ---- 1. Cholesky S = U^H U
  {
    dim3 g2((n + 255) / 256, n);
    hipLaunchKernelGGL(k_zprep, g2, dim3(256), 0, s, n, c.d_hs.p, c.d_hsdiag.p, Ur, Ui, Hr, Hi);
    W2K_STAGE(s, "k_zprep");
    // W: U^-1 (zero below the diagonal), formed with U and conjugated by the
    // reduction
    hipLaunchKernelGGL(k_zfill, dim3((unsigned)((2 * nn + 255) / 256)), dim3(256), 0, s, 2 * nn, Wr, 0.0);
    W2K_STAGE(s, "k_zfill");
    // S = U^H U and W = U^-1 together, by recursive halving over the
    // NB-blocks: with S = [S11 S12; . S22], U11 and W11 from the first half,
    // U12 = W11^H S12, S22 -= U12^H U12, U22 and W22 from the second half and
    // W12 = -W11 U12 W22.  The block kernels factor and invert one NB-block;
    // the rest is a few large products instead of one thin product per
    // block row.
    std::function<int(int, int)> chol = [&](int lo, int hi) -> int {
      const int nblk = (hi - lo + NB - 1) / NB;
      if (nblk <= 1) {
        const int kb = hi - lo;
        hipLaunchKernelGGL(k_zpotf2, dim3(1), dim3(256), 0, s, n, lo, kb, Ur, Ui, B.info);
        W2K_STAGE(s, "k_zpotf2");
        hipLaunchKernelGGL(k_ztrtri_diag, dim3(1), dim3(256), 0, s, n, lo, kb, Ur, Ui, 0, Wr, Wi);
        W2K_STAGE(s, "k_ztrtri_diag");
        return 0;
      }
      const int mid = lo + (nblk / 2) * NB;
      const int s1 = mid - lo, s2 = hi - mid;
      const size_t o11 = (size_t)lo + (size_t)n * lo, o12 = (size_t)lo + (size_t)n * mid;
      const size_t o22 = (size_t)mid + (size_t)n * mid;
      int r;
      if ((r = chol(lo, mid))) return r;
      // U12 = W11^H S12 (W11^H lower triangular), through X
      if ((r = zgemm3(c, s, 'C', 'N', s1, s2, s1, 1.0, Wr + o11, Wi + o11, n, Ur + o12, Ui + o12, n, 0.0, Xr, Xi, n, 0,
                      2, 0, ws, wsn)))
        return r;
      W2K_HIP(hipMemcpy2DAsync(Ur + o12, sizeof(double) * n, Xr, sizeof(double) * n, sizeof(double) * s1, s2,
                               hipMemcpyDeviceToDevice, s));
      W2K_HIP(hipMemcpy2DAsync(Ui + o12, sizeof(double) * n, Xi, sizeof(double) * n, sizeof(double) * s1, s2,
                               hipMemcpyDeviceToDevice, s));
      // S22 -= U12^H U12 (upper triangle)
      if ((r = zgemm3(c, s, 'C', 'N', s2, s2, s1, -1.0, Ur + o12, Ui + o12, n, Ur + o12, Ui + o12, n, 1.0, Ur + o22,
                      Ui + o22, n, 2)))
        return r;
      if ((r = chol(mid, hi))) return r;
      // W12 = -(W11 U12) W22 (W11, W22 upper triangular)
      if ((r = zgemm3(c, s, 'N', 'N', s1, s2, s1, 1.0, Wr + o11, Wi + o11, n, Ur + o12, Ui + o12, n, 0.0, Xr, Xi, n, 0,
                      1, 0, ws, wsn)))
        return r;
      return zgemm3(c, s, 'N', 'N', s1, s2, s2, -1.0, Xr, Xi, n, Wr + o22, Wi + o22, n, 0.0, Wr + o12, Wi + o12, n, 0,
                    0, 1, ws, wsn);
    };
    if ((rc = chol(0, n))) return rc;
  }
  W2K_HIP(hipEventRecord(c.ev[7], s));
  int hinfo[4];
  W2K_HIP(hipMemcpyAsync(hinfo, B.info, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  if (hinfo[0] != 0) return hinfo[0];  // ZPOTRF INFO, seclr4.F:200-205
  // ---- 2. reduction with V = conj(U): W = V^-1, X = H W, C = W^H X
  {
    // V^-1 = conj(U^-1), the Cholesky's W conjugated
    hipLaunchKernelGGL(k_negate, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nn, Wi);
    W2K_STAGE(s, "k_negate");
    if ((rc = zgemm3(c, s, 'N', 'N', n, n, n, 1.0, Hr, Hi, n, Wr, Wi, n, 0.0, Xr, Xi, n, 0, 0, 1))) return rc;
    if ((rc = zgemm3(c, s, 'C', 'N', n, n, n, 1.0, Wr, Wi, n, Xr, Xi, n, 0.0, Hr, Hi, n, 1, 2, 0))) return rc;
    dim3 g2((n + 255) / 256, n);
    hipLaunchKernelGGL(k_zmirror_lower, g2, dim3(256), 0, s, n, Hr, Hi);
    W2K_STAGE(s, "k_zmirror_lower");
  }
  W2K_HIP(hipEventRecord(c.ev[8], s));
  // ---- 3. Hermitian tridiagonalisation of C (in H)
  const int nts = B.nt;
  ZTriWork g;
  g.n = n; g.ar = Hr; g.ai = Hi;
  g.Pr = B.P; g.Pi = B.P + (size_t)3 * n * TNB;
  g.d = B.d; g.e = B.e;
  g.taur = B.tau; g.taui = B.tau + n;
  g.scr = B.scal; g.sci = B.scal + n;
  g.part = B.part;
  g.dotr = B.tdots; g.doti = B.tdots + (size_t)nts * 2 * TNB;
  g.vy = B.tvy; g.norm = B.tnorm; g.counter = B.ticket;
  // optional per-launch events around the SYMV (kernel-level roofline)
  const bool kt = c.ktime != 0;
  if (kt && c.kev.size() < (size_t)2 * n) {
    const size_t old = c.kev.size();
    c.kev.resize((size_t)2 * n);
    for (size_t q = old; q < c.kev.size(); q++) W2K_HIP(hipEventCreate(&c.kev[q]));
  }
  int nsymv = 0;
  for (int p0 = 0; p0 < n; p0 += TNB) {
    const int nb = std::min(TNB, n - p0);
    g.p0 = p0;
    g.nb = nb;
    hipLaunchKernelGGL(k_ztri_col, dim3((n - p0 + ZCOL_ROWS - 1) / ZCOL_ROWS), dim3(64 * ZCOL_WAVES), 0, s, g, -1, p0, 0,
                       0, 0);
    W2K_STAGE(s, "k_ztri_col");
    int nref = 0;
    for (int k = 0; k < nb; k++) {
      const int i = p0 + k;
      if (i >= n - 1) break;
      nref = k + 1;
      const int rows = n - (i + 1);
      const int nt = (rows + SYMV_T - 1) / SYMV_T, nblk = nt * (nt + 1) / 2;
      // timed launches carry their events in the dispatch itself
      if (kt) hipExtLaunchKernelGGL(k_ztri_symv, dim3(nblk + 4 * nt), dim3(SYMV_T), 0, s, c.kev[2 * nsymv],
                                    c.kev[2 * nsymv + 1], 0, g, i, nt);
      else hipLaunchKernelGGL(k_ztri_symv, dim3(nblk + 4 * nt), dim3(SYMV_T), 0, s, g, i, nt);
      W2K_STAGE(s, "k_ztri_symv");
      nsymv++;
      const int next = (k + 1 < nb) ? i + 1 : -1;
      hipLaunchKernelGGL(k_ztri_col, dim3((rows + ZCOL_ROWS - 1) / ZCOL_ROWS), dim3(64 * ZCOL_WAVES), 0, s, g, i,
                         next, nt, nt, nblk);
      W2K_STAGE(s, "k_ztri_col");
    }
    if (nref > 0) {
      hipLaunchKernelGGL(k_ztri_store, dim3((n + 255) / 256, nref), dim3(256), 0, s, g, nref);
      W2K_STAGE(s, "k_ztri_store");
    }
    const int p1 = p0 + nb;
    if (p1 < n) {
      // A(p1:n, p1:n) -= [V W] [W V]^H, lower triangle
      const int r = n - p1;
      const size_t o1 = (size_t)p1 + (size_t)n * p1, ow = (size_t)n * nb + p1;
      if ((rc = zgemm(s, 'N', 'C', r, r, 2 * nb, -1.0, g.Pr + p1, g.Pi + p1, n, g.Pr + ow, g.Pi + ow, n, 1.0, Hr + o1,
                      Hi + o1, n, 1)))
        return rc;
    }
  }
  W2K_HIP(hipEventRecord(c.ev[9], s));
  if (kt) {
    // [10] summed SYMV kernel time, [11] number of SYMV launches
    W2K_HIP(hipStreamSynchronize(s));
    double tot = 0.0;
    int cntl = 0;
    for (int q = 0; q < nsymv; q++) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, c.kev[2 * q], c.kev[2 * q + 1]) != hipSuccess) {
        (void)hipGetLastError();
        break;
      }
      tot += ms;
      cntl++;
    }
    c.last_ms[10] = tot;
    c.last_ms[11] = cntl;
  }
  // ---- 4./5. real tridiagonal eigenpairs (shared with the real build)
  double *Zt = Xr;  // n x m real, then the real part of Q Zt
  double *Zi = Xi;  // imaginary part of Q Zt
  int m = 0, nbelw = 0;
  if ((rc = tridiag_eigen(c, B, el, eu, nume, Zt, &m, &nbelw))) return rc;
  *nbelw_out = nbelw;
  c.nbelw = nbelw;
  if ((rc = c.d_z.alloc((size_t)2 * n * (m > 0 ? m : 1)))) return rc;
  W2K_HIP(hipEventRecord(c.ev[10], s));
  if (m > 0) {
    // ---- 6. Z = Q Zt with complex reflectors, sort, Z = V^-1 Z
    const size_t zc = (size_t)n * m;
    hipLaunchKernelGGL(k_zfill, dim3((unsigned)((zc + 255) / 256)), dim3(256), 0, s, zc, Zi, 0.0);
    W2K_STAGE(s, "k_zfill");
    // blocks of ZBT reflectors (the launches of a block, not its arithmetic,
    // set the pace at these orders: 64-reflector blocks halve them)
    constexpr int ZBT = 64;
    // The reflector panel V, its Gram matrix and T depend on the reflectors
    // only: they are formed on the second stream up to two blocks ahead, in
    // two buffers, while this stream applies the blocks to Z (Y1 = V^H Z,
    // Y2 = T Y1, Z -= V Y2).  bt_ready[u]: block in buffer u formed;
    // bt_used[u]: its products with Z done.
    // w_bt: V (re, im: n x ZBT each) x 2, T (re, im) x 2, the Gram matrix,
    // Y1 and Y2 (re, im: ZBT x m each), the Gram products' split-K workspace
    constexpr size_t GWS = (size_t)64 * ZBT * ZBT;
    if ((rc = c.w_bt.alloc((size_t)2 * (2 * n + 3 * ZBT + 2 * m) * ZBT + GWS))) return rc;
    double *vr[2], *vi[2], *tr[2], *ti[2];
    double *q = c.w_bt.p;
    for (int u = 0; u < 2; u++) {
      vr[u] = q; vi[u] = q + (size_t)n * ZBT; q += (size_t)2 * n * ZBT;
    }
    for (int u = 0; u < 2; u++) {
      tr[u] = q; ti[u] = q + ZBT * ZBT; q += 2 * ZBT * ZBT;
    }
    double *gmr = q, *gmi = q + ZBT * ZBT;
    q += 2 * ZBT * ZBT;
    double *y1r = q, *y1i = y1r + (size_t)ZBT * m, *y2r = y1i + (size_t)ZBT * m, *y2i = y2r + (size_t)ZBT * m;
    double *gws = y2i + (size_t)ZBT * m;
    const int nref = n - 1;
    const int nblk = nref > 0 ? (nref + ZBT - 1) / ZBT : 0;
    hipStream_t sa = c.aux;
    auto p0_of = [&](int k) { return (nblk - 1 - k) * ZBT; };
    auto prep = [&](int k) -> int {
      const int u = k & 1, p0 = p0_of(k), nb = std::min(ZBT, nref - p0);
      if (k >= 2) W2K_HIP(hipStreamWaitEvent(sa, c.bt_used[u], 0));
      hipLaunchKernelGGL(k_zvpanel, dim3((n + 255) / 256, nb), dim3(256), 0, sa, n, p0, nb, Hr, Hi, vr[u], vi[u]);
      W2K_STAGE(sa, "k_zvpanel");
      // rows 0 .. p0 of the block's reflectors are zero: the products start at
      // row p0+1 rounded down to 16 rows, which keeps the operands' 128-byte alignment
      const int o = (p0 + 1) & ~15, nr = n - o;
      int r2;
      if ((r2 = zgemm(sa, 'C', 'N', nb, nb, nr, 1.0, vr[u] + o, vi[u] + o, n, vr[u] + o, vi[u] + o, n, 0.0, gmr, gmi,
                      nb, 0, 0, 0, gws, GWS)))
        return r2;
      hipLaunchKernelGGL(k_zlarft, dim3(1), dim3(256), 0, sa, nb, gmr, gmi, g.taur + p0, g.taui + p0, tr[u],
                         ti[u]);
      W2K_STAGE(sa, "k_zlarft");
      W2K_HIP(hipEventRecord(c.bt_ready[u], sa));
      return 0;
    };
    if (nblk > 0) {
      // the second stream starts once the reflectors and Zt are in place
      W2K_HIP(hipEventRecord(c.bt_start, s));
      W2K_HIP(hipStreamWaitEvent(sa, c.bt_start, 0));
      for (int k = 0; k < std::min(nblk, 2); k++)
        if ((rc = prep(k))) return rc;
    }
    for (int k = 0; k < nblk; k++) {
      const int u = k & 1, p0 = p0_of(k), nb = std::min(ZBT, nref - p0);
      const int o = (p0 + 1) & ~15, nr = n - o;  // the block's reflectors are zero above row p0+1
      W2K_HIP(hipStreamWaitEvent(s, c.bt_ready[u], 0));
      if ((rc = zgemm(s, 'C', 'N', nb, m, nr, 1.0, vr[u] + o, vi[u] + o, n, Zt + o, Zi + o, n, 0.0, y1r, y1i, nb, 0,
                      0, 0, ws, wsn)))
        return rc;
      if ((rc = zgemm(s, 'N', 'N', nb, m, nb, 1.0, tr[u], ti[u], nb, y1r, y1i, nb, 0.0, y2r, y2i, nb, 0, 1, 0)))
        return rc;
      if ((rc = zgemm(s, 'N', 'N', nr, m, nb, -1.0, vr[u] + o, vi[u] + o, n, y2r, y2i, nb, 1.0, Zt + o, Zi + o, n)))
        return rc;
      if (k + 2 < nblk) {
        W2K_HIP(hipEventRecord(c.bt_used[u], s));
        if ((rc = prep(k + 2))) return rc;
      }
    }
    if ((rc = sort_eigen(c, B, m))) return rc;
    hipLaunchKernelGGL(k_zgather_cols, dim3((n + 255) / 256, m), dim3(256), 0, s, n, m, B.perm, Zt, Zi, Ur, Ui);
    W2K_STAGE(s, "k_zgather_cols");
    // Z = W Zs into the (free) H storage, then interleaved into d_z
    if ((rc = zgemm3(c, s, 'N', 'N', n, m, n, 1.0, Wr, Wi, n, Ur, Ui, n, 0.0, Hr, Hi, n, 0, 1, 0))) return rc;
    hipLaunchKernelGGL(k_interleave, dim3((unsigned)((zc + 255) / 256)), dim3(256), 0, s, zc, Hr, Hi, c.d_z.p);
    W2K_STAGE(s, "k_interleave");
    W2K_HIP(hipMemcpyAsync(c.d_eig.p, B.wv, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
  }
  W2K_HIP(hipEventRecord(c.ev[11], s));
  W2K_HIP(hipStreamSynchronize(s));
  take_ms(c, 3, 6, 7);
  take_ms(c, 4, 7, 8);
  take_ms(c, 5, 8, 9);
  take_ms(c, 6, 9, 10);
  take_ms(c, 7, 10, 11);
  c.ne = m;
  *ne_out = m;
  return 0;
}

}  // namespace w2k
