// SECLR4 on the GPU (seclr4.F:149-358): the generalized eigenproblem
// H z = e S z for the eigenvalues in (EL, EU], at most NUME of them.
//
// Input: HS/HSDIAG as CALKPT leaves them after HNS -- H in the lower triangle
// (diagonal included), S in the strict upper triangle, diag(S) in HSDIAG.
//
// Stages (all on the context stream):
//   1. Cholesky S = U^T U (POTRF 'U', seclr4.F:196) together with
//      W = U^-1: recursive halving on 64-column boundaries, 64x64 leaves
//      factored and inverted by one wave in registers, the TRSM as a GEMM
//      with the leading half's inverse, the SYRK as an MFMA GEMM on the upper
//      triangle, W12 = -(W11 U12) W22.
//   2. Reduction to standard form C = U^-T H U^-1 (DSCGST, dscgst.f:99-142):
//      X = H W and C = W^T X as MFMA GEMMs (lower triangle of C; the tridiagonalisation reads and
//      updates only the lower triangle).
//   3. Householder tridiagonalisation (DSYTRD2, dsytrd2.f, lower; DLATRD
//      panels of 32 columns, trailing rank-64 update on MFMA), DLARFG signs.
//   4. Bisection (DSTEBZ2, RANGE='V', ORDER='B', ABSTOL=0, dstebz2.f:301-652):
//      split points, per-block Gershgorin intervals, NWL (eigenvalues <= EL),
//      block-ordered list capped at NUME; one thread per eigenvalue.
//   5. Inverse iteration (DSTEIN): one thread per eigenvalue, close
//      eigenvalues perturbed as DSTEIN does, then vectors of one cluster
//      (gap <= 1e-3 |T|_1) orthogonalised in order (block classical
//      Gram-Schmidt, twice, on MFMA) and normalised with the largest
//      component positive.
//   6. Z = Q Zt with the reflectors in compact WY form (DORMTR), then the
//      selection sort of dsyevx2.f:381-400, then Z = U^-1 Z (DTRSM,
//      seclr4.F:354) as a triangular MFMA GEMM with W = U^-1.
// The arithmetic here is real FP64 throughout; the complex build (lapw1c)
// runs the same stages in seclr4_cplx.hip.
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include <functional>
#include <utility>
#include <vector>

#include <hip/hip_ext.h>

#include "seclr4_shared.h"

namespace w2k {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ double block_sum(double v, double *red)
{
  // sum over the block (blockDim.x <= 1024, multiple of 64)
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  if (threadIdx.x < 64) {
    t = (threadIdx.x < nw) ? red[threadIdx.x] : 0.0;
    for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    if (threadIdx.x == 0) red[32] = t;
  }
  __syncthreads();
  return red[32];
}

// 1. Sw (upper incl. diagonal) and Hf (full symmetric, leading dimension ldh)
// from HS/HSDIAG
__global__ void k_prep(int n, int ldh, const double *__restrict__ hs, const double *__restrict__ hsdiag,
                       double *__restrict__ sw, double *__restrict__ hf)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const size_t o = (size_t)i + (size_t)n * j, oh = (size_t)i + (size_t)ldh * j;
  const double v = hs[o];
  if (i < j) {
    sw[o] = v;
    hf[oh] = hs[(size_t)j + (size_t)n * i];
  } else if (i == j) {
    sw[o] = hsdiag[i];
    hf[oh] = v;
  } else {
    sw[o] = 0.0;
    hf[oh] = v;
  }
}

// 1/sqrt(x), x > 0: the hardware estimate refined by two Newton steps
// y <- y + y (1/2 - (x/2) y^2) (a few ulp)
__device__ __forceinline__ double rsqrt_nr(double x)
{
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  double e = fma(-h * y, y, 0.5);
  y = fma(y, e, y);
  e = fma(-h * y, y, 0.5);
  return fma(y, e, y);
}

// v of lane `src` (a wave-uniform lane index) in every lane
__device__ __forceinline__ double lane_bcast(double v, int src)
{
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), src);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)((unsigned long long)b >> 32), src);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// X = M^-1 of an N x N (N <= 64) upper triangular M held in LDS (M(r,q) at
// sm[q LD + r]) by one wave, with dinv[r] = 1/M(r,r): lane c < N ends with
// column c of X in x[0..N-1] (x[r] = 0 for r > c).  Rows from the bottom:
// X(r,c) = -dinv[r] sum_{r<q<=c} M(r,q) X(q,c), each row of M read in one
// batch of broadcast LDS loads; the newest term (q = r+1) is added last.
template <int N, int LD>
__device__ __forceinline__ void wave_trtri(const double *sm, const double *dinv, int lane, double *x)
{
#pragma unroll
  for (int r = 0; r < N; r++) x[r] = 0.0;
#pragma unroll
  for (int r = N - 1; r >= 0; r--) {
    double ur[N];
#pragma unroll
    for (int q = r + 1; q < N; q++) ur[q] = sm[q * LD + r];
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int q = N - 1; q > r + 1; q--) {
      if (q & 1) s1 = fma(ur[q], x[q], s1);
      else s0 = fma(ur[q], x[q], s0);
    }
    double sum = s0 + s1;
    if (r + 1 < N) sum = fma(ur[r + 1], x[r + 1], sum);
    const double d = dinv[r];
    x[r] = (lane == r) ? d : -sum * d;
  }
}

// Cholesky S_kk = U^T U of the kb x kb diagonal block at (k0,k0) (upper) and
// the inverse of its factor, by one wave: U over the block of a, U^-1 into
// the same block of w (zeros below the diagonal).  Lane c holds column c in
// registers; step j takes the pivot from lane j, scales row j by the inverse
// square root of the pivot and broadcasts the row through LDS for the rank-1
// update (rows below the diagonal of a column are carried along unused).  A
// block smaller than 64 is padded with the identity.  A non-positive pivot
// sets *info = k0 + j + 1 for the first such j (DPOTRF's INFO) and is taken
// as 1 so that the arithmetic stays finite.  Loads and stores go through LDS
// so that they are coalesced along the columns.
__global__ __launch_bounds__(64) void k_chol_diag(int n, int k0, int kb, double *__restrict__ a,
                                                  double *__restrict__ w, int *__restrict__ info)
{
  constexpr int LD = NB + 1;
  __shared__ double sm[NB * LD];  // sm[c LD + r] = block element (r, c)
  __shared__ double urow[2][NB];
  __shared__ double dinv[NB];
  const int lane = threadIdx.x;
  {
    // all 64 loads in flight together (clamped addresses, masked values)
    double ld[NB];
    const int rl = min(lane, kb - 1);
#pragma unroll
    for (int c = 0; c < NB; c++) ld[c] = a[(size_t)(k0 + rl) + (size_t)n * (k0 + min(c, kb - 1))];
#pragma unroll
    for (int c = 0; c < NB; c++) sm[c * LD + lane] = (lane <= c && c < kb) ? ld[c] : (lane == c ? 1.0 : 0.0);
  }
  __syncthreads();
  double v[NB];
#pragma unroll
  for (int r = 0; r < NB; r++) v[r] = sm[lane * LD + r];
  int first = 0;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    const double ajj = lane_bcast(v[j], j);
    const bool ok = ajj > 0.0;
    first = (!ok && first == 0) ? j + 1 : first;
    v[j] *= rsqrt_nr(ok ? ajj : 1.0);
    urow[j & 1][lane] = v[j];
    __syncthreads();
    double u[NB];
#pragma unroll
    for (int r = j + 1; r < NB; r++) u[r] = urow[j & 1][r];
#pragma unroll
    for (int r = j + 1; r < NB; r++) v[r] = fma(-u[r], v[j], v[r]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NB; r++) sm[lane * LD + r] = (r <= lane) ? v[r] : 0.0;
  __syncthreads();
  for (int c = 0; c < kb; c++)
    if (lane <= c) a[(size_t)(k0 + lane) + (size_t)n * (k0 + c)] = sm[c * LD + lane];
  dinv[lane] = 1.0 / sm[lane * LD + lane];
  __syncthreads();
  double x[NB];
  wave_trtri<NB, LD>(sm, dinv, lane, x);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NB; r++) sm[lane * LD + r] = (r <= lane) ? x[r] : 0.0;
  __syncthreads();
  if (lane < kb)
    for (int c = 0; c < kb; c++) w[(size_t)(k0 + lane) + (size_t)n * (k0 + c)] = sm[c * LD + lane];
  if (lane == 0 && first != 0 && *info == 0) *info = k0 + first;
}

__global__ void k_fill(size_t cnt, double *p, double v)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = v;
}

// ------------------------------------------------------------------ tridiagonalisation
// Panel of TNB columns (DLATRD, lower): the panel buffer P holds, for panel
// column k, the reflector V(:,k) at column k, W(:,k) at column nb+k and a
// second copy of V(:,k) at column 2nb+k, so that the trailing update
// A -= V W^T + W V^T is one GEMM with operands P(:,0:2nb) and P(:,nb:3nb).
// Per column i of the panel two multi-workgroup kernels run:
//   k_tri_symv : DLARFG of column i (beta, tau, scal) from the norm partials,
//                y = A(i+1:n, i+1:n) v as partial sums per column chunk, the
//                dot products V^T v, W^T v and v^T y as partials, and V(:,k)
//                written out (v = A(:,i) * scal, v(i+1) = 1);
//   k_tri_col  : w = tau (y - W V^T v - V W^T v) - (tau/2)(w^T v) v written
//                to W(:,k); the next column updated with the panel so far and
//                the squared norm of its sub-subdiagonal part written as one
//                partial per block.
struct TriWork {
  int n, p0, nb;
  int lda;            // leading dimension of a (a multiple of 16: 128-byte aligned columns)
  double *a;          // n x n, both triangles
  double *P;          // n x 3 TNB panel buffer
  double *d, *e, *tau, *scal;
  double *part;       // nchunk x n SYMV partial sums
  double *dots;       // tiles x 2 TNB partial V^T v, W^T v
  double *vy;         // tiles x chunks partial v^T y
  double *norm;       // blocks of the column update: partial squared norms
  int *counter;       // unused by the real kernels
};


constexpr int COL_ROWS = 64;  // rows per block of the column update (one per lane)
constexpr int COL_WAVES = 8;  // waves per block of the column update

// DLARFG of column i on (A(i+1,i), A(i+2:n,i)), from the squared norm
// partials of the column-update kernel; every wave of the SYMV evaluates it
// with identical arithmetic (fixed-order sums), so no grid-wide step is
// needed between the two kernels.  tri_norm_load issues the loads (first in
// the wave, so that the reflector is formed while the tile is still on its
// way); tri_reflector returns scal = 1/(alpha - beta).
constexpr int NORM_B = 3;  // partials per lane loaded up front (n <= 64 NORM_B COL_ROWS)
__device__ __forceinline__ void tri_norm_load(const TriWork &g, int i, int lane, double *x, double &alpha)
{
  const int nnorm = (g.n - i + COL_ROWS - 1) / COL_ROWS;  // blocks of the update of column i
#pragma unroll
  for (int u = 0; u < NORM_B; u++) x[u] = g.norm[min(lane + 64 * u, nnorm - 1)];
  alpha = g.a[(size_t)g.lda * i + i + 1];
}

__device__ __forceinline__ double tri_reflector(const TriWork &g, int i, int lane, bool store, const double *x,
                                                double alpha)
{
  const int nnorm = (g.n - i + COL_ROWS - 1) / COL_ROWS;
  double xs = 0.0;
#pragma unroll
  for (int u = 0; u < NORM_B; u++) xs += (lane + 64 * u < nnorm) ? x[u] : 0.0;
  for (int b = lane + 64 * NORM_B; b < nnorm; b += 64) xs += g.norm[b];
  xs = wave_sum(xs);
  xs = __shfl(xs, 0, 64);
  const double xnorm = sqrt(xs);
  double tau = 0.0, beta = alpha, scal = 0.0;
  if (xnorm != 0.0) {
    beta = -copysign(sqrt(alpha * alpha + xnorm * xnorm), alpha);
    tau = (beta - alpha) / beta;
    scal = 1.0 / (alpha - beta);
  }
  if (store && lane == 0) {
    g.e[i] = beta;
    g.tau[i] = tau;
    g.scal[i] = scal;
  }
  return scal;
}

// y = A(r0:n, r0:n) v from the lower triangle only (A is symmetric): one
// wave per 64 x 64 tile (I, J), J <= I, adds A_IJ v_J to the rows of tile I
// and, for J < I, A_IJ^T v_I to the rows of tile J (column sums by a
// wave-wide reduce-scatter, no LDS); a diagonal tile adds its lower part by
// rows and its strict lower part by columns.  Row r of tile R receives
// exactly one partial per tile index: from (R, J) at slot J <= R and from
// (I, R) at slot I > R, so y(r) = sum over slots in a fixed order.  A tile
// wave writes its partials into a block of its own, 128 doubles at bid * 128
// (row sums, then column sums), so that consecutive waves write consecutive
// memory (tiles x n vectors put every 512-byte segment in another place,
// which costs the kernel's read stream more HBM time).  Every tile wave
// writes its partial v^T y.  The nt waves after the
// tiles each take one row block I: they write V(:,k) (v = A(:,i) * scal,
// v(i+1) = 1) and the partial dot products V^T v, W^T v of the block's rows.
__global__ __launch_bounds__(64, 3) void k_tri_symv(TriWork g, int i, int nt)
{
  const int n = g.n, lda = g.lda, r0 = i + 1, ii = i - g.p0, nb = g.nb;
  const int base = r0 & ~15;  // tile grid origin: rows on 128-byte boundaries
  const int lane = threadIdx.x;
  const int nblk = nt * (nt + 1) / 2;
  const int bid = blockIdx.x;
  const double *ai = g.a + (size_t)lda * i;
  double xn[NORM_B], alpha;
  tri_norm_load(g, i, lane, xn, alpha);
  if (bid >= nblk) {
    // row block I, two waves: V column and V^T v (half 0), W^T v (half 1)
    const int I = (bid - nblk) >> 1, half = (bid - nblk) & 1;
    const int R = base + SYMV_T * I + lane;
    const bool live = R >= r0 && R < n;
    const int Rs = live ? R : r0;
    const double a_r = ai[Rs];
    double t[TNB];  // 32 values per wave keep the branch's registers at the tile loop's level
    if (ii > 0) {
#pragma unroll
      for (int k = 0; k < TNB; k++) t[k] = g.P[(size_t)n * (half * nb + k) + Rs];
    }
    const double sc = tri_reflector(g, i, lane, false, xn, alpha);
    const double vr = live ? (R == r0 ? 1.0 : a_r * sc) : 0.0;
    if (live && half == 0) {
      g.P[(size_t)n * ii + R] = vr;
      g.P[(size_t)n * (2 * nb + ii) + R] = vr;
    }
    if (ii > 0) {
      // t[k] = X(R,k) v(R) for k < ii, summed over the lanes: lanes 2q and
      // 2q + 1 end with the sum of column q
#pragma unroll
      for (int k = 0; k < TNB; k++) {
        const bool use = live && k < ii;
        t[k] = use ? t[k] * vr : 0.0;
      }
      const double dq = wave_transpose_sum32x(t, lane);  // lanes 2k, 2k + 1: column k
      if (!(lane & 1)) g.dots[(size_t)I * 2 * TNB + half * TNB + (lane >> 1)] = dq;
    }
    return;
  }
  // triangular block index -> (I, J), J <= I
  int I = (int)((sqrt(8.0 * (double)bid + 1.0) - 1.0) * 0.5);
  while ((I + 1) * (I + 2) / 2 <= bid) I++;
  while (I * (I + 1) / 2 > bid) I--;
  const int J = bid - I * (I + 1) / 2;
  const int R = base + SYMV_T * I + lane;
  const int C0 = base + SYMV_T * J;
  const bool live = R >= r0 && R < n;
  const bool off = (I != J);
  const int m = min(SYMV_T, n - C0);
  const int Rl = min(R, n - 1);  // the lane's row for the loads (in bounds)
  // every load of the tile and of v is issued before the first use: clamped
  // (in-bounds) addresses, values masked afterwards
  const double a_c = ai[min(C0 + lane, n - 1)];
  const double a_r = ai[Rl];
  const double *ap = g.a + (size_t)Rl + (size_t)lda * C0;
  double t[SYMV_T];
#pragma unroll
  for (int j = 0; j < SYMV_T; j++) t[j] = ap[(size_t)lda * min(j, m - 1)];
  const double sc = tri_reflector(g, i, lane, bid == 0, xn, alpha);
  const int c = C0 + lane;
  const double vc = (c >= r0 && c < n) ? (c == r0 ? 1.0 : a_c * sc) : 0.0;  // v of the lane's column
  const double vr = live ? (R == r0 ? 1.0 : a_r * sc) : 0.0;
  if (m < SYMV_T || !live) {
#pragma unroll
    for (int j = 0; j < SYMV_T; j++) t[j] = (live && j < m) ? t[j] : 0.0;
  }
  // only the lower triangle of A is kept current (the trailing updates skip
  // the upper one): a diagonal tile uses its lower part, by rows and (strictly
  // below the diagonal) by columns
  if (!off) {
#pragma unroll
    for (int j = 0; j < SYMV_T; j++) t[j] = (j <= lane) ? t[j] : 0.0;
  }
  // row sums: v(C0 + j) broadcast from lane j through scalar registers
  const long long vcb = __double_as_longlong(vc);
  const int vlo = (int)(unsigned)(vcb & 0xffffffffll), vhi = (int)(unsigned)((unsigned long long)vcb >> 32);
  double y0 = 0.0, y1 = 0.0, y2 = 0.0, y3 = 0.0;
#pragma unroll
  for (int j = 0; j < SYMV_T; j += 4) {
    double vj[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const unsigned lo = (unsigned)__builtin_amdgcn_readlane(vlo, j + u);
      const unsigned hi = (unsigned)__builtin_amdgcn_readlane(vhi, j + u);
      vj[u] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
    y0 += t[j] * vj[0];
    y1 += t[j + 1] * vj[1];
    y2 += t[j + 2] * vj[2];
    y3 += t[j + 3] * vj[3];
  }
  double yrow = (y0 + y1) + (y2 + y3);
  if (!off) {
#pragma unroll
    for (int j = 0; j < SYMV_T; j++) t[j] = (j < lane) ? t[j] * vr : 0.0;
    yrow += wave_transpose_sum(t, lane);  // column C0 + lane = row R of the strict lower part
  }
  if (live) g.part[(size_t)bid * 128 + lane] = yrow;
  // v^T y over the tile: an off-diagonal tile adds v_I^T A_IJ v_J through its
  // rows and the same number again through its columns
  double vy = wave_sum(yrow * vr);
  if (lane == 0) g.vy[bid] = off ? 2.0 * vy : vy;
  if (off) {
    double *pc = g.part + (size_t)bid * 128 + 64 + lane;
    const bool st = lane < m && c >= r0;
#pragma unroll
    for (int j = 0; j < SYMV_T; j++) t[j] *= vr;
    const double cs = wave_transpose_sum(t, lane);  // column C0 + lane of A_IJ^T v_I
    if (st) *pc = cs;
  }
}

// Finish W(:, fin - p0) (fin >= 0) and update column col (col >= 0) with the
// panel so far (DLATRD); write the squared-norm partial of its
// sub-subdiagonal part per block (DLARFG itself is evaluated by the next
// SYMV's waves).  A block takes COL_ROWS = 64 consecutive rows, one per lane,
// in each of its COL_WAVES waves, so that every load of a SYMV partial or a
// panel column is one contiguous 512-byte row segment; the waves split the
// SYMV partials (chunk w, w + 8, ...) and the panel columns (four each) and
// their row sums meet in LDS.  Loads are issued in batches with clamped
// indices and masked values, so that a batch is in flight at once.
// offset of partial c of the rows of SYMV row block I (the lane's place in
// the block at + l): the row sums of tile (I, c) for c <= I, else the column
// sums of tile (c, I)
__device__ __forceinline__ size_t tpart_off(int I, int c)
{
  return c <= I ? (size_t)(I * (I + 1) / 2 + c) * 128 : (size_t)(c * (c + 1) / 2 + I) * 128 + 64;
}

__global__ __launch_bounds__(64 * COL_WAVES) void k_tri_col(TriWork g, int fin, int col, int ntile, int nvy)
{
  constexpr int NT = 64 * COL_WAVES;   // threads
  constexpr int KW = TNB / COL_WAVES;  // panel columns per wave
  constexpr int YB = 16;               // SYMV-partial chunks per wave and batch
  constexpr int VB = 10;               // first batch of v^T y pairs per thread (nvy <= 2 NT VB)
  constexpr int DG = NT / 32;          // groups of the V^T v, W^T v sums (32 column pairs each)
  constexpr int DB = 8;                // first batch of tiles per group (ntile <= DG DB)
  const int n = g.n, p0 = g.p0, nb = g.nb, t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nch = ntile;
  const int rbase = fin >= 0 ? fin + 1 : col;
  const int r = rbase + blockIdx.x * COL_ROWS + lane;
  const bool live = r < n;
  const int rs = live ? r : rbase;  // rbase < n: a safe row for dead lanes
  const int ii = fin >= 0 ? fin - p0 : -1;
  const int jj = col - p0;  // columns 0..jj-1 of the panel update column col
  __shared__ double ab[2 * TNB];  // a_k = V_k^T v, b_k = W_k^T v
  __shared__ double wc[TNB], vc[TNB];
  __shared__ double dred[DG][2 * TNB];
  __shared__ double red[3][COL_WAVES][COL_ROWS];  // per wave: y, z (W a + V b), a0
  __shared__ double sh[2][COL_WAVES];
  const double *V = g.P, *W = g.P + (size_t)n * nb;
  double *ac = g.a + (size_t)g.lda * (col >= 0 ? col : 0);
  // ---- per-block scalars: row col of the panel, v^T y, V^T v and W^T v
  if (col >= 0 && t < jj) {
    vc[t] = V[(size_t)n * t + col];
    if (t < ii || fin < 0) wc[t] = W[(size_t)n * t + col];
  }
  if (fin >= 0) {
    {
      const int vlast = (nvy - 1) & ~1;  // the last pair (vy has one spare entry)
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
    if (col >= 0) {
      double p = 0.0;
      const int cl = col - (rbase & ~15);  // row col in row block 0 of the SYMV's grid
      for (int c = t; c < nch; c += NT) p += g.part[tpart_off(0, c) + cl];
      p = wave_sum(p);
      if (lane == 0) sh[1][w] = p;
    }
    {
      const int dq = t & 31, dp = t >> 5;
      double s0 = 0.0, s1 = 0.0;
      {
        double2 xd[DB];
#pragma unroll
        for (int u = 0; u < DB; u++)
          xd[u] = *reinterpret_cast<const double2 *>(g.dots + (size_t)min(dp + DG * u, ntile - 1) * 2 * TNB + 2 * dq);
#pragma unroll
        for (int u = 0; u < DB; u++) {
          const bool k = dp + DG * u < ntile;
          s0 += k ? xd[u].x : 0.0;
          s1 += k ? xd[u].y : 0.0;
        }
      }
      for (int b = dp + DG * DB; b < ntile; b += DG) {
        const double2 x = *reinterpret_cast<const double2 *>(g.dots + (size_t)b * 2 * TNB + 2 * dq);
        s0 += x.x;
        s1 += x.y;
      }
      dred[dp][2 * dq] = s0;
      dred[dp][2 * dq + 1] = s1;
    }
  }
  // ---- this wave's share of the rows' sums
  double kv[KW], kw[KW];  // V(r, k), W(r, k), k = KW w + u
#pragma unroll
  for (int u = 0; u < KW; u++) {
    kv[u] = V[(size_t)n * (KW * w + u) + rs];
    kw[u] = W[(size_t)n * (KW * w + u) + rs];
  }
  double y = 0.0;
  if (fin >= 0) {
    // the row's block and place in the SYMV's tile grid (origin rbase & ~15)
    const int rg = rs - (rbase & ~15), I = rg >> 6, l = rg & 63;
    for (int c0 = w; c0 < nch; c0 += COL_WAVES * YB) {
      double xp[YB];
#pragma unroll
      for (int u = 0; u < YB; u++) xp[u] = g.part[tpart_off(I, min(c0 + COL_WAVES * u, nch - 1)) + l];
#pragma unroll
      for (int u = 0; u < YB; u++) y += (live && c0 + COL_WAVES * u < nch) ? xp[u] : 0.0;  // rows past n: never written
    }
  }
  __syncthreads();
  if (fin >= 0 && t < 2 * TNB) {
    double q = 0.0;
#pragma unroll
    for (int p = 0; p < DG; p++) q += dred[p][t];
    ab[t] = ((t & (TNB - 1)) < ii) ? q : 0.0;
  }
  __syncthreads();
  {
    double z = 0.0, a0 = 0.0;
#pragma unroll
    for (int u = 0; u < KW; u++) {
      const int k = KW * w + u;
      if (fin >= 0 && k < ii) z += kw[u] * ab[k] + kv[u] * ab[TNB + k];
      // V(r,k) W(col,k) + W(r,k) V(col,k), k < jj, k != ii; panel columns
      // k >= jj hold values of an earlier panel: selected away
      const bool use = col >= 0 && k < jj && k != ii;
      a0 += (use ? kv[u] : 0.0) * (use ? wc[min(k, TNB - 1)] : 0.0) +
            (use ? kw[u] : 0.0) * (use ? vc[min(k, TNB - 1)] : 0.0);
    }
    red[0][w][lane] = y;
    red[1][w][lane] = z;
    red[2][w][lane] = a0;
  }
  __syncthreads();
  // W(q, ii) = 0 above this kernel's rows
  if (fin >= 0)
    for (int q = blockIdx.x * NT + t; q < rbase; q += gridDim.x * NT) g.P[(size_t)n * (nb + ii) + q] = 0.0;
  if (w != 0) return;
  // ---- wave 0 finishes its 64 rows
  double sy[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    double sum = 0.0;
#pragma unroll
    for (int p = 0; p < COL_WAVES; p++) sum += red[a][p][lane];
    sy[a] = sum;
  }
  double wr = 0.0, vii = 0.0, wcii = 0.0;  // W(r, ii), V(r, ii), W(col, ii)
  if (fin >= 0) {
    const double tauf = g.tau[fin];
    vii = V[(size_t)n * ii + rs];
    double vty = 0.0, scol = 0.0;
#pragma unroll
    for (int p = 0; p < COL_WAVES; p++) {
      vty += sh[0][p];
      scol += sh[1][p];
    }
    double ssum = 0.0;
#pragma unroll
    for (int k = 0; k < TNB; k++) ssum += ab[k] * ab[TNB + k];
    const double alpha2 = -0.5 * tauf * tauf * (vty - 2.0 * ssum);
    wr = tauf * (sy[0] - sy[1]) + alpha2 * vii;
    if (live) g.P[(size_t)n * (nb + ii) + r] = wr;
    if (col >= 0) {
      // row col = fin + 1 of the same formula (v(col) = 1), in every lane
      double corr = 0.0;
      for (int k = 0; k < ii; k++) corr += wc[k] * ab[k] + vc[k] * ab[TNB + k];
      wcii = tauf * (scol - corr) + alpha2;
    }
  }
  if (col < 0) return;
  double a = ac[rs] - sy[2];
  if (fin >= 0) a -= vii * wcii + wr * vc[ii];
  double nrm = 0.0;
  if (live) {
    ac[r] = a;
    if (r == col) {
      g.d[col] = a;
      if (col == n - 1) {
        g.tau[col] = 0.0;
        g.scal[col] = 0.0;
      }
    }
    if (r >= col + 2) nrm = a * a;
  }
  nrm = wave_sum(nrm);
  if (lane == 0) g.norm[blockIdx.x] = nrm;
}

// reflectors of panel columns into A below the subdiagonal (read by the back-transform)
__global__ void k_tri_store(TriWork g, int ncols)
{
  const int n = g.n;
  const int k = blockIdx.y;
  const int i = g.p0 + k;
  const int r = i + 2 + blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncols || r >= n) return;
  g.a[(size_t)r + (size_t)g.lda * i] = g.P[(size_t)n * k + r];
}

// ------------------------------------------------------------------ bisection
struct BisArgs {
  int n;
  const double *d, *e;
  double *e2;          // squared off-diagonals, 0 at split points
  double el, eu;
  int nume;
  // outputs
  int *cnt;            // [0] = M, [1] = NWL, [2] = number of blocks
  int *tblk;           // task: block index
  int *tidx;           // task: index within block (or -1 for a 1x1 block)
  int *bstart, *bend;  // block ranges
  double *bgl, *bgu, *batol, *bnrm;  // per block interval, tolerance, one-norm
  double *pivmin;
};

constexpr int SB = 16;  // rows per batch of loads ahead of a serial recurrence

// One row of the Sturm recurrence t <- (d - e^2 / t) - x with PIVMIN
// (dlaebz); the quotient through the hardware reciprocal refined by two
// Newton steps (|t| >= PIVMIN >= DBL_MIN and e^2 / PIVMIN <= 1/DBL_MIN, so
// neither the reciprocal nor the product leaves the normal range).
__device__ __forceinline__ void sturm_row(double &t, int &c, double dj, double e2, double x, double pivmin)
{
  double r = __builtin_amdgcn_rcp(t);
  r = fma(r, fma(-t, r, 1.0), r);
  r = fma(r, fma(-t, r, 1.0), r);
  t = (dj - e2 * r) - x;
  if (fabs(t) < pivmin) t = -pivmin;
  c += (t <= 0.0);
}

// Number of eigenvalues <= x (the lane's own shift) of an unreduced block
// (d, e) of order n, by one wave (dlaebz's Sturm count), the squared
// off-diagonals formed as e*e (inside a block they equal the E2 of the
// splitting step).  The recurrence is serial; d and e are read 64 rows at a
// time, one row per lane and a chunk ahead, and handed to the recurrence by
// readlane, four rows per step.
__device__ __forceinline__ int sturm_count_wave(int n, const double *d, const double *e, double x, double pivmin,
                                                int lane)
{
  double t = d[0] - x;
  if (fabs(t) < pivmin) t = -pivmin;
  int c = (t <= 0.0);
  const int elast = max(n - 2, 0);
  double dc = d[min(1 + lane, n - 1)];  // chunk row j0 + lane: d(j), e(j-1)
  double ec = e[min(lane, elast)];
  for (int j0 = 1; j0 < n; j0 += 64) {
    const double dx = d[min(j0 + 64 + lane, n - 1)];
    const double ex = e[min(j0 + 63 + lane, elast)];
    const int rows = min(64, n - j0);
    int u = 0;
    for (; u + 4 <= rows; u += 4) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const double ej = lane_bcast(ec, u + v);
        sturm_row(t, c, lane_bcast(dc, u + v), ej * ej, x, pivmin);
      }
    }
    for (; u < rows; u++) {
      const double ej = lane_bcast(ec, u);
      sturm_row(t, c, lane_bcast(dc, u), ej * ej, x, pivmin);
    }
    dc = dx;
    ec = ex;
  }
  return c;
}

__device__ __forceinline__ double wave_max(double v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v)
{
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// one wave: split points, blocks, counts, task list (dstebz2.f:301-540).
// Pass 1 forms E2 and PIVMIN over all rows; pass 2 finds the split points
// again in order and handles each block as it closes (one-norm, Gershgorin
// interval, Sturm counts at its ends, its tasks).  Rows are spread over the
// lanes; the block list and the running counts are wave-uniform and lane 0
// writes them.
__global__ __launch_bounds__(64) void k_bisect_prep(BisArgs g)
{
  if (blockIdx.x != 0) return;
  const int n = g.n, lane = threadIdx.x;
  const double safemn = DBL_MIN, ulp = DBL_EPSILON, fudge = 2.1;
  auto split_at = [&](int j) {  // E(j-1) negligible: a block ends before row j
    const double ej = g.e[j - 1];
    const double t = ej * ej;
    return fabs(g.d[j] * g.d[j - 1]) * ulp * ulp + safemn > t;
  };
  double pm = 1.0;
  for (int j0 = 1; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    if (j < n) {
      const double ej = g.e[j - 1];
      const double t = ej * ej;
      const bool sp = split_at(j);
      g.e2[j - 1] = sp ? 0.0 : t;
      if (!sp) pm = fmax(pm, t);
    }
  }
  const double pivmin = wave_max(pm) * safemn;
  if (lane == 0) *g.pivmin = pivmin;
  int nb = 0, m = 0, nwl = 0;
  auto block = [&](int ib, int ie) {
    const int b = nb++;
    const int in = ie - ib;
    if (lane == 0) {
      g.bstart[b] = ib;
      g.bend[b] = ie;
    }
    double on = 0.0, gu = -DBL_MAX, gl = DBL_MAX;
    for (int i0 = ib; i0 < ie; i0 += 64) {
      const int i = i0 + lane;
      if (i < ie) {
        const double di = g.d[i];
        const double em = (i > ib) ? fabs(g.e[i - 1]) : 0.0;
        const double ep = (i < ie - 1) ? fabs(g.e[i]) : 0.0;
        on = fmax(on, fabs(di) + em + ep);
        gu = fmax(gu, di + em + ep);
        gl = fmin(gl, di - em - ep);
      }
    }
    on = wave_max(on);
    if (lane == 0) g.bnrm[b] = on;
    const double dib = g.d[ib];
    if (in == 1) {
      if (g.el >= dib - pivmin) nwl++;
      if (g.el < dib - pivmin && g.eu >= dib - pivmin) {
        if (lane == 0) {
          g.tblk[m] = b;
          g.tidx[m] = -1;
        }
        m++;
      }
      return;
    }
    gu = wave_max(gu);
    gl = wave_min(gl);
    const double bnorm = fmax(fabs(gl), fabs(gu));
    gl = gl - fudge * bnorm * ulp * in - fudge * pivmin;
    gu = gu + fudge * bnorm * ulp * in + fudge * pivmin;
    if (lane == 0) g.batol[b] = ulp * fmax(fabs(gl), fabs(gu));
    if (gu < g.el) {
      nwl += in;
      return;
    }
    gl = fmax(gl, g.el);
    gu = fmin(gu, g.eu);
    if (lane == 0) {
      g.bgl[b] = gl;
      g.bgu[b] = gu;
    }
    if (gl >= gu) return;
    const int cl = sturm_count_wave(in, g.d + ib, g.e + ib, gl, pivmin, lane);
    const int cu = sturm_count_wave(in, g.d + ib, g.e + ib, gu, pivmin, lane);
    nwl += cl;
    for (int k = cl + lane; k < cu; k += 64) {
      const int q = m + (k - cl);
      if (q < n) {
        g.tblk[q] = b;
        g.tidx[q] = k;
      }
    }
    m += cu - cl;
  };
  int start = 0;
  for (int j0 = 1; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    unsigned long long mask = __ballot(j < n && split_at(j));
    while (mask) {
      const int bit = __ffsll((long long)mask) - 1;
      mask &= mask - 1;
      block(start, j0 + bit);
      start = j0 + bit;
    }
  }
  block(start, n);
  if (lane == 0) {
    g.cnt[0] = m < g.nume ? m : g.nume;
    g.cnt[1] = nwl;
    g.cnt[2] = nb;
  }
}

// multisection: one wave per eigenvalue; each round the 64 lanes take the
// Sturm count at 64 interior points of (lo, hi], and the bracket shrinks to the
// sub-interval where the count first exceeds the eigenvalue's index (65x per
// round instead of 2x); stopping test of dlaebz (ATOL, PIVMIN, RELTOL)
__global__ __launch_bounds__(256) void k_bisect(BisArgs g, int m, double *__restrict__ w)
{
  const int lane = threadIdx.x & 63;
  const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
  if (t >= m) return;  // wave-uniform
  const int b = g.tblk[t], k = g.tidx[t];
  const int ib = g.bstart[b], in = g.bend[b] - ib;
  if (k < 0) {
    if (lane == 0) w[t] = g.d[ib];
    return;
  }
  const double pivmin = *g.pivmin, atoli = g.batol[b], rtoli = 2.0 * DBL_EPSILON;
  double lo = g.bgl[b], hi = g.bgu[b];
  for (int it = 0; it < 64; it++) {
    const double w1 = fabs(hi - lo), w2 = fmax(fabs(hi), fabs(lo));
    if (w1 < fmax(atoli, fmax(pivmin, rtoli * w2))) break;
    const double h = (hi - lo) / 65.0;
    double x = lo + h * (double)(lane + 1);
    if (lane == 63 || !(x < hi)) x = hi;  // the last point is hi itself
    const int cnt = sturm_count_wave(in, g.d + ib, g.e + ib, x, pivmin, lane);
    const unsigned long long above = __ballot(cnt > k);
    if (above == 0ull) {
      lo = hi;  // cannot happen for a valid bracket; stop narrowing
      break;
    }
    const int f = __ffsll((long long)above) - 1;  // first point with count > k
    const double xf = __shfl(x, f, 64);
    const double xl = f > 0 ? __shfl(x, f > 0 ? f - 1 : 0, 64) : lo;
    hi = xf;
    lo = xl;
  }
  if (lane == 0) w[t] = 0.5 * (lo + hi);
}

// ------------------------------------------------------------------ inverse iteration
// single thread: DSTEIN shifts (close eigenvalues separated by 10 ulp) and clusters
__global__ void k_stein_prep(BisArgs g, int m, const double *__restrict__ w, double *__restrict__ xs,
                             int *__restrict__ cstart)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int prevb = -1, gp = 0;
  double xjm = 0.0;
  for (int t0 = 0; t0 < m; t0 += SB) {
    int bb[SB];
    double wb[SB], nb[SB];
#pragma unroll
    for (int u = 0; u < SB; u++) {
      bb[u] = g.tblk[min(t0 + u, m - 1)];
      wb[u] = w[min(t0 + u, m - 1)];
    }
#pragma unroll
    for (int u = 0; u < SB; u++) nb[u] = g.bnrm[bb[u]];
#pragma unroll
    for (int u = 0; u < SB; u++) {
      const int t = t0 + u;
      if (t >= m) continue;
      const int b = bb[u];
      double xj = wb[u];
      if (b != prevb) {
        gp = t;
      } else {
        const double pertol = 10.0 * fabs(DBL_EPSILON * xj);
        if (xj - xjm < pertol) xj = xjm + pertol;
        if (xj - xjm > 1.0e-3 * nb[u]) gp = t;
      }
      xs[t] = xj;
      cstart[t] = gp;
      xjm = xj;
      prevb = b;
    }
  }
}

// tridiagonal LU with partial pivoting (DLAGTF) and solves (DLAGTS), one
// thread per eigenvalue.  Factors and the work vector are stored as
// [row pair][eigenvalue][2]: one 16-byte load brings two rows of a thread's
// eigenvalue and a wave's loads coalesce, so a group of SG rows needs half
// the load instructions (a wave holds at most 63 loads in flight, and with
// one wave per 64 eigenvalues that count sets the pace).  Every running
// value is carried in registers so that no pass waits on its own stores,
// and U's diagonal is kept as its reciprocal so the back-substitutions
// multiply.  The planes hold rows 0 .. n (n + 1 rows rounded up to pairs).
constexpr int SG = 32;  // rows per group of the solves: the group's loads are issued before its arithmetic

__global__ __launch_bounds__(64) void k_stein(BisArgs g, int m, const double *__restrict__ xs,
                                              double *__restrict__ z, int ldz, double *__restrict__ u0,
                                              double *__restrict__ u1, double *__restrict__ u2,
                                              double *__restrict__ mul, double *__restrict__ x,
                                              int *__restrict__ piv, double dtscale)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= m) return;
  const int b = g.tblk[t];
  const int ib = g.bstart[b], ie = g.bend[b], bs = ie - ib;
  double *zc = z + (size_t)ldz * t;
  if (bs == 1) {
    zc[ib] = 1.0;
    return;
  }
  const double lam = xs[t];
  const double *d = g.d + ib, *e = g.e + ib;
  const size_t M = (size_t)m;
  // element (row k, this eigenvalue) and pair index (even row k) of a plane
  auto at = [&](int k) -> size_t { return ((size_t)(k >> 1) * M + t) * 2 + (k & 1); };
  auto pr = [&](int k) -> size_t { return (size_t)(k >> 1) * M + t; };
  const double2 *X2 = reinterpret_cast<const double2 *>(x);
  const double2 *R2 = reinterpret_cast<const double2 *>(u0);
  const double2 *A12 = reinterpret_cast<const double2 *>(u1);
  const double2 *A22 = reinterpret_cast<const double2 *>(u2);
  const double2 *M2 = reinterpret_cast<const double2 *>(mul);
  const int2 *P2 = reinterpret_cast<const int2 *>(piv);
  const double onenrm = g.bnrm[b];
  const double tol = fmax(fabs(lam), onenrm) * DBL_EPSILON;
  // factorisation; small pivots of U perturbed to +-tol for the solves
  double a0 = d[0] - lam, a1 = e[0];
  for (int k0 = 0; k0 < bs - 1; k0 += SB) {
    double dbt[SB], ebt[SB + 1];  // d(k0+1 ..), e(k0 ..) of this batch of rows
#pragma unroll
    for (int u = 0; u < SB; u++) dbt[u] = d[min(k0 + 1 + u, bs - 1)];
#pragma unroll
    for (int u = 0; u <= SB; u++) ebt[u] = e[min(k0 + u, bs - 2)];
#pragma unroll
    for (int kq = 0; kq < SB; kq++) {
      const int k = k0 + kq;
      if (k < bs - 1) {
        const double sub = ebt[kq];
        const double dn = dbt[kq] - lam, en = (k + 1 < bs - 1) ? ebt[kq + 1] : 0.0;
        const size_t o = at(k);
        double p0, p1, p2, mm;
        int pv;
        if (fabs(a0) >= fabs(sub)) {
          pv = 0; p0 = a0; p1 = a1; p2 = 0.0;
          mm = (a0 != 0.0) ? sub / a0 : 0.0;
          a0 = dn - mm * a1;
          a1 = en;
        } else {
          pv = 1; p0 = sub; p1 = dn; p2 = en;
          mm = a0 / sub;
          const double na0 = a1 - mm * dn, na1 = -mm * en;
          a0 = na0;
          a1 = na1;
        }
        if (fabs(p0) < tol) p0 = (p0 < 0 ? -tol : tol);
        piv[o] = pv; u0[o] = 1.0 / p0; u1[o] = p1; u2[o] = p2; mul[o] = mm;
      }
    }
  }
  double ulast;  // |U(bs,bs)| after the perturbation (DSTEIN's scale of the start vector)
  {
    const size_t o = at(bs - 1);
    double p0 = a0;
    if (fabs(p0) < tol) p0 = (p0 < 0 ? -tol : tol);
    u0[o] = 1.0 / p0; u1[o] = 0.0; u2[o] = 0.0;
    ulast = fabs(p0);
  }
  // deterministic start vector in [-1, 1), then solves until DSTEIN's test
  // (dstein.f): the input scaled to 1-norm bs |T|_1 max(eps, |U(bs,bs)|), a
  // solve passes when the largest entry of its solution reaches
  // sqrt(0.1 / bs), and the vector is accepted after EXTRA + 1 = 3 passing
  // solves; at most MAXITS = 5 solves, after which the vector is counted as
  // not converged (INFO > 0, dsyevx2.f:352-354).  A first solve that passes
  // (the usual case) gives exactly three solves.
  constexpr int MAXITS = 5, EXTRA = 2;
  const double dtpcrt = sqrt(0.1 / (double)bs) * dtscale;
  unsigned long long seed = 0x9E3779B97F4A7C15ull ^ ((unsigned long long)(t + 1) * 0xD1B54A32D192ED03ull);
  auto rnd = [&]() {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    return ((double)(seed >> 11) / 9007199254740992.0) * 2.0 - 1.0;
  };
  double sc = 1.0;
  int nrmchk = 0;
  bool conv = false;
  for (int it = 0; it < MAXITS; it++) {
    const bool first = (it == 0);
    // forward: L^-1 P x, x scaled by 1/max|x| of the previous solve; the
    // groups start at even rows, and row k reads x(k+1) before writing x(k)
    double xk = first ? rnd() : sc * x[at(0)];
    double l1 = fabs(xk);  // 1-norm of this solve's input
    int k = 0;
    for (; k + SG <= bs - 1; k += SG) {
      int2 pg[SG / 2];
      double2 mg[SG / 2], xg[SG / 2 + 1];  // xg: x rows k .. k+SG+1
#pragma unroll
      for (int q = 0; q < SG / 2; q++) {
        pg[q] = P2[pr(k + 2 * q)];
        mg[q] = M2[pr(k + 2 * q)];
      }
      if (!first) {
#pragma unroll
        for (int q = 0; q <= SG / 2; q++) xg[q] = X2[pr(k + 2 * q)];
      }
#pragma unroll
      for (int q = 0; q < SG; q++) {
        const double xr = (q & 1) ? xg[(q + 1) / 2].x : xg[q / 2].y;  // row k+q+1
        double xn = first ? rnd() : sc * xr;
        l1 += fabs(xn);
        const int pv = (q & 1) ? pg[q / 2].y : pg[q / 2].x;
        const double mm = (q & 1) ? mg[q / 2].y : mg[q / 2].x;
        if (pv) { const double tmp = xk; xk = xn; xn = tmp; }
        x[at(k + q)] = xk;
        xk = xn - mm * xk;
      }
    }
    for (; k < bs - 1; k++) {
      double xn = first ? rnd() : sc * x[at(k + 1)];
      l1 += fabs(xn);
      if (piv[at(k)]) { const double tmp = xk; xk = xn; xn = tmp; }
      x[at(k)] = xk;
      xk = xn - mul[at(k)] * xk;
    }
    x[at(bs - 1)] = xk;
    // backward: U^-1, groups of SG rows ending at an odd row.  A solve that
    // may be the last (two passing solves behind it, or the last allowed)
    // writes Z; one that may be followed by another writes x.
    double mx = 0.0, x1 = 0.0, x2 = 0.0;
    const bool wz = (nrmchk == EXTRA) || (it == MAXITS - 1);
    const bool wx = !wz;
    auto row = [&](int kk, double xv, double r0, double b1, double b2) {
      double v = xv;
      if (kk + 1 < bs) v -= b1 * x1;
      if (kk + 2 < bs) v -= b2 * x2;
      v *= r0;
      if (wz) zc[ib + kk] = v;
      if (wx) x[at(kk)] = v;
      mx = fmax(mx, fabs(v));
      x2 = x1;
      x1 = v;
    };
    k = bs - 1;
    if (!(k & 1)) {
      row(k, x[at(k)], u0[at(k)], 0.0, 0.0);
      k--;
    }
    for (; k - SG + 1 >= 0; k -= SG) {
      double2 xg[SG / 2], rg[SG / 2], bg1[SG / 2], bg2[SG / 2];
#pragma unroll
      for (int q = 0; q < SG / 2; q++) {
        const size_t p = pr(k - SG + 1 + 2 * q);
        xg[q] = X2[p];
        rg[q] = R2[p];
        bg1[q] = A12[p];
        bg2[q] = A22[p];
      }
#pragma unroll
      for (int q = SG / 2 - 1; q >= 0; q--) {
        const int kk = k - SG + 1 + 2 * q;  // even row of pair q
        row(kk + 1, xg[q].y, rg[q].y, bg1[q].y, bg2[q].y);
        row(kk, xg[q].x, rg[q].x, bg1[q].x, bg2[q].x);
      }
    }
    for (; k >= 0; k--) row(k, x[at(k)], u0[at(k)], u1[at(k)], u2[at(k)]);
    // DSTEIN's growth test on the solution of the scaled input
    const double scl = (double)bs * onenrm * fmax(DBL_EPSILON, ulast) / l1;
    if (mx * scl >= dtpcrt && ++nrmchk >= EXTRA + 1) {
      conv = true;
      break;
    }
    if (wz && it < MAXITS - 1) {
      // a candidate last solve failed: it wrote only Z, and the next solve
      // reads its input back from there (the same values)
      for (int kk = 0; kk < bs; kk++) x[at(kk)] = zc[ib + kk];
    }
    sc = 1.0 / mx;
  }
  if (!conv) atomicAdd(g.cnt + 3, 1);
}

// orthonormalise columns c0..c0+cw-1 of z in order (MGS), one work-group
__global__ __launch_bounds__(256) void k_mgs_chunk(int n, double *z, int ldz, int c0, int cw)
{
  __shared__ double red[40];
  for (int j = 0; j < cw; j++) {
    double *zj = z + (size_t)ldz * (c0 + j);
    for (int k = 0; k < j; k++) {
      const double *zk = z + (size_t)ldz * (c0 + k);
      double s = 0.0;
      for (int r = threadIdx.x; r < n; r += blockDim.x) s += zk[r] * zj[r];
      s = block_sum(s, red);
      for (int r = threadIdx.x; r < n; r += blockDim.x) zj[r] -= s * zk[r];
      __syncthreads();
    }
    double s = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) s += zj[r] * zj[r];
    s = block_sum(s, red);
    const double sc = 1.0 / sqrt(s);
    for (int r = threadIdx.x; r < n; r += blockDim.x) zj[r] *= sc;
    __syncthreads();
  }
}

// unit 2-norm, largest |component| positive (first maximum), per column
__global__ __launch_bounds__(256) void k_normsign(int n, double *z, int ldz)
{
  __shared__ double red[40];
  __shared__ double bestv[256];
  __shared__ int besti[256];
  double *zc = z + (size_t)ldz * blockIdx.x;
  double s = 0.0, bv = -1.0;
  int bi = 0x7fffffff;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    const double v = zc[r];
    s += v * v;
    if (fabs(v) > bv) { bv = fabs(v); bi = r; }
  }
  bestv[threadIdx.x] = bv;
  besti[threadIdx.x] = bi;
  s = block_sum(s, red);
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = -1.0;
    int idx = 0;
    for (int k = 0; k < (int)blockDim.x; k++)
      if (bestv[k] > b || (bestv[k] == b && besti[k] < idx)) { b = bestv[k]; idx = besti[k]; }
    red[34] = zc[idx];
  }
  __syncthreads();
  double sc = 1.0 / sqrt(s);
  if (red[34] < 0.0) sc = -sc;
  for (int r = threadIdx.x; r < n; r += blockDim.x) zc[r] *= sc;
}

// ------------------------------------------------------------------ back-transforms
// explicit panel of reflectors p0..p0+nb-1 (v(i+1)=1, v(i+2:n) from A) into vp
__global__ void k_vpanel(int n, int lda, int p0, int nb, const double *__restrict__ a, double *__restrict__ vp)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (r >= n || k >= nb) return;
  const int i = p0 + k;
  double v = 0.0;
  if (r == i + 1) v = 1.0;
  else if (r > i + 1) v = a[(size_t)r + (size_t)lda * i];
  vp[(size_t)r + (size_t)n * k] = v;
}

// T (nb x nb upper, nb <= BTNB) of the compact WY form I - V T V^T of
// H_0 H_1 ... H_(nb-1) from G = V^T V (DLARFT 'F','C').  DLARFT's recurrence
// T(0:i,i) = -tau_i T(0:i,0:i) G(0:i,i), T(i,i) = tau_i makes T the inverse of
// the upper triangular M = diag(1/tau) + (strict upper part of G), which is
// formed here by blocks: X11 = M11^-1 and X22 = M22^-1 (64 x 64, one wave
// each), then T12 = -X11 (M12 X22) (four waves).  A reflector with tau = 0
// (H_i = I) has a zero row and column in DLARFT's T; here its row and column
// of M are those of the identity and T(i,i) is set to 0 afterwards.
constexpr int BTNB = 128;
__global__ __launch_bounds__(256) void k_larft_inv(int nb, const double *__restrict__ g,
                                                   const double *__restrict__ tau, double *__restrict__ tm)
{
  constexpr int LDM = BTNB + 1;
  __shared__ double ms[BTNB * LDM];  // ms[c LDM + r] = M(r, c), then T(r, c)
  __shared__ double dinv[BTNB], taus[BTNB];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  if (t < BTNB) {
    const double tt = (t < nb) ? tau[t] : 0.0;
    taus[t] = tt;
    dinv[t] = (tt != 0.0) ? tt : 1.0;
  }
  __syncthreads();
  {
    // thread t: row r = t % 128 of columns c = t / 128 + 2u; every load of G
    // in flight together (clamped addresses, masked values)
    const int r = t & (BTNB - 1), c0 = t >> 7;
    double gv[BTNB / 2];
#pragma unroll
    for (int u = 0; u < BTNB / 2; u++) gv[u] = g[min(r, nb - 1) + (size_t)nb * min(c0 + 2 * u, nb - 1)];
    const double taur = taus[r];
#pragma unroll
    for (int u = 0; u < BTNB / 2; u++) {
      const int c = c0 + 2 * u;
      double v = 0.0;
      if (r == c) {
        v = (taur != 0.0) ? 1.0 / taur : 1.0;
      } else if (r < c && c < nb) {
        v = (taur != 0.0 && taus[c] != 0.0) ? gv[u] : 0.0;
      }
      ms[c * LDM + r] = v;
    }
  }
  __syncthreads();
  double x[NB];
  if (wid < 2 && (wid == 0 || nb > NB)) wave_trtri<NB, LDM>(ms + wid * NB * (LDM + 1), dinv + wid * NB, lane, x);
  __syncthreads();
  if (wid < 2 && (wid == 0 || nb > NB)) {
    double *blk = ms + wid * NB * (LDM + 1);
#pragma unroll
    for (int r = 0; r < NB; r++) blk[lane * LDM + r] = (r <= lane) ? x[r] : 0.0;
  }
  __syncthreads();
  if (nb > NB) {
    // thread (c = lane, rows 16 wid .. 16 wid + 15): P = M12 X22, then T12 = -X11 P
    const int c = lane, r0 = 16 * wid;
    double acc[16];
#pragma unroll
    for (int u = 0; u < 16; u++) acc[u] = 0.0;
    for (int q = 0; q < NB; q++) {
      const double xq = ms[(NB + c) * LDM + NB + q];  // X22(q, c)
#pragma unroll
      for (int u = 0; u < 16; u++) acc[u] = fma(ms[(NB + q) * LDM + r0 + u], xq, acc[u]);  // M12(r, q)
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; u++) ms[(NB + c) * LDM + r0 + u] = acc[u];  // P(r, c) over M12
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; u++) acc[u] = 0.0;
    for (int q = 0; q < NB; q++) {
      const double pq = ms[(NB + c) * LDM + q];  // P(q, c)
#pragma unroll
      for (int u = 0; u < 16; u++) acc[u] = fma(ms[q * LDM + r0 + u], pq, acc[u]);  // X11(r, q)
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; u++) ms[(NB + c) * LDM + r0 + u] = -acc[u];
    __syncthreads();
  }
  for (int idx = t; idx < nb * nb; idx += blockDim.x) {
    const int q = idx % nb, c = idx / nb;
    double v = (q <= c) ? ms[c * LDM + q] : 0.0;
    if (q == c && taus[q] == 0.0) v = 0.0;
    tm[idx] = v;
  }
}

// Cholesky G = R^T R of a small Gram matrix (cw <= N <= 64) and R^-1
// (upper), by one wave in registers as k_chol_diag (lane c holds column c;
// columns past cw and the lanes above N carry the identity); *fail set when
// G is not numerically positive definite (a non-positive or non-finite
// pivot, taken as 1 so that the arithmetic stays finite)
template <int N>
__global__ __launch_bounds__(64) void k_chol_rinv(int cw, const double *__restrict__ g, double *__restrict__ rinv,
                                                  int *__restrict__ fail)
{
  static_assert(N <= 64, "one lane per column");
  constexpr int LD = N + 1;
  __shared__ double sm[64 * LD];  // sm[c LD + r] = R(r, c), then R^-1(r, c)
  __shared__ double urow[2][64];
  __shared__ double dinv[64];
  const int c = threadIdx.x;
  double v[N];
  const int cl = min(c, cw - 1);
#pragma unroll
  for (int r = 0; r < N; r++) v[r] = g[min(r, cw - 1) + (size_t)cw * cl];  // clamped: all loads in flight
#pragma unroll
  for (int r = 0; r < N; r++) v[r] = (r <= c && c < cw) ? v[r] : (r == c ? 1.0 : 0.0);
  bool bad = false;
#pragma unroll
  for (int j = 0; j < N; j++) {
    const double ajj = lane_bcast(v[j], j);
    const bool ok = ajj > 0.0 && ajj <= DBL_MAX;
    bad = bad || !ok;
    v[j] *= rsqrt_nr(ok ? ajj : 1.0);
    urow[j & 1][c] = v[j];
    __syncthreads();
    double u[N];
#pragma unroll
    for (int r = j + 1; r < N; r++) u[r] = urow[j & 1][r];
#pragma unroll
    for (int r = j + 1; r < N; r++) v[r] = fma(-u[r], v[j], v[r]);
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < N; r++) sm[c * LD + r] = (r <= c) ? v[r] : 0.0;
  __syncthreads();
  dinv[c] = (c < N) ? 1.0 / sm[c * LD + c] : 1.0;
  __syncthreads();
  double x[N];
  wave_trtri<N, LD>(sm, dinv, c, x);
  if (c < cw) {
#pragma unroll
    for (int r = 0; r < N; r++)
      if (r < cw) rinv[r + (size_t)cw * c] = (r <= c) ? x[r] : 0.0;
  }
  if (c == 0 && bad) *fail = 1;
}

__global__ void k_gather_cols(int n, int m, const int *__restrict__ perm, const double *__restrict__ src,
                              double *__restrict__ dst)
{
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = blockIdx.y;
  if (r >= n || c >= m) return;
  dst[(size_t)r + (size_t)n * c] = src[(size_t)r + (size_t)n * perm[c]];
}

// ------------------------------------------------------------------ driver
int seclr4_buffers(Ctx &c, int n, int cplx, Seclr4Bufs &B)
{
  int rc;
  const int nt = (n + SYMV_T - 1) / SYMV_T;
  const size_t z = cplx ? 2 : 1;
  const size_t nts = (size_t)nt;
  const size_t nnorm = nts + (size_t)n / COL_ROWS + 64;  // partial norms per block of 64 (complex) or 32 rows
  // SYMV partials: 128 (real) or 256 (complex) per tile of the lower triangle
  const size_t npart = cplx ? std::max(2 * nts * n, (size_t)128 * nts * (nts + 1))
                            : std::max(nts * n, (size_t)64 * nts * (nts + 1));
  const size_t small = (size_t)12 * n + z * 2 * n + z * 3 * (size_t)n * TNB + npart + z * nts * 2 * TNB +
                       nts * (nts + 1) / 2 + 2 + nnorm + 64 + 32 * 24;
  if ((rc = c.w_small.alloc(small))) return rc;
  if ((rc = c.w_int.alloc((size_t)6 * n + 32))) return rc;
  B.n = n;
  B.nt = nt;
  B.cplx = cplx;
  double *p = c.w_small.p;
  auto take = [&](size_t cnt) {  // 256-byte aligned pieces
    double *q = p;
    p += (cnt + 31) & ~(size_t)31;
    return q;
  };
  B.d = take(n); B.e = take(n); B.e2 = take(n);
  B.tau = take(z * n); B.scal = take(z * n);
  B.wv = take(n);
  B.P = take(z * 3 * (size_t)n * TNB);
  B.part = take(npart);
  B.tdots = take(z * nts * 2 * TNB);
  B.tvy = take(nts * (nts + 1) / 2 + 2);
  B.tnorm = take(nnorm);
  B.bgl = take(n); B.bgu = take(n); B.batol = take(n); B.bnrm = take(n); B.xs = take(n);
  B.pivmin = take(1);
  // ints: [0,4) INFO, [4,12) counts, [12,16) grid-reduction tickets, then task arrays
  B.info = c.w_int.p;
  B.cnt = B.info + 4;
  B.ticket = B.info + 12;
  B.tblk = B.info + 16;
  B.tidx = B.tblk + n;
  B.bstart = B.tidx + n;
  B.bend = B.bstart + n;
  B.cstart = B.bend + n;
  B.perm = B.cstart + n;
  W2K_HIP(hipMemsetAsync(c.w_int.p, 0, sizeof(int) * 16, c.stream));
  return 0;
}

int tridiag_eigen(Ctx &c, Seclr4Bufs &B, double el, double eu, int nume, double *Zt, int *m_out, int *nbelw_out)
{
  const int n = B.n;
  hipStream_t s = c.stream;
  int rc;
  *m_out = 0;
  // ---- bisection
  BisArgs b;
  b.n = n; b.d = B.d; b.e = B.e; b.e2 = B.e2; b.el = el; b.eu = eu; b.nume = nume;
  b.cnt = B.cnt; b.tblk = B.tblk; b.tidx = B.tidx; b.bstart = B.bstart; b.bend = B.bend;
  b.bgl = B.bgl; b.bgu = B.bgu; b.batol = B.batol; b.bnrm = B.bnrm; b.pivmin = B.pivmin;
  hipLaunchKernelGGL(k_bisect_prep, dim3(1), dim3(64), 0, s, b);
  W2K_STAGE(s, "k_bisect_prep");
  int hc[4];
  W2K_HIP(hipMemcpyAsync(hc, B.cnt, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  const int m = hc[0];
  *nbelw_out = hc[1];
  if ((rc = c.d_eig.alloc(m > 0 ? m : 1))) return rc;
  if (m == 0) return 0;
  double *w = B.wv;
  hipLaunchKernelGGL(k_bisect, dim3((m + 3) / 4), dim3(256), 0, s, b, m, w);
  W2K_STAGE(s, "k_bisect");
  // ---- inverse iteration + cluster orthogonalisation
  hipLaunchKernelGGL(k_stein_prep, dim3(1), dim3(64), 0, s, b, m, w, B.xs, B.cstart);
  W2K_STAGE(s, "k_stein_prep");
  const size_t zc = (size_t)n * m;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((zc + 255) / 256)), dim3(256), 0, s, zc, Zt, 0.0);
  W2K_STAGE(s, "k_fill");
  // factor planes of rows 0 .. n in row pairs (16-byte aligned: m pairs per row pair)
  const size_t zp = (size_t)((n + 2) / 2) * 2 * m;
  if ((rc = c.w_fac.alloc((size_t)5 * zp)) || (rc = c.w_piv.alloc(zp))) return rc;
  {
    // W2K_TEST_STEIN_FAIL=1 (tests only) raises DSTEIN's growth threshold
    // out of reach, so that every vector takes the non-convergence path
    const bool force_fail = std::getenv("W2K_TEST_STEIN_FAIL") && std::atoi(std::getenv("W2K_TEST_STEIN_FAIL"));
    double *f = c.w_fac.p;
    hipLaunchKernelGGL(k_stein, dim3((m + 63) / 64), dim3(64), 0, s, b, m, B.xs, Zt, n, f, f + zp, f + 2 * zp,
                       f + 3 * zp, f + 4 * zp, c.w_piv.p, force_fail ? 1.0e300 : 1.0);
    W2K_STAGE(s, "k_stein");
  }
  std::vector<int> cs(m);
  int nfail = 0;
  W2K_HIP(hipMemcpyAsync(cs.data(), B.cstart, sizeof(int) * m, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipMemcpyAsync(&nfail, B.cnt + 3, sizeof(int), hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  if (nfail > 0) {
    // DSTEIN's INFO: the number of vectors that did not converge in MAXITS
    // solves (dsyevx2.f:352-354); SECLR4 stops on it (seclr4.F:304-309)
    set_last_error("SYEVX: %d eigenvectors failed to converge in inverse iteration", nfail);
    return nfail;
  }
  // clusters [t0, t1): chunks of CW vectors, each orthogonalised against
  // the cluster's earlier chunks (classical Gram-Schmidt, twice, on MFMA)
  // and then within itself by CholQR2 (Z = Q R, R from the Cholesky of
  // Z^T Z), which equals in-order Gram-Schmidt.  All chunks run without a
  // host round trip, each with a flag that its Cholesky sets when the Gram
  // matrix is not numerically definite; the flags are read once, and a
  // cluster with a flagged chunk is restored from a copy taken before from
  // that chunk on and redone chunk by chunk, a flagged chunk then falling
  // back to modified Gram-Schmidt
  constexpr int CW = 64;
  std::vector<std::pair<int, int>> cl;
  size_t nch = 0;
  for (int t0 = 0; t0 < m;) {
    int t1 = t0 + 1;
    while (t1 < m && cs[t1] == t0) t1++;
    if (t1 - t0 > 1) {
      cl.emplace_back(t0, t1);
      nch += (size_t)(t1 - t0 + CW - 1) / CW;
    }
    t0 = t1;
  }
  if (!cl.empty()) {
    if ((rc = c.w_cl.alloc((size_t)2 * n * CW + 2 * CW * CW + (size_t)m * CW + nch + 1))) return rc;
    double *csave = c.w_cl.p, *ctmp = csave + (size_t)n * CW, *cg = ctmp + (size_t)n * CW;
    double *crinv = cg + CW * CW, *cy = crinv + CW * CW;
    int *flags = reinterpret_cast<int *>(cy + (size_t)m * CW);
    double *zbak = c.w_fac.p;  // the factors are dead; 5 zp >= n m
    auto cgs2 = [&](int t0, int c0, int cw) -> int {
      const int np = c0 - t0;
      double *Q = Zt + (size_t)n * t0, *Zc = Zt + (size_t)n * c0;
      for (int pass = 0; pass < 2 && np > 0; pass++) {
        int r2;
        if ((r2 = dgemm(s, 'T', 'N', np, cw, n, 1.0, Q, n, Zc, n, 0.0, cy, np, 0, 0, 0, c.w_gk.p, c.w_gk.n)))
          return r2;
        if ((r2 = dgemm(s, 'N', 'N', n, cw, np, -1.0, Q, n, cy, np, 1.0, Zc, n))) return r2;
      }
      return 0;
    };
    // two passes, Zc -> ctmp -> Zc
    auto cholqr2 = [&](double *Zc, int cw, int *flag) -> int {
      double *src = Zc, *dst = ctmp;
      for (int pass = 0; pass < 2; pass++) {
        int r2;
        if ((r2 = dgemm(s, 'T', 'N', cw, cw, n, 1.0, src, n, src, n, 0.0, cg, cw, 0, 0, 0, c.w_gk.p, c.w_gk.n)))
          return r2;
        hipLaunchKernelGGL(k_chol_rinv<CW>, dim3(1), dim3(64), 0, s, cw, cg, crinv, flag);
        W2K_STAGE(s, "k_chol_rinv");
        if ((r2 = dgemm(s, 'N', 'N', n, cw, cw, 1.0, src, n, crinv, cw, 0.0, dst, n, 0, 0, 1))) return r2;
        std::swap(src, dst);
      }
      return 0;
    };
    W2K_HIP(hipMemcpyAsync(zbak, Zt, sizeof(double) * zc, hipMemcpyDeviceToDevice, s));
    W2K_HIP(hipMemsetAsync(flags, 0, sizeof(int) * nch, s));
    size_t fi = 0;
    for (auto &tc : cl)
      for (int c0 = tc.first; c0 < tc.second; c0 += CW, fi++) {
        const int cw = std::min(CW, tc.second - c0);
        if ((rc = cgs2(tc.first, c0, cw)) || (rc = cholqr2(Zt + (size_t)n * c0, cw, flags + fi))) return rc;
      }
    std::vector<int> hf(nch);
    W2K_HIP(hipMemcpyAsync(hf.data(), flags, sizeof(int) * nch, hipMemcpyDeviceToHost, s));
    W2K_HIP(hipStreamSynchronize(s));
    // W2K_TEST_CLUSTER_FALLBACK (tests only): 1 flags every chunk, so that
    // every cluster is restored and redone; 2 also sends every redone chunk
    // to modified Gram-Schmidt
    const char *fe = std::getenv("W2K_TEST_CLUSTER_FALLBACK");
    const int force = fe ? std::atoi(fe) : 0;
    if (force >= 1) std::fill(hf.begin(), hf.end(), 1);
    fi = 0;
    for (auto &tc : cl) {
      const int t0 = tc.first, t1 = tc.second;
      const size_t nc = (size_t)(t1 - t0 + CW - 1) / CW;
      size_t jf = 0;
      while (jf < nc && !hf[fi + jf]) jf++;
      fi += nc;
      if (jf == nc) continue;
      const int cf = t0 + (int)jf * CW;
      W2K_HIP(hipMemcpyAsync(Zt + (size_t)n * cf, zbak + (size_t)n * cf, sizeof(double) * n * (t1 - cf),
                             hipMemcpyDeviceToDevice, s));
      int *cfail = flags + nch;
      for (int c0 = cf; c0 < t1; c0 += CW) {
        const int cw = std::min(CW, t1 - c0);
        double *Zc = Zt + (size_t)n * c0;
        W2K_HIP(hipMemcpyAsync(csave, Zc, sizeof(double) * n * cw, hipMemcpyDeviceToDevice, s));
        W2K_HIP(hipMemsetAsync(cfail, 0, sizeof(int), s));
        if ((rc = cgs2(t0, c0, cw)) || (rc = cholqr2(Zc, cw, cfail))) return rc;
        int hfail = 0;
        W2K_HIP(hipMemcpyAsync(&hfail, cfail, sizeof(int), hipMemcpyDeviceToHost, s));
        W2K_HIP(hipStreamSynchronize(s));
        if (hfail || force >= 2) {
          W2K_HIP(hipMemcpyAsync(Zc, csave, sizeof(double) * n * cw, hipMemcpyDeviceToDevice, s));
          if ((rc = cgs2(t0, c0, cw))) return rc;
          hipLaunchKernelGGL(k_mgs_chunk, dim3(1), dim3(256), 0, s, n, Zt, n, c0, cw);
          W2K_STAGE(s, "k_mgs_chunk");
        }
      }
    }
  }
  hipLaunchKernelGGL(k_normsign, dim3(m), dim3(256), 0, s, n, Zt, n);
  W2K_STAGE(s, "k_normsign");
  W2K_HIP(hipStreamSynchronize(s));
  *m_out = m;
  return 0;
}

int sort_eigen(Ctx &c, Seclr4Bufs &B, int m)
{
  hipStream_t s = c.stream;
  std::vector<double> hw(m);
  std::vector<int> hp(m);
  W2K_HIP(hipMemcpyAsync(hw.data(), B.wv, sizeof(double) * m, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < m; j++) hp[j] = j;
  for (int j = 0; j < m - 1; j++) {
    int ii = -1;
    double tmin = hw[j];
    for (int jj = j + 1; jj < m; jj++)
      if (hw[jj] < tmin) { ii = jj; tmin = hw[jj]; }
    if (ii >= 0) {
      hw[ii] = hw[j];
      hw[j] = tmin;
      std::swap(hp[ii], hp[j]);
    }
  }
  W2K_HIP(hipMemcpyAsync(B.perm, hp.data(), sizeof(int) * m, hipMemcpyHostToDevice, s));
  W2K_HIP(hipMemcpyAsync(B.wv, hw.data(), sizeof(double) * m, hipMemcpyHostToDevice, s));
  W2K_HIP(hipStreamSynchronize(s));  // the host vectors above go out of scope
  return 0;
}

// Launch mode of the tridiagonalisation (2n kernel launches and n/32 GEMMs
// with no host read-back): direct launches, or one HIP graph captured from
// them and replayed.  The choice is made per shape by measurement: the first
// call with a given key (order, leading dimension and the buffers the
// launches address) launches directly, the second captures the launches into
// a graph and replays it, and later calls replay it only if the first replay
// was faster than the direct launches (the phase time of the previous call
// decides).  Kernel timing, W2K_SYNC_CHECK or a failed capture keep the
// direct launches.
template <class F>
static int tri_run(Ctx &c, hipStream_t s, int n, int lda, const double *a, bool kt, F &launch)
{
  Ctx::TriGraph &tg = c.tri_graph;
  const bool same = tg.n == n && tg.lda == lda && tg.a == a && tg.small == c.w_small.p && tg.ints == c.w_int.p;
  if (same) {
    // the phase time of the previous call, by the mode it used
    if (tg.mode == 0) tg.direct_ms = c.last_ms[5];
    if (tg.mode == 1) {
      tg.graph_ms = c.last_ms[5];
      if (tg.direct_ms > 0.0 && tg.graph_ms >= tg.direct_ms) tg.use = 0;
    }
  } else {
    if (tg.exec) (void)hipGraphExecDestroy(tg.exec);
    tg = Ctx::TriGraph();
    tg.n = n; tg.lda = lda; tg.a = a; tg.small = c.w_small.p; tg.ints = c.w_int.p;
  }
  tg.seen++;
  tg.mode = -1;
  const bool allowed = c.graphs && !kt && !sync_checks_enabled() && tg.use;
  if (allowed && tg.exec) {
    W2K_HIP(hipGraphLaunch(tg.exec, s));
    tg.mode = c.last_ms[12] = 1;
    return 0;
  }
  if (allowed && tg.seen >= 2) {
    if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
      hipGraph_t graph = nullptr;
      const int rc = launch();
      const hipError_t e = hipStreamEndCapture(s, &graph);
      hipGraphExec_t ex = nullptr;
      if (rc == 0 && e == hipSuccess && hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess) {
        (void)hipGraphDestroy(graph);
        tg.exec = ex;
        W2K_HIP(hipGraphLaunch(ex, s));
        tg.mode = c.last_ms[12] = 2;
        return 0;
      }
      if (graph) (void)hipGraphDestroy(graph);
    }
    // nothing of a failed capture ran: clear the error and launch directly
    (void)hipGetLastError();
    tg.use = 0;
  }
  const int rc = launch();
  if (rc) return rc;
  tg.mode = c.last_ms[12] = 0;
  return 0;
}

static int seclr4_real(Ctx &c, double el, double eu, int nume, int *ne_out, int *nbelw_out)
{
  const int n = c.n;
  hipStream_t s = c.stream;
  int rc;
  const size_t nn = (size_t)n * n;
  // workspaces: w_a = S/U, w_b = H full then C, w_c = W = U^-1, w_x = X.
  // H/C has its leading dimension rounded up to 16 so that its columns start
  // on 128-byte boundaries (the tridiagonalisation SYMV reads aligned rows)
  const int lda = (n + 15) & ~15;
  if ((rc = c.w_a.alloc(nn)) || (rc = c.w_b.alloc((size_t)lda * n)) || (rc = c.w_c.alloc(nn))) return rc;
  double *U = c.w_a.p, *Hf = c.w_b.p, *W = c.w_c.p;
  DBuf<double> &X = c.w_x;
  if ((rc = X.alloc(nn))) return rc;
  if ((rc = c.w_gk.alloc((size_t)1 << 22))) return rc;  // split-K partial products
  Seclr4Bufs B;
  if ((rc = seclr4_buffers(c, n, 0, B))) return rc;
  double *d = B.d, *e = B.e, *tau = B.tau, *scal = B.scal, *P = B.P;
  int *info = B.info;

  W2K_HIP(hipEventRecord(c.ev[6], s));
  // ---- 1. Cholesky S = U^T U and W = U^-1 together, by recursive halving
  // on 64-column boundaries: factor the leading half (U11, W11), form
  // U12 = W11^T S12 and S22 -= U12^T U12 (upper triangle), factor the
  // trailing half (U22, W22), then W12 = -(W11 U12) W22.  The 64x64 leaves
  // are factored and inverted by one wave each; the products at the top
  // levels are square.  W is zero below the diagonal.
  {
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nn, W, 0.0);
    W2K_STAGE(s, "k_fill");
    dim3 g2((n + 255) / 256, n);
    hipLaunchKernelGGL(k_prep, g2, dim3(256), 0, s, n, lda, c.d_hs.p, c.d_hsdiag.p, U, Hf);
    W2K_STAGE(s, "k_prep");
    std::function<int(int, int)> rchol = [&](int j0, int nb) -> int {
      if (nb <= NB) {
        hipLaunchKernelGGL(k_chol_diag, dim3(1), dim3(64), 0, s, n, j0, nb, U, W, info);
        W2K_STAGE(s, "k_chol_diag");
        return 0;
      }
      const int n1 = ((nb / 2 + NB - 1) / NB) * NB, n2 = nb - n1;
      int r;
      if ((r = rchol(j0, n1))) return r;
      double *S12 = U + (size_t)j0 + (size_t)n * (j0 + n1), *S22 = U + (size_t)(j0 + n1) + (size_t)n * (j0 + n1);
      const double *W11 = W + (size_t)j0 + (size_t)n * j0;
      // X is free until the reduction: T (n1 x n2, leading dimension n1) at
      // its start, the rest the split-K workspace of the products
      double *T = X.p, *ws = X.p + (size_t)n1 * n2;
      const size_t wsn = nn - (size_t)n1 * n2;
      if ((r = dgemm(s, 'T', 'N', n1, n2, n1, 1.0, W11, n, S12, n, 0.0, T, n1, 0, 2, 0, ws, wsn))) return r;
      W2K_HIP(hipMemcpy2DAsync(S12, sizeof(double) * n, T, sizeof(double) * n1, sizeof(double) * n1, n2,
                               hipMemcpyDeviceToDevice, s));
      if ((r = dgemm(s, 'T', 'N', n2, n2, n1, -1.0, S12, n, S12, n, 1.0, S22, n, 2, 0, 0, X.p, nn))) return r;
      if ((r = rchol(j0 + n1, n2))) return r;
      if ((r = dgemm(s, 'N', 'N', n1, n2, n1, 1.0, W11, n, S12, n, 0.0, T, n1, 0, 1, 0, ws, wsn))) return r;
      return dgemm(s, 'N', 'N', n1, n2, n2, -1.0, T, n1, W + (size_t)(j0 + n1) + (size_t)n * (j0 + n1), n, 0.0,
                   W + (size_t)j0 + (size_t)n * (j0 + n1), n, 0, 0, 1, ws, wsn);
    };
    if ((rc = rchol(0, n))) return rc;
  }
  W2K_HIP(hipEventRecord(c.ev[7], s));
  int hinfo[4];
  W2K_HIP(hipMemcpyAsync(hinfo, info, sizeof(int) * 4, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  if (hinfo[0] != 0) return hinfo[0];  // POTRF INFO, seclr4.F:200-205
  // ---- 2. reduction: X = H W, C = W^T X (lower triangle only)
  {
    if ((rc = dgemm(s, 'N', 'N', n, n, n, 1.0, Hf, lda, W, n, 0.0, X.p, n, 0, 0, 1))) return rc;
    if ((rc = dgemm(s, 'T', 'N', n, n, n, 1.0, W, n, X.p, n, 0.0, Hf, lda, 1, 2, 0))) return rc;
  }
  W2K_HIP(hipEventRecord(c.ev[8], s));
  // ---- 3. tridiagonalisation of C (in Hf)
  double *A = Hf;
  int nsymv = 0;
  {
    // optional per-launch events around the SYMV (kernel-level roofline)
    const bool kt = c.ktime != 0;
    if (kt && c.kev.size() < (size_t)2 * n) {
      const size_t old = c.kev.size();
      c.kev.resize((size_t)2 * n);
      for (size_t q = old; q < c.kev.size(); q++) W2K_HIP(hipEventCreate(&c.kev[q]));
    }
    TriWork g;
    g.n = n; g.lda = lda; g.a = A; g.P = P; g.d = d; g.e = e; g.tau = tau; g.scal = scal;
    g.part = B.part; g.dots = B.tdots; g.vy = B.tvy; g.norm = B.tnorm; g.counter = B.ticket;
    // every launch of the phase (no host read-back in between)
    auto launches = [&]() -> int {
      nsymv = 0;
      for (int p0 = 0; p0 < n; p0 += TNB) {
        const int nb = std::min(TNB, n - p0);
        g.p0 = p0;
        g.nb = nb;
        // reflector of the panel's first column (already updated by earlier panels)
        hipLaunchKernelGGL(k_tri_col, dim3((n - p0 + COL_ROWS - 1) / COL_ROWS), dim3(64 * COL_WAVES), 0, s, g, -1, p0, 0,
                           0);
        W2K_STAGE(s, "k_tri_col");
        int nref = 0;  // panel columns that carry a reflector
        for (int k = 0; k < nb; k++) {
          const int i = p0 + k;
          if (i >= n - 1) break;
          nref = k + 1;
          const int rows = n - (i + 1);
          const int base = (i + 1) & ~15;  // the SYMV's aligned tile origin
          const int nt = (n - base + SYMV_T - 1) / SYMV_T, nblk = nt * (nt + 1) / 2;
          // timed launches carry their events in the dispatch itself, so the
          // duration is the kernel's own (no marker packets between launches)
          if (kt) hipExtLaunchKernelGGL(k_tri_symv, dim3(nblk + 2 * nt), dim3(SYMV_T), 0, s, c.kev[2 * nsymv],
                                        c.kev[2 * nsymv + 1], 0, g, i, nt);
          else hipLaunchKernelGGL(k_tri_symv, dim3(nblk + 2 * nt), dim3(SYMV_T), 0, s, g, i, nt);
          W2K_STAGE(s, "k_tri_symv");
          nsymv++;
          const int next = (k + 1 < nb) ? i + 1 : -1;
          hipLaunchKernelGGL(k_tri_col, dim3((rows + COL_ROWS - 1) / COL_ROWS), dim3(64 * COL_WAVES), 0, s, g, i,
                             next, nt, nblk);
          W2K_STAGE(s, "k_tri_col");
        }
        if (nref > 0) {
          hipLaunchKernelGGL(k_tri_store, dim3((n + 255) / 256, nref), dim3(256), 0, s, g, nref);
          W2K_STAGE(s, "k_tri_store");
        }
        const int p1 = p0 + nb;
        if (p1 < n) {
          // A(p1:n, p1:n) -= [V W] [W V]^T, lower triangle
          const int r = n - p1;
          double *A1 = A + (size_t)p1 + (size_t)lda * p1;
          int r2;
          if ((r2 = dgemm(s, 'N', 'T', r, r, 2 * nb, -1.0, P + p1, n, P + (size_t)n * nb + p1, n, 1.0, A1, lda, 1)))
            return r2;
        }
      }
      return 0;
    };
    if ((rc = tri_run(c, s, n, lda, A, kt, launches))) return rc;
  }
  W2K_HIP(hipEventRecord(c.ev[9], s));
  if (c.ktime) {
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
  // ---- 4./5. bisection, inverse iteration, cluster orthogonalisation
  double *Zt = X.p;  // n x m
  int m = 0, nbelw = 0;
  if ((rc = tridiag_eigen(c, B, el, eu, nume, Zt, &m, &nbelw))) return rc;
  *nbelw_out = nbelw;
  c.nbelw = nbelw;
  if ((rc = c.d_z.alloc((size_t)n * (m > 0 ? m : 1)))) return rc;
  W2K_HIP(hipEventRecord(c.ev[10], s));
  if (m > 0) {
    // ---- 6. Z = Q Zt (reflectors of step 3, last panel first), sort, Z = W Z
    // Blocks of BTNB reflectors, last block first.  The reflector panel V,
    // its Gram matrix and T (I - V T V^T) depend on the reflectors only, so
    // they are formed on the second stream up to two blocks ahead, in two
    // buffers, while this stream applies the blocks to Z:
    //   Y1 = V^T Z,  Y2 = T Y1,  Z -= V Y2.
    // bt_ready[u]: block in buffer u formed; bt_used[u]: its products with Z
    // done, so the buffer may take the block after next.
    // w_bt: V (n x BTNB) x 2, T (BTNB x BTNB) x 2, the Gram matrix, Y1 and
    // Y2 (BTNB x m each), and the split-K workspace of the Gram products
    constexpr size_t GWS = (size_t)32 * BTNB * BTNB;
    if ((rc = c.w_bt.alloc((size_t)(2 * n + 2 * m + 3 * BTNB) * BTNB + GWS))) return rc;
    double *vbuf[2], *tbuf[2];
    vbuf[0] = c.w_bt.p;
    vbuf[1] = vbuf[0] + (size_t)n * BTNB;
    tbuf[0] = vbuf[1] + (size_t)n * BTNB;
    tbuf[1] = tbuf[0] + BTNB * BTNB;
    double *gm = tbuf[1] + BTNB * BTNB, *y1 = gm + BTNB * BTNB, *y2 = y1 + (size_t)BTNB * m;
    double *gws = y2 + (size_t)BTNB * m;
    const int nref = n - 1;  // reflectors 0..n-2
    const int nblk = nref > 0 ? (nref + BTNB - 1) / BTNB : 0;
    hipStream_t sa = c.aux;
    auto p0_of = [&](int q) { return (nblk - 1 - q) * BTNB; };
    auto prep = [&](int q) -> int {
      const int u = q & 1, p0 = p0_of(q), nb = std::min(BTNB, nref - p0);
      if (q >= 2) W2K_HIP(hipStreamWaitEvent(sa, c.bt_used[u], 0));
      hipLaunchKernelGGL(k_vpanel, dim3((n + 255) / 256, nb), dim3(256), 0, sa, n, lda, p0, nb, A, vbuf[u]);
      W2K_STAGE(sa, "k_vpanel");
      int r2;
      // rows 0 .. p0 of the block's reflectors are zero: the products start at
      // row p0+1 rounded down to 16 rows, which keeps the operands' 128-byte alignment
      const int o = (p0 + 1) & ~15, nr = n - o;
      if ((r2 = dgemm(sa, 'T', 'N', nb, nb, nr, 1.0, vbuf[u] + o, n, vbuf[u] + o, n, 0.0, gm, nb, 0, 0, 0, gws,
                      GWS)))
        return r2;
      hipLaunchKernelGGL(k_larft_inv, dim3(1), dim3(256), 0, sa, nb, gm, tau + p0, tbuf[u]);
      W2K_STAGE(sa, "k_larft_inv");
      W2K_HIP(hipEventRecord(c.bt_ready[u], sa));
      return 0;
    };
    if (nblk > 0) {
      // the second stream starts once the reflectors and Zt are in place
      W2K_HIP(hipEventRecord(c.bt_start, s));
      W2K_HIP(hipStreamWaitEvent(sa, c.bt_start, 0));
      for (int q = 0; q < std::min(nblk, 2); q++)
        if ((rc = prep(q))) return rc;
    }
    for (int q = 0; q < nblk; q++) {
      const int u = q & 1, p0 = p0_of(q), nb = std::min(BTNB, nref - p0);
      W2K_HIP(hipStreamWaitEvent(s, c.bt_ready[u], 0));
      const int o = (p0 + 1) & ~15, nr = n - o;  // the block's reflectors are zero above row p0+1
      if ((rc = dgemm(s, 'T', 'N', nb, m, nr, 1.0, vbuf[u] + o, n, Zt + o, n, 0.0, y1, nb, 0, 0, 0, c.w_gk.p,
                      c.w_gk.n)))
        return rc;
      if ((rc = dgemm(s, 'N', 'N', nb, m, nb, 1.0, tbuf[u], nb, y1, nb, 0.0, y2, nb, 0, 1, 0))) return rc;
      if ((rc = dgemm(s, 'N', 'N', nr, m, nb, -1.0, vbuf[u] + o, n, y2, nb, 1.0, Zt + o, n))) return rc;
      if (q + 2 < nblk) {
        W2K_HIP(hipEventRecord(c.bt_used[u], s));
        if ((rc = prep(q + 2))) return rc;
      }
    }
    if ((rc = sort_eigen(c, B, m))) return rc;
    double *Zs = U;  // U is not needed after the inversion
    hipLaunchKernelGGL(k_gather_cols, dim3((n + 255) / 256, m), dim3(256), 0, s, n, m, B.perm, Zt, Zs);
    W2K_STAGE(s, "k_gather_cols");
    if ((rc = dgemm(s, 'N', 'N', n, m, n, 1.0, W, n, Zs, n, 0.0, c.d_z.p, n, 0, 1, 0))) return rc;
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

int seclr4(Ctx &c, double el, double eu, int nume, int *ne_out, int *nbelw_out)
{
  *ne_out = 0;
  *nbelw_out = 0;
  if (el >= eu) { el = -1.0e70; eu = 1.0e70; }  // seclr4.F:149-152
  if (nume <= 0) nume = c.n;
  c.ne = 0;
  if (c.run().is_complex) return seclr4_complex(c, el, eu, nume, ne_out, nbelw_out);
  return seclr4_real(c, el, eu, nume, ne_out, nbelw_out);
}

}  // namespace w2k
