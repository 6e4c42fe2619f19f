// HNS on the GPU (hns.F:150-488): non-spherical V_LM terms added to the lower
// triangle of HS (diagonal included).
//
// For every atom (all NDF equivalent atoms) the augmentation coefficients
//   ALM/BLM/CLM(i, L0M0) = 4 pi R^2 sqrt(1/Omega) i^L0 e^{2 pi i G_i.tau}
//                          conj Y_L0M0(R k_i) x {AL, BL, 0} (PW rows) or
//                          x {ALO, BLO, CLO} (LO rows of that atom/l/jlo)
// (hns.F:160-259) are contracted with the per-atom coupling coefficients
// (Gaunt x radial integrals, hns.F:266-434, precomputed once per run) into
//   A2 = sum_p TUU conj(ALM_p) + TDU conj(BLM_p) + TUU21 conj(CLM_p), etc.
// and stored as the two operands of ONE rank-K update over all atoms:
//   REAL:    HS_low += [A2R,-A2I,B2R,-B2I] . [A1R,A1I,B1R,B1I]^T
//            (+ [C2R,-C2I] . [C1R,C1I]^T on the LO x LO block)
//   COMPLEX: HS_low += A2 A1^T + B2 B1^T (+ C2 C1^T), "T" not "H" (hns.F:451),
//            as two real GEMMs for the real and imaginary parts
// which replaces the per-atom DSYR2M/ZHER2M calls (dsyr2m.f:66-77) with a
// single MFMA GEMM of inner dimension 4*sum(LINDEX).
#include <algorithm>
#include <vector>

#include "special.h"
#include "w2k_internal.h"

namespace w2k {

constexpr double PI_H = 3.141592653589793238462643;

struct HnsArgs {
  int n, nv, nnsl, lt, ndf;
  const double *kv, *cs, *al, *bl;
  const int *atom_of, *lindex, *koff;  // koff[e]: first column of atom e (in units of LINDEX columns)
  const int *lm0_of;                   // [nat][25]
  const double *tcoef;                 // [nat][25(k)][25(p)][9][2]
  const int *lorow;                    // [nat][LOMAXP][NLOAT][2]: first/last 0-based LO row (or -1)
  const double *loabc;                 // [nat][LOMAXP][NLOAT][3]
  const double *rotij, *rotloc, *rmt;
  double vi;
  int ktot;                            // sum over atoms of LINDEX
  double *L, *R, *R2;                  // n x (4 ktot), R2 only for the complex build
  double *LC, *RC, *RC2;               // nlo x (2 ktot)
  int nlo;
};

// ALM/BLM/CLM(i, L0M0) of row i and equivalent atom e (hns.F:160-259):
// 4 pi R^2 sqrt(1/Omega) i^L0 e^{2 pi i G_i.tau} conj Y_L0M0(R k_i) times
// {AL, BL, 0} (plane-wave rows) or {ALO, BLO, CLO} (the LO rows of that
// atom, l and jlo)
__device__ __forceinline__ void hns_alm(const HnsArgs &g, int i, int e, int a, cd *alm, cd *blm, cd *clm)
{
  const int n = g.n;
  const double r = g.rmt[a];
  const double rhoatm = 4.0 * PI_H * sqrt(g.vi) * r * r;
  // phase and Y_lm of the rotated k+G, hns.F:171-183
  const double phr = g.cs[2 * ((size_t)e * n + i)], phi = g.cs[2 * ((size_t)e * n + i) + 1];
  double v[3] = {g.kv[4 * i], g.kv[4 * i + 1], g.kv[4 * i + 2]}, r1[3], r2[3];
  rotate3(v, g.rotij + 9 * e, r1);
  rotate3(r1, g.rotloc + 9 * a, r2);
  cd yl[W2K_NSL2];
  ylm(r2, g.nnsl - 1, yl);
  int lm = 0;
  double cilr = 1.0, cili = 0.0;
  for (int l0 = 0; l0 < g.nnsl; l0++) {
    double fa = 0.0, fb = 0.0, fc = 0.0;
    bool on = false;
    if (i < g.nv) {
      fa = g.al[i + (size_t)n * (l0 + (size_t)g.lt * a)];
      fb = g.bl[i + (size_t)n * (l0 + (size_t)g.lt * a)];
      on = true;
    } else if (l0 <= W2K_LOMAX) {
      for (int jlo = 0; jlo < W2K_NLOAT; jlo++) {
        const int *rr = g.lorow + 2 * (l0 + W2K_LOMAXP * (jlo + W2K_NLOAT * a));
        const int qq = i - g.nv;
        if (rr[0] >= 0 && qq >= rr[0] && qq <= rr[1]) {
          const double *ab = g.loabc + 3 * (l0 + W2K_LOMAXP * (jlo + W2K_NLOAT * a));
          fa = ab[0]; fb = ab[1]; fc = ab[2];
          on = true;
        }
      }
    }
    for (int m0 = -l0; m0 <= l0; m0++, lm++) {
      if (!on) { alm[lm] = {0, 0}; blm[lm] = {0, 0}; clm[lm] = {0, 0}; continue; }
      // (CIL*PHS) * conj(YL)
      const cd cp = cmul({cilr, cili}, {phr, phi});
      const cd t = cmul(cp, conj(yl[lm]));
      const double ar = rhoatm * t.re, ai = rhoatm * t.im;
      alm[lm] = {fa * ar, fa * ai};
      blm[lm] = {fb * ar, fb * ai};
      clm[lm] = {fc * ar, fc * ai};
    }
    const double t = cilr;  // CIL = CIL * i
    cilr = -cili;
    cili = t;
  }
}

// one thread per (row i, equivalent atom e)
template <bool CPX>
__global__ __launch_bounds__(64) void k_hns_prep(HnsArgs g)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y;
  if (i >= g.n) return;
  const int a = g.atom_of[e];
  const int lidx = g.lindex[a];
  if (lidx == 0) return;
  const int n = g.n;
  cd alm[W2K_NSL2], blm[W2K_NSL2], clm[W2K_NSL2];
  hns_alm(g, i, e, a, alm, blm, clm);
  const bool lorow = (i >= g.nv);
  const int q = i - g.nv;
  const int kb = g.koff[e];
  const double *T = g.tcoef + (size_t)a * W2K_NSL2 * W2K_NSL2 * 18;
  const size_t ldL = (size_t)n, ldC = (size_t)g.nlo;
  const int KT = g.ktot;
  for (int k = 0; k < lidx; k++) {
    const int lk = g.lm0_of[a * W2K_NSL2 + k];
    double a2r = 0, a2i = 0, b2r = 0, b2i = 0, c2r = 0, c2i = 0;
    const int np = g.nnsl * g.nnsl;
    for (int p = 0; p < np; p++) {
      const double *t = T + ((size_t)k * W2K_NSL2 + p) * 18;
      const cd A = conj(alm[p]), B = conj(blm[p]), Cc = conj(clm[p]);
      // A2 += TUU A + TDU B + TUU21 C ; B2 += TUD A + TDD B + TUD21 C ; C2 += TUU12 A + TDU12 B + TUU22 C
      a2r += t[0] * A.re - t[1] * A.im + t[2] * B.re - t[3] * B.im + t[4] * Cc.re - t[5] * Cc.im;
      a2i += t[0] * A.im + t[1] * A.re + t[2] * B.im + t[3] * B.re + t[4] * Cc.im + t[5] * Cc.re;
      b2r += t[6] * A.re - t[7] * A.im + t[8] * B.re - t[9] * B.im + t[10] * Cc.re - t[11] * Cc.im;
      b2i += t[6] * A.im + t[7] * A.re + t[8] * B.im + t[9] * B.re + t[10] * Cc.im + t[11] * Cc.re;
      c2r += t[12] * A.re - t[13] * A.im + t[14] * B.re - t[15] * B.im + t[16] * Cc.re - t[17] * Cc.im;
      c2i += t[12] * A.im + t[13] * A.re + t[14] * B.im + t[15] * B.re + t[16] * Cc.im + t[17] * Cc.re;
    }
    const cd a1 = alm[lk], b1 = blm[lk], c1 = clm[lk];
    // columns: block 0 A, block 1 A (imag), block 2 B, block 3 B (imag)
    const size_t c0 = (size_t)(kb + k);
    g.L[i + ldL * (c0 + 0 * (size_t)KT)] = a2r;
    g.L[i + ldL * (c0 + 1 * (size_t)KT)] = -a2i;
    g.L[i + ldL * (c0 + 2 * (size_t)KT)] = b2r;
    g.L[i + ldL * (c0 + 3 * (size_t)KT)] = -b2i;
    g.R[i + ldL * (c0 + 0 * (size_t)KT)] = a1.re;
    g.R[i + ldL * (c0 + 1 * (size_t)KT)] = a1.im;
    g.R[i + ldL * (c0 + 2 * (size_t)KT)] = b1.re;
    g.R[i + ldL * (c0 + 3 * (size_t)KT)] = b1.im;
    if (CPX) {
      g.R2[i + ldL * (c0 + 0 * (size_t)KT)] = a1.im;
      g.R2[i + ldL * (c0 + 1 * (size_t)KT)] = -a1.re;
      g.R2[i + ldL * (c0 + 2 * (size_t)KT)] = b1.im;
      g.R2[i + ldL * (c0 + 3 * (size_t)KT)] = -b1.re;
    }
    if (lorow) {
      g.LC[q + ldC * (c0 + 0 * (size_t)KT)] = c2r;
      g.LC[q + ldC * (c0 + 1 * (size_t)KT)] = -c2i;
      g.RC[q + ldC * (c0 + 0 * (size_t)KT)] = c1.re;
      g.RC[q + ldC * (c0 + 1 * (size_t)KT)] = c1.im;
      if (CPX) {
        g.RC2[q + ldC * (c0 + 0 * (size_t)KT)] = c1.im;
        g.RC2[q + ldC * (c0 + 1 * (size_t)KT)] = -c1.re;
      }
    }
  }
}

// HS (complex interleaved, lower incl. diagonal) += Re + i Im from planar
// temporaries (only the lower triangle of the temporaries is valid)
__global__ void k_add_cplx_lower(int n, int off, int ld, const double *__restrict__ re, const double *__restrict__ im,
                                 int ldt, double *__restrict__ hs)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n || i < j) return;
  const size_t o = (size_t)(i + off) + (size_t)ld * (j + off);
  hs[2 * o] += re[i + (size_t)ldt * j];
  hs[2 * o + 1] += im[i + (size_t)ldt * j];
}

// kernel arguments of the HNS stage over the context's buffers
static HnsArgs hns_args(Ctx &c, int ktot, const int *d_koff)
{
  RunDev &R = c.run();
  const int n = c.n, nlo = c.nlo, ndf = R.ndf;
  const int cpx = R.is_complex;
  const size_t kcols = (size_t)4 * ktot;
  const size_t lcsz = (size_t)(nlo > 0 ? nlo : 1) * 2 * ktot;
  HnsArgs g;
  g.n = n; g.nv = c.nv; g.nnsl = R.nnsl; g.lt = R.ltop; g.ndf = ndf;
  g.kv = c.d_kvec.p; g.cs = c.d_cs.p; g.al = c.d_al.p; g.bl = c.d_bl.p;
  g.atom_of = R.d_atom_of.p; g.lindex = R.d_lindex.p; g.koff = d_koff; g.lm0_of = R.d_lm0_of.p;
  g.tcoef = R.d_tcoef.p; g.lorow = R.d_lorow.p; g.loabc = R.d_loabc.p;
  g.rotij = R.d_rotij.p; g.rotloc = R.d_rotloc.p; g.rmt = R.d_rmt.p; g.vi = R.vi;
  g.ktot = ktot;
  g.L = c.d_hnsL.p;
  g.R = c.d_hnsR.p;
  g.R2 = cpx ? c.d_hnsR.p + (size_t)n * kcols : nullptr;
  g.LC = c.h_small.p;
  g.RC = c.h_small.p + lcsz;
  g.RC2 = cpx ? c.h_small.p + 2 * lcsz : nullptr;
  g.nlo = nlo;
  return g;
}

int hns(Ctx &c)
{
  RunDev &R = c.run();
  if (!c.have_hamilt) return W2K_ESTATE;
  const int n = c.n, nlo = c.nlo, ndf = R.ndf;
  const int cpx = R.is_complex;
  hipStream_t s = c.stream;
  int ktot = 0;
  std::vector<int> koff(ndf);
  for (int e = 0; e < ndf; e++) {
    koff[e] = ktot;
    ktot += R.lindex[R.atom_of[e]];
  }
  if (ktot == 0) return 0;
  int rc;
  const size_t kcols = (size_t)4 * ktot;
  if ((rc = c.d_hnsL.alloc((size_t)n * kcols))) return rc;
  if ((rc = c.d_hnsR.alloc((size_t)n * kcols * (cpx ? 2 : 1)))) return rc;
  const size_t lcsz = (size_t)(nlo > 0 ? nlo : 1) * 2 * ktot;
  if ((rc = c.h_small.alloc(lcsz * (cpx ? 3 : 2)))) return rc;
  int *d_koff;
  if ((rc = c.h_int.alloc(ndf))) return rc;
  d_koff = c.h_int.p;
  W2K_HIP(hipMemcpyAsync(d_koff, koff.data(), sizeof(int) * ndf, hipMemcpyHostToDevice, s));
  W2K_HIP(hipStreamSynchronize(s));  // koff is a host temporary
  HnsArgs g = hns_args(c, ktot, d_koff);
  // atoms without non-spherical terms leave their columns untouched: clear
  W2K_HIP(hipMemsetAsync(c.d_hnsL.p, 0, sizeof(double) * (size_t)n * kcols, s));
  W2K_HIP(hipMemsetAsync(c.d_hnsR.p, 0, sizeof(double) * (size_t)n * kcols * (cpx ? 2 : 1), s));
  W2K_HIP(hipMemsetAsync(c.h_small.p, 0, sizeof(double) * lcsz * (cpx ? 3 : 2), s));
  dim3 grid((n + 63) / 64, ndf);
  if (cpx) hipLaunchKernelGGL(k_hns_prep<true>, grid, dim3(64), 0, s, g);
  else hipLaunchKernelGGL(k_hns_prep<false>, grid, dim3(64), 0, s, g);
  W2K_STAGE(s, "k_hns_prep");
  const int K4 = (int)kcols, K2 = 2 * ktot;
  const int sl = c.hns_slices;
  if (sl > 0 && (rc = c.w_dig.alloc(split_gemm_ws_bytes(n, n, K4, sl)))) return rc;
  // the rank-K product L R^T (lower triangle): FP64 MFMA, or INT8 digit
  // slices with FP64 recombination in split precision
  auto rank_k = [&](const double *Rop, double *C, int accumulate) -> int {
    if (sl > 0) return split_gemm_nt(s, n, n, K4, g.L, n, Rop, n, C, n, 1, accumulate, sl, c.w_dig.p);
    return dgemm(s, 'N', 'T', n, n, K4, 1.0, g.L, n, Rop, n, accumulate ? 1.0 : 0.0, C, n, 1);
  };
  if (!cpx) {
    W2K_HIP(hipEventRecord(c.ev[14], s));
    if ((rc = rank_k(g.R, c.d_hs.p, 1))) return rc;
    W2K_HIP(hipEventRecord(c.ev[15], s));
    if (nlo > 0) {
      double *hsl = c.d_hs.p + (size_t)c.nv + (size_t)n * c.nv;
      if ((rc = dgemm(s, 'N', 'T', nlo, nlo, K2, 1.0, g.LC, nlo, g.RC, nlo, 1.0, hsl, n, 1))) return rc;
    }
  } else {
    // planar temporaries for the real and imaginary parts
    if ((rc = c.h_re.alloc((size_t)n * n))) return rc;
    if ((rc = c.h_im.alloc((size_t)n * n))) return rc;
    W2K_HIP(hipEventRecord(c.ev[14], s));
    if ((rc = rank_k(g.R, c.h_re.p, 0))) return rc;
    if ((rc = rank_k(g.R2, c.h_im.p, 0))) return rc;
    W2K_HIP(hipEventRecord(c.ev[15], s));
    hipLaunchKernelGGL(k_add_cplx_lower, dim3((n + 255) / 256, n), dim3(256), 0, s, n, 0, n, c.h_re.p, c.h_im.p, n,
                       c.d_hs.p);
    W2K_STAGE(s, "k_add_cplx_lower");
    if (nlo > 0) {
      if ((rc = dgemm(s, 'N', 'T', nlo, nlo, K2, 1.0, g.LC, nlo, g.RC, nlo, 0.0, c.h_re.p, nlo, 1))) return rc;
      if ((rc = dgemm(s, 'N', 'T', nlo, nlo, K2, 1.0, g.LC, nlo, g.RC2, nlo, 0.0, c.h_im.p, nlo, 1))) return rc;
      hipLaunchKernelGGL(k_add_cplx_lower, dim3((nlo + 255) / 256, nlo), dim3(256), 0, s, nlo, c.nv, n, c.h_re.p,
                         c.h_im.p, nlo, c.d_hs.p);
      W2K_STAGE(s, "k_add_cplx_lower");
    }
  }
  c.have_hns_ops = 1;
  return 0;
}

// ------------------------------------------------------------------ adjoint
// Band-energy adjoint of the HNS update with respect to the VNS radial
// integrals (an extended form with no counterpart in the reference).  The
// stored lower triangle gains L(I,J) += sum_k A2(I,k) A1(J,k) + B2 B1 (+ C2 C1
// on the LO block), with A2(I,k) = sum_p T_c(k,p) conj XLM_c(I,p) (hns.F:266-
// 434); the real build keeps the real part.  With E = Re sum_{I>=J} w Gamma L
// (w = 2 off the diagonal, 1 on it) this gives
//   E = Re sum_{c,k,p} T_c(k,p) X_c(k,p),
//   X_c(k,p) = sum_I conj XLM_x(I,p) M_m(I,k),  M_m = (w Gamma)_lower Y1_m,
// for the coefficient c = 3 m + x (x: ALM/BLM/CLM, m: A1/B1/C1), and every
// coupling term T_c += GFAC * VNS adds Re(GFAC X_c) to dE/dVNS.

// XLM tables: xlm[((e * 3 + x) * np + p) * n + i], complex interleaved
__global__ __launch_bounds__(64) void k_alm_store(HnsArgs g, int np, double *__restrict__ xlm)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y;
  if (i >= g.n) return;
  const int a = g.atom_of[e];
  cd t[3][W2K_NSL2];
  hns_alm(g, i, e, a, t[0], t[1], t[2]);
  const size_t n = (size_t)g.n;
  for (int x = 0; x < 3; x++)
    for (int p = 0; p < np; p++) {
      const size_t o = (((size_t)e * 3 + x) * np + p) * n + i;
      xlm[2 * o] = t[x][p].re;
      xlm[2 * o + 1] = t[x][p].im;
    }
}

// w Gamma on the lower triangle: off-diagonal doubled, the diagonal's
// imaginary part dropped (the solver reads the real part), upper zeroed
__global__ void k_gamma_weight(int n, double *__restrict__ gr, double *__restrict__ gi)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  const size_t o = (size_t)i + (size_t)n * j;
  if (i > j) {
    gr[o] *= 2.0;
    if (gi) gi[o] *= 2.0;
  } else {
    if (i < j) gr[o] = 0.0;
    if (gi) gi[o] = 0.0;
  }
}

struct AdjXArgs {
  int n, nv, nlo, np, ktot;
  const int *atom_of, *lindex, *koff;
  const double *xlm;              // XLM tables
  const double *p1, *p2;          // (w Gamma_r) R, (w Gamma_i) R (p2 complex build only), n x 4 ktot
  const double *pc1, *pc2;        // the same for the LO block, nlo x 2 ktot
  int cpx;
  double *x;                      // X[e][m][x][k][p], complex interleaved
};

// M_m(I, k) of atom e (column c0 = koff(e) + k), complex
__device__ __forceinline__ cd adj_m(const AdjXArgs &g, int m, int I, int c0)
{
  const int KT = g.ktot;
  if (m < 2) {
    const double *p1 = g.p1 + (size_t)g.n * (c0 + 2 * m * KT);
    const double *p1i = g.p1 + (size_t)g.n * (c0 + (2 * m + 1) * KT);
    if (!g.cpx) return {p1[I], p1i[I]};
    const double *p2 = g.p2 + (size_t)g.n * (c0 + 2 * m * KT);
    const double *p2i = g.p2 + (size_t)g.n * (c0 + (2 * m + 1) * KT);
    return {p1[I] - p2i[I], p1i[I] + p2[I]};
  }
  if (I < g.nv) return {0.0, 0.0};
  const int q = I - g.nv;
  const double *c1 = g.pc1 + (size_t)g.nlo * c0, *c1i = g.pc1 + (size_t)g.nlo * (c0 + KT);
  if (!g.cpx) return {c1[q], c1i[q]};
  const double *c2 = g.pc2 + (size_t)g.nlo * c0, *c2i = g.pc2 + (size_t)g.nlo * (c0 + KT);
  return {c1[q] - c2i[q], c1i[q] + c2[q]};
}

// one block per (atom e, m, x): X(k,p) = sum_I conj XLM_x(I,p) M_m(I,k),
// rows in chunks of 32 through LDS, each sum in row order
__global__ __launch_bounds__(256) void k_adj_x(AdjXArgs g)
{
  const int e = blockIdx.x, m = blockIdx.y / 3, x = blockIdx.y % 3;
  const int t = threadIdx.x;
  const int a = g.atom_of[e];
  const int lk = g.lindex[a], np = g.np;
  double *out = g.x + 2 * (((size_t)e * 9 + m * 3 + x) * W2K_NSL2 * W2K_NSL2);
  __shared__ double xr[W2K_NSL2][33], xi[W2K_NSL2][33], mr[W2K_NSL2][33], mi[W2K_NSL2][33];
  double accr[3] = {0, 0, 0}, acci[3] = {0, 0, 0};
  const int i0 = (m == 2 || x == 2) ? g.nv : 0;  // CLM and M_C vanish on plane-wave rows
  const int c0 = g.koff[e];
  const double *xb = g.xlm + 2 * (((size_t)e * 3 + x) * np) * g.n;
  for (int I0 = i0; I0 < g.n; I0 += 32) {
    __syncthreads();
    for (int q = t; q < 32 * W2K_NSL2; q += 256) {
      const int r = q & 31, col = q >> 5, I = I0 + r;
      double ar = 0, ai = 0, br = 0, bi = 0;
      if (I < g.n) {
        if (col < np) {
          ar = xb[2 * ((size_t)col * g.n + I)];
          ai = xb[2 * ((size_t)col * g.n + I) + 1];
        }
        if (col < lk) {
          const cd v = adj_m(g, m, I, c0 + col);
          br = v.re;
          bi = v.im;
        }
      }
      xr[col][r] = ar; xi[col][r] = ai;
      mr[col][r] = br; mi[col][r] = bi;
    }
    __syncthreads();
    for (int u = 0; u < 3; u++) {
      const int pair = t + 256 * u;
      if (pair >= lk * np) continue;
      const int k = pair / np, p = pair % np;
      double sr = accr[u], si = acci[u];
      for (int r = 0; r < 32; r++) {
        // conj(X) * M
        sr += xr[p][r] * mr[k][r] + xi[p][r] * mi[k][r];
        si += xr[p][r] * mi[k][r] - xi[p][r] * mr[k][r];
      }
      accr[u] = sr;
      acci[u] = si;
    }
  }
  for (int u = 0; u < 3; u++) {
    const int pair = t + 256 * u;
    if (pair >= lk * np) continue;
    const int k = pair / np, p = pair % np;
    out[2 * (k * W2K_NSL2 + p)] = accr[u];
    out[2 * (k * W2K_NSL2 + p) + 1] = acci[u];
  }
}

int adjoint_vns(Ctx &c, const double *f_host, double *const gout[6])
{
  RunDev &R = c.run();
  const int n = c.n, nv = c.nv, nlo = c.nlo, ne = c.ne, ndf = R.ndf, nat = R.nat;
  const int cpx = R.is_complex;
  hipStream_t s = c.stream;
  const int LOP2 = W2K_LOMAX + 2;
  const size_t nvns = (size_t)W2K_NSLMAX * R.lmmxd * W2K_NSLMAX * nat;
  const size_t nvlo = (size_t)LOP2 * R.lmmxd * W2K_NSLMAX * nat;
  for (int q = 0; q < 6; q++)
    if (gout[q]) std::fill(gout[q], gout[q] + (q < 3 ? nvns : nvlo), 0.0);
  int ktot = 0;
  std::vector<int> koff(ndf);
  for (int e = 0; e < ndf; e++) {
    koff[e] = ktot;
    ktot += R.lindex[R.atom_of[e]];
  }
  if (ktot == 0 || ne == 0) return 0;
  if (!c.have_hns_ops) return W2K_ESTATE;
  int rc;
  const int np = R.nnsl * R.nnsl;
  const size_t kcols = (size_t)4 * ktot;
  // every buffer is sized before the first launch (w_x serves both the
  // planar eigenvectors of Gamma and, afterwards, the XLM tables)
  const size_t gsz = (size_t)n * n;
  const size_t xlmsz = (size_t)2 * ndf * 3 * np * n;
  if ((rc = c.w_a.alloc(gsz)) || (cpx && (rc = c.w_b.alloc(gsz)))) return rc;
  if ((rc = c.w_x.alloc(std::max(xlmsz, (size_t)4 * n * ne)))) return rc;
  double *Gr = c.w_a.p, *Gi = cpx ? c.w_b.p : nullptr;
  if ((rc = c.w_v.alloc((size_t)ne))) return rc;
  W2K_HIP(hipMemcpyAsync(c.w_v.p, f_host, sizeof(double) * ne, hipMemcpyHostToDevice, s));
  if ((rc = adjoint_gamma(c, c.w_v.p, n, Gr, Gi))) return rc;
  hipLaunchKernelGGL(k_gamma_weight, dim3((n + 255) / 256, n), dim3(256), 0, s, n, Gr, Gi);
  W2K_STAGE(s, "k_gamma_weight");
  // M = (w Gamma) [A1 B1] (and the LO block with C1)
  const size_t psz = (size_t)n * kcols, pcsz = (size_t)(nlo > 0 ? nlo : 1) * 2 * ktot;
  if ((rc = c.w_c.alloc((cpx ? 2 : 1) * (psz + pcsz)))) return rc;
  double *P1 = c.w_c.p, *P2 = cpx ? P1 + psz : nullptr;
  double *PC1 = P1 + (cpx ? 2 : 1) * psz, *PC2 = cpx ? PC1 + pcsz : nullptr;
  int *d_koff;
  if ((rc = c.w_int.alloc(ndf))) return rc;
  d_koff = c.w_int.p;
  W2K_HIP(hipMemcpyAsync(d_koff, koff.data(), sizeof(int) * ndf, hipMemcpyHostToDevice, s));
  W2K_HIP(hipStreamSynchronize(s));  // koff is a host temporary
  HnsArgs g = hns_args(c, ktot, d_koff);
  if ((rc = dgemm(s, 'N', 'N', n, (int)kcols, n, 1.0, Gr, n, g.R, n, 0.0, P1, n, 0, 2, 0))) return rc;
  if (cpx && (rc = dgemm(s, 'N', 'N', n, (int)kcols, n, 1.0, Gi, n, g.R, n, 0.0, P2, n, 0, 2, 0))) return rc;
  if (nlo > 0) {
    const size_t olo = (size_t)nv + (size_t)n * nv;
    if ((rc = dgemm(s, 'N', 'N', nlo, 2 * ktot, nlo, 1.0, Gr + olo, n, g.RC, nlo, 0.0, PC1, nlo, 0, 2, 0))) return rc;
    if (cpx && (rc = dgemm(s, 'N', 'N', nlo, 2 * ktot, nlo, 1.0, Gi + olo, n, g.RC, nlo, 0.0, PC2, nlo, 0, 2, 0)))
      return rc;
  }
  // XLM tables of every atom (w_x is free again once the Gamma GEMMs ran:
  // same stream)
  hipLaunchKernelGGL(k_alm_store, dim3((n + 63) / 64, ndf), dim3(64), 0, s, g, np, c.w_x.p);
  W2K_STAGE(s, "k_alm_store");
  // X per (atom, m, x)
  const size_t xsz = (size_t)2 * ndf * 9 * W2K_NSL2 * W2K_NSL2;
  if ((rc = c.w_fac.alloc(xsz))) return rc;
  W2K_HIP(hipMemsetAsync(c.w_fac.p, 0, sizeof(double) * xsz, s));
  AdjXArgs xa;
  xa.n = n; xa.nv = nv; xa.nlo = nlo; xa.np = np; xa.ktot = ktot;
  xa.atom_of = R.d_atom_of.p; xa.lindex = R.d_lindex.p; xa.koff = d_koff;
  xa.xlm = c.w_x.p; xa.p1 = P1; xa.p2 = P2; xa.pc1 = PC1; xa.pc2 = PC2; xa.cpx = cpx; xa.x = c.w_fac.p;
  hipLaunchKernelGGL(k_adj_x, dim3(ndf, 9), dim3(256), 0, s, xa);
  W2K_STAGE(s, "k_adj_x");
  std::vector<double> hx(xsz);
  W2K_HIP(hipMemcpyAsync(hx.data(), c.w_fac.p, sizeof(double) * xsz, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  // X per atom type (equivalent atoms in order), then the coupling terms
  const size_t xe = (size_t)2 * 9 * W2K_NSL2 * W2K_NSL2;
  std::vector<double> xa_t((size_t)nat * xe, 0.0);
  for (int e = 0; e < ndf; e++) {
    const int a = R.atom_of[e];
    for (size_t q = 0; q < xe; q++) xa_t[a * xe + q] += hx[e * xe + q];
  }
  for (const RunDev::HnsTerm &h : R.hns_terms) {
    double *gdst = gout[h.arr];
    if (!gdst) continue;
    const int m = h.c / 3, x = h.c % 3;
    const double *X = &xa_t[h.a * xe + 2 * (((size_t)m * 3 + x) * W2K_NSL2 * W2K_NSL2 + (size_t)h.k * W2K_NSL2 + h.p)];
    const size_t lead = h.arr < 3 ? W2K_NSLMAX : LOP2;
    gdst[h.x + lead * (h.lmx + (size_t)R.lmmxd * (h.y + W2K_NSLMAX * (size_t)h.a))] += h.gr * X[0] - h.gi * X[1];
  }
  return 0;
}

}  // namespace w2k
