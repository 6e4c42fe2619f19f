// HAMILT on the GPU (hamilt.F:142-818): spherical muffin-tin H/S, warped
// interstitial potential and the local-orbital columns, written into HS/HSDIAG
// in the reference layout (S in the upper triangle, H in the strict lower
// triangle, diag(H) in HSDIAG).
//
// Kernels (one launch each per k-point):
//   k_cs         cos/sin(2 pi KZZ_n . tau_e), N x NDF   (all phase factors)
//   k_albl       a_l, b_l (hamilt.F:160-175)           (spherical Bessel)
//   k_colcoef    per column i, atom a, l: the four radial coefficients of
//                loops 230 (hamilt.F:386-424) times (2l+1)*VIFPR4
//   k_hamilt_pw  the pair loop 260.  One wave = 64 rows x C columns.  The
//                Legendre polynomials of its pairs stay in registers, the
//                per-column coefficients are wave-uniform (scalar loads), the
//                per-row tables (AL, BL, cos/sin) are coalesced vector loads.
//                The step function Theta(G_J-G_I) (loops 160-180) is summed per
//                group of equal sphere radii, so the pair loop evaluates one
//                sincos per distinct RMT rather than one per atom.
//   k_hamilt_lo  the LO columns (hamilt.F:475-815), one thread per (row, LO column)
#include "special.h"
#include "w2k_internal.h"

namespace w2k {

constexpr double PI = 3.141592653589793238462643;

// ------------------------------------------------------------------ tables
__global__ void k_cs(int n, const int *__restrict__ kzz, const double *__restrict__ pos, double *__restrict__ cs)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int e = blockIdx.y;
  if (i >= n) return;
  const double *p = pos + 3 * e;
  const double arg = p[0] * kzz[3 * i] + p[1] * kzz[3 * i + 1] + p[2] * kzz[3 * i + 2];
  double s, c;
  sincos(2.0 * PI * arg, &s, &c);
  cs[2 * ((size_t)e * n + i)] = c;
  cs[2 * ((size_t)e * n + i) + 1] = s;
}

// AL/BL (n, l, a) column-major; rows >= nv (LO vectors) are zero
__global__ void k_albl(int n, int nv, int lt, const double *__restrict__ kv, const double *__restrict__ rmt,
                       const double *__restrict__ rad, const int *__restrict__ lapw, double *__restrict__ al,
                       double *__restrict__ bl)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (i >= n) return;
  double fj[W2K_LMAX + 1], dfj[W2K_LMAX + 1];
  const double rkn = kv[4 * i + 3];
  const double r = rmt[a];
  if (i < nv) {
    sphbes(lt - 1, r * rkn, fj);
    dvbes1(fj, dfj, rkn, r, lt);
  }
  const double *U = rad + (size_t)a * 6 * W2K_LMAX;
  for (int l = 0; l < lt; l++) {
    double va = 0.0, vb = 0.0;
    if (i < nv) {
      const double u = U[l], up = U[W2K_LMAX + l], du = U[2 * W2K_LMAX + l], dup = U[3 * W2K_LMAX + l];
      if (lapw[a * W2K_LMAX + l]) {
        va = rkn * dfj[l] * up - fj[l] * dup;
        vb = fj[l] * du - rkn * dfj[l] * u;
      } else {
        va = fj[l] / u / (r * r);
      }
    }
    const size_t o = i + (size_t)n * (l + (size_t)lt * a);
    al[o] = va;
    bl[o] = vb;
  }
}

// coef[((a*lt + l)*nv + i)*4 + {0..3}] = (2l+1) VIFPR4_a * {s_a, s_b, h_a, h_b}
// S += P_l (AL_j s_a + BL_j s_b), H += P_l (AL_j h_a + BL_j h_b) reproduces
// loops j2/j3 of hamilt.F:386-424 (APW: BL = 0).
__global__ void k_colcoef(int n, int nv, int lt, double vi, const double *__restrict__ rmt,
                          const double *__restrict__ rad, const int *__restrict__ lapw,
                          const double *__restrict__ al, const double *__restrict__ bl, double *__restrict__ coef)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (i >= nv) return;
  const double r = rmt[a];
  const double r2 = r * r;
  const double vifpr4 = 4.0 * PI * vi * (r2 * r2);
  const double *U = rad + (size_t)a * 6 * W2K_LMAX;
  for (int l = 0; l < lt; l++) {
    const double u = U[l], up = U[W2K_LMAX + l], du = U[2 * W2K_LMAX + l], dup = U[3 * W2K_LMAX + l];
    const double chn = U[4 * W2K_LMAX + l], e = U[5 * W2K_LMAX + l];
    const double ai = al[i + (size_t)n * (l + (size_t)lt * a)];
    const double bi = bl[i + (size_t)n * (l + (size_t)lt * a)];
    const double f = (double)(2 * l + 1) * vifpr4;
    double sa, sb, ha, hb;
    if (!lapw[a * W2K_LMAX + l]) {
      const double atmp1 = ai * u * du * r2;
      sa = ai;
      sb = 0.0;
      ha = e * ai + atmp1;
      hb = 0.0;
    } else {
      const double btmp = bi * chn;
      const double atmp1 = r2 * (ai * u * du + bi * up * du);
      const double btmp1 = r2 * (bi * up * dup + ai * u * dup);
      sa = ai;
      sb = btmp;
      ha = e * ai + atmp1;
      hb = ai + e * btmp + btmp1;
    }
    double *c = coef + (((size_t)a * lt + l) * nv + i) * 4;
    c[0] = f * sa;
    c[1] = f * sb;
    c[2] = f * ha;
    c[3] = f * hb;
  }
}

// ------------------------------------------------------------------ pair loop
struct PwArgs {
  int n, nv, nat, lt;
  const int *kzz;          // [n][3]
  const double *kv;        // [n][4]
  const double *cs;        // [ndf][n][2]
  const double *al, *bl;   // (n, lt, nat)
  const double *coef;      // [nat][lt][nv][4]
  const int *aorder;       // atom types sorted by radius group
  const int *grp_end;      // [ngrp]: end index (exclusive) in aorder of each group
  const double *grp_r;     // [ngrp]: radius of the group
  int ngrp;
  const int *mult, *first_eq;
  const double *v3;        // 3*V_a
  double theta0;           // Theta(0) = 1 - sum_a V_a MULT_a
  const double *warp;      // complex box
  int w1, w2, w3;
  double *hs, *hsdiag;
};

template <int LT, int C, bool CPX>
__global__ __launch_bounds__(256) void k_hamilt_pw(PwArgs g)
{
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i0 = (blockIdx.y * 4 + wave) * C;
  const int jb = blockIdx.x * 64;
  if (i0 >= g.nv || jb > i0 + C - 1) return;  // wave-uniform
  const int j = jb + lane;
  const bool rowok = j < g.nv;
  const int jj = rowok ? j : 0;
  const int n = g.n;

  const double xj = g.kv[4 * jj], yj = g.kv[4 * jj + 1], zj = g.kv[4 * jj + 2], rj = g.kv[4 * jj + 3];
  const int kj0 = g.kzz[3 * jj], kj1 = g.kzz[3 * jj + 1], kj2 = g.kzz[3 * jj + 2];

  double P[C][LT];
  double x2[C], slen[C];
  double Sr[C], Si[C], Hr[C], Hi[C], thr[C], thi[C];
  bool ok[C];
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int i = min(i0 + c, g.nv - 1);
    ok[c] = rowok && (i0 + c < g.nv) && (j <= i0 + c);
    const double xi = g.kv[4 * i], yi = g.kv[4 * i + 1], zi = g.kv[4 * i + 2], ri = g.kv[4 * i + 3];
    // Legendre polynomials, hamilt.F:250-281
    double x = xi * xj + yi * yj + zi * zj;
    const double den = ri * rj;
    if (den > 1.0e-8) x = x / den;
    double pplx = x * x, pply = pplx - 1.0;
    P[c][0] = 1.0;
    P[c][1] = x;
    const bool low = fabs(x) < 0.9999999;
#pragma unroll
    for (int k = 2; k < LT; k++) {
      if (low) {
        const double xl = (double)(k - 1) / (double)k;
        P[c][k] = pplx + xl * pply;
        pplx = x * P[c][k];
        pply = pplx - P[c][k - 1];
      } else {
        P[c][k] = x * P[c][k - 1];
      }
    }
    x2[c] = xj * xi + yj * yi + zj * zi;
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    slen[c] = sqrt(dx * dx + dy * dy + dz * dz);
    // warped interstitial potential WARP(KZZ_J - KZZ_I), warpin.f:31-37
    const int q1 = kj0 - g.kzz[3 * i], q2 = kj1 - g.kzz[3 * i + 1], q3 = kj2 - g.kzz[3 * i + 2];
    double wr = 0.0, wi = 0.0;
    if (abs(q1) <= g.w1 && abs(q2) <= g.w2 && abs(q3) <= g.w3) {
      const long b = (q1 + g.w1) + (long)(2 * g.w1 + 1) * ((q2 + g.w2) + (long)(2 * g.w2 + 1) * (q3 + g.w3));
      wr = g.warp[2 * b];
      wi = g.warp[2 * b + 1];
    }
    Sr[c] = 0.0; Si[c] = 0.0;
    Hr[c] = wr;
    Hi[c] = CPX ? wi : 0.0;
    thr[c] = 0.0; thi[c] = 0.0;
  }

  int pa = 0;
  for (int gi = 0; gi < g.ngrp; gi++) {
    const int pend = g.grp_end[gi];
    double pr[C], pim[C];
#pragma unroll
    for (int c = 0; c < C; c++) { pr[c] = 0.0; pim[c] = 0.0; }
    for (; pa < pend; pa++) {
      const int a = g.aorder[pa];
      // phase factors PHASE(J,a) = sum_eq exp(2 pi i (G_J - G_I).tau), hamilt.F:199-243
      double phr[C], phi[C];
#pragma unroll
      for (int c = 0; c < C; c++) { phr[c] = 0.0; phi[c] = 0.0; }
      const int e0 = g.first_eq[a], me = g.mult[a];
      for (int e = e0; e < e0 + me; e++) {
        const double cj = g.cs[2 * ((size_t)e * n + jj)], sj = g.cs[2 * ((size_t)e * n + jj) + 1];
#pragma unroll
        for (int c = 0; c < C; c++) {
          const int i = min(i0 + c, g.nv - 1);
          const double ci = g.cs[2 * ((size_t)e * n + i)], si = g.cs[2 * ((size_t)e * n + i) + 1];
          phr[c] += cj * ci + sj * si;
          if (CPX) phi[c] += sj * ci - cj * si;
        }
      }
      const double v3 = g.v3[a];
#pragma unroll
      for (int c = 0; c < C; c++) { pr[c] += v3 * phr[c]; if (CPX) pim[c] += v3 * phi[c]; }
      // muffin-tin radial sums, loops 230/240
      double t1[C], t2[C];
#pragma unroll
      for (int c = 0; c < C; c++) { t1[c] = 0.0; t2[c] = 0.0; }
      const double *alr = g.al + (size_t)n * g.lt * a;
      const double *blr = g.bl + (size_t)n * g.lt * a;
      const double *cf = g.coef + (size_t)a * g.lt * g.nv * 4;
#pragma unroll
      for (int l = 0; l < LT; l++) {
        const double aj = alr[(size_t)n * l + jj], bj = blr[(size_t)n * l + jj];
        const double *cl = cf + ((size_t)l * g.nv + i0) * 4;
#pragma unroll
        for (int c = 0; c < C; c++) {
          const int cc = (i0 + c < g.nv) ? c : 0;
          const double sa = cl[4 * cc], sb = cl[4 * cc + 1], ha = cl[4 * cc + 2], hb = cl[4 * cc + 3];
          t2[c] += P[c][l] * (aj * sa + bj * sb);
          t1[c] += P[c][l] * (aj * ha + bj * hb);
        }
      }
#pragma unroll
      for (int c = 0; c < C; c++) {
        Sr[c] += phr[c] * t2[c];
        Hr[c] += phr[c] * t1[c];
        if (CPX) { Si[c] += phi[c] * t2[c]; Hi[c] += phi[c] * t1[c]; }
      }
    }
    // step function contribution of this radius group, loop 170
    const double rg = g.grp_r[gi];
#pragma unroll
    for (int c = 0; c < C; c++) {
      const double h1 = slen[c] * rg;
      double s, co;
      sincos(h1, &s, &co);
      const double besr = (h1 * co - s) / (h1 * h1 * h1);
      thr[c] += pr[c] * besr;
      if (CPX) thi[c] += pim[c] * besr;
    }
  }
  // interstitial: S = Theta, H = X2 Theta + WARP (loops 160-190); store
#pragma unroll
  for (int c = 0; c < C; c++) {
    if (!ok[c]) continue;
    const int i = i0 + c;
    double tr = thr[c], ti = thi[c];
    if (i == j) { tr = g.theta0; ti = 0.0; }
    const double sr = Sr[c] + tr, si = Si[c] + ti;
    const double hr = Hr[c] + x2[c] * tr, hi = Hi[c] + x2[c] * ti;
    if (CPX) {
      g.hs[2 * ((size_t)j + (size_t)n * i)] = sr;
      g.hs[2 * ((size_t)j + (size_t)n * i) + 1] = si;
      if (j < i) {
        g.hs[2 * ((size_t)i + (size_t)n * j)] = hr;
        g.hs[2 * ((size_t)i + (size_t)n * j) + 1] = hi;
      } else {
        g.hsdiag[2 * i] = hr;
        g.hsdiag[2 * i + 1] = hi;
      }
    } else {
      g.hs[(size_t)j + (size_t)n * i] = sr;
      if (j < i) g.hs[(size_t)i + (size_t)n * j] = hr;
      else g.hsdiag[i] = hr;
    }
  }
}

// ------------------------------------------------------------------ LO columns
struct LoArgs {
  int n, nv, nlo, lt;
  const int *kzz;
  const double *kv;
  const double *cs;
  const double *al, *bl;
  const int *locol;        // [nlo][8]: a, l, jlo, jeqo(1-based), mo1, nnlo(1-based), first_eq, mult
  const double *lopar;     // [nlo][8]: c1r,c2r,c3r,c11r,c12r,c13r,akinlor,pi4r2v
  const double *loabc;     // [nat][LOMAXP][NLOAT][3]: alo, blo, clo
  const double *rad;       // radial params (u, up)
  const double *rotij, *rotloc;
  double *hs, *hsdiag;
};

template <bool CPX>
__global__ __launch_bounds__(256) void k_hamilt_lo(LoArgs g)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int q = blockIdx.y;  // LO column number 0..nlo-1
  const int *lc = g.locol + 8 * q;
  const int a = lc[0], l = lc[1], jlo = lc[2], jeqo = lc[3], mo1 = lc[4], nnlo = lc[5], e0 = lc[6], mu = lc[7];
  const int col = g.nv + nnlo - 1;
  if (j > col) return;
  const int n = g.n;
  const double *lp = g.lopar + 8 * q;
  const double c1r = lp[0], c2r = lp[1], c3r = lp[2], c11r = lp[3], c12r = lp[4], c13r = lp[5];
  const double akinlor = lp[6], pi4r2v = lp[7];
  const double u = g.rad[(size_t)a * 6 * W2K_LMAX + l], up = g.rad[(size_t)a * 6 * W2K_LMAX + W2K_LMAX + l];
  double sr = 0.0, si = 0.0, hr = 0.0, hi = 0.0;
  // LO-LO row ranges (hamilt.F:678-680), Fortran indices with NV+NNLO the column
  const int NVNNLO = g.nv + nnlo;
  const int J = j + 1;
  cd yc[7], yj[7];
  cd ybuf[16];
  for (int jq = 0; jq < mu; jq++) {
    const int e = e0 + jq;
    // phase PHSC(J) = conj(exp(2 pi i (KZZ_col - KZZ_J).tau_e))
    const double cc = g.cs[2 * ((size_t)e * n + col)], sc = g.cs[2 * ((size_t)e * n + col) + 1];
    const double cj = g.cs[2 * ((size_t)e * n + j)], sj = g.cs[2 * ((size_t)e * n + j) + 1];
    const double phr = cc * cj + sc * sj, phi = -(sc * cj - cc * sj);
    // Y_lm of the LO vector and of the row vector, rotated (ROTIJ then ROTLOC)
    double v[3], r1[3], r2[3];
    v[0] = g.kv[4 * col]; v[1] = g.kv[4 * col + 1]; v[2] = g.kv[4 * col + 2];
    rotate3(v, g.rotij + 9 * e, r1);
    rotate3(r1, g.rotloc + 9 * a, r2);
    ylm(r2, l, ybuf);
    for (int m = 0; m <= 2 * l; m++) yc[m] = ybuf[l * l + m];
    v[0] = g.kv[4 * j]; v[1] = g.kv[4 * j + 1]; v[2] = g.kv[4 * j + 2];
    rotate3(v, g.rotij + 9 * e, r1);
    rotate3(r1, g.rotloc + 9 * a, r2);
    ylm(r2, l, ybuf);
    for (int m = 0; m <= 2 * l; m++) yj[m] = ybuf[l * l + m];
    for (int mo2 = -l; mo2 <= l; mo2++) {
      const cd yn = yc[mo2 + l];
      const double clylnr = yn.re, clylni = -yn.im;
      const double c1rup = c1r * clylnr * pi4r2v, c1iup = c1r * clylni * pi4r2v;
      const double c2rup = c2r * clylnr * pi4r2v, c2iup = c2r * clylni * pi4r2v;
      const double c3rup = c3r * clylnr * pi4r2v, c3iup = c3r * clylni * pi4r2v;
      const double c11rup = c11r * clylnr * pi4r2v, c11iup = c11r * clylni * pi4r2v;
      const double c12rup = c12r * clylnr * pi4r2v, c12iup = c12r * clylni * pi4r2v;
      const double c13rup = c13r * clylnr * pi4r2v, c13iup = c13r * clylni * pi4r2v;
      const double akr = akinlor * clylnr * pi4r2v, aki = akinlor * clylni * pi4r2v;
      const cd y = yj[mo2 + l];
      const double clylr = y.re, clyli = -y.im;
      if (j < g.nv) {
        const double alj = g.al[j + (size_t)n * (l + (size_t)g.lt * a)];
        const double blj = g.bl[j + (size_t)n * (l + (size_t)g.lt * a)];
        const double c6r = clylr * alj * pi4r2v, c6i = clyli * alj * pi4r2v;
        const double c7r = clylr * blj * pi4r2v, c7i = clyli * blj * pi4r2v;
        double cabr = c6r * c1rup + c7r * c2rup + c6i * c1iup + c7i * c2iup;
        double cabi = -c6i * c1rup - c7i * c2rup + c6r * c1iup + c7r * c2iup;
        sr += cabr * phr - cabi * phi;
        if (CPX) si += cabr * phi + cabi * phr;
        cabr = c6r * c11rup + c7r * c12rup + c6i * c11iup + c7i * c12iup;
        cabi = -c6i * c11rup - c7i * c12rup + c6r * c11iup + c7r * c12iup;
        const double ta = c6r * u + c7r * up, tb = c6i * u + c7i * up;
        const double akinbr = ta * akr + tb * aki;
        const double akinbi = ta * aki - tb * akr;
        hr += cabr * phr - cabi * phi;
        if (CPX) hi += cabr * phi + cabi * phr;
        hr += akinbr * phr - akinbi * phi;
        if (CPX) hi += akinbr * phi + akinbi * phr;
      }
      for (int jlop = 1; jlop <= jlo; jlop++) {
        const int d = jlo - jlop;
        const int jst = NVNNLO - (mo1 + l) - (2 * l + 1) * (jeqo - 1) - d * (2 * l + 1) * mu;
        const int jen = NVNNLO - d * (mo1 + l + 1) - d * (jeqo - 1) * (2 * l + 1);
        if (J < jst || J > jen) continue;
        const double *ab = g.loabc + 3 * (l + W2K_LOMAXP * ((jlop - 1) + W2K_NLOAT * a));
        const double c6r = clylr * ab[0] * pi4r2v, c6i = clyli * ab[0] * pi4r2v;
        const double c7r = clylr * ab[1] * pi4r2v, c7i = clyli * ab[1] * pi4r2v;
        const double c8r = clylr * ab[2] * pi4r2v, c8i = clyli * ab[2] * pi4r2v;
        double cr = c6r * c1rup + c6i * c1iup + c7r * c2rup + c7i * c2iup + c8r * c3rup + c8i * c3iup;
        double ci = -c6i * c1rup + c6r * c1iup - c7i * c2rup + c7r * c2iup - c8i * c3rup + c8r * c3iup;
        sr += cr * phr - ci * phi;
        if (CPX) si += cr * phi + ci * phr;
        cr = c6r * c11rup + c6i * c11iup + c7r * c12rup + c7i * c12iup + c8r * c13rup + c8i * c13iup;
        ci = -c6i * c11rup + c6r * c11iup - c7i * c12rup + c7r * c12iup - c8i * c13rup + c8r * c13iup;
        hr += cr * phr - ci * phi;
        if (CPX) hi += cr * phi + ci * phr;
      }
    }
  }
  if (CPX) {
    g.hs[2 * ((size_t)j + (size_t)n * col)] = sr;
    g.hs[2 * ((size_t)j + (size_t)n * col) + 1] = si;
    if (j < col) {
      g.hs[2 * ((size_t)col + (size_t)n * j)] = hr;
      g.hs[2 * ((size_t)col + (size_t)n * j) + 1] = hi;
    } else {
      g.hsdiag[2 * col] = hr;
      g.hsdiag[2 * col + 1] = hi;
    }
  } else {
    g.hs[(size_t)j + (size_t)n * col] = sr;
    if (j < col) g.hs[(size_t)col + (size_t)n * j] = hr;
    else g.hsdiag[col] = hr;
  }
}

// ------------------------------------------------------------------ driver
template <int LT, bool CPX>
static void launch_pw(const PwArgs &a, hipStream_t s)
{
  // columns per wave: the largest count whose registers still allow three
  // waves per SIMD (measured on the 64-atom cell: C = 3 takes 9.4 ms, C = 4
  // at two waves per SIMD 13.3 ms)
  constexpr int C = CPX ? 2 : 3;
  dim3 grid((a.nv + 63) / 64, (a.nv + 4 * C - 1) / (4 * C));
  hipLaunchKernelGGL((k_hamilt_pw<LT, C, CPX>), grid, dim3(256), 0, s, a);
}

template <bool CPX>
static int dispatch_pw(const PwArgs &a, hipStream_t s)
{
  switch (a.lt) {
    case 3: launch_pw<3, CPX>(a, s); break;
    case 4: launch_pw<4, CPX>(a, s); break;
    case 5: launch_pw<5, CPX>(a, s); break;
    case 6: launch_pw<6, CPX>(a, s); break;
    case 7: launch_pw<7, CPX>(a, s); break;
    case 8: launch_pw<8, CPX>(a, s); break;
    case 9: launch_pw<9, CPX>(a, s); break;
    case 10: launch_pw<10, CPX>(a, s); break;
    case 11: launch_pw<11, CPX>(a, s); break;
    case 12: launch_pw<12, CPX>(a, s); break;
    case 13: launch_pw<13, CPX>(a, s); break;
    default: return W2K_EINVAL;
  }
  return 0;
}

int hamilt(Ctx &c)
{
  RunDev &R = c.run();
  const int n = c.n, nv = c.nv, lt = R.ltop, nat = R.nat, ndf = R.ndf;
  const int cpx = R.is_complex;
  hipStream_t s = c.stream;
  int rc;
  if ((rc = c.d_cs.alloc((size_t)2 * ndf * n))) return rc;
  if ((rc = c.d_al.alloc((size_t)n * lt * nat))) return rc;
  if ((rc = c.d_bl.alloc((size_t)n * lt * nat))) return rc;
  if ((rc = c.d_colcoef.alloc((size_t)4 * nat * lt * (nv > 0 ? nv : 1)))) return rc;
  hipLaunchKernelGGL(k_cs, dim3((n + 255) / 256, ndf), dim3(256), 0, s, n, c.d_kzz.p, R.d_pos.p, c.d_cs.p);
  W2K_STAGE(s, "k_cs");
  hipLaunchKernelGGL(k_albl, dim3((n + 127) / 128, nat), dim3(128), 0, s, n, nv, lt, c.d_kvec.p, R.d_rmt.p,
                     R.d_rad.p, R.d_lapw.p, c.d_al.p, c.d_bl.p);
  W2K_STAGE(s, "k_albl");
  if (nv > 0) {
    hipLaunchKernelGGL(k_colcoef, dim3((nv + 127) / 128, nat), dim3(128), 0, s, n, nv, lt, R.vi, R.d_rmt.p,
                       R.d_rad.p, R.d_lapw.p, c.d_al.p, c.d_bl.p, c.d_colcoef.p);
    W2K_STAGE(s, "k_colcoef");
    PwArgs a;
    a.n = n; a.nv = nv; a.nat = nat; a.lt = lt;
    a.kzz = c.d_kzz.p; a.kv = c.d_kvec.p; a.cs = c.d_cs.p; a.al = c.d_al.p; a.bl = c.d_bl.p;
    a.coef = c.d_colcoef.p;
    a.aorder = R.d_aorder.p; a.grp_end = R.d_grp_end.p; a.grp_r = R.d_grp_r.p; a.ngrp = R.ngrp;
    a.mult = R.d_mult.p; a.first_eq = R.d_first_eq.p; a.v3 = R.d_v3.p; a.theta0 = R.theta0;
    a.warp = R.d_warp.p; a.w1 = R.warp1; a.w2 = R.warp2; a.w3 = R.warp3;
    a.hs = c.d_hs.p; a.hsdiag = c.d_hsdiag.p;
    W2K_HIP(hipEventRecord(c.ev[12], s));
    rc = cpx ? dispatch_pw<true>(a, s) : dispatch_pw<false>(a, s);
    if (rc) return rc;
    W2K_STAGE(s, "k_hamilt_pw");
    W2K_HIP(hipEventRecord(c.ev[13], s));
  }
  if (c.nlo > 0) {
    if (c.nlo != R.nlo_cols) return W2K_EINVAL;
    LoArgs b;
    b.n = n; b.nv = nv; b.nlo = c.nlo; b.lt = lt;
    b.kzz = c.d_kzz.p; b.kv = c.d_kvec.p; b.cs = c.d_cs.p; b.al = c.d_al.p; b.bl = c.d_bl.p;
    b.locol = R.d_locol.p; b.lopar = R.d_lopar.p; b.loabc = R.d_loabc.p; b.rad = R.d_rad.p;
    b.rotij = R.d_rotij.p; b.rotloc = R.d_rotloc.p; b.hs = c.d_hs.p; b.hsdiag = c.d_hsdiag.p;
    dim3 grid((n + 127) / 128, c.nlo);
    if (cpx) hipLaunchKernelGGL(k_hamilt_lo<true>, grid, dim3(128), 0, s, b);
    else hipLaunchKernelGGL(k_hamilt_lo<false>, grid, dim3(128), 0, s, b);
    W2K_STAGE(s, "k_hamilt_lo");
  }
  c.have_hamilt = 1;
  return 0;
}

}  // namespace w2k
