/*
 * ORACLE -- test infrastructure only.  CPU restatement of SECLR4
 * (seclr4.F:149-358) with the LAPACK steps it calls written out unblocked:
 *
 *   conjugation of HS (complex)            seclr4.F:159-166
 *   diag swap, POTRF('U') on S             seclr4.F:183-205 (DPOTF2 algorithm)
 *   diag swap, DSCGST/ZHCGST               seclr4.F:210-235, dscgst.f:99-142
 *       (A := inv(U^H) A inv(U) on the lower triangle)
 *   DSYEVX2/ZHEEVX2 ('V','V','L')          dsyevx2.f:297-400, zheevx2.f:305-367
 *       Householder tridiagonalisation (DSYTD2/ZHETD2, DLARFG/ZLARFG signs)
 *       DSTEBZ2 bisection on (VL,VU], ABSTOL=0, block order, NUME cap, NWL
 *       (dstebz2.f:301-652; Sturm counts as DLAEBZ)
 *       DSTEIN-style inverse iteration with reorthogonalisation inside
 *       clusters (gap < 1e-3 |T|_1) and the sign convention "largest
 *       component positive"
 *       DORMTR/ZUNMTR back-application of the reflectors
 *       selection sort of the eigenpairs (dsyevx2.f:381-400)
 *   diag swap, conjugated U (complex), TRSM  seclr4.F:325-357
 *
 * The arithmetic runs in complex form for both builds (a real matrix has
 * zero imaginary parts, which leaves the real results unchanged).
 *
 * Pinned by tests/test_oracle.py: eigenvalues and eigenvectors against the
 * reference golden vectors (oracle/_ref), and against scipy.linalg.eigh.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../include/w2k_lapw1.h"

typedef double complex cplx;

static cplx ld_get(const double *a, int ld, int cpx, int i, int j)
{
  size_t o = (size_t)i + (size_t)ld * j;
  return cpx ? CMPLX(a[2 * o], a[2 * o + 1]) : CMPLX(a[o], 0.0);
}
static void ld_set(double *a, int ld, int cpx, int i, int j, cplx v)
{
  size_t o = (size_t)i + (size_t)ld * j;
  if (cpx) { a[2 * o] = creal(v); a[2 * o + 1] = cimag(v); } else a[o] = creal(v);
}

/* Sturm count: number of eigenvalues of T <= x (DLAEBZ IJOB=1/2 recurrence). */
static int sturm(int n, const double *d, const double *e2, double x, double pivmin)
{
  double t = d[0] - x;
  if (fabs(t) < pivmin) t = -pivmin;
  int c = (t <= 0.0);
  for (int j = 1; j < n; j++) {
    t = d[j] - e2[j - 1] / t - x;
    if (fabs(t) < pivmin) t = -pivmin;
    if (t <= 0.0) c++;
  }
  return c;
}

/* Solve (T - lam I) x = b for one tridiagonal block by Gaussian elimination
 * with partial pivoting (the DLAGTF/DLAGTS factorisation: U with two
 * super-diagonals); pivots smaller than tol are replaced by +-tol. */
static void tri_solve(int n, const double *d, const double *e, double lam, double *b, double tol,
                      double *u0, double *u1, double *u2, double *mul, int *piv)
{
  double a0 = d[0] - lam, a1 = (n > 1) ? e[0] : 0.0;
  for (int k = 0; k < n - 1; k++) {
    double sub = e[k];
    double dn = d[k + 1] - lam, en = (k + 1 < n - 1) ? e[k + 1] : 0.0;
    if (fabs(a0) >= fabs(sub)) {
      piv[k] = 0; u0[k] = a0; u1[k] = a1; u2[k] = 0.0;
      double m = (a0 != 0.0) ? sub / a0 : 0.0;
      mul[k] = m;
      a0 = dn - m * a1;
      a1 = en;
    } else {
      piv[k] = 1; u0[k] = sub; u1[k] = dn; u2[k] = en;
      double m = a0 / sub;
      mul[k] = m;
      double na0 = a1 - m * dn, na1 = -m * en;
      a0 = na0;
      a1 = na1;
    }
  }
  u0[n - 1] = a0; u1[n - 1] = 0.0; u2[n - 1] = 0.0;
  for (int k = 0; k < n; k++) if (fabs(u0[k]) < tol) u0[k] = (u0[k] < 0 ? -tol : tol);
  for (int k = 0; k < n - 1; k++) {
    if (piv[k]) { double t = b[k]; b[k] = b[k + 1]; b[k + 1] = t; }
    b[k + 1] = b[k + 1] - mul[k] * b[k];
  }
  for (int k = n - 1; k >= 0; k--) {
    double s = b[k];
    if (k + 1 < n) s -= u1[k] * b[k + 1];
    if (k + 2 < n) s -= u2[k] * b[k + 2];
    b[k] = s / u0[k];
  }
}

/*
 * oracle_seclr4: hs (ld x n, real or complex interleaved) and hsdiag in the
 * state CALKPT leaves them after HNS (H diag in HS, S diag in HSDIAG).
 * On return eigval[0..ne-1], z (ldz x ne).  Returns INFO (0 ok, >0 POTRF
 * minor, W2K_ENUME when more than nume eigenvalues would be needed is not
 * raised: the count is capped at nume as dstebz2.f:647-649 does).
 */
int oracle_seclr4(int n, int cpx, double *hs, int ld, double *hsdiag, double el, double eu, int nume,
                  int *ne_out, int *nbelw_out, double *eigval, double *z, int ldz)
{
  if (n < 1 || ld < n || ldz < n) return W2K_EINVAL;
  if (el >= eu) { el = -1.0e70; eu = 1.0e70; }
  cplx *A = malloc(sizeof(cplx) * (size_t)n * n);
  cplx *dg = malloc(sizeof(cplx) * n);
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < n; i++) A[i + (size_t)n * j] = ld_get(hs, ld, cpx, i, j);
    dg[j] = cpx ? CMPLX(hsdiag[2 * j], hsdiag[2 * j + 1]) : CMPLX(hsdiag[j], 0.0);
  }
#define AA(i, j) A[(i) + (size_t)n * (j)]
  if (cpx) {
    for (size_t k = 0; k < (size_t)n * n; k++) A[k] = conj(A[k]);
    for (int j = 0; j < n; j++) dg[j] = conj(dg[j]);
  }
  /* swap: S diag into HS */
  for (int j = 0; j < n; j++) { cplx t = AA(j, j); AA(j, j) = dg[j]; dg[j] = t; }
  /* POTRF('U'): S = U^H U, U in the upper triangle */
  for (int j = 0; j < n; j++) {
    double ajj = creal(AA(j, j));
    for (int k = 0; k < j; k++) ajj -= creal(conj(AA(k, j)) * AA(k, j));
    if (ajj <= 0.0 || isnan(ajj)) { free(A); free(dg); return j + 1; }
    ajj = sqrt(ajj);
    AA(j, j) = ajj;
    for (int i = j + 1; i < n; i++) {
      cplx s = AA(j, i);
      for (int k = 0; k < j; k++) s -= conj(AA(k, j)) * AA(k, i);
      AA(j, i) = s / ajj;
    }
  }
  /* swap: H diag back into HS, U diag into dg */
  for (int j = 0; j < n; j++) { cplx t = AA(j, j); AA(j, j) = dg[j]; dg[j] = t; }
  /* reduction (DSCGST/ZHCGST): the panel copies rows of U into columns of
   * the lower factor without conjugation (zhcgst.f PANELB = U^T), so
   * C = inv(U^T) H inv(conj U), i.e. C = V^{-H} H V^{-1} with V = conj(U);
   * H Hermitian from its lower triangle */
  cplx *Hf = malloc(sizeof(cplx) * (size_t)n * n);
#define HF(i, j) Hf[(i) + (size_t)n * (j)]
#define UU(i, j) conj((i) == (j) ? dg[i] : AA(i, j))
  for (int j = 0; j < n; j++)
    for (int i = j; i < n; i++) { HF(i, j) = AA(i, j); HF(j, i) = conj(AA(i, j)); }
  for (int j = 0; j < n; j++) HF(j, j) = creal(HF(j, j));
  /* X = H V^{-1}: solve X V = H row by row (column sweep) */
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < n; i++) {
      cplx s = HF(i, j);
      for (int k = 0; k < j; k++) s -= HF(i, k) * UU(k, j);
      HF(i, j) = s / UU(j, j);
    }
  }
  /* C = V^{-H} X: solve V^H C = X (forward substitution) */
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < n; i++) {
      cplx s = HF(i, j);
      for (int k = 0; k < i; k++) s -= conj(UU(k, i)) * HF(k, j);
      HF(i, j) = s / conj(UU(i, i));
    }
  }
  /* Hermitian part in the lower triangle */
  for (int j = 0; j < n; j++)
    for (int i = j; i < n; i++) AA(i, j) = (i == j) ? creal(HF(i, i)) : HF(i, j);
  free(Hf);

  /* ---------------- tridiagonalisation, lower (DSYTD2/ZHETD2) -------------- */
  double *d = malloc(sizeof(double) * n), *e = malloc(sizeof(double) * n);
  cplx *tau = calloc(n, sizeof(cplx));
  cplx *w = malloc(sizeof(cplx) * n);
  AA(0, 0) = creal(AA(0, 0));
  for (int i = 0; i < n - 1; i++) {
    /* reflector H(i) annihilating A(i+2:n, i) */
    cplx alpha = AA(i + 1, i);
    int m = n - i - 1;
    double xnorm = 0.0;
    for (int k = i + 2; k < n; k++) xnorm += creal(conj(AA(k, i)) * AA(k, i));
    xnorm = sqrt(xnorm);
    double alphr = creal(alpha), alphi = cimag(alpha);
    cplx taui;
    if (xnorm == 0.0 && alphi == 0.0) {
      taui = 0.0;
    } else {
      double beta = -copysign(sqrt(alphr * alphr + alphi * alphi + xnorm * xnorm), alphr);
      taui = CMPLX((beta - alphr) / beta, -alphi / beta);
      cplx sc = 1.0 / (alpha - beta);
      for (int k = i + 2; k < n; k++) AA(k, i) *= sc;
      alpha = beta;
    }
    (void)m;
    e[i] = creal(alpha);
    if (taui != 0.0) {
      AA(i + 1, i) = 1.0;
      /* w = taui * A22 * v, A22 Hermitian from its lower triangle */
      for (int r = i + 1; r < n; r++) {
        cplx s = 0.0;
        for (int c = i + 1; c < n; c++) {
          cplx arc = (r >= c) ? AA(r, c) : conj(AA(c, r));
          if (r == c) arc = creal(arc);
          s += arc * AA(c, i);
        }
        w[r] = taui * s;
      }
      /* alpha2 = -1/2 taui (w^H v) */
      cplx dot = 0.0;
      for (int r = i + 1; r < n; r++) dot += conj(w[r]) * AA(r, i);
      cplx alpha2 = -0.5 * taui * dot;
      for (int r = i + 1; r < n; r++) w[r] += alpha2 * AA(r, i);
      /* A22 -= v w^H + w v^H (lower) */
      for (int c = i + 1; c < n; c++)
        for (int r = c; r < n; r++) AA(r, c) -= AA(r, i) * conj(w[c]) + w[r] * conj(AA(c, i));
      for (int c = i + 1; c < n; c++) AA(c, c) = creal(AA(c, c));
    } else {
      AA(i + 1, i + 1) = creal(AA(i + 1, i + 1));
    }
    AA(i + 1, i) = e[i];
    d[i] = creal(AA(i, i));
    tau[i] = taui;
  }
  d[n - 1] = creal(AA(n - 1, n - 1));

  /* ---------------- bisection (DSTEBZ2, RANGE='V', ORDER='B') ------------- */
  const double safemn = DBL_MIN, ulp = DBL_EPSILON; /* DLAMCH('S'), DLAMCH('P') */
  const double rtoli = ulp * 2.0, fudge = 2.1;
  double *e2 = malloc(sizeof(double) * n);
  int *isplit = malloc(sizeof(int) * (n + 1));
  int nsplit = 0;
  double pivmin = 1.0;
  for (int j = 1; j < n; j++) {
    double t = e[j - 1] * e[j - 1];
    if (fabs(d[j] * d[j - 1]) * ulp * ulp + safemn > t) {
      isplit[nsplit++] = j; /* block ends after index j-1 (1-based end j) */
      e2[j - 1] = 0.0;
    } else {
      e2[j - 1] = t;
      if (t > pivmin) pivmin = t;
    }
  }
  isplit[nsplit++] = n;
  pivmin *= safemn;
  double *wv = malloc(sizeof(double) * n);
  int *iblock = malloc(sizeof(int) * n);
  int mtot = 0, nwl = 0;
  int iend = 0;
  for (int jb = 0; jb < nsplit; jb++) {
    int ibeg = iend;
    iend = isplit[jb];
    int in = iend - ibeg;
    if (in == 1) {
      if (el >= d[ibeg] - pivmin) nwl++;
      if (el < d[ibeg] - pivmin && eu >= d[ibeg] - pivmin) { wv[mtot] = d[ibeg]; iblock[mtot++] = jb; }
      continue;
    }
    double gu = d[ibeg], gl = d[ibeg], t1 = 0.0;
    for (int j = ibeg; j < iend - 1; j++) {
      double t2 = fabs(e[j]);
      gu = fmax(gu, d[j] + t1 + t2);
      gl = fmin(gl, d[j] - t1 - t2);
      t1 = t2;
    }
    gu = fmax(gu, d[iend - 1] + t1);
    gl = fmin(gl, d[iend - 1] - t1);
    double bnorm = fmax(fabs(gl), fabs(gu));
    gl = gl - fudge * bnorm * ulp * in - fudge * pivmin;
    gu = gu + fudge * bnorm * ulp * in + fudge * pivmin;
    double atoli = ulp * fmax(fabs(gl), fabs(gu));
    if (gu < el) { nwl += in; continue; }
    gl = fmax(gl, el);
    gu = fmin(gu, eu);
    if (gl >= gu) continue;
    int cl = sturm(in, d + ibeg, e2 + ibeg, gl, pivmin);
    int cu = sturm(in, d + ibeg, e2 + ibeg, gu, pivmin);
    nwl += cl;
    for (int idx = cl; idx < cu; idx++) {
      /* bisect for the (idx+1)-th eigenvalue of the block in (gl, gu] */
      double lo = gl, hi = gu;
      for (int it = 0; it < 200; it++) {
        double tmp1 = fabs(hi - lo), tmp2 = fmax(fabs(hi), fabs(lo));
        if (tmp1 < fmax(atoli, fmax(pivmin, rtoli * tmp2))) break;
        double mid = 0.5 * (lo + hi);
        if (sturm(in, d + ibeg, e2 + ibeg, mid, pivmin) > idx) hi = mid; else lo = mid;
      }
      wv[mtot] = 0.5 * (lo + hi);
      iblock[mtot++] = jb;
    }
  }
  int m = mtot > nume ? nume : mtot;
  *nbelw_out = nwl;

  /* ---------------- inverse iteration (DSTEIN-like) ----------------------- */
  double *zt = calloc((size_t)n * (m > 0 ? m : 1), sizeof(double));
  double *w1 = malloc(sizeof(double) * n), *w2 = malloc(sizeof(double) * n), *w3 = malloc(sizeof(double) * n);
  double *w4 = malloc(sizeof(double) * n), *bb = malloc(sizeof(double) * n);
  int *piv = malloc(sizeof(int) * n);
  unsigned long long seed = 0x9E3779B97F4A7C15ull;
  iend = 0;
  int jstart = 0;
  for (int jb = 0; jb < nsplit; jb++) {
    int ibeg = iend;
    iend = isplit[jb];
    int bs = iend - ibeg;
    double onenrm = fabs(d[ibeg]) + (bs > 1 ? fabs(e[ibeg]) : 0.0);
    for (int i = ibeg + 1; i < iend; i++) {
      double v = fabs(d[i]) + fabs(e[i - 1]) + (i < iend - 1 ? fabs(e[i]) : 0.0);
      if (v > onenrm) onenrm = v;
    }
    double ortol = 1.0e-3 * onenrm;
    double dtpcrt = sqrt(0.1 / bs);
    int gpind = -1;
    double xjm = 0.0;
    for (int j = jstart; j < m && iblock[j] == jb; j++) {
      double xj = wv[j];
      if (bs == 1) { zt[ibeg + (size_t)n * j] = 1.0; jstart = j + 1; continue; }
      if (gpind < 0) gpind = j;
      else {
        double eps1 = fabs(ulp * xj), pertol = 10.0 * eps1;
        if (xj - xjm < pertol) xj = xjm + pertol;
        if (xj - xjm > ortol) gpind = j;
      }
      for (int i = 0; i < bs; i++) {
        seed = seed * 6364136223846793005ull + 1442695040888963407ull;
        bb[i] = ((double)(seed >> 11) / 9007199254740992.0) * 2.0 - 1.0;
      }
      double tol = fmax(fabs(xj) * ulp, onenrm * ulp);
      for (int its = 0; its < 5; its++) {
        double nrm = 0.0;
        for (int i = 0; i < bs; i++) nrm = fmax(nrm, fabs(bb[i]));
        for (int i = 0; i < bs; i++) bb[i] /= nrm;
        tri_solve(bs, d + ibeg, e + ibeg, xj, bb, tol, w1, w2, w3, w4, piv);
        for (int g = gpind; g < j; g++) {
          double dt = 0.0;
          for (int i = 0; i < bs; i++) dt += bb[i] * zt[ibeg + i + (size_t)n * g];
          for (int i = 0; i < bs; i++) bb[i] -= dt * zt[ibeg + i + (size_t)n * g];
        }
        double amax = 0.0;
        for (int i = 0; i < bs; i++) amax = fmax(amax, fabs(bb[i]));
        double s2 = 0.0;
        for (int i = 0; i < bs; i++) s2 += bb[i] * bb[i];
        if (amax / sqrt(s2) >= dtpcrt && its >= 2) break;
      }
      double s2 = 0.0;
      for (int i = 0; i < bs; i++) s2 += bb[i] * bb[i];
      double scl = 1.0 / sqrt(s2);
      int jmax = 0;
      for (int i = 1; i < bs; i++) if (fabs(bb[i]) > fabs(bb[jmax])) jmax = i;
      if (bb[jmax] < 0.0) scl = -scl;
      for (int i = 0; i < bs; i++) zt[ibeg + i + (size_t)n * j] = bb[i] * scl;
      xjm = xj;
      jstart = j + 1;
    }
  }
  /* ---------------- back-apply the reflectors: Z = Q Zt ------------------- */
  cplx *Zc = malloc(sizeof(cplx) * (size_t)n * (m > 0 ? m : 1));
  for (int j = 0; j < m; j++) for (int i = 0; i < n; i++) Zc[i + (size_t)n * j] = zt[i + (size_t)n * j];
  for (int i = n - 2; i >= 0; i--) {
    cplx t = tau[i];
    if (t == 0.0) continue;
    /* v(i+1) = 1, v(i+2:n) = A(i+2:n, i) */
    for (int j = 0; j < m; j++) {
      cplx s = Zc[i + 1 + (size_t)n * j];
      for (int k = i + 2; k < n; k++) s += conj(AA(k, i)) * Zc[k + (size_t)n * j];
      s *= t;
      Zc[i + 1 + (size_t)n * j] -= s;
      for (int k = i + 2; k < n; k++) Zc[k + (size_t)n * j] -= s * AA(k, i);
    }
  }
  /* ---------------- selection sort ascending ------------------------------ */
  for (int j = 0; j < m - 1; j++) {
    int ii = -1;
    double t = wv[j];
    for (int jj = j + 1; jj < m; jj++) if (wv[jj] < t) { ii = jj; t = wv[jj]; }
    if (ii >= 0) {
      wv[ii] = wv[j];
      wv[j] = t;
      for (int k = 0; k < n; k++) {
        cplx s = Zc[k + (size_t)n * ii];
        Zc[k + (size_t)n * ii] = Zc[k + (size_t)n * j];
        Zc[k + (size_t)n * j] = s;
      }
    }
  }
  /* ---------------- back-transform: diag swap, conj(U) (complex), TRSM ---- */
  for (int j = 0; j < n; j++) { cplx t = AA(j, j); AA(j, j) = dg[j]; dg[j] = t; }
  if (cpx) for (int j = 0; j < n; j++) for (int i = 0; i <= j; i++) AA(i, j) = conj(AA(i, j));
  for (int c = 0; c < m; c++) {
    for (int i = n - 1; i >= 0; i--) {
      cplx s = Zc[i + (size_t)n * c];
      for (int k = i + 1; k < n; k++) s -= AA(i, k) * Zc[k + (size_t)n * c];
      Zc[i + (size_t)n * c] = s / AA(i, i);
    }
  }
  for (int j = 0; j < m; j++) {
    eigval[j] = wv[j];
    for (int i = 0; i < n; i++) ld_set(z, ldz, cpx, i, j, Zc[i + (size_t)n * j]);
  }
  /* HS on exit: lower = reduced matrix, upper = (conj) U, as the reference */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) ld_set(hs, ld, cpx, i, j, AA(i, j));
  *ne_out = m;
  free(A); free(dg); free(d); free(e); free(tau); free(w); free(e2); free(isplit); free(wv);
  free(iblock); free(zt); free(w1); free(w2); free(w3); free(w4); free(bb); free(piv); free(Zc);
#undef AA
#undef HF
#undef UU
  return 0;
}
