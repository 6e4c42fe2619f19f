// ATPAR's radial part on the device (atpar.f:212-330, 470-561): the radial
// functions u_l(E_l), their energy derivatives and the local-orbital functions
// of every atom type by outward integration of the scalar-relativistic radial
// equation (OUTWIN, outwin.f:1-113), their boundary values and norms, and the
// non-spherical radial integrals VNS1..3 / VNS1LO..3LO (RINT13, rint13.f:1-64).
// The functions stay on the device for the adjoint with respect to V_LM(r).
//
// Arithmetic follows the reference step by step: OUTWIN's predictor-corrector
// over the logarithmic mesh r_k = RO exp(DH (k-1)), the normalisation by
// 1/sqrt(RINT13(u, u)), the energy derivative by the difference of the
// functions at E -+ DELE (DELE = 2e-3 Hartree, scaled by 0.25/DELE), its
// orthogonalisation to u, and RINT13's Simpson sum with the mesh points
// multiplied out one by one.
#include <algorithm>
#include <cmath>
#include <tuple>
#include <vector>

#include "w2k_internal.h"

namespace w2k {

struct AtparArgs {
  int nat, nt, nslmax, lomax, jrid, lmaxd;
  int rel;
  const int *jrj;
  const double *ro, *dh, *zz, *vr;    // vr: [a][jrid], r V in Hartree
  const double *e;                     // [a][lmaxd] energies (Ry)
  const double *elo;                   // [a][LOMAXP]: energy of the LO radial function (Ry)
  const int *loor;                     // [a][LOMAXP]
  double *fn;                          // [a][l < nslmax][4][jrid]: A1, B1, AE1, BE1
  double *fnlo;                        // [a][l <= lomax][2][jrid]: A1LO, B1LO
  double *work;                        // [a][l][2][jrid]: A, B of OUTWIN
  double *out;                         // [a][l][6]: P, DP, PE, DPE, PEI, E used
  double *outlo;                       // [a][l][4]: PLO, DPLO, PI12LO, PE12LO
};

// OUTWIN (outwin.f:20-112): A (large) and B (small, times c/2) components,
// value and slope of u/r at the sphere boundary
__device__ void outwin(bool rel, const double *v, double rnot, double dh, int jri, double eh, double fl, double z,
                       double *A, double *B, double &val, double &slo)
{
  const double e = eh * 2.0;
  const double zz = z + z;
  double c = 2.0 * 137.0359895;
  if (!rel) c = 1.0e10;
  const double fllp1 = fl * (fl + 1.0);
  const double r83sq = 64.0 / 9.0, r1 = 1.0 / 9.0, r2 = -5.0 * r1, r3 = 19.0 * r1, h83 = 8.0 / 3.0;
  const double g0 = 1.0;
  double s, sf, f0;
  if (z < 0.9) {
    s = fl + 1.0;
    sf = fl;
    f0 = fl / c;
  } else {
    const double aa = zz / c;
    s = sqrt(fllp1 + 1.0 - aa * aa);
    sf = s;
    f0 = g0 * (s - 1.0) / aa;
  }
  double d1[3], d2[3];
  for (int k = 0; k < 3; k++) {
    const double r = rnot * exp(dh * (double)k);
    const double drdi = dh * r;
    A[k] = pow(r, s) * g0;
    B[k] = pow(r, sf) * f0;
    d1[k] = drdi * A[k] * s / r;
    d2[k] = drdi * B[k] * sf / r;
  }
  double dg1 = d1[0], dg2 = d1[1], dg3 = d1[2], df1 = d2[0], df2 = d2[1], df3 = d2[2];
  double a_prev = A[2], b_prev = B[2];
  for (int k = 3; k < jri; k++) {
    const double r = rnot * exp(dh * (double)k);
    const double drdi = dh * r;
    const double phi = (e - 2.0 * v[k] / r) * drdi / c;
    const double u = drdi * c + phi;
    const double x = -drdi / r;
    const double y = -fllp1 * x * x / u + phi;
    const double det = r83sq - x * x + u * y;
    const double b11 = a_prev * h83 + r1 * dg1 + r2 * dg2 + r3 * dg3;
    const double b22 = b_prev * h83 + r1 * df1 + r2 * df2 + r3 * df3;
    const double ak = (b11 * (h83 - x) + b22 * u) / det;
    const double bk = (b22 * (h83 + x) - b11 * y) / det;
    A[k] = ak;
    B[k] = bk;
    dg1 = dg2; dg2 = dg3; dg3 = u * bk - x * ak;
    df1 = df2; df2 = df3; df3 = x * bk - y * ak;
    a_prev = ak;
    b_prev = bk;
  }
  for (int k = 0; k < jri; k++) B[k] = B[k] * c / 2.0;
  const double rj = rnot * exp(dh * (double)(jri - 1));
  val = A[jri - 1] / rj;
  slo = dg3 / (dh * rj);
  slo = (slo - val) / rj;
}

// RINT13 (rint13.f:28-62): integral of c1 a x + c2 b y over the mesh
template <class F>
__device__ double rint13_of(double rnot, double dx, int jri, F f)
{
  const double d = exp(dx);
  int j = 3 - (jri % 2);  // 1-based
  const int j1 = j - 1;
  double r = rnot * pow(d, (double)(j - 1));
  const double r1 = r / d;
  double z4 = 0.0, z2 = 0.0;
  for (;;) {
    z4 = z4 + r * f(j - 1);
    if (z4 == 0.0) z4 = 2.0e-55;
    r = r * d;
    j = j + 1;
    if (j >= jri) break;
    z2 = z2 + r * f(j - 1);
    r = r * d;
    j = j + 1;
  }
  const double p1 = rnot * f(0);
  const double p2 = r1 * f(j1 - 1);
  double s = 2 * z2 + 4 * z4 + r * f(j - 1) + p2;
  s = (dx * s + p1) / 3.0;
  if (j1 > 1) s = s + 0.5 * dx * (p1 + p2);
  return s;
}

__device__ double rint13(double c1, double c2, const double *a, const double *b, const double *x, const double *y,
                         double rnot, double dx, int jri)
{
  return rint13_of(rnot, dx, jri, [&](int k) { return c1 * a[k] * x[k] + c2 * b[k] * y[k]; });
}

// one thread per (atom type, l): atpar.f:212-289 and, for l <= LOMAX with a
// local-orbital function, atpar.f:297-330.  The functions of l < NSLMAX are
// kept (fn, fnlo); those of higher l go to the thread's scratch rows.
__global__ void k_atpar_l(AtparArgs g, double *__restrict__ scratch)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = t / g.nt, l = t % g.nt;
  if (a >= g.nat) return;
  const bool rel = g.rel != 0;
  const double c1 = 1.0, c2 = rel ? 1.0 / (137.0359895 * 137.0359895) : 1.0e-22;
  const double dele = 2.0e-3, delei = 0.25 / dele;
  const int jri = g.jrj[a];
  const double rnot = g.ro[a], dx = g.dh[a], z = g.zz[a];
  const double *v = g.vr + (size_t)a * g.jrid;
  double *A = g.work + ((size_t)(a * g.nt + l) * 2) * g.jrid, *B = A + g.jrid;
  double *A1 = (l < g.nslmax) ? g.fn + ((size_t)(a * g.nslmax + l) * 4) * g.jrid
                              : scratch + ((size_t)(a * g.nt + l) * 4) * g.jrid;
  double *B1 = A1 + g.jrid, *AE = A1 + 2 * (size_t)g.jrid, *BE = A1 + 3 * (size_t)g.jrid;
  const double fl = (double)l;
  double ei = g.e[(size_t)a * g.lmaxd + l] / 2.0;
  double P = 0, DP = 0, PE = 0, DPE = 0, PEI = 0;
  for (int pass = 0; pass < 2; pass++) {
    double uvb, duvb, uve, duve, uv, duv;
    outwin(rel, v, rnot, dx, jri, ei - dele, fl, z, A, B, uvb, duvb);
    double ovlp = rint13(c1, c2, A, B, A, B, rnot, dx, jri);
    double trx = 1.0 / sqrt(ovlp);
    for (int m = 0; m < jri; m++) { AE[m] = trx * A[m]; BE[m] = trx * B[m]; }
    uvb = trx * uvb;
    duvb = trx * duvb;
    outwin(rel, v, rnot, dx, jri, ei + dele, fl, z, A, B, uve, duve);
    ovlp = rint13(c1, c2, A, B, A, B, rnot, dx, jri);
    trx = 1.0 / sqrt(ovlp);
    uve = delei * (trx * uve - uvb);
    duve = delei * (trx * duve - duvb);
    for (int m = 0; m < jri; m++) {
      AE[m] = delei * (trx * A[m] - AE[m]);
      BE[m] = delei * (trx * B[m] - BE[m]);
    }
    outwin(rel, v, rnot, dx, jri, ei, fl, z, A, B, uv, duv);
    ovlp = rint13(c1, c2, A, B, A, B, rnot, dx, jri);
    trx = 1.0 / sqrt(ovlp);
    P = trx * uv;
    DP = trx * duv;
    for (int m = 0; m < jri; m++) { A1[m] = trx * A[m]; B1[m] = trx * B[m]; }
    const double cross = rint13(c1, c2, A1, B1, AE, BE, rnot, dx, jri);
    const double tr = -cross;
    for (int m = 0; m < jri; m++) { AE[m] = AE[m] + tr * A1[m]; BE[m] = BE[m] + tr * B1[m]; }
    PE = uve + tr * P;
    DPE = duve + tr * DP;
    PEI = rint13(c1, c2, AE, BE, AE, BE, rnot, dx, jri);
    // a function almost zero at the boundary: the energy is lowered by
    // 0.05 Ry and the l is done again, once (atpar.f:281-290)
    if (pass == 0 && fabs(P) < 1.0e-3) {
      ei = (g.e[(size_t)a * g.lmaxd + l] - 0.05) / 2.0;
      continue;
    }
    break;
  }
  double *o = g.out + (size_t)(a * g.nt + l) * 6;
  o[0] = P; o[1] = DP; o[2] = PE; o[3] = DPE; o[4] = PEI; o[5] = 2.0 * ei;
  if (l <= g.lomax && g.loor[a * (g.lomax + 1) + l]) {
    double uv, duv;
    outwin(rel, v, rnot, dx, jri, g.elo[a * (g.lomax + 1) + l] / 2.0, fl, z, A, B, uv, duv);
    const double ovlp = rint13(c1, c2, A, B, A, B, rnot, dx, jri);
    const double trx = 1.0 / sqrt(ovlp);
    double *ALO = g.fnlo + ((size_t)(a * (g.lomax + 1) + l) * 2) * g.jrid, *BLO = ALO + g.jrid;
    for (int m = 0; m < jri; m++) { ALO[m] = trx * A[m]; BLO[m] = trx * B[m]; }
    double *q = g.outlo + (size_t)(a * (g.lomax + 1) + l) * 4;
    q[0] = trx * uv;
    q[1] = trx * duv;
    q[2] = rint13(c1, c2, A1, B1, ALO, BLO, rnot, dx, jri);
    q[3] = rint13(c1, c2, AE, BE, ALO, BLO, rnot, dx, jri);
  }
}

// one thread per (atom type, l0, lm, lp) triple that the coupling list reads:
// VNS1..3 and, for l0 <= LOMAX, the local-orbital integrals (atpar.f:489-552)
struct VnsTriple {
  int a, l0, lm, lp;
};

__global__ void k_atpar_vns(int ntrip, const VnsTriple *__restrict__ trip, AtparArgs g, const double *__restrict__ vlm,
                            int lmmxd, double *__restrict__ res)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip) return;
  const VnsTriple q = trip[t];
  const bool rel = g.rel != 0;
  const double c1 = 1.0, c2 = rel ? 1.0 / (137.0359895 * 137.0359895) : 1.0e-22;
  const int jri = g.jrj[q.a];
  const double rnot = g.ro[q.a], dx = g.dh[q.a];
  const double *V = vlm + ((size_t)q.a * lmmxd + q.lm) * g.jrid;
  const double *f0 = g.fn + ((size_t)(q.a * g.nslmax + q.l0) * 4) * g.jrid;
  const double *fp = g.fn + ((size_t)(q.a * g.nslmax + q.lp) * 4) * g.jrid;
  const double *A1 = f0, *B1 = f0 + g.jrid, *AE1 = f0 + 2 * (size_t)g.jrid, *BE1 = f0 + 3 * (size_t)g.jrid;
  const double *A1p = fp, *B1p = fp + g.jrid, *AE1p = fp + 2 * (size_t)g.jrid, *BE1p = fp + 3 * (size_t)g.jrid;
  double *o = res + (size_t)t * 6;
  // VNS1 = <u_l0 V u_lp>, VNS2 = <ue V ue>, VNS3 = <u_l0 V ue_lp>
  o[0] = rint13_of(rnot, dx, jri, [&](int k) { return c1 * (A1[k] * V[k]) * A1p[k] + c2 * (B1[k] * V[k]) * B1p[k]; });
  o[1] = rint13_of(rnot, dx, jri,
                   [&](int k) { return c1 * (AE1[k] * V[k]) * AE1p[k] + c2 * (BE1[k] * V[k]) * BE1p[k]; });
  o[2] = rint13_of(rnot, dx, jri, [&](int k) { return c1 * (A1[k] * V[k]) * AE1p[k] + c2 * (B1[k] * V[k]) * BE1p[k]; });
  o[3] = o[4] = o[5] = 0.0;
  if (q.l0 <= g.lomax && g.loor[q.a * (g.lomax + 1) + q.l0]) {
    const double *L0 = g.fnlo + ((size_t)(q.a * (g.lomax + 1) + q.l0) * 2) * g.jrid, *L0b = L0 + g.jrid;
    o[3] = rint13_of(rnot, dx, jri, [&](int k) { return c1 * (L0[k] * V[k]) * A1p[k] + c2 * (L0b[k] * V[k]) * B1p[k]; });
    o[4] =
        rint13_of(rnot, dx, jri, [&](int k) { return c1 * (L0[k] * V[k]) * AE1p[k] + c2 * (L0b[k] * V[k]) * BE1p[k]; });
    if (q.lp <= g.lomax && g.loor[q.a * (g.lomax + 1) + q.lp]) {
      const double *Lp = g.fnlo + ((size_t)(q.a * (g.lomax + 1) + q.lp) * 2) * g.jrid, *Lpb = Lp + g.jrid;
      o[5] =
          rint13_of(rnot, dx, jri, [&](int k) { return c1 * (L0[k] * V[k]) * Lp[k] + c2 * (L0b[k] * V[k]) * Lpb[k]; });
    }
  }
}

// RINT13's weight of mesh point k (0-based): the integral is sum_k W_k f_k
// with W as RINT13 forms it (Simpson sums over alternate points, end
// corrections); the mesh points are multiplied out as in rint13.f
__device__ double rint13_weight(double rnot, double dx, int jri, int k)
{
  const double d = exp(dx);
  const int j0 = 3 - (jri % 2);  // first Z4 point (1-based)
  const int kk = k + 1;          // 1-based
  double w = 0.0;
  if (kk >= j0 && kk <= jri - 1) {
    double r = rnot * pow(d, (double)(j0 - 1));
    for (int j = j0; j < kk; j++) r = r * d;
    w = (((kk - j0) % 2 == 0) ? 4.0 : 2.0) * r * dx / 3.0;
  }
  if (kk == jri) {
    double r = rnot * pow(d, (double)(j0 - 1));
    for (int j = j0; j < jri; j++) r = r * d;
    w += r * dx / 3.0;
  }
  const int j1 = j0 - 1;
  const double r1 = rnot * pow(d, (double)(j0 - 1)) / d;
  if (kk == 1) w += rnot / 3.0 + (j1 > 1 ? 0.5 * dx * rnot : 0.0);
  if (kk == j1) w += dx * r1 / 3.0 + (j1 > 1 ? 0.5 * dx * r1 : 0.0);
  return w;
}

// dE/dV_LM(r_k) for every (atom type, lm, k): sum over the integrals of that
// (atom, lm) of dE/dVNS times the weight times the product of the two radial
// functions at r_k
struct VlmTerm {
  int a, lm, l0, lp, kind;  // kind 0..5: VNS1, VNS2, VNS3, VNS1LO, VNS2LO, VNS3LO
  double gv;                // dE/dVNS of that integral
};

__global__ void k_adjoint_vlm(int nterm, const VlmTerm *__restrict__ terms, const int *__restrict__ tstart, int ngroup,
                              AtparArgs g, double *__restrict__ gvlm, int lmmxd)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int grp = blockIdx.y;
  if (grp >= ngroup) return;
  const int t0 = tstart[grp], t1 = tstart[grp + 1];
  if (t0 >= t1) return;
  const int a = terms[t0].a, lm = terms[t0].lm;
  const int jri = g.jrj[a];
  if (k >= jri) return;
  const bool rel = g.rel != 0;
  const double c1 = 1.0, c2 = rel ? 1.0 / (137.0359895 * 137.0359895) : 1.0e-22;
  double acc = 0.0;
  for (int t = t0; t < t1; t++) {
    const VlmTerm q = terms[t];
    const double *f0 = g.fn + ((size_t)(a * g.nslmax + q.l0) * 4) * g.jrid;
    const double *fp = g.fn + ((size_t)(a * g.nslmax + q.lp) * 4) * g.jrid;
    double x0a, x0b, xpa, xpb;
    if (q.kind >= 3) {
      const double *L0 = g.fnlo + ((size_t)(a * (g.lomax + 1) + q.l0) * 2) * g.jrid;
      x0a = L0[k];
      x0b = L0[g.jrid + k];
    } else if (q.kind == 1) {
      x0a = f0[2 * (size_t)g.jrid + k];
      x0b = f0[3 * (size_t)g.jrid + k];
    } else {
      x0a = f0[k];
      x0b = f0[g.jrid + k];
    }
    if (q.kind == 0 || q.kind == 3) {
      xpa = fp[k];
      xpb = fp[g.jrid + k];
    } else if (q.kind == 5) {
      const double *Lp = g.fnlo + ((size_t)(a * (g.lomax + 1) + q.lp) * 2) * g.jrid;
      xpa = Lp[k];
      xpb = Lp[g.jrid + k];
    } else {
      xpa = fp[2 * (size_t)g.jrid + k];
      xpb = fp[3 * (size_t)g.jrid + k];
    }
    acc += q.gv * (c1 * x0a * xpa + c2 * x0b * xpb);
  }
  gvlm[((size_t)a * lmmxd + lm) * g.jrid + k] = acc * rint13_weight(g.ro[a], g.dh[a], jri, k);
}

// ------------------------------------------------------------------ host
template <class T>
static int put(DBuf<T> &b, const T *src, size_t count, hipStream_t s)
{
  int rc = b.alloc(count > 0 ? count : 1);
  if (rc) return rc;
  if (count) W2K_HIP(hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
  return 0;
}

static AtparArgs atpar_args(RunDev &R)
{
  AtparArgs g;
  g.nat = R.nat;
  g.nt = R.at_nt;
  g.nslmax = W2K_NSLMAX;
  g.lomax = W2K_LOMAX;
  g.jrid = R.at_jrid;
  g.lmaxd = W2K_LMAX;
  g.rel = R.at_rel;
  g.jrj = R.d_at_jrj.p;
  g.ro = R.d_at_ro.p;
  g.dh = R.d_at_dh.p;
  g.zz = R.d_at_zz.p;
  g.vr = nullptr;
  g.e = g.elo = nullptr;
  g.loor = nullptr;
  g.fn = R.d_at_fn.p;
  g.fnlo = R.d_at_fnlo.p;
  g.work = g.out = g.outlo = nullptr;
  return g;
}

// the unique (atom, l0, lm, lp) triples of the run's HNS coupling terms
static std::vector<VnsTriple> vns_triples(const RunDev &R)
{
  std::vector<VnsTriple> t;
  for (const RunDev::HnsTerm &h : R.hns_terms) t.push_back({h.a, h.x, h.lmx, h.y});
  std::sort(t.begin(), t.end(), [](const VnsTriple &p, const VnsTriple &q) {
    return std::make_tuple(p.a, p.l0, p.lm, p.lp) < std::make_tuple(q.a, q.l0, q.lm, q.lp);
  });
  t.erase(std::unique(t.begin(), t.end(),
                      [](const VnsTriple &p, const VnsTriple &q) {
                        return p.a == q.a && p.l0 == q.l0 && p.lm == q.lm && p.lp == q.lp;
                      }),
          t.end());
  return t;
}

int atpar(Ctx &c, const w2k_radial_desc *rd, double *const out[16])
{
  RunDev &R = c.run();
  hipStream_t s = c.stream;
  const int nat = R.nat, nt = R.ltop;
  if (!rd || !rd->jrj || !rd->ro || !rd->dh || !rd->zz || !rd->vr || rd->jrid < 5) return W2K_EINVAL;
  for (int a = 0; a < nat; a++)
    if (rd->jrj[a] < 5 || rd->jrj[a] > rd->jrid || !(rd->dh[a] > 0.0) || !(rd->ro[a] > 0.0)) return W2K_EINVAL;
  if (rd->vlm) {
    for (const RunDev::HnsTerm &h : R.hns_terms)
      if (h.lmx < 0 || h.lmx >= rd->lmmxd) return W2K_EINVAL;
  }
  const int jrid = rd->jrid;
  int rc;
  R.at_ready = 0;
  R.at_rel = rd->rel ? 1 : 0;
  R.at_jrid = jrid;
  R.at_nt = nt;
  // LO energies ELO(l, ILO(l)) of the l that carry a local-orbital function
  std::vector<double> elo((size_t)nat * W2K_LOMAXP, 0.0);
  std::vector<int> loor((size_t)nat * W2K_LOMAXP, 0);
  for (int a = 0; a < nat; a++)
    for (int l = 0; l <= W2K_LOMAX; l++) {
      const int il = R.ilo[l + W2K_LOMAXP * a];
      loor[a * W2K_LOMAXP + l] = R.loor[l + W2K_LOMAXP * a] ? 1 : 0;
      if (il > 0) elo[a * W2K_LOMAXP + l] = R.elo[l + W2K_LOMAXP * ((il - 1) + W2K_NLOAT * (size_t)a)];
    }
  DBuf<double> d_vr, d_e, d_elo, d_work, d_out, d_outlo, d_scr;
  DBuf<int> d_loor;
  if ((rc = put(R.d_at_jrj, rd->jrj, nat, s)) || (rc = put(R.d_at_ro, rd->ro, nat, s)) ||
      (rc = put(R.d_at_dh, rd->dh, nat, s)) || (rc = put(R.d_at_zz, rd->zz, nat, s)) ||
      (rc = put(d_vr, rd->vr, (size_t)jrid * nat, s)) || (rc = put(d_e, R.e.data(), R.e.size(), s)) ||
      (rc = put(d_elo, elo.data(), elo.size(), s)) || (rc = put(d_loor, loor.data(), loor.size(), s)))
    return rc;
  if ((rc = R.d_at_fn.alloc((size_t)nat * W2K_NSLMAX * 4 * jrid)) ||
      (rc = R.d_at_fnlo.alloc((size_t)nat * W2K_LOMAXP * 2 * jrid)) ||
      (rc = d_work.alloc((size_t)nat * nt * 2 * jrid)) || (rc = d_scr.alloc((size_t)nat * nt * 4 * jrid)) ||
      (rc = d_out.alloc((size_t)nat * nt * 6)) || (rc = d_outlo.alloc((size_t)nat * W2K_LOMAXP * 4)))
    return rc;
  W2K_HIP(hipMemsetAsync(R.d_at_fnlo.p, 0, sizeof(double) * R.d_at_fnlo.n, s));
  W2K_HIP(hipMemsetAsync(d_outlo.p, 0, sizeof(double) * d_outlo.n, s));
  AtparArgs g = atpar_args(R);
  g.vr = d_vr.p;
  g.e = d_e.p;
  g.elo = d_elo.p;
  g.loor = d_loor.p;
  g.work = d_work.p;
  g.out = d_out.p;
  g.outlo = d_outlo.p;
  const int nth = nat * nt;
  hipLaunchKernelGGL(k_atpar_l, dim3((nth + 63) / 64), dim3(64), 0, s, g, d_scr.p);
  W2K_STAGE(s, "k_atpar_l");
  std::vector<double> o((size_t)nth * 6), olo((size_t)nat * W2K_LOMAXP * 4);
  W2K_HIP(hipMemcpyAsync(o.data(), d_out.p, sizeof(double) * o.size(), hipMemcpyDeviceToHost, s));
  W2K_HIP(hipMemcpyAsync(olo.data(), d_outlo.p, sizeof(double) * olo.size(), hipMemcpyDeviceToHost, s));
  // non-spherical integrals of the coupling list
  std::vector<VnsTriple> trip;
  std::vector<double> vres;
  if (rd->vlm && !R.hns_terms.empty()) {
    trip = vns_triples(R);
    DBuf<int> d_trip;
    DBuf<double> d_vlm, d_res;
    if ((rc = put(d_trip, reinterpret_cast<const int *>(trip.data()), trip.size() * 4, s)) ||
        (rc = put(d_vlm, rd->vlm, (size_t)jrid * rd->lmmxd * nat, s)) || (rc = d_res.alloc(trip.size() * 6)))
      return rc;
    const int ntr = (int)trip.size();
    hipLaunchKernelGGL(k_atpar_vns, dim3((ntr + 63) / 64), dim3(64), 0, s, ntr,
                       reinterpret_cast<const VnsTriple *>(d_trip.p), g, d_vlm.p, rd->lmmxd, d_res.p);
    W2K_STAGE(s, "k_atpar_vns");
    vres.resize(trip.size() * 6);
    W2K_HIP(hipMemcpyAsync(vres.data(), d_res.p, sizeof(double) * vres.size(), hipMemcpyDeviceToHost, s));
  }
  W2K_HIP(hipStreamSynchronize(s));
  for (int a = 0; a < nat; a++)
    for (int l = 0; l < nt; l++)
      for (int q = 0; q < 6; q++)
        if (out[q]) out[q][l + W2K_LMAX * (size_t)a] = o[((size_t)a * nt + l) * 6 + q];
  for (int a = 0; a < nat; a++)
    for (int l = 0; l <= W2K_LOMAX; l++)
      if (loor[a * W2K_LOMAXP + l])
        for (int q = 0; q < 4; q++)
          if (out[6 + q]) out[6 + q][l + W2K_LOMAXP * (size_t)a] = olo[((size_t)a * W2K_LOMAXP + l) * 4 + q];
  const size_t LOP2 = W2K_LOMAX + 2;
  for (size_t t = 0; t < trip.size(); t++) {
    const VnsTriple &q = trip[t];
    const double *v = &vres[t * 6];
    const size_t o3 = q.l0 + W2K_NSLMAX * (q.lm + (size_t)R.lmmxd * (q.lp + W2K_NSLMAX * (size_t)q.a));
    for (int k = 0; k < 3; k++)
      if (out[10 + k]) out[10 + k][o3] = v[k];
    if (q.l0 <= W2K_LOMAX) {
      const size_t ol = q.l0 + LOP2 * (q.lm + (size_t)R.lmmxd * (q.lp + W2K_NSLMAX * (size_t)q.a));
      // atpar.f:509-552: VNS1LO, VNS2LO zero unless LOOR(l0); VNS3LO only
      // written for LOOR(l0) and lp <= LOMAX (zero unless LOOR(lp) too)
      if (out[13]) out[13][ol] = v[3];
      if (out[14]) out[14][ol] = v[4];
      if (loor[q.a * W2K_LOMAXP + q.l0] && q.lp <= W2K_LOMAX && out[15]) out[15][ol] = v[5];
    }
  }
  R.at_ready = 1;
  return 0;
}

int adjoint_vlm(Ctx &c, const double *f_host, int lmmxd, double *g_vlm)
{
  RunDev &R = c.run();
  if (!R.at_ready) return W2K_ESTATE;
  hipStream_t s = c.stream;
  const int nat = R.nat;
  const size_t nlo3 = (size_t)(W2K_LOMAX + 2) * R.lmmxd * W2K_NSLMAX * nat;
  const size_t n3 = (size_t)W2K_NSLMAX * R.lmmxd * W2K_NSLMAX * nat;
  std::vector<double> gv[6];
  double *g6[6];
  for (int k = 0; k < 6; k++) {
    gv[k].assign(k < 3 ? n3 : nlo3, 0.0);
    g6[k] = gv[k].data();
  }
  int rc = adjoint_vns(c, f_host, g6);
  if (rc) return rc;
  // one term per distinct VNS entry the coupling list reads
  std::vector<VlmTerm> terms;
  for (const RunDev::HnsTerm &h : R.hns_terms) {
    const size_t ld = h.arr < 3 ? W2K_NSLMAX : W2K_LOMAX + 2;
    const size_t o = h.x + ld * (h.lmx + (size_t)R.lmmxd * (h.y + W2K_NSLMAX * (size_t)h.a));
    terms.push_back({h.a, h.lmx, h.x, h.y, h.arr, gv[h.arr][o]});
  }
  std::sort(terms.begin(), terms.end(), [](const VlmTerm &p, const VlmTerm &q) {
    return std::make_tuple(p.a, p.lm, p.kind, p.l0, p.lp) < std::make_tuple(q.a, q.lm, q.kind, q.l0, q.lp);
  });
  terms.erase(std::unique(terms.begin(), terms.end(),
                          [](const VlmTerm &p, const VlmTerm &q) {
                            return p.a == q.a && p.lm == q.lm && p.kind == q.kind && p.l0 == q.l0 && p.lp == q.lp;
                          }),
              terms.end());
  for (const VlmTerm &t : terms)
    if (t.lm >= lmmxd) return W2K_EINVAL;
  std::vector<int> start;
  for (size_t t = 0; t < terms.size(); t++)
    if (t == 0 || terms[t].a != terms[t - 1].a || terms[t].lm != terms[t - 1].lm) start.push_back((int)t);
  const int ngroup = (int)start.size();
  start.push_back((int)terms.size());
  const size_t ng = (size_t)R.at_jrid * lmmxd * nat;
  std::fill(g_vlm, g_vlm + ng, 0.0);
  if (ngroup == 0) return 0;
  DBuf<int> d_terms, d_start;
  DBuf<double> d_g;
  static_assert(sizeof(VlmTerm) % sizeof(int) == 0, "VlmTerm as ints");
  if ((rc = put(d_terms, reinterpret_cast<const int *>(terms.data()), terms.size() * sizeof(VlmTerm) / sizeof(int), s)) ||
      (rc = put(d_start, start.data(), start.size(), s)) || (rc = d_g.alloc(ng)))
    return rc;
  W2K_HIP(hipMemsetAsync(d_g.p, 0, sizeof(double) * ng, s));
  AtparArgs g = atpar_args(R);
  hipLaunchKernelGGL(k_adjoint_vlm, dim3((R.at_jrid + 255) / 256, ngroup), dim3(256), 0, s, (int)terms.size(),
                     reinterpret_cast<const VlmTerm *>(d_terms.p), d_start.p, ngroup, g, d_g.p, lmmxd);
  W2K_STAGE(s, "k_adjoint_vlm");
  W2K_HIP(hipMemcpyAsync(g_vlm, d_g.p, sizeof(double) * ng, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace w2k
