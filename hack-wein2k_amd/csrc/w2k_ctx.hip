// Context, per-run tables and the extern "C" entry points of libw2klapw1.so.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>

#include "w2k_internal.h"
#include "../../include/w2k_lapw1_test.h"

// Scratch of w2k_seclr4_noalloc: every SECLR4 workspace of that entry point
// is carved from this array (each device holds its own copy for the life of
// the process).  Host code never names it except to look up its device
// address once per device (noalloc_scratch_init); the kernels receive that
// address through the workspaces carved from it.
static __device__ __align__(256) unsigned char scratch[W2K_NOALLOC_SCRATCH_BYTES];

namespace w2k {

static thread_local char g_last_error[512];

int w2k_hip_fail(hipError_t e, const char *what, const char *file, int line)
{
  const char *base = strrchr(file, '/');
  snprintf(g_last_error, sizeof g_last_error, "%s failed (%s) at %s:%d", what, hipGetErrorString(e),
           base ? base + 1 : file, line);
  if (getenv("W2K_VERBOSE")) fprintf(stderr, "w2k: %s\n", g_last_error);
  return e == hipErrorOutOfMemory ? W2K_ENOMEM : W2K_EDEVICE;
}

void set_last_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
  va_end(ap);
  if (getenv("W2K_VERBOSE")) fprintf(stderr, "w2k: %s\n", g_last_error);
}

// W2K_SYNC_CHECK=1: synchronise after every kernel so a device fault is
// attributed to the kernel that caused it (slow; for diagnosis only)
static int sync_checks()
{
  static const int on = getenv("W2K_SYNC_CHECK") != nullptr && atoi(getenv("W2K_SYNC_CHECK")) != 0;
  return on;
}
int sync_checks_enabled() { return sync_checks(); }

int stage_check(hipStream_t s, const char *name, const char *file, int line)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return w2k_hip_fail(e, name, file, line);
  if (sync_checks()) {
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return w2k_hip_fail(e, name, file, line);
  }
  return 0;
}

thread_local Arena *g_arena = nullptr;

template <class T>
int DBuf<T>::alloc(size_t count)
{
  if (count <= n && p) return 0;
  release();
  if (count == 0) count = 1;
  if (g_arena) {
    const size_t bytes = (sizeof(T) * count + 255) & ~(size_t)255;
    if (g_arena->used + bytes > g_arena->size) {
      set_last_error("scratch of the allocation-free entry point exhausted (%zu of %zu bytes in use, %zu more asked)",
                     g_arena->used, g_arena->size, bytes);
      return W2K_ENOMEM;
    }
    p = reinterpret_cast<T *>(g_arena->base + g_arena->used);
    g_arena->used += bytes;
    n = count;
    view = true;
    return 0;
  }
  if (hipMalloc(&p, sizeof(T) * count) != hipSuccess) {
    p = nullptr;
    n = 0;
    return W2K_ENOMEM;
  }
  n = count;
  return 0;
}
template <class T>
void DBuf<T>::release()
{
  if (p && !view) (void)hipFree(p);
  p = nullptr;
  n = 0;
  view = false;
}
template struct DBuf<double>;
template struct DBuf<int>;
template struct DBuf<signed char>;

// host -> device copy of setup data; waits for the copy, so the host source
// (often a temporary) may be released as soon as this returns
template <class T>
static int upload(DBuf<T> &b, const T *src, size_t count, hipStream_t s)
{
  int rc = b.alloc(count);
  if (rc) return rc;
  if (count) {
    W2K_HIP(hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
    W2K_HIP(hipStreamSynchronize(s));
  }
  return 0;
}

// ---------------------------------------------------------------- per-run upload
struct GEntry {
  int l0, m0, lp, mp, lmx;
  double gr, gi;
};

// LMSORT (lmsort.f): stable order by (L0,M0,LP,MP,LMX), L0=LQNS(3)-1,
// M0=LQNS(6), LP=LQNS(1)-1, MP=LQNS(5), LMX=LQNS(4); duplicates keep the first.
static std::vector<GEntry> lmsort(const w2k_run_desc *r, int a)
{
  std::vector<GEntry> v;
  const int nq = r->lqind[a];
  for (int q = 0; q < nq; q++) {
    const int *lq = r->lqns + 6 * (q + (size_t)r->ngaud * a);
    GEntry g{lq[2] - 1, lq[5], lq[0] - 1, lq[4], lq[3], r->gfac[2 * (q + (size_t)r->ngaud * a)],
             r->gfac[2 * (q + (size_t)r->ngaud * a) + 1]};
    v.push_back(g);
  }
  auto key = [](const GEntry &x) { return std::make_tuple(x.l0, x.m0, x.lp, x.mp, x.lmx); };
  std::stable_sort(v.begin(), v.end(), [&](const GEntry &x, const GEntry &y) { return key(x) < key(y); });
  std::vector<GEntry> u;
  for (auto &g : v)
    if (u.empty() || key(u.back()) != key(g)) u.push_back(g);
  return u;
}

int run_upload(Ctx &c, const w2k_run_desc *r)
{
  RunDev &R = c.run();
  hipStream_t s = c.stream;
  if (!r || r->nat < 1 || r->ndf < r->nat || r->ltop < 3 || r->ltop > W2K_LMAX || r->lmaxd != W2K_LMAX ||
      r->nnsl < 1 || r->nnsl > W2K_NSLMAX || r->nnsl > r->ltop)
    return W2K_EINVAL;
  // the HNS and LO kernels read AL/BL at l < NNSL and at every LO's l: both
  // must lie inside the (n, ltop, nat) table
  for (int a = 0; a < r->nat; a++)
    for (int l = 0; l <= W2K_LOMAX; l++)
      if (r->ilo[l + (W2K_LOMAX + 1) * a] < 0 || r->ilo[l + (W2K_LOMAX + 1) * a] > W2K_NLOAT ||
          (r->ilo[l + (W2K_LOMAX + 1) * a] > 0 && l >= r->ltop))
        return W2K_EINVAL;
  R.is_complex = r->is_complex ? 1 : 0;
  // host copies that ATPAR's radial part reads (energies, local orbitals)
  R.e.assign(r->e, r->e + (size_t)W2K_LMAX * r->nat);
  R.elo.assign(r->elo, r->elo + (size_t)W2K_LOMAXP * W2K_NLOAT * r->nat);
  R.ilo.assign(r->ilo, r->ilo + (size_t)W2K_LOMAXP * r->nat);
  R.loor.assign(r->loor, r->loor + (size_t)W2K_LOMAXP * r->nat);
  // radial functions of an earlier w2k_atpar stay valid for a run of the
  // same atom types (the V_LM adjoint reads them)
  if (r->nat != R.nat) R.at_ready = 0;
  R.nat = r->nat;
  R.ndf = r->ndf;
  R.ltop = r->ltop;
  R.nnsl = r->nnsl;
  R.ins = r->ins;
  R.nlo = r->nlo;
  R.vi = r->vi;
  const int nat = r->nat, ndf = r->ndf;
  R.mult.assign(r->mult, r->mult + nat);
  R.first_eq.resize(nat);
  R.atom_of.resize(ndf);
  int e = 0;
  for (int a = 0; a < nat; a++) {
    R.first_eq[a] = e;
    for (int m = 0; m < R.mult[a]; m++) {
      if (e >= ndf) return W2K_EINVAL;
      R.atom_of[e++] = a;
    }
  }
  if (e != ndf) return W2K_EINVAL;
  int rc;
  if ((rc = upload(R.d_pos, r->pos, (size_t)3 * ndf, s))) return rc;
  if ((rc = upload(R.d_rmt, r->rmt, (size_t)nat, s))) return rc;
  if ((rc = upload(R.d_v, r->v, (size_t)nat, s))) return rc;
  if ((rc = upload(R.d_rotloc, r->rotloc, (size_t)9 * nat, s))) return rc;
  if ((rc = upload(R.d_rotij, r->rotij, (size_t)9 * ndf, s))) return rc;
  if ((rc = upload(R.d_mult, R.mult.data(), (size_t)nat, s))) return rc;
  if ((rc = upload(R.d_atom_of, R.atom_of.data(), (size_t)ndf, s))) return rc;
  if ((rc = upload(R.d_first_eq, R.first_eq.data(), (size_t)nat, s))) return rc;
  // radial parameters [a][6][LMAX]
  std::vector<double> rad((size_t)nat * 6 * W2K_LMAX);
  const double *src[6] = {r->u, r->up, r->du, r->dup, r->chn, r->e};
  for (int a = 0; a < nat; a++)
    for (int k = 0; k < 6; k++)
      for (int l = 0; l < W2K_LMAX; l++) rad[((size_t)a * 6 + k) * W2K_LMAX + l] = src[k][l + W2K_LMAX * a];
  if ((rc = upload(R.d_rad, rad.data(), rad.size(), s))) return rc;
  if ((rc = upload(R.d_lapw, r->lapw, (size_t)nat * W2K_LMAX, s))) return rc;
  // warped interstitial potential
  if (r->ins == 0 && r->warp) {
    R.warp1 = r->warp1; R.warp2 = r->warp2; R.warp3 = r->warp3;
    const size_t nw = (size_t)(2 * R.warp1 + 1) * (2 * R.warp2 + 1) * (2 * R.warp3 + 1);
    if ((rc = upload(R.d_warp, r->warp, 2 * nw, s))) return rc;
  } else {
    R.warp1 = R.warp2 = R.warp3 = 0;
    double z[2] = {0.0, 0.0};
    if ((rc = upload(R.d_warp, z, 2, s))) return rc;
  }
  // step function: radius groups, 3 V_a, Theta(0) (hamilt.F:322-329)
  {
    std::vector<int> order(nat);
    for (int a = 0; a < nat; a++) order[a] = a;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return r->rmt[x] < r->rmt[y]; });
    std::vector<int> gend;
    std::vector<double> gr;
    for (int k = 0; k < nat; k++) {
      if (k > 0 && r->rmt[order[k]] != r->rmt[order[k - 1]]) gend.push_back(k);
      if (gr.empty() || r->rmt[order[k]] != gr.back()) gr.push_back(r->rmt[order[k]]);
    }
    gend.push_back(nat);
    R.ngrp = (int)gend.size();
    std::vector<double> v3(nat);
    double th0 = 0.0;
    for (int a = 0; a < nat; a++) {
      v3[a] = 3.0 * r->v[a];
      th0 = th0 + 1.0 / (double)nat - r->v[a] * (double)r->mult[a];
    }
    R.theta0 = th0;
    if ((rc = upload(R.d_aorder, order.data(), order.size(), s))) return rc;
    if ((rc = upload(R.d_grp_end, gend.data(), gend.size(), s))) return rc;
    if ((rc = upload(R.d_grp_r, gr.data(), gr.size(), s))) return rc;
    if ((rc = upload(R.d_v3, v3.data(), v3.size(), s))) return rc;
  }
  // local orbitals
  {
    const double PI = 4.0 * atan(1.0);
    auto LO = [&](const double *arr, int l, int jlo, int a) { return arr[l + W2K_LOMAXP * (jlo + W2K_NLOAT * a)]; };
    auto LO2 = [&](const double *arr, int l, int a) { return arr[l + W2K_LOMAXP * a]; };
    auto RAD = [&](const double *arr, int l, int a) { return arr[l + W2K_LMAX * a]; };
    std::vector<double> loabc((size_t)nat * W2K_LOMAXP * W2K_NLOAT * 3, 0.0);
    std::vector<int> lorow((size_t)nat * W2K_LOMAXP * W2K_NLOAT * 2, -1);
    std::vector<int> locol;
    std::vector<double> lopar;
    int nnlo = 0, cnt = 0;
    for (int a = 0; a < nat; a++) {
      const double rr = r->rmt[a];
      const double pi4r2v = 4.0 * PI * sqrt(r->vi) * rr * rr;
      for (int l = 0; l <= W2K_LOMAX; l++) {
        const int nil = r->ilo[l + W2K_LOMAXP * a];
        for (int jlo = 0; jlo < nil; jlo++) {
          const double alo = LO(r->alo, l, jlo, a), blo = LO(r->blo, l, jlo, a), clo = LO(r->clo, l, jlo, a);
          const double elo = LO(r->elo, l, jlo, a);
          double *ab = &loabc[3 * (l + W2K_LOMAXP * (jlo + W2K_NLOAT * a))];
          ab[0] = alo; ab[1] = blo; ab[2] = clo;
          // LO rows of (a, l, jlo): the LOHNS range (lohns.f), 0-based relative to NV
          int *rw = &lorow[2 * (l + W2K_LOMAXP * (jlo + W2K_NLOAT * a))];
          rw[0] = cnt;
          rw[1] = cnt + (2 * l + 1) * R.mult[a] - 1;
          cnt += (2 * l + 1) * R.mult[a];
          const double pi12 = LO2(r->pi12lo, l, a), pe12 = LO2(r->pe12lo, l, a), dplo = LO2(r->dplo, l, a);
          const double el = RAD(r->e, l, a), chn = RAD(r->chn, l, a);
          const double HALF = 0.5;
          // hamilt.F:520-548
          const double c1r = alo + pi12 * clo;
          const double c2r = blo * chn + pe12 * clo;
          const double c11r = alo * el + HALF * (blo + clo * pi12 * (elo + el));
          const double c12r = blo * chn * el + HALF * (alo + clo * pi12 + clo * pe12 * (elo + el));
          const double c3r = clo + pe12 * blo + alo * pi12;
          const double c13r = clo * elo + HALF * (blo * pi12 + (elo + el) * (alo * pi12 + pe12 * blo));
          const double akinlor = HALF * (rr * rr) * (alo * RAD(r->du, l, a) + blo * RAD(r->dup, l, a) + clo * dplo);
          for (int jeqo = 1; jeqo <= R.mult[a]; jeqo++) {
            for (int mo1 = -l; mo1 <= l; mo1++) {
              nnlo++;
              int cc[8] = {a, l, jlo + 1, jeqo, mo1, nnlo, R.first_eq[a], R.mult[a]};
              locol.insert(locol.end(), cc, cc + 8);
              double pp[8] = {c1r, c2r, c3r, c11r, c12r, c13r, akinlor, pi4r2v};
              lopar.insert(lopar.end(), pp, pp + 8);
            }
          }
        }
      }
    }
    R.nlo_cols = nnlo;
    R.locol = locol;
    if ((rc = upload(R.d_loabc, loabc.data(), loabc.size(), s))) return rc;
    if ((rc = upload(R.d_lorow, lorow.data(), lorow.size(), s))) return rc;
    if ((rc = upload(R.d_locol, locol.data(), locol.size(), s))) return rc;
    if ((rc = upload(R.d_lopar, lopar.data(), lopar.size(), s))) return rc;
  }
  // non-spherical coupling coefficients (hns.F:266-380), per atom type
  {
    R.lindex.assign(nat, 0);
    R.lm0_of.assign((size_t)nat * W2K_NSL2, 0);
    R.hns_terms.clear();
    R.lmmxd = r->lmmxd;
    std::vector<double> T((size_t)nat * W2K_NSL2 * W2K_NSL2 * 18, 0.0);
    const int LOP2 = W2K_LOMAX + 2;  // leading dimension of VNS*LO
    auto VNS = [&](const double *arr, int x, int lmx, int y, int a) {
      return arr[x + W2K_NSLMAX * (lmx + (size_t)r->lmmxd * (y + W2K_NSLMAX * (size_t)a))];
    };
    auto VNSLO = [&](const double *arr, int x, int lmx, int y, int a) {
      return arr[x + LOP2 * (lmx + (size_t)r->lmmxd * (y + W2K_NSLMAX * (size_t)a))];
    };
    for (int a = 0; a < nat; a++) {
      std::vector<GEntry> gl = lmsort(r, a);
      int q = 0, k = 0;
      while (q < (int)gl.size()) {
        const int L0 = gl[q].l0, M0 = gl[q].m0;
        if (k >= W2K_NSL2 || L0 >= r->nnsl) return W2K_EINVAL;
        R.lm0_of[a * W2K_NSL2 + k] = L0 * (L0 + 1) + M0;
        while (q < (int)gl.size() && gl[q].l0 == L0 && gl[q].m0 == M0) {
          const int LP = gl[q].lp, MP = gl[q].mp;
          const int p = LP * (LP + 1) + MP;
          double t[18] = {0};
          // t: TUU, TDU, TUU21, TUD, TDD, TUD21, TUU12, TDU12, TUU22 (re, im)
          while (q < (int)gl.size() && gl[q].l0 == L0 && gl[q].m0 == M0 && gl[q].lp == LP && gl[q].mp == MP) {
            const int lmx = gl[q].lmx - 1;
            const double gr = gl[q].gr, gi = gl[q].gi;
            // coefficient cc (0..8) += GFAC * VNS array arr at (x, lmx, y)
            auto add = [&](int cc, int arr, const double *base, bool lo, int x, int y) {
              const double v = lo ? VNSLO(base, x, lmx, y, a) : VNS(base, x, lmx, y, a);
              t[2 * cc] += gr * v;
              t[2 * cc + 1] += gi * v;
              R.hns_terms.push_back({a, k, p, cc, arr, x, lmx, y, gr, gi});
            };
            add(0, 0, r->vns1, false, LP, L0);   // TUU
            add(1, 2, r->vns3, false, L0, LP);   // TDU
            add(3, 2, r->vns3, false, LP, L0);   // TUD
            add(4, 1, r->vns2, false, LP, L0);   // TDD
            int loorl0 = 0;
            if (L0 <= W2K_LOMAX) {
              loorl0 = r->loor[L0 + W2K_LOMAXP * a];
              if (loorl0) {
                add(6, 3, r->vns1lo, true, L0, LP);  // TUU12
                add(7, 4, r->vns2lo, true, L0, LP);  // TDU12
              }
            }
            if (LP <= W2K_LOMAX && r->loor[LP + W2K_LOMAXP * a]) {
              if (loorl0) add(8, 5, r->vns3lo, true, LP, L0);  // TUU22
              add(2, 3, r->vns1lo, true, LP, L0);                // TUU21
              add(5, 4, r->vns2lo, true, LP, L0);                // TUD21
            }
            q++;
          }
          double *dst = &T[(((size_t)a * W2K_NSL2 + k) * W2K_NSL2 + p) * 18];
          for (int z = 0; z < 18; z++) dst[z] = t[z];
        }
        k++;
      }
      R.lindex[a] = k;
    }
    if ((rc = upload(R.d_tcoef, T.data(), T.size(), s))) return rc;
    if ((rc = upload(R.d_lm0_of, R.lm0_of.data(), R.lm0_of.size(), s))) return rc;
    if ((rc = upload(R.d_lindex, R.lindex.data(), R.lindex.size(), s))) return rc;
  }
  W2K_HIP(hipStreamSynchronize(s));
  return 0;
}

__global__ void k_swap_diag(int n, int cpx, double *hs, double *hsdiag)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t o = (size_t)i * (n + 1);
  if (cpx) {
    double r = hs[2 * o], im = hs[2 * o + 1];
    hs[2 * o] = hsdiag[2 * i];
    hs[2 * o + 1] = hsdiag[2 * i + 1];
    hsdiag[2 * i] = r;
    hsdiag[2 * i + 1] = im;
  } else {
    double t = hs[o];
    hs[o] = hsdiag[i];
    hsdiag[i] = t;
  }
}

__global__ void k_diag_copy(int n, int cpx, int dir, double *hs, double *buf)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t o = (size_t)i * (n + 1);
  const int w = cpx ? 2 : 1;
  for (int k = 0; k < w; k++) {
    if (dir == 0) buf[w * i + k] = hs[w * o + k];
    else hs[w * o + k] = buf[w * i + k];
  }
}

// elapsed time between two recorded events into last_ms[slot]; a failed
// query leaves the slot at 0 and does not leave an error behind
void take_ms(Ctx &c, int slot, int e0, int e1)
{
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c.ev[e0], c.ev[e1]) != hipSuccess) {
    (void)hipGetLastError();
    ms = 0.f;
  }
  c.last_ms[slot] = ms;
}

// every per-k-point buffer and cached key of a context (not the per-run
// tables, streams or events); the streams must be idle
void drop_kpoint_state(Ctx &c)
{
  DBuf<double> *bufs[] = {&c.d_kvec, &c.d_hs, &c.d_hsdiag, &c.d_al, &c.d_bl, &c.d_cs, &c.d_theta, &c.d_colcoef,
                          &c.d_hnsL, &c.d_hnsR, &c.h_small, &c.h_re, &c.h_im, &c.d_eig, &c.d_z, &c.w_a,
                          &c.w_b, &c.w_c, &c.w_v, &c.w_x, &c.w_small, &c.w_fac, &c.w_bt, &c.w_cl, &c.w_gk, &c.w_3m};
  for (DBuf<double> *b : bufs) b->release();
  c.d_kzz.release();
  c.h_int.release();
  c.w_int.release();
  c.w_piv.release();
  c.w_dig.release();
  if (c.tri_graph.exec) (void)hipGraphExecDestroy(c.tri_graph.exec);
  c.tri_graph = Ctx::TriGraph();
  c.have_basis = c.have_hamilt = c.have_hns_ops = 0;
  c.ne = c.nbelw = 0;
}

static int timed_begin(Ctx &c, int slot) { W2K_HIP(hipEventRecord(c.ev[2 * slot], c.stream)); return 0; }
static int timed_end(Ctx &c, int slot)
{
  W2K_HIP(hipEventRecord(c.ev[2 * slot + 1], c.stream));
  return 0;
}

int basis_buffers(Ctx &c, int nv, int nlo)
{
  const int n = nv + nlo;
  if (n > c.n_cap) {
    // a basis larger than any this context has held: drain both streams and
    // release every per-k buffer and cached key, so that the k-point runs
    // exactly as on a fresh context (smaller bases reuse the buffers, as the
    // batched path's streams do)
    W2K_HIP(hipStreamSynchronize(c.stream));
    W2K_HIP(hipStreamSynchronize(c.aux));
    drop_kpoint_state(c);
    c.n_cap = n;
  }
  c.nv = nv;
  c.nlo = nlo;
  c.n = n;
  c.have_basis = 0;
  c.have_hamilt = 0;
  c.ne = 0;
  c.have_hns_ops = 0;
  int rc;
  if ((rc = c.d_kzz.alloc((size_t)3 * n)) || (rc = c.d_kvec.alloc((size_t)4 * n))) return rc;
  const int w = c.run().is_complex ? 2 : 1;
  if ((rc = c.d_hs.alloc((size_t)w * n * n))) return rc;
  if ((rc = c.d_hsdiag.alloc((size_t)w * n))) return rc;
  W2K_HIP(hipMemsetAsync(c.d_hs.p, 0, sizeof(double) * w * (size_t)n * n, c.stream));
  W2K_HIP(hipMemsetAsync(c.d_hsdiag.p, 0, sizeof(double) * w * (size_t)n, c.stream));
  return 0;
}

}  // namespace w2k

using namespace w2k;

extern "C" {

const char *w2k_version(void) { return "w2k-mi355x 0.1 (gfx950)"; }

const char *w2k_last_error(void) { return g_last_error; }

int w2k_set_kernel_timing(w2k_ctx *x, int on)
{
  if (!x) return W2K_EINVAL;
  x->c.ktime = on ? 1 : 0;
  return 0;
}

int w2k_set_graphs(w2k_ctx *x, int on)
{
  if (!x) return W2K_EINVAL;
  x->c.graphs = on ? 1 : 0;
  for (w2k_ctx *w : x->c.workers) w->c.graphs = x->c.graphs;
  return 0;
}

int w2k_set_split_precision(w2k_ctx *x, int slices)
{
  if (!x) return W2K_EINVAL;
  if (slices != 0 && (slices < 3 || slices > 7)) return W2K_EINVAL;
  x->c.hns_slices = slices;
  return 0;
}

// a context without per-run data: device, stream and events
static int ctx_open(w2k_ctx **out, int device)
{
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) return W2K_ENODEV;
  if (hipSetDevice(device) != hipSuccess) return W2K_ENODEV;
  w2k_ctx *x = new (std::nothrow) w2k_ctx;
  if (!x) return W2K_ENOMEM;
  Ctx &c = x->c;
  c.device = device;
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete x;
    return W2K_EDEVICE;
  }
  if (hipStreamCreateWithFlags(&c.aux, hipStreamNonBlocking) != hipSuccess) {
    (void)hipStreamDestroy(c.stream);
    delete x;
    return W2K_EDEVICE;
  }
  for (int k = 0; k < W2K_NEVENTS; k++) (void)hipEventCreate(&c.ev[k]);
  hipEvent_t *bt[5] = {&c.bt_start, &c.bt_ready[0], &c.bt_ready[1], &c.bt_used[0], &c.bt_used[1]};
  for (hipEvent_t *e : bt) (void)hipEventCreateWithFlags(e, hipEventDisableTiming);
  *out = x;
  return 0;
}

// device address of `scratch` per device, looked up once per device with
// that device current (at the first context created on it)
static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_ptr;

static int noalloc_scratch_init(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return 0;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return w2k_hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
  g_scratch_ptr[device] = ptr;
  return 0;
}

static void *noalloc_scratch(int device)
{
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

// an upper bound of the device memory (bytes) SECLR4 takes from the arena
// for order n and at most m eigenpairs: every workspace of seclr4_real /
// seclr4_complex, seclr4_buffers and tridiag_eigen, each rounded to 256 bytes
static size_t noalloc_bytes(size_t n, size_t m, int cplx)
{
  const size_t z = cplx ? 2 : 1, lda = (n + 15) & ~(size_t)15, nt = (n + 63) / 64;
  auto r = [](size_t doubles) { return (doubles * 8 + 255) & ~(size_t)255; };
  size_t b = 0;
  b += cplx ? 4 * r(2 * n * n) : 3 * r(n * n) + r(lda * n);                  // U, H/C, W, X
  b += r((size_t)1 << 22);                                                   // split-K partials
  const size_t npart = cplx ? std::max(2 * nt * n, 128 * nt * (nt + 1)) : std::max(nt * n, 64 * nt * (nt + 1));
  b += r(12 * n + z * 2 * n + z * 3 * n * 32 + npart + z * nt * 64 + nt * (nt + 1) / 2 + 2 + nt + n / 32 +
         64 + 64 + 768);                                                     // tridiagonal work
  b += r((6 * n + 32 + 1) / 2);                                              // ints
  b += r(m + 1);                                                             // eigenvalues
  const size_t zp = ((n + 2) / 2) * 2 * m;
  b += r(5 * zp) + r((zp + 1) / 2);                                          // inverse iteration
  b += r(2 * n * 64 + 2 * 64 * 64 + m * 64 + m + 1);                         // clusters
  b += r(z * n * (m + 1));                                                   // Z
  b += cplx ? r(2 * (2 * n + 3 * 64 + 2 * m) * 64 + 64 * 64 * 64) : r((2 * n + 2 * m + 3 * 128) * 128 + 32 * 128 * 128);
  return b;
}

int w2k_create(w2k_ctx **out, int device, const w2k_run_desc *run)
{
  if (!out || !run) return W2K_EINVAL;
  *out = nullptr;
  w2k_ctx *x;
  int rc = ctx_open(&x, device);
  if (rc) return rc;
  rc = run_upload(x->c, run);
  if (!rc) rc = noalloc_scratch_init(device);
  if (rc) {
    w2k_destroy(x);
    return rc;
  }
  *out = x;
  return 0;
}

static void drop_workers(Ctx &c)
{
  for (w2k_ctx *w : c.workers) w2k_destroy(w);
  c.workers.clear();
}

int w2k_set_run(w2k_ctx *x, const w2k_run_desc *run)
{
  if (!x) return W2K_EINVAL;
  if (hipSetDevice(x->c.device) != hipSuccess) return W2K_ENODEV;
  drop_workers(x->c);  // they read the tables being replaced
  x->c.have_hamilt = 0;
  return run_upload(x->c, run);
}

void w2k_destroy(w2k_ctx *x)
{
  if (!x) return;
  drop_workers(x->c);
  (void)hipSetDevice(x->c.device);
  if (x->c.stream) (void)hipStreamSynchronize(x->c.stream);
  if (x->c.aux) (void)hipStreamSynchronize(x->c.aux);
  if (x->c.tri_graph.exec) (void)hipGraphExecDestroy(x->c.tri_graph.exec);
  for (int k = 0; k < W2K_NEVENTS; k++) (void)hipEventDestroy(x->c.ev[k]);
  for (hipEvent_t e : {x->c.bt_start, x->c.bt_ready[0], x->c.bt_ready[1], x->c.bt_used[0], x->c.bt_used[1]})
    if (e) (void)hipEventDestroy(e);
  if (x->c.aux) (void)hipStreamDestroy(x->c.aux);
  for (hipEvent_t e : x->c.kev) (void)hipEventDestroy(e);
  if (x->c.stream) (void)hipStreamDestroy(x->c.stream);
  x->c.stream = nullptr;
  delete x;
}

int w2k_set_basis(w2k_ctx *x, int nv, int nlo, const int *kzz, const double *xk, const double *yk, const double *zk,
                  const double *rk)
{
  if (!x || nv < 0 || nlo < 0 || nv + nlo < 1 || !kzz || !xk || !yk || !zk || !rk) return W2K_EINVAL;
  Ctx &c = x->c;
  if (nlo != 0 && nlo != c.run().nlo_cols) return W2K_EINVAL;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const int n = nv + nlo;
  int rc;
  if ((rc = basis_buffers(c, nv, nlo))) return rc;
  std::vector<double> kv((size_t)4 * n);
  for (int i = 0; i < n; i++) {
    kv[4 * i] = xk[i];
    kv[4 * i + 1] = yk[i];
    kv[4 * i + 2] = zk[i];
    kv[4 * i + 3] = rk[i];
  }
  c.h_kzz.assign(kzz, kzz + (size_t)3 * n);
  W2K_HIP(hipMemcpyAsync(c.d_kzz.p, kzz, sizeof(int) * 3 * (size_t)n, hipMemcpyHostToDevice, c.stream));
  W2K_HIP(hipMemcpyAsync(c.d_kvec.p, kv.data(), sizeof(double) * kv.size(), hipMemcpyHostToDevice, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  c.have_basis = 1;
  return 0;
}

int w2k_coors(w2k_ctx *x, const w2k_lattice *lat, double sx, double sy, double sz, int *nv, int *nlo)
{
  if (!x || !lat) return W2K_EINVAL;
  Ctx &c = x->c;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return coors(c, lat, sx, sy, sz, nv, nlo);
}

int w2k_get_basis(w2k_ctx *x, int *kzz, double *xk, double *yk, double *zk, double *rk)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return get_basis(c, kzz, xk, yk, zk, rk);
}

int w2k_hamilt(w2k_ctx *x)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  int rc = timed_begin(c, 0);
  if (!rc) rc = w2k::hamilt(c);
  if (!rc) rc = timed_end(c, 0);
  if (!rc) {
    W2K_HIP(hipStreamSynchronize(c.stream));
    take_ms(c, 0, 0, 1);
    take_ms(c, 8, 12, 13);
  }
  return rc;
}

int w2k_swap_diag(w2k_ctx *x)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  hipLaunchKernelGGL(k_swap_diag, dim3((c.n + 255) / 256), dim3(256), 0, c.stream, c.n, c.run().is_complex, c.d_hs.p,
                     c.d_hsdiag.p);
  W2K_STAGE(c.stream, "k_swap_diag");
  return 0;
}

int w2k_hns(w2k_ctx *x)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  int rc = timed_begin(c, 1);
  if (!rc) rc = w2k::hns(c);
  if (!rc) rc = timed_end(c, 1);
  if (!rc) {
    W2K_HIP(hipStreamSynchronize(c.stream));
    take_ms(c, 1, 2, 3);
    take_ms(c, 9, 14, 15);
  }
  return rc;
}

int w2k_seclr4(w2k_ctx *x, double el, double eu, int nume, int *ne, int *nbelw)
{
  if (!x || !ne || !nbelw || nume < 0) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  int rc = timed_begin(c, 2);
  if (!rc) rc = w2k::seclr4(c, el, eu, nume, ne, nbelw);
  int rc2 = timed_end(c, 2);
  if (!rc) rc = rc2;
  W2K_HIP(hipStreamSynchronize(c.stream));
  take_ms(c, 2, 4, 5);
  return rc;
}

int w2k_seclr4_noalloc(w2k_ctx *x, int n, double *hs, double *hsdiag, double el, double eu, int nume, int *ne,
                       int *nbelw, double *eigval, double *z, int ldz)
{
  if (!x || n < 1 || !hs || !hsdiag || !ne || !nbelw || !eigval || !z || ldz < n || nume < 0) return W2K_EINVAL;
  Ctx &c = x->c;
  const int cplx = c.run().is_complex;
  const size_t m = (size_t)(nume > 0 ? std::min(nume, n) : n);
  const size_t need = noalloc_bytes((size_t)n, m, cplx);
  if (n > W2K_NOALLOC_MAXN || need > (size_t)W2K_NOALLOC_SCRATCH_BYTES) {
    set_last_error("w2k_seclr4_noalloc: order %d with up to %zu eigenpairs needs %zu bytes of scratch; the scratch holds "
                   "%zu bytes (orders up to %d always fit)", n, m, need, (size_t)W2K_NOALLOC_SCRATCH_BYTES,
                   W2K_NOALLOC_MAXN);
    return W2K_EINVAL;
  }
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  void *base = noalloc_scratch(c.device);
  if (!base) return W2K_ESTATE;
  // the context's own per-k buffers and graph key are set aside (not freed)
  // and restored afterwards; HS/HSDIAG are the caller's buffers, every
  // workspace and the eigenpairs come from the scratch
  DBuf<double> *dbufs[] = {&c.d_hs, &c.d_hsdiag, &c.d_eig, &c.d_z, &c.w_a, &c.w_b, &c.w_c, &c.w_v, &c.w_x,
                           &c.w_small, &c.w_fac, &c.w_bt, &c.w_cl, &c.w_gk};
  DBuf<int> *ibufs[] = {&c.w_int, &c.w_piv};
  constexpr int ND = sizeof(dbufs) / sizeof(dbufs[0]), NI = sizeof(ibufs) / sizeof(ibufs[0]);
  DBuf<double> dsave[ND];
  DBuf<int> isave[NI];
  for (int q = 0; q < ND; q++) dbufs[q]->swap(dsave[q]);
  for (int q = 0; q < NI; q++) ibufs[q]->swap(isave[q]);
  Ctx::TriGraph tg_save = c.tri_graph;
  c.tri_graph = Ctx::TriGraph();
  const int n_save = c.n, ne_save = c.ne, nbelw_save = c.nbelw, graphs_save = c.graphs, kt_save = c.ktime;
  const int hb = c.have_basis, hh = c.have_hamilt, ho = c.have_hns_ops;
  c.d_hs.p = hs; c.d_hs.n = (size_t)(cplx ? 2 : 1) * n * n; c.d_hs.view = true;
  c.d_hsdiag.p = hsdiag; c.d_hsdiag.n = (size_t)(cplx ? 2 : 1) * n; c.d_hsdiag.view = true;
  c.n = n;
  c.graphs = 0;
  c.ktime = 0;
  Arena arena;
  arena.base = static_cast<char *>(base);
  arena.size = W2K_NOALLOC_SCRATCH_BYTES;
  g_arena = &arena;
  int rc = w2k::seclr4(c, el, eu, nume, ne, nbelw);
  if (rc == 0 && *ne > 0) {
    const size_t es = sizeof(double) * (cplx ? 2 : 1);
    hipError_t e = hipMemcpy2DAsync(z, es * ldz, c.d_z.p, es * n, es * n, *ne, hipMemcpyDeviceToDevice, c.stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(eigval, c.d_eig.p, sizeof(double) * *ne, hipMemcpyDeviceToDevice, c.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    if (e != hipSuccess) rc = w2k_hip_fail(e, "w2k_seclr4_noalloc: result copy", __FILE__, __LINE__);
  }
  if (rc != 0) {
    // nothing may still read the scratch when the call returns
    (void)hipStreamSynchronize(c.stream);
    (void)hipStreamSynchronize(c.aux);
  }
  g_arena = nullptr;
  for (int q = 0; q < ND; q++) {
    dbufs[q]->release();  // views only: nothing is freed
    dbufs[q]->swap(dsave[q]);
  }
  for (int q = 0; q < NI; q++) {
    ibufs[q]->release();
    ibufs[q]->swap(isave[q]);
  }
  c.tri_graph = tg_save;
  c.n = n_save;
  c.ne = ne_save;
  c.nbelw = nbelw_save;
  c.graphs = graphs_save;
  c.ktime = kt_save;
  c.have_basis = hb;
  c.have_hamilt = hh;
  c.have_hns_ops = ho;
  return rc;
}

int w2k_calkpt(w2k_ctx *x, int nohns, double el, double eu, int nume, int *ne, int *nbelw)
{
  int rc = w2k_hamilt(x);
  if (rc) return rc;
  if ((rc = w2k_swap_diag(x))) return rc;
  if (!nohns && (rc = w2k_hns(x))) return rc;
  return w2k_seclr4(x, el, eu, nume, ne, nbelw);
}

int w2k_calkpt_batched(w2k_ctx *x, int nk, const w2k_basis *bases, int nohns, const double *el, const double *eu,
                       int nume, int nworkers, int *ne, int *nbelw, int *status, double *eigval, int ldeig,
                       double *const *z, const int *ldz)
{
  if (!x || nk < 0 || nume < 1 || ldeig < nume) return W2K_EINVAL;
  if (nk == 0) return 0;
  if (!bases || !el || !eu || !ne || !nbelw || !eigval) return W2K_EINVAL;
  Ctx &c = x->c;
  int nmax = 0;
  for (int q = 0; q < nk; q++) {
    const w2k_basis &b = bases[q];
    const int n = b.nv + b.nlo;
    if (b.nv < 0 || b.nlo < 0 || n < 1 || !b.kzz || !b.xk || !b.yk || !b.zk || !b.rk) return W2K_EINVAL;
    if (b.nlo != 0 && b.nlo != c.run().nlo_cols) return W2K_EINVAL;
    if (z && z[q] && (!ldz || ldz[q] < n)) return W2K_EINVAL;
    nmax = std::max(nmax, n);
  }
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  // concurrent streams: enough to keep the chip busy with small orders
  int P = nworkers > 0 ? nworkers : (nmax <= 600 ? 8 : nmax <= 1500 ? 6 : nmax <= 3000 ? 4 : 2);
  P = std::min(std::min(P, nk), 16);
  while ((int)c.workers.size() < P) {
    w2k_ctx *w;
    const int rc = ctx_open(&w, c.device);
    if (rc) return rc;
    w->c.runp = &c.own_run;
    c.workers.push_back(w);
  }
  for (int w = 0; w < P; w++) {
    c.workers[w]->c.hns_slices = c.hns_slices;
    c.workers[w]->c.graphs = c.graphs;
  }
  // longest first: k-points in decreasing N, each to the least loaded stream
  // (load ~ N^3); a stream's list is then in decreasing N, so its buffers
  // are sized by its first k-point
  std::vector<int> order(nk);
  std::iota(order.begin(), order.end(), 0);
  auto nof = [&](int q) { return bases[q].nv + bases[q].nlo; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nof(a) > nof(b); });
  std::vector<std::vector<int>> plan(P);
  std::vector<double> load(P, 0.0);
  for (int q : order) {
    const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    plan[w].push_back(q);
    const double n = nof(q);
    load[w] += n * n * n;
  }
  std::vector<int> st(nk, 0);
  std::vector<std::string> err(nk);
  auto work = [&](int w) {
    (void)hipSetDevice(c.device);
    w2k_ctx *wx = c.workers[w];
    for (int q : plan[w]) {
      const w2k_basis &b = bases[q];
      int kne = 0, knb = 0;
      int rc = w2k_set_basis(wx, b.nv, b.nlo, b.kzz, b.xk, b.yk, b.zk, b.rk);
      if (!rc) rc = w2k_calkpt(wx, nohns, el[q], eu[q], nume, &kne, &knb);
      const bool wz = z && z[q];
      if (!rc) rc = w2k_get_eig(wx, eigval + (size_t)q * ldeig, wz ? z[q] : nullptr, wz ? ldz[q] : nof(q));
      ne[q] = rc ? 0 : kne;
      nbelw[q] = rc ? 0 : knb;
      st[q] = rc;
      if (rc) err[q] = w2k_last_error();
    }
  };
  std::vector<std::thread> th;
  for (int w = 1; w < P; w++) th.emplace_back(work, w);
  work(0);
  for (auto &t : th) t.join();
  if (status) std::copy(st.begin(), st.end(), status);
  for (int q = 0; q < nk; q++)
    if (st[q]) {
      snprintf(g_last_error, sizeof g_last_error, "k-point %d: %s", q, err[q].c_str());
      return st[q];
    }
  return 0;
}

int w2k_adjoint_warp(w2k_ctx *x, const double *f, double *e_band, double *g_warp)
{
  if (!x || (!f && x->c.ne > 0)) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return adjoint_warp(c, f, e_band, g_warp);
}

int w2k_adjoint_vns(w2k_ctx *x, const double *f, double *g_vns1, double *g_vns2, double *g_vns3, double *g_vns1lo,
                    double *g_vns2lo, double *g_vns3lo)
{
  if (!x || (!f && x->c.ne > 0)) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  double *const g[6] = {g_vns1, g_vns2, g_vns3, g_vns1lo, g_vns2lo, g_vns3lo};
  return adjoint_vns(c, f, g);
}

int w2k_set_stars(w2k_ctx *x, const w2k_stars *st)
{
  if (!x || !st) return W2K_EINVAL;
  Ctx &c = x->c;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  drop_workers(c);  // they read the box being replaced
  c.have_hamilt = 0;
  return set_stars(c, st);
}

int w2k_get_warp(w2k_ctx *x, int *w1, int *w2, int *w3, double *warp)
{
  if (!x || !w1 || !w2 || !w3) return W2K_EINVAL;
  Ctx &c = x->c;
  RunDev &R = c.run();
  *w1 = R.warp1;
  *w2 = R.warp2;
  *w3 = R.warp3;
  if (!warp) return 0;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const size_t nbox = (size_t)(2 * R.warp1 + 1) * (2 * R.warp2 + 1) * (2 * R.warp3 + 1);
  W2K_HIP(hipMemcpyAsync(warp, R.d_warp.p, sizeof(double) * 2 * nbox, hipMemcpyDeviceToHost, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_adjoint_stars(w2k_ctx *x, const double *f, double *e_band, double *g_potk)
{
  if (!x || (!f && x->c.ne > 0)) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return adjoint_stars(c, f, e_band, g_potk);
}

int w2k_atpar(w2k_ctx *x, const w2k_radial_desc *rad, double *p, double *dp, double *pe, double *dpe, double *pei,
              double *e_used, double *plo, double *dplo, double *pi12lo, double *pe12lo, double *vns1, double *vns2,
              double *vns3, double *vns1lo, double *vns2lo, double *vns3lo)
{
  if (!x || !rad) return W2K_EINVAL;
  Ctx &c = x->c;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  double *const out[16] = {p, dp, pe, dpe, pei, e_used, plo, dplo, pi12lo, pe12lo, vns1, vns2, vns3, vns1lo, vns2lo,
                           vns3lo};
  int rc = atpar(c, rad, out);
  if (!rc) c.at_lmmxd = rad->lmmxd;
  return rc;
}

int w2k_adjoint_vlm(w2k_ctx *x, const double *f, double *g_vlm)
{
  if (!x || !g_vlm || (!f && x->c.ne > 0)) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return adjoint_vlm(c, f, c.at_lmmxd, g_vlm);
}

int w2k_write_vector(w2k_ctx *x, const char *path, int append, const char *kname, double sx, double sy, double sz,
                     double weight, const char *ipgr)
{
  if (!x || !path) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return write_vector(c, path, append, kname, sx, sy, sz, weight, ipgr);
}

int w2k_density_interstitial(w2k_ctx *x, const double *f, double *rho)
{
  if (!x || !rho || (!f && x->c.ne > 0)) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  return density_interstitial(c, f, rho);
}

int w2k_get_hs(w2k_ctx *x, double *hs, int ld, double *hsdiag)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis || ld < c.n) return W2K_EINVAL;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const int w = c.run().is_complex ? 2 : 1;
  if (hs)
    W2K_HIP(hipMemcpy2DAsync(hs, sizeof(double) * w * (size_t)ld, c.d_hs.p, sizeof(double) * w * (size_t)c.n,
                             sizeof(double) * w * (size_t)c.n, c.n, hipMemcpyDeviceToHost, c.stream));
  if (hsdiag)
    W2K_HIP(hipMemcpyAsync(hsdiag, c.d_hsdiag.p, sizeof(double) * w * (size_t)c.n, hipMemcpyDeviceToHost, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_put_hs(w2k_ctx *x, const double *hs, int ld, const double *hsdiag)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis || ld < c.n) return W2K_EINVAL;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const int w = c.run().is_complex ? 2 : 1;
  if (hs)
    W2K_HIP(hipMemcpy2DAsync(c.d_hs.p, sizeof(double) * w * (size_t)c.n, hs, sizeof(double) * w * (size_t)ld,
                             sizeof(double) * w * (size_t)c.n, c.n, hipMemcpyHostToDevice, c.stream));
  if (hsdiag)
    W2K_HIP(hipMemcpyAsync(c.d_hsdiag.p, hsdiag, sizeof(double) * w * (size_t)c.n, hipMemcpyHostToDevice, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_get_diag(w2k_ctx *x, double *hs_diag, double *hsdiag)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const int w = c.run().is_complex ? 2 : 1;
  int rc;
  if ((rc = c.w_v.alloc((size_t)w * c.n))) return rc;
  if (hs_diag) {
    hipLaunchKernelGGL(k_diag_copy, dim3((c.n + 255) / 256), dim3(256), 0, c.stream, c.n, c.run().is_complex, 0,
                       c.d_hs.p, c.w_v.p);
    W2K_STAGE(c.stream, "k_diag_copy");
    W2K_HIP(hipMemcpyAsync(hs_diag, c.w_v.p, sizeof(double) * w * c.n, hipMemcpyDeviceToHost, c.stream));
  }
  if (hsdiag) W2K_HIP(hipMemcpyAsync(hsdiag, c.d_hsdiag.p, sizeof(double) * w * c.n, hipMemcpyDeviceToHost, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_put_diag(w2k_ctx *x, const double *hs_diag, const double *hsdiag)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_basis) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const int w = c.run().is_complex ? 2 : 1;
  int rc;
  if ((rc = c.w_v.alloc((size_t)w * c.n))) return rc;
  if (hs_diag) {
    W2K_HIP(hipMemcpyAsync(c.w_v.p, hs_diag, sizeof(double) * w * c.n, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_diag_copy, dim3((c.n + 255) / 256), dim3(256), 0, c.stream, c.n, c.run().is_complex, 1,
                       c.d_hs.p, c.w_v.p);
    W2K_STAGE(c.stream, "k_diag_copy");
  }
  if (hsdiag) W2K_HIP(hipMemcpyAsync(c.d_hsdiag.p, hsdiag, sizeof(double) * w * c.n, hipMemcpyHostToDevice, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_get_albl(w2k_ctx *x, double *al, double *bl)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (!c.have_hamilt) return W2K_ESTATE;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const size_t cnt = (size_t)c.n * c.run().ltop * c.run().nat;
  if (al) W2K_HIP(hipMemcpyAsync(al, c.d_al.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c.stream));
  if (bl) W2K_HIP(hipMemcpyAsync(bl, c.d_bl.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_get_eig(w2k_ctx *x, double *eigval, double *z, int ldz)
{
  if (!x) return W2K_EINVAL;
  Ctx &c = x->c;
  if (ldz < c.n) return W2K_EINVAL;
  if (hipSetDevice(c.device) != hipSuccess) return W2K_ENODEV;
  const int w = c.run().is_complex ? 2 : 1;
  if (c.ne > 0) {
    if (eigval) W2K_HIP(hipMemcpyAsync(eigval, c.d_eig.p, sizeof(double) * c.ne, hipMemcpyDeviceToHost, c.stream));
    if (z)
      W2K_HIP(hipMemcpy2DAsync(z, sizeof(double) * w * (size_t)ldz, c.d_z.p, sizeof(double) * w * (size_t)c.n,
                               sizeof(double) * w * (size_t)c.n, c.ne, hipMemcpyDeviceToHost, c.stream));
  }
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int w2k_device_buffers(w2k_ctx *x, void **hs, void **hsdiag, void **eigval, void **z)
{
  if (!x) return W2K_EINVAL;
  if (hs) *hs = x->c.d_hs.p;
  if (hsdiag) *hsdiag = x->c.d_hsdiag.p;
  if (eigval) *eigval = x->c.d_eig.p;
  if (z) *z = x->c.d_z.p;
  return 0;
}

void *w2k_stream(w2k_ctx *x) { return x ? (void *)x->c.stream : nullptr; }

int w2k_sync(w2k_ctx *x)
{
  if (!x) return W2K_EINVAL;
  if (hipSetDevice(x->c.device) != hipSuccess) return W2K_ENODEV;
  W2K_HIP(hipStreamSynchronize(x->c.stream));
  return 0;
}

int w2k_last_timings(w2k_ctx *x, double *ms, int n)
{
  if (!x || !ms) return W2K_EINVAL;
  for (int k = 0; k < n && k < W2K_NTIMINGS; k++) ms[k] = x->c.last_ms[k];
  return 0;
}

}  // extern "C"

static int test_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda, const double *B,
                      int ldb, double beta, double *C, int ldc, int tri, int atri, int btri, size_t wsn,
                      int variant = -1)
{
  if (hipSetDevice(0) != hipSuccess) return W2K_ENODEV;
  const size_t na = (size_t)lda * ((ta == 'N' || ta == 'n') ? k : m);
  const size_t nb = (size_t)ldb * ((tb == 'N' || tb == 'n') ? n : k);
  const size_t nc = (size_t)ldc * n;
  DBuf<double> dA, dB, dC, dW;
  int rc;
  if ((rc = dA.alloc(na)) || (rc = dB.alloc(nb)) || (rc = dC.alloc(nc))) return rc;
  if (wsn && (rc = dW.alloc(wsn))) return rc;
  W2K_HIP(hipMemcpy(dA.p, A, sizeof(double) * na, hipMemcpyHostToDevice));
  W2K_HIP(hipMemcpy(dB.p, B, sizeof(double) * nb, hipMemcpyHostToDevice));
  W2K_HIP(hipMemcpy(dC.p, C, sizeof(double) * nc, hipMemcpyHostToDevice));
  if (variant >= 0)
    rc = dgemm_tiles(nullptr, ta, tb, m, n, k, alpha, dA.p, lda, dB.p, ldb, beta, dC.p, ldc, tri, atri, btri, variant);
  else
    rc = dgemm(nullptr, ta, tb, m, n, k, alpha, dA.p, lda, dB.p, ldb, beta, dC.p, ldc, tri, atri, btri,
               wsn ? dW.p : nullptr, wsn);
  if (rc) return rc;
  W2K_HIP(hipDeviceSynchronize());
  W2K_HIP(hipMemcpy(C, dC.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int w2k_test_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                              const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri)
{
  return test_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, atri, btri, 0);
}

extern "C" int w2k_test_dgemm_tiles(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                                    const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri,
                                    int btri, int variant)
{
  if (variant != 0 && variant != 1) return W2K_EINVAL;
  return test_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, atri, btri, 0, variant);
}

extern "C" int w2k_test_dgemm_ws(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                                 const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri,
                                 int btri, long wsn)
{
  if (wsn < 0) return W2K_EINVAL;
  return test_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, tri, atri, btri, (size_t)wsn);
}

extern "C" int w2k_test_split_gemm(int m, int n, int k, const double *A, int lda, const double *B, int ldb, double *C,
                                   int ldc, int tri, int accumulate, int slices)
{
  if (m <= 0 || n <= 0 || k <= 0 || lda < m || ldb < n || ldc < m) return W2K_EINVAL;
  if (slices < 3 || slices > 7 || (tri != 0 && tri != 1)) return W2K_EINVAL;
  if (hipSetDevice(0) != hipSuccess) return W2K_ENODEV;
  const size_t na = (size_t)lda * k, nb = (size_t)ldb * k, nc = (size_t)ldc * n;
  DBuf<double> dA, dB, dC;
  DBuf<signed char> ws;
  int rc;
  if ((rc = dA.alloc(na)) || (rc = dB.alloc(nb)) || (rc = dC.alloc(nc))) return rc;
  if ((rc = ws.alloc(split_gemm_ws_bytes(m, n, k, slices)))) return rc;
  W2K_HIP(hipMemcpy(dA.p, A, sizeof(double) * na, hipMemcpyHostToDevice));
  W2K_HIP(hipMemcpy(dB.p, B, sizeof(double) * nb, hipMemcpyHostToDevice));
  W2K_HIP(hipMemcpy(dC.p, C, sizeof(double) * nc, hipMemcpyHostToDevice));
  rc = split_gemm_nt(nullptr, m, n, k, dA.p, lda, dB.p, ldb, dC.p, ldc, tri, accumulate, slices, ws.p);
  if (rc) return rc;
  W2K_HIP(hipDeviceSynchronize());
  W2K_HIP(hipMemcpy(C, dC.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
  return 0;
}

// plain device-to-device copy, the calibration kernel for the HBM byte
// counters (its algorithmic traffic is exactly 8 n bytes read + 8 n written)
__global__ void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, size_t n2)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) dst[i] = src[i];
}

extern "C" int w2k_test_copy(long n, int reps, double *ms_per_copy)
{
  if (n <= 0 || (n & 1) || reps <= 0) return W2K_EINVAL;
  if (hipSetDevice(0) != hipSuccess) return W2K_ENODEV;
  DBuf<double> a, b;
  int rc;
  if ((rc = a.alloc((size_t)n)) || (rc = b.alloc((size_t)n))) return rc;
  const size_t n2 = (size_t)n / 2;
  const unsigned grid = (unsigned)((n2 + 255) / 256);
  W2K_HIP(hipMemset(a.p, 0, sizeof(double) * (size_t)n));
  // one untimed copy first (first-touch page mapping)
  hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, (const double2 *)a.p, (double2 *)b.p, n2);
  W2K_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  W2K_HIP(hipEventCreate(&e0));
  W2K_HIP(hipEventCreate(&e1));
  W2K_HIP(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, nullptr, (const double2 *)a.p, (double2 *)b.p, n2);
  W2K_HIP(hipGetLastError());
  W2K_HIP(hipEventRecord(e1, nullptr));
  W2K_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  W2K_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_copy) *ms_per_copy = ms / reps;
  return 0;
}
