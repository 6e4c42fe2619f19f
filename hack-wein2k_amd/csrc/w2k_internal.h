// Internal declarations of the MI355X lapw1 library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/w2k_lapw1.h"

#define W2K_LOMAXP (W2K_LOMAX + 1)
#define W2K_NSL2 (W2K_NSLMAX * W2K_NSLMAX)
#define W2K_NEVENTS 16
#define W2K_NTIMINGS 13

namespace w2k {

// ---------------------------------------------------------------- errors
struct Status {
  int code = 0;
};

#define W2K_HIP(call)                                       \
  do {                                                      \
    hipError_t e_ = (call);                                 \
    if (e_ != hipSuccess) return w2k_hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)

int w2k_hip_fail(hipError_t e, const char *what, const char *file, int line);
// message for w2k_last_error() of a call that returns a positive INFO
void set_last_error(const char *fmt, ...);
// W2K_SYNC_CHECK=1 in the environment: synchronise after every kernel
int sync_checks_enabled();

// launch check after kernel `name` (and a stream synchronisation when the
// environment sets W2K_SYNC_CHECK=1)
int stage_check(hipStream_t s, const char *name, const char *file, int line);
#define W2K_STAGE(s, name)                                       \
  do {                                                           \
    int rc_ = stage_check((s), (name), __FILE__, __LINE__);      \
    if (rc_) return rc_;                                         \
  } while (0)

// ---------------------------------------------------------------- device buffers
// Bump arena of the allocation-free entry point (w2k_seclr4_noalloc): while
// g_arena is set on a thread, DBuf::alloc on that thread carves its buffer
// from the arena (256-byte pieces) instead of calling hipMalloc, and fails
// with W2K_ENOMEM when the arena is exhausted.
struct Arena {
  char *base = nullptr;
  size_t size = 0, used = 0;
};
extern thread_local Arena *g_arena;

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  bool view = false;  // memory not owned (arena piece or caller's buffer): never freed
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  int alloc(size_t count);
  void release();
  // exchange the buffers of two handles (nothing allocated or freed)
  void swap(DBuf &o)
  {
    std::swap(p, o.p);
    std::swap(n, o.n);
    std::swap(view, o.view);
  }
  ~DBuf() { release(); }
};

// ---------------------------------------------------------------- per-run data (device)
struct RunDev {
  int is_complex = 0, nat = 0, ndf = 0, ltop = 0, nnsl = 0, ins = 0, nlo = 0;
  double vi = 0;
  std::vector<int> mult, atom_of;  // atom_of[e] = inequivalent index of equivalent atom e
  std::vector<int> first_eq;       // first equivalent-atom index of each type
  // host copies (small) used to derive device tables
  std::vector<double> pos, rmt, v, rotloc, rotij;
  std::vector<double> u, up, du, dup, chn, e;  // [a][l] (LMAX stride)
  std::vector<int> lapw, ilo, loor;
  std::vector<double> alo, blo, clo, elo, pi12lo, pe12lo, dplo;
  int warp1 = 0, warp2 = 0, warp3 = 0;
  // device copies
  DBuf<double> d_pos, d_rmt, d_v, d_rotloc, d_rotij;
  DBuf<int> d_mult, d_atom_of, d_first_eq;
  DBuf<double> d_rad;     // [a][6][LMAX]: u, up, du, dup, chn, e
  DBuf<int> d_lapw;       // [a][LMAX]
  DBuf<double> d_warp;    // complex box, interleaved
  // interstitial stars (w2k_set_stars): members of every star from STERN
  int nkk = 0, iord = 0;
  DBuf<int> d_kkk;        // [nkk][iord][3] member vectors (first nst[s] valid)
  DBuf<double> d_taup;    // [nkk][iord][2] phase factors TAUP
  DBuf<int> d_nst;        // [nkk] star sizes
  DBuf<double> d_potk;    // [nkk][2] POTK
  // ATPAR's radial functions (w2k_atpar), kept for the V_LM(r) adjoint
  int at_ready = 0, at_rel = 0, at_jrid = 0, at_nt = 0;
  DBuf<int> d_at_jrj;
  DBuf<double> d_at_ro, d_at_dh, d_at_zz;
  DBuf<double> d_at_fn;    // [a][l < NSLMAX][4][jrid]: A1, B1, AE1, BE1
  DBuf<double> d_at_fnlo;  // [a][l <= LOMAX][2][jrid]: A1LO, B1LO
  // HNS coupling recipe: every term GFAC * VNS that run_upload adds into a
  // coupling coefficient T_c(a,k,p), kept for the band-energy adjoint
  // (dE/dVNS = Re(GFAC * dE/dT) summed over the terms that read it)
  struct HnsTerm {
    int a, k, p, c;       // atom type, unique (L0,M0) index, LP*(LP+1)+MP, coefficient 0..8
    int arr, x, lmx, y;   // VNS array (0..2: VNS1..3, 3..5: VNS1LO..3LO) and its indices
    double gr, gi;        // GFAC
  };
  std::vector<HnsTerm> hns_terms;
  int lmmxd = 0;
  // HNS: per inequivalent atom, the sorted unique (L0,M0) list and the
  // coupling coefficients T[k][p][9] (complex), k < lindex[a], p < 25
  std::vector<int> lindex;          // LINDEX per atom type
  std::vector<int> lm0_of;          // [a][25]: L0*(L0+1)+M0 of the k-th unique (L0,M0)
  DBuf<int> d_lm0_of;
  DBuf<double> d_tcoef;             // [a][25][25][9][2]
  DBuf<int> d_lindex;
  // LO columns in HAMILT order (hamilt.F:495-552): a, l, jlo, jeqo, mo1, nnlo, first_eq, mult
  int nlo_cols = 0;
  std::vector<int> locol;
  DBuf<int> d_locol;                // [nlo][8]
  DBuf<double> d_lopar;             // [nlo][8]: c1r,c2r,c3r,c11r,c12r,c13r,akinlor,pi4r2v
  DBuf<double> d_loabc;             // [nat][LOMAXP][NLOAT][3]: alo, blo, clo
  DBuf<int> d_lorow;                // [nat][LOMAXP][NLOAT][2]: LO rows (0-based, relative to NV) of (a,l,jlo)
  // atom types ordered by sphere radius (groups share one sincos per pair)
  int ngrp = 0;
  DBuf<int> d_aorder, d_grp_end;
  DBuf<double> d_grp_r, d_v3;
  double theta0 = 0;
};

// ---------------------------------------------------------------- context
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // per-run tables: this context's own, or (worker contexts of a batched
  // call) the parent's, read-only while k-points run
  RunDev own_run;
  RunDev *runp = &own_run;
  RunDev &run() { return *runp; }
  // worker contexts of batched k-points (owned; share own_run)
  std::vector<w2k_ctx *> workers;
  // basis
  int nv = 0, nlo = 0, n = 0;
  int n_cap = 0;  // largest basis since the per-k buffers were last dropped
  int have_basis = 0, have_hamilt = 0;
  DBuf<int> d_kzz;          // [n][3]
  std::vector<int> h_kzz;   // host copy of KZZ (index bounds for the adjoint)
  DBuf<double> d_kvec;      // [n][4]: xk, yk, zk, rk
  DBuf<double> d_hs;        // n*n (complex: 2*n*n)
  DBuf<double> d_hsdiag;    // n (complex: 2n)
  DBuf<double> d_al, d_bl;  // [n][ltop][nat] Fortran order (n fastest)
  DBuf<double> d_cs;        // [ndf][n][2]: cos, sin of 2 pi KZZ.tau
  DBuf<double> d_theta;     // step-function box (complex interleaved)
  int tb1 = 0, tb2 = 0, tb3 = 0;
  DBuf<double> d_colcoef;   // HAMILT per-column MT coefficients
  // HNS work
  DBuf<double> d_hnsL, d_hnsR;  // rank-update operands
  DBuf<double> h_small;         // LO-block operands
  DBuf<int> h_int;              // per-atom column offsets
  DBuf<double> h_re, h_im;      // complex build: planar products
  int have_hns_ops = 0;         // L/R operands of the last HNS are valid (same basis)
  int hns_slices = 0;           // 0: FP64 rank-K update; 3..7: INT8 digit slices (split precision)
  int at_lmmxd = 0;             // second dimension of V_LM of the last w2k_atpar
  DBuf<signed char> w_dig;      // split precision: digit planes and row exponents
  // eigen results
  int ne = 0, nbelw = 0;
  DBuf<double> d_eig, d_z;
  // SECLR4 work (grown on demand)
  DBuf<double> w_a, w_b, w_c, w_v, w_x;
  DBuf<double> w_small;
  DBuf<int> w_int;
  DBuf<double> w_fac;   // inverse-iteration factors, [plane][row][eigenvalue]
  DBuf<int> w_piv;      // inverse-iteration pivots
  DBuf<double> w_bt;    // back-transform panel work
  DBuf<double> w_cl;    // cluster orthogonalisation work
  DBuf<double> w_gk;    // split-K GEMM partial products
  DBuf<double> w_3m;    // three-product complex GEMM: Re + Im of both operands, one real product
  // launch mode of the tridiagonalisation, chosen per shape by measurement:
  // direct launches or one captured HIP graph replayed (seclr4.hip)
  int graphs = 0;  // 1: capture and replay where it measures faster (off by default)
  struct TriGraph {
    int n = -1, lda = 0;  // key: order, leading dimension and the buffers the launches address
    const void *a = nullptr, *small = nullptr, *ints = nullptr;
    hipGraphExec_t exec = nullptr;
    int seen = 0;      // calls with this key
    int mode = -1;     // last call: 0 direct, 1 replay, 2 capture and replay, -1 none
    int use = 1;       // 0 once a replay was slower than direct launches (or capture failed)
    double direct_ms = 0.0, graph_ms = 0.0;
  } tri_graph;
  // second stream of the back-transform: the reflector panel, Gram matrix
  // and T factor of the next blocks are formed there, ahead of the products
  // with Z on `stream` (default priority, like `stream`)
  hipStream_t aux = nullptr;
  hipEvent_t bt_start = nullptr, bt_ready[2] = {nullptr, nullptr}, bt_used[2] = {nullptr, nullptr};
  // timing
  hipEvent_t ev[W2K_NEVENTS];
  double last_ms[W2K_NTIMINGS] = {0};
  // per-launch timing of the tridiagonalisation SYMV (off unless requested)
  int ktime = 0;
  std::vector<hipEvent_t> kev;
};

// kernels / stages (implemented in the .hip files)
int run_upload(Ctx &c, const w2k_run_desc *r);
int hamilt(Ctx &c);
int hns(Ctx &c);
int seclr4(Ctx &c, double el, double eu, int nume, int *ne, int *nbelw);
// band-energy adjoint with respect to the warped interstitial box (after SECLR4)
int adjoint_warp(Ctx &c, const double *f_host, double *e_band, double *g_warp);
// the same gradient left on the device (*d_box: interleaved box of the run's
// layout in a context workspace, or nullptr when every entry is zero)
int adjoint_warp_box(Ctx &c, const double *f_host, double *e_band, const double **d_box);
// STERN + SETWAR on the device: the run's warped box from its stars
int set_stars(Ctx &c, const w2k_stars *st);
// band-energy gradient with respect to the star coefficients POTK (after SECLR4)
int adjoint_stars(Ctx &c, const double *f_host, double *e_band, double *g_potk);
// interstitial band density rho(Q) on the run's box (host, interleaved)
int density_interstitial(Ctx &c, const double *f_host, double *rho_host);
// the vector-file record of the k-point (TAPEWF's layout) appended to / written at path
int write_vector(Ctx &c, const char *path, int append, const char *kname, double sx, double sy, double sz,
                 double weight, const char *ipgr);
// ATPAR's radial part on the device; outputs in the reference layouts
// (host, any may be NULL): out[0..5] P, DP, PE, DPE, PEI, E (LMAX, NAT);
// out[6..9] PLO, DPLO, PI12LO, PE12LO (LOMAX+1, NAT); out[10..15] VNS1..3,
// VNS1LO..3LO (run descriptor layouts)
int atpar(Ctx &c, const w2k_radial_desc *rd, double *const out[16]);
// band-energy gradient with respect to V_LM(r) (after a CALKPT with HNS and
// w2k_atpar on this run): g (jrid, lmmxd, NAT)
int adjoint_vlm(Ctx &c, const double *f_host, int lmmxd, double *g_vlm);
// Gamma = Z diag(f) Z^H on the lower triangle of the leading `rows` rows,
// planar (Gi only for the complex build); d_f on the device
int adjoint_gamma(Ctx &c, const double *d_f, int rows, double *Gr, double *Gi);
// band-energy adjoint with respect to the VNS radial integrals (after a
// CALKPT that ran HNS); g[0..5] = dE/dVNS1..3, dE/dVNS1LO..3LO (host, desc layout)
int adjoint_vns(Ctx &c, const double *f_host, double *const g[6]);
// release every per-k-point buffer and cached key (streams idle)
void drop_kpoint_state(Ctx &c);
// per-k buffers of a basis of NV + NLO rows (KZZ, k+G, zeroed HS/HSDIAG), as
// w2k_set_basis leaves them before the rows are filled in
int basis_buffers(Ctx &c, int nv, int nlo);
// COORS + LOPW on the device into the context's basis (w2k_coors)
int coors(Ctx &c, const w2k_lattice *L, double sx, double sy, double sz, int *nv, int *nlo);
// the context's basis on the host (any pointer may be NULL)
int get_basis(Ctx &c, int *kzz, double *xk, double *yk, double *zk, double *rk);
// elapsed milliseconds between events e0 and e1 into last_ms[slot]
void take_ms(Ctx &c, int slot, int e0, int e1);

// split precision C (+)= A B^T (A: m x k, B: n x k, column-major, FP64) on the
// INT8 matrix cores with `slices` 7-bit digits per element (3..7); tri 0
// full, 1 lower; ws must hold split_gemm_ws_bytes(m, n, k, slices) bytes
size_t split_gemm_ws_bytes(int m, int n, int k, int slices);
int split_gemm_nt(hipStream_t s, int m, int n, int k, const double *A, int lda, const double *B, int ldb, double *C,
                  int ldc, int tri, int accumulate, int slices, void *ws);

// generic FP64 GEMM on the device: C = alpha * op(A) * op(B) + beta * C
// tri: 0 full, 1 lower (i >= j) only, 2 upper (i <= j) only (tiles outside skipped)
// atri/btri: 0 dense, 1 A (resp. B) upper triangular, 2 lower triangular (zero
// blocks skipped; the caller guarantees the structure)
// ws/wsn: optional workspace (doubles) that lets a product with few output
// tiles and a long K split K over the chip (partials reduced in fixed order)
int dgemm(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
          const double *B, int ldb, double beta, double *C, int ldc, int tri = 0, int atri = 0, int btri = 0,
          double *ws = nullptr, size_t wsn = 0);
// K from which a complex product onto few output tiles is split (slices of
// half of it); a smaller value in a test build puts small products on that path
#ifndef W2K_ZSPLIT_KMIN
#define W2K_ZSPLIT_KMIN 512
#endif
// planar complex C = alpha op(A) op(B) + beta C in one kernel (alpha, beta
// real; op 'N', 'T', 'C' conjugate transpose, 'R' conjugate); tri as dgemm.
// With a workspace ws (wsn doubles), few output tiles (under 256 of 64 x 64)
// and K >= W2K_ZSPLIT_KMIN on a full output, K is split over up to 1024 / tiles slices
// whose partial sums a second kernel adds in a fixed order.
int zgemm_fused(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *Ar, const double *Ai,
                int lda, const double *Br, const double *Bi, int ldb, double beta, double *Cr, double *Ci, int ldc,
                int tri, double *ws = nullptr, size_t wsn = 0);
// the same product on a given tile kernel (0: 64x64, 1: 128x128; no split K)
int dgemm_tiles(hipStream_t s, char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                const double *B, int ldb, double beta, double *C, int ldc, int tri, int atri, int btri, int variant);

}  // namespace w2k

struct w2k_ctx {
  w2k::Ctx c;
};
