/*
 * w2k_lapw1.h -- C-ABI of the MI355X lapw1 hot path (HAMILT, HNS, SECLR4).
 *
 * The three Fortran routines that CALKPT calls per k-point
 *   HAMILT  (SRC_lapw1/hamilt.F:1-2, called at calkpt.F:155-156)
 *   HNS     (SRC_lapw1/hns.F:1-3,    called at calkpt.F:169-172)
 *   SECLR4  (SRC_lapw1/seclr4.F:3,   called at calkpt.F:196)
 * are replaced by a context that lives on one GPU.  The Fortran side reaches
 * these entry points through ISO_C_BINDING (hack-wein2k_amd/fortran/
 * w2k_dropin.F90 provides HAMILT/HNS/SECLR4 with the reference signatures).
 *
 * Storage conventions follow the reference's sequential build exactly
 * (modules.F:441-481): one N x N column-major array HS with the overlap S in
 * the upper triangle (diagonal included) and the Hamiltonian H in the strict
 * lower triangle (complex build: conjugated), HSDIAG holding diag(H) after
 * HAMILT.  Complex arrays are interleaved (re, im) double pairs, i.e. the
 * Fortran COMPLEX*16 layout.
 *
 * Every entry point returns an int status: 0 on success, > 0 a LAPACK-style
 * INFO from the eigensolver (the reference prints it and STOPs,
 * seclr4.F:200-205, 230-235, 304-320), < 0 an argument or device error
 * (W2K_E* below).  A context is not thread-safe; distinct contexts are
 * independent (one HIP stream each).
 */
#ifndef W2K_LAPW1_H
#define W2K_LAPW1_H

#ifdef __cplusplus
extern "C" {
#endif

/* compile-time dimensions of the reference (param.inc_r:4-24) */
#define W2K_LMAX   13
#define W2K_LOMAX  3
#define W2K_NLOAT  3
#define W2K_NSLMAX 5

/* error codes (< 0) */
#define W2K_EINVAL   -1   /* bad argument / inconsistent sizes            */
#define W2K_ENODEV   -2   /* no HIP device / device selection failed      */
#define W2K_ENOMEM   -3   /* device or host allocation failed             */
#define W2K_EDEVICE  -4   /* a HIP call or kernel launch failed           */
#define W2K_ESTATE   -5   /* call out of order (e.g. HNS before HAMILT)   */
#define W2K_ENUME    -6   /* more eigenvalues in the window than NUME     */

/*
 * Per-run (k-independent) data: everything HAMILT and HNS read from their
 * argument lists and from the modules albl/loabc/lolog/atspdt/loint/rotmat/
 * struk/out (hamilt.F:5-27, hns.F:5-17).  All arrays
 * keep the reference's Fortran layouts and leading dimensions.  The pointers
 * are read during w2k_create / w2k_set_run only (the data is copied).
 */
typedef struct w2k_run_desc {
  int is_complex;        /* 0: lapw1 (REAL, inversion-symmetric), 1: lapw1c     */
  int nat;               /* NAT, inequivalent atoms                             */
  int ndf;               /* NDF, all atoms                                      */
  int ltop;              /* LTOP (= NT = lmax+1), hamilt.F:1                    */
  int lmaxd;             /* leading dim of radial arrays: LMAX (13)             */
  int nnsl;              /* NNSL = LNSMAX+1, hns.F:1                            */
  int lmmxd;             /* 2nd dim of VNS* arrays (LMMX=120, or fewer)         */
  int ngaud;             /* leading dim of LQNS/GFAC (NGAU=2350, or fewer)      */
  int ins;               /* INS: 0 = warped interstitial (hamilt.F:289-347)     */
  const int *mult;       /* MULT(NAT)                                           */
  const double *pos;     /* POS(3,NDF), fractional coordinates                  */
  const double *rmt;     /* R(NAT)                                              */
  const double *v;       /* V(NAT) = 4 pi R^3 / (3 Omega)                       */
  double vi;             /* VOLUME = 1/Omega                                    */
  /* radial parameters (0:LMAX-1, NAT): hamilt.F:52-55 U,UP,DU,DUP,CHN,E    */
  const double *u, *up, *du, *dup, *chn, *e;
  const int *lapw;       /* LAPW(0:LMAX-1,NAT) as 0/1 (lolog)                   */
  /* local orbitals (loabc, lolog; modules.F:720-785, 875-898) */
  int nlo;               /* NLO                                                 */
  const int *ilo;        /* ILO(0:LOMAX,NAT)                                    */
  const int *loor;       /* LOOR(0:LOMAX,NAT) as 0/1                            */
  const double *alo, *blo, *clo, *elo;   /* (0:LOMAX, NLOAT, NAT)               */
  const double *pi12lo, *pe12lo, *dplo;  /* (0:LOMAX, NAT)                      */
  const double *rotloc;  /* ROTLOC(3,3,NAT)                                     */
  const double *rotij;   /* ROTIJ(3,3,NDF)                                      */
  /* warped interstitial potential box (module out, warpin.f:31-37) */
  int warp1, warp2, warp3;
  const double *warp;    /* COMPLEX WARP(-w1:w1,-w2:w2,-w3:w3), interleaved     */
  /* non-spherical potential (atspdt, loint; hns.F:25-37) */
  const int *lqind;      /* LQIND(NAT)                                          */
  const int *lqns;       /* LQNS(6, ngaud, NAT)                                 */
  const double *gfac;    /* COMPLEX GFAC(ngaud, NAT), interleaved               */
  const double *vns1, *vns2, *vns3;        /* (0:NSLMAX-1, lmmxd, 0:NSLMAX-1, NAT) */
  /* (LOMAX+2, lmmxd, NSLMAX, NAT): the allocation of init_loint(LOMAX+1,..)
   * (atpar.f:75 with modules.F:1188-1195), indexed VNS1LO(L+1, LMX, L'+1) */
  const double *vns1lo, *vns2lo, *vns3lo;
} w2k_run_desc;

typedef struct w2k_ctx w2k_ctx;

/* Create a context on HIP device `device` and upload the per-run data.
 * Replaces the per-run state HAMILT/HNS read from modules (no reference
 * counterpart as a call: INILPW/ATPAR fill those modules once per run). */
int w2k_create(w2k_ctx **ctx, int device, const w2k_run_desc *run);
/* Replace the per-run data of an existing context. */
int w2k_set_run(w2k_ctx *ctx, const w2k_run_desc *run);
void w2k_destroy(w2k_ctx *ctx);

/* Per-k basis (COORS/LOPW output, coors.f:1-236, lopw.f:1-144): NV plane
 * waves followed by NLO local-orbital vectors; KZZ(3,NV+NLO) integer
 * reciprocal coordinates, XK/YK/ZK/RK(NV+NLO) cartesian k+G and |k+G|.
 * Allocates the device HS (order NV+NLO) and HSDIAG.  Host pointers. */
int w2k_set_basis(w2k_ctx *ctx, int nv, int nlo, const int *kzz,
                  const double *xk, const double *yk, const double *zk,
                  const double *rk);

/* Reciprocal lattice and cut-off that COORS reads from its argument list and
 * the modules gener/struk/orth/char (coors.f:1-18). */
typedef struct w2k_lattice {
  double br2[9];  /* BR2(3,3), Fortran order: column j = reciprocal vector j  */
  double pia[3];  /* PIA(3) = 2 pi / (a, b, c)                                */
  int ortho;      /* ORTHO (orthogonal axes: KZZ from PIA, coors.f:123-129)    */
  int cxz;        /* LATTIC(1:3) == 'CXZ' (coors.f:135-139)                    */
  double rkm;     /* RKM = RKmax / min RMT: the radius of the k+G sphere      */
  int nmatmax;    /* NMATMAX (param.inc_r:16): largest order NV + NLO          */
} w2k_lattice;

/* COORS + LOPW on the device (coors.f:1-236 called at calkpt.F:143, with
 * LOPW at coors.f:208): the plane waves k+G, k = (sx, sy, sz), with
 * |k+G| <= RKM in the reference's order (by length; equal lengths in the
 * order of the box walk JA, JB, JC), cut back at NMATMAX - NLO to the last
 * gap wider than 1e-5 in length, followed by the NLO local-orbital vectors
 * LOPW picks by Gram-Schmidt over the plane-wave list.  Becomes the context's
 * basis as w2k_set_basis would with the same KZZ/XK/YK/ZK/RK (bit for bit the
 * reference's; RK of the local-orbital rows is what the reference leaves in
 * those slots).  *nv and *nlo receive NV and NLO (either may be NULL).
 * LOPW running out of plane waves (lopw.f:139) returns W2K_EINVAL. */
int w2k_coors(w2k_ctx *ctx, const w2k_lattice *lat, double sx, double sy, double sz,
              int *nv, int *nlo);
/* The context's basis on the host: KZZ(3, NV+NLO), XK, YK, ZK, RK(NV+NLO)
 * (any pointer may be NULL). */
int w2k_get_basis(w2k_ctx *ctx, int *kzz, double *xk, double *yk, double *zk,
                  double *rk);

/* HAMILT (hamilt.F:1-845): spherical MT + warped interstitial + LO block into
 * the device HS/HSDIAG, in the reference layout.  Also keeps the AL/BL
 * tables (module albl) for HNS.  */
int w2k_hamilt(w2k_ctx *ctx);
/* The DSWAP of calkpt.F:160-166 (diag(HS) <-> HSDIAG) on the device copy. */
int w2k_swap_diag(w2k_ctx *ctx);
/* HNS (hns.F:1-526): non-spherical V_LM terms into the lower triangle of
 * the device HS (diagonal included).  Requires w2k_hamilt on this basis. */
int w2k_hns(w2k_ctx *ctx);
/* SECLR4 (seclr4.F:3-471): generalized eigensolve of the device HS/HSDIAG
 * for the eigenvalues in (el, eu] (el >= eu: all, seclr4.F:149-152), at most
 * `nume` of them (dstebz2.f:647-649).  On return *ne eigenpairs, *nbelw =
 * number of eigenvalues <= el.  HS is destroyed (as in the reference).
 * Eigenvalues/vectors stay on the device; fetch with w2k_get_eig. */
int w2k_seclr4(w2k_ctx *ctx, double el, double eu, int nume, int *ne, int *nbelw);
/* Allocation-free SECLR4 for callers that cannot call hipMalloc (code under
 * HIP graph capture elsewhere in their program, tight serving loops,
 * harnesses that forbid allocation).  It has no counterpart in the reference:
 * it is w2k_seclr4 on caller-owned device buffers, with every workspace
 * carved from a static device array of W2K_NOALLOC_SCRATCH_BYTES that each
 * device holds for the life of the process; it calls no hipMalloc, hipFree
 * or allocating library routine (the context itself is created beforehand
 * with w2k_create).  hs (N x N, ld = N) and hsdiag are the device HS/HSDIAG in
 * the reference layout (complex interleaved) and are destroyed, as by
 * w2k_seclr4; on return eigval[0..*ne) and Z(1:N, 1:*ne) at z (leading
 * dimension ldz >= N) are written on the device.  Results equal w2k_seclr4's
 * on the same HS.  Orders up to W2K_NOALLOC_MAXN always fit the scratch; a
 * larger order, or one whose workspaces would not fit, returns W2K_EINVAL
 * before any work is queued (w2k_last_error says why).  Every call on a
 * device shares that device's scratch, so calls that overlap on one device
 * (from different threads or contexts) race; calls on one context are
 * serial.  The context's own HS and eigenpairs are left untouched. */
#define W2K_NOALLOC_SCRATCH_BYTES (128u << 20)
#define W2K_NOALLOC_MAXN 768
int w2k_seclr4_noalloc(w2k_ctx *ctx, int n, double *hs, double *hsdiag, double el, double eu, int nume,
                       int *ne, int *nbelw, double *eigval, double *z, int ldz);
/* The whole CALKPT hot path on the device: HAMILT, diag swap, HNS (unless
 * nohns != 0), SECLR4 (calkpt.F:155-196). */
int w2k_calkpt(w2k_ctx *ctx, int nohns, double el, double eu, int nume,
               int *ne, int *nbelw);

/* One k-point's basis for the batched call (the arguments of w2k_set_basis). */
typedef struct w2k_basis {
  int nv, nlo;
  const int *kzz;                              /* (3, nv+nlo) */
  const double *xk, *yk, *zk, *rk;             /* (nv+nlo) */
} w2k_basis;

/* Batched k-points: the CALKPT hot path (as w2k_calkpt) for nk k-points of
 * one run, each with its own basis (ragged N) and window (el[q], eu[q]],
 * at most nume eigenpairs each, run concurrently on `nworkers` streams of
 * the context's device (0: chosen from the largest N).  The k-points of one
 * stream are taken in decreasing N, assigned longest-first to the least
 * loaded stream.  Every k-point's result is identical to a w2k_calkpt call
 * on its own.  Outputs, per k-point q: ne[q], nbelw[q], status[q] (0 or the
 * code w2k_calkpt would return; may be NULL), EIGVAL in
 * eigval[q*ldeig .. q*ldeig+ne[q]) (ldeig >= nume), and, when z and z[q] are
 * non-NULL, Z(1:N_q, 1:ne[q]) at z[q] with leading dimension ldz[q] >= N_q
 * (complex interleaved).  Returns 0 when every k-point succeeded, else the
 * status of the first failed k-point in index order.  The worker streams and
 * their workspaces persist in the context for later batches.  An extended
 * form with no counterpart in the reference (lapw1 runs k-points in a
 * serial loop, lapw1.F:50-66). */
int w2k_calkpt_batched(w2k_ctx *ctx, int nk, const w2k_basis *bases, int nohns, const double *el,
                       const double *eu, int nume, int nworkers, int *ne, int *nbelw, int *status,
                       double *eigval, int ldeig, double *const *z, const int *ldz);

/* Band-energy adjoint at the k-point of the last w2k_seclr4 / w2k_calkpt
 * (an extended form with no counterpart in the reference).  With the NE
 * weights f[0..NE) held fixed, E_band = sum_n f[n] EIGVAL(n) is written to
 * *e_band, and, when g_warp is non-NULL, its gradient with respect to every
 * entry of the run's warped interstitial box WARP(Q) (w2k_run_desc.warp,
 * hamilt.F:335-345): g = dE/dRe WARP(Q) + i dE/dIm WARP(Q), interleaved, in
 * the box's own layout ((2 warp1+1) x (2 warp2+1) x (2 warp3+1), Fortran
 * order).  By Hellmann-Feynman g(Q) = sum over plane-wave pairs I > J with
 * KZZ_J - KZZ_I = Q of 2 conj(Gamma(I,J)), plus sum_I Gamma(I,I) at Q = 0,
 * Gamma = sum_n f_n z_n z_n^H (the real build: real parts only).  The
 * radial V_LM part is w2k_adjoint_vns. */
int w2k_adjoint_warp(w2k_ctx *ctx, const double *f, double *e_band, double *g_warp);

/* Band-energy adjoint with respect to the non-spherical radial integrals
 * (the second half of the adjoint; no counterpart in the reference): the
 * gradient of E_band = sum_n f[n] EIGVAL(n) (weights held fixed) with
 * respect to every entry of VNS1..3 and VNS1LO..3LO, written in the layouts
 * of w2k_run_desc (vns1..3: (NSLMAX, lmmxd, NSLMAX, NAT); vns1lo..3lo:
 * (LOMAX+2, lmmxd, NSLMAX, NAT)); any output may be NULL.  Needs the last
 * solve of this basis to have run HNS (w2k_calkpt with nohns = 0).  The
 * integrals enter H only through the HNS coupling coefficients
 * T = sum GFAC VNS (hns.F:266-380), so
 * dE/dVNS = sum over the terms that read it of Re(GFAC X), X = sum_I conj
 * XLM(I,p) [(w Gamma) Y1](I,k), Gamma = sum_n f_n z_n z_n^H. */
int w2k_adjoint_vns(w2k_ctx *ctx, const double *f, double *g_vns1, double *g_vns2, double *g_vns3,
                    double *g_vns1lo, double *g_vns2lo, double *g_vns3lo);

/* Interstitial potential stars (module totpot, read by rdswar.f:19-36) and
 * the space group (module out: IORD, IMAT, TAU). */
typedef struct w2k_stars {
  int nkk;                   /* NKK stars                                       */
  const int *ja, *jb, *jc;   /* JA/JB/JC(NKK): representative of each star      */
  const double *potk;        /* COMPLEX POTK(NKK), interleaved                  */
  int iord;                  /* IORD symmetry operations (<= 48)                */
  const int *imat;           /* IMAT(3,3,IORD)                                  */
  const double *tau;         /* TAU(3,IORD), non-primitive translations         */
} w2k_stars;
/* SETWAR (setwar.f:40-74) with STERN (stern.F:1-97) on the device: the
 * stars are expanded into the warped interstitial box of the context's run,
 * WARP(K) = POTK_s conj(TAUP(K)) for every member K of star s, replacing the
 * box of w2k_run_desc (its extent is that of the star members).  The run
 * must have INS = 0.  Also keeps the star tables for w2k_adjoint_stars. */
int w2k_set_stars(w2k_ctx *ctx, const w2k_stars *stars);
/* The run's warped box as HAMILT reads it: *w1..*w3 its half-widths and, when
 * warp is non-NULL, the (2 w1+1)(2 w2+1)(2 w3+1) complex entries (Fortran
 * order, interleaved).  No counterpart in the reference (module out). */
int w2k_get_warp(w2k_ctx *ctx, int *w1, int *w2, int *w3, double *warp);
/* Band-energy adjoint with respect to the star coefficients (an extended
 * form with no counterpart in the reference), after a solved k-point on a
 * run whose box came from w2k_set_stars: E_band = sum_n f[n] EIGVAL(n) to
 * *e_band and, when g_potk is non-NULL, dE/dRe POTK_s + i dE/dIm POTK_s for
 * every star (interleaved; the real build reads the box through its real
 * part, so its imaginary parts are 0) = sum over the members K of g(K)
 * TAUP(K), g the warped-box gradient of w2k_adjoint_warp. */
int w2k_adjoint_stars(w2k_ctx *ctx, const double *f, double *e_band, double *g_potk);
/* Radial input of ATPAR per atom type (modules potnlc and radial,
 * atpar.f:89-101 and 352-375). */
typedef struct w2k_radial_desc {
  int rel;                /* REL: scalar-relativistic (1) or not (0)         */
  int jrid;               /* leading dimension of vr and vlm (>= every JRJ)  */
  const int *jrj;         /* JRJ(NAT): radial mesh points                    */
  const double *ro;       /* RO(NAT): first mesh point                       */
  const double *dh;       /* DH(NAT): logarithmic step                       */
  const double *zz;       /* nuclear charge of every atom type               */
  const double *vr;       /* VR(jrid, NAT): r V_00(r) in Hartree             */
  int lmmxd;              /* second dimension of vlm                         */
  const double *vlm;      /* VLM(jrid, lmmxd, NAT): V_LM(r) in Ry, LM order of
                           * LQNS(4,:,:) (after CBCOMB's combination for
                           * cubic sites); NULL: no non-spherical integrals  */
} w2k_radial_desc;
/* ATPAR's radial part on the device (atpar.f:212-330, 470-561; OUTWIN
 * outwin.f:1-113; RINT13 rint13.f:1-64): for every atom type and l < LTOP
 * the radial function at E_l (the run's E, Ry), its energy derivative and
 * their boundary values and norms; the local-orbital functions at ELO(l,
 * ILO(l)) where LOOR(l); and, with vlm, the non-spherical integrals that the
 * run's HNS coupling list reads.  Outputs (host, any may be NULL) in the
 * reference layouts: P, DP, PE, DPE, PEI and the energy used (ATPAR lowers E_l
 * by 0.05 Ry once when P is almost zero) as (LMAX, NAT); PLO, DPLO, PI12LO,
 * PE12LO as (LOMAX+1, NAT); VNS1..3 and VNS1LO..3LO in the w2k_run_desc
 * layouts (only the entries the coupling list reads are written).  The
 * functions stay on the device for w2k_adjoint_vlm. */
int w2k_atpar(w2k_ctx *ctx, const w2k_radial_desc *rad, double *p, double *dp, double *pe, double *dpe,
              double *pei, double *e_used, double *plo, double *dplo, double *pi12lo, double *pe12lo,
              double *vns1, double *vns2, double *vns3, double *vns1lo, double *vns2lo, double *vns3lo);
/* Band-energy adjoint with respect to the non-spherical potential V_LM(r)
 * (an extended form with no counterpart in the reference), after a solved
 * k-point with HNS on a run whose radial functions came from w2k_atpar:
 * g_vlm(k, lm, a) = dE_band/dV_LM(r_k) of atom type a, layout (jrid, lmmxd,
 * NAT) of the last w2k_atpar call, = sum over the integrals VNS_x(l, lm, l')
 * of dE/dVNS_x times RINT13's weight of r_k times the product of the two
 * radial functions at r_k (large components, and small components times
 * 1/c^2 when relativistic). */
int w2k_adjoint_vlm(w2k_ctx *ctx, const double *f, double *g_vlm);
/* The vector-file record of the last solved k-point exactly as TAPEWF writes
 * it to unit 10 (tapewf.F:45-56, called from prtkpt.F:163-166): Fortran
 * unformatted sequential records with 4-byte length markers: (SX, SY, SZ,
 * KNAME(10 chars), NV+NLO, NE, WEIGHT, IPGR(3 chars)), (KZZ(1:3,I), I=1,NV+NLO)
 * and per eigenpair (I, EIGVAL(I)) and Z(1:NV+NLO, I) (conjugated in the
 * complex build).  Written to `path`, or appended when append != 0, so that
 * lapw2 can read the file the reference would have written.  No counterpart
 * in the hot path itself (lapw1 writes it from the host after SECLR4). */
int w2k_write_vector(w2k_ctx *ctx, const char *path, int append, const char *kname, double sx, double sy,
                     double sz, double weight, const char *ipgr);
/* The interstitial band density of the last solved k-point in reciprocal
 * space, as lapw2 forms it (SRC_lapw2/fourir.F through |psi_n(r)|^2 on an FFT
 * grid), here by pair sums on the device: rho(Q) = sum_n f[n] sum over plane
 * waves I, J with KZZ_J - KZZ_I = Q of conj z(I,n) z(J,n), for every Q of the
 * run's warped box (layout of w2k_get_warp, interleaved; the real build
 * keeps the real part).  No counterpart in lapw1. */
int w2k_density_interstitial(w2k_ctx *ctx, const double *f, double *rho);
/* Host <-> device transfers of the reference-layout arrays (ld >= N).  */
int w2k_get_hs(w2k_ctx *ctx, double *hs, int ld, double *hsdiag);
int w2k_put_hs(w2k_ctx *ctx, const double *hs, int ld, const double *hsdiag);
/* diag(HS) and HSDIAG only: what the host touches between the calls of
 * CALKPT (the DSWAP), so a Fortran caller can keep HS device-resident. */
int w2k_get_diag(w2k_ctx *ctx, double *hs_diag, double *hsdiag);
int w2k_put_diag(w2k_ctx *ctx, const double *hs_diag, const double *hsdiag);
/* AL/BL tables (N, 0:LTOP-1, NAT), Fortran order, for inspection. */
int w2k_get_albl(w2k_ctx *ctx, double *al, double *bl);
/* EIGVAL(1:ne) and Z(1:N,1:ne) (ldz >= N); complex Z interleaved. */
int w2k_get_eig(w2k_ctx *ctx, double *eigval, double *z, int ldz);

/* Device-resident variant for callers that already hold device memory:
 * the device addresses of HS (N*N, ld = N), HSDIAG, EIGVAL, Z (ld = N). */
int w2k_device_buffers(w2k_ctx *ctx, void **hs, void **hsdiag, void **eigval, void **z);
/* Stream of the context (hipStream_t as void*); all work is queued on it. */
void *w2k_stream(w2k_ctx *ctx);
/* Wait for all queued work of the context. */
int w2k_sync(w2k_ctx *ctx);

/* Kernel timing of the most recent calls, in milliseconds, measured with HIP
 * events on the context stream: [0] HAMILT, [1] HNS, [2] SECLR4 total,
 * [3] Cholesky, [4] reduction to standard form, [5] tridiagonalisation,
 * [6] bisection + inverse iteration, [7] back-transforms, [8] the HAMILT
 * plane-wave pair kernel alone, [9] the HNS rank-K GEMM alone, [10] the
 * summed time of the tridiagonalisation SYMV kernels and [11] their number of
 * launches (both only while per-launch timing is on), [12] the launch mode of
 * the last tridiagonalisation (0 direct launches, 1 replay of a captured HIP
 * graph, 2 capture and replay; see w2k_set_graphs).  At most n values are
 * written. */
int w2k_last_timings(w2k_ctx *ctx, double *ms, int n);

/* Per-launch HIP events around the tridiagonalisation SYMV kernels (slots
 * [10], [11] of w2k_last_timings); off by default since the events add a
 * little time between launches.  No counterpart in the reference. */
int w2k_set_kernel_timing(w2k_ctx *ctx, int on);

/* Launch mode of the tridiagonalisation (off by default): with on = 1 the
 * library captures the phase's launches (two kernels per column and one GEMM
 * per 32-column panel, no host read-back) into a HIP graph on the second
 * call with the same order and buffers, replays it on the third, and keeps
 * replaying only if that replay measured faster than the direct launches;
 * on = 0 always launches directly.  Results are the same either way.  The
 * setting is passed on to the streams of w2k_calkpt_batched.  No counterpart
 * in the reference. */
int w2k_set_graphs(w2k_ctx *ctx, int on);

/* Split precision for the HNS rank-K update (hns.F:443-488, which the
 * reference does in FP64 with DSYR2M / ZHER2M): slices = 0 keeps the FP64
 * product (default); slices = 3..7 computes it on the INT8 matrix cores with
 * that many signed 7-bit digits per element and exact INT32 sums, recombined
 * in FP64.  The error of each element is at most
 * K (slices + 2) 2^(E_i + F_j - 7 slices), with 2^E_i, 2^F_j bounding the
 * rows of the two operands.  An extended form with no counterpart in the
 * reference; returns W2K_EINVAL for other values. */
int w2k_set_split_precision(w2k_ctx *ctx, int slices);

/* Library version string. */
const char *w2k_version(void);

/* Text of the last failed device call on this thread: the call or kernel,
 * the runtime's error string and the source position (empty when none).
 * Setting W2K_SYNC_CHECK=1 in the environment synchronises after every kernel
 * so that a device fault is attributed to the kernel that caused it. */
const char *w2k_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* W2K_LAPW1_H */
