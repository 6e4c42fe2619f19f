// Pieces of the device SECLR4 shared by the real (seclr4.hip) and complex
// (seclr4_cplx.hip) drivers: block sizes, the work-buffer layout and the
// stages that act on the real tridiagonal matrix (bisection, inverse
// iteration, cluster orthogonalisation, the ascending sort).
#pragma once

#include "w2k_internal.h"

namespace w2k {

constexpr int NB = 64;      // Cholesky / inversion block
constexpr int TNB = 32;     // tridiagonalisation panel
constexpr int SYMV_T = 64;  // tile of the symmetric / Hermitian SYMV (one wave)

__device__ __forceinline__ double wave_sum(double v)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// device work buffers of one SECLR4 call (carved from Ctx::w_small / w_int);
// complex quantities are planar: real parts, then imaginary parts at +stride
struct Seclr4Bufs {
  int n = 0, nt = 0;        // order, SYMV tiles per dimension
  int cplx = 0;
  double *d, *e, *e2;       // tridiagonal (real in both builds), squared off-diagonals
  double *tau, *scal;       // reflector scalars (complex: re[n], im[n])
  double *wv;               // eigenvalues (block order, then sorted)
  double *P;                // panel buffer n x 3 TNB (complex: twice)
  double *part;             // SYMV partials nt x n (complex: twice)
  double *tdots;            // nt x 2 TNB (complex: twice)
  double *tvy;              // nt (nt + 1) / 2
  double *tnorm;            // nt
  double *bgl, *bgu, *batol, *bnrm, *xs, *pivmin;
  int *info, *cnt, *ticket, *tblk, *tidx, *bstart, *bend, *cstart, *perm;
};

// allocate and carve the work buffers; zero INFO, counts and tickets
int seclr4_buffers(Ctx &c, int n, int cplx, Seclr4Bufs &B);

// bisection (window (el, eu], NUME cap, NBELW) on B.d/B.e, inverse iteration
// and orthogonalisation of close eigenvalues: m eigenvalues in B.wv (block
// order) and their tridiagonal eigenvectors in Zt (n x m, real).  Zt must
// hold n * n doubles.  Allocates c.d_eig for m values.
int tridiag_eigen(Ctx &c, Seclr4Bufs &B, double el, double eu, int nume, double *Zt, int *m_out, int *nbelw_out);

// selection sort of dsyevx2.f:381-400 on the host: B.wv ascending, B.perm the
// permutation applied to the eigenvector columns
int sort_eigen(Ctx &c, Seclr4Bufs &B, int m);

int seclr4_complex(Ctx &c, double el, double eu, int nume, int *ne, int *nbelw);

}  // namespace w2k
