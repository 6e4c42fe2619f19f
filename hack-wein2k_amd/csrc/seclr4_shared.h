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

// v from the partner lane of a halving step with stride S: any lane that has
// bit S flipped and the bits above S unchanged will do, so the short strides
// use row-local DPP moves (S = 1, 2: quad permutes; S = 4, 8: half-row and
// row mirrors), S = 16 a swizzle within 32 lanes and S = 32 a bpermute.
template <int S>
__device__ __forceinline__ double bfly_partner(double v)
{
  if constexpr (S == 32) {
    return __shfl_xor(v, 32, 64);
  } else {
    const long long b = __double_as_longlong(v);
    int lo = (int)(unsigned)(b & 0xffffffffll), hi = (int)(unsigned)((unsigned long long)b >> 32);
    if constexpr (S == 16) {
      lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
      hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
    } else {
      constexpr int ctrl = S == 1 ? 0xB1 : S == 2 ? 0x4E : S == 4 ? 0x141 : 0x140;
      lo = __builtin_amdgcn_update_dpp(0, lo, ctrl, 0xF, 0xF, false);
      hi = __builtin_amdgcn_update_dpp(0, hi, ctrl, 0xF, 0xF, false);
    }
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
  }
}

// One halving step of a wave-wide reduce-scatter: lanes with bit S set keep
// entries S..2S-1, the others keep 0..S-1, and each adds the half that its
// partner lane gives away.
template <int S>
__device__ __forceinline__ void bfly_step(double *t, int lane)
{
  const bool hi = (lane & S) != 0;
#pragma unroll
  for (int j = 0; j < S; j++) {
    const double keep = hi ? t[j + S] : t[j];
    const double send = hi ? t[j] : t[j + S];
    t[j] = keep + bfly_partner<S>(send);
  }
}

// t[0..63] per lane -> returns, in lane l, the sum over the 64 lanes of t[l]
// (a transpose-and-sum in six halving steps, no LDS).  Overwrites t.
__device__ __forceinline__ double wave_transpose_sum(double *t, int lane)
{
  bfly_step<32>(t, lane);
  bfly_step<16>(t, lane);
  bfly_step<8>(t, lane);
  bfly_step<4>(t, lane);
  bfly_step<2>(t, lane);
  bfly_step<1>(t, lane);
  return t[0];
}

// t[0..31] per lane -> returns, in lane l, the sum over the 64 lanes of
// t[l % 32] (five halving steps within each 32-lane half, then the halves
// added).  Overwrites t.
__device__ __forceinline__ double wave_transpose_sum32(double *t, int lane)
{
  bfly_step<16>(t, lane);
  bfly_step<8>(t, lane);
  bfly_step<4>(t, lane);
  bfly_step<2>(t, lane);
  bfly_step<1>(t, lane);
  return t[0] + __shfl_xor(t[0], 32, 64);
}

// t[0..15] per lane -> returns, in lane l, the sum over the 64 lanes of
// t[l % 16] (four halving steps within each 16-lane group, then the groups
// added).  Overwrites t.
__device__ __forceinline__ double wave_transpose_sum16(double *t, int lane)
{
  bfly_step<8>(t, lane);
  bfly_step<4>(t, lane);
  bfly_step<2>(t, lane);
  bfly_step<1>(t, lane);
  const double s = t[0] + __shfl_xor(t[0], 16, 64);
  return s + __shfl_xor(s, 32, 64);
}

// One halving step with the lane bit S of a permlane swap (S = 16: gfx950
// v_permlane16_swap, S = 32: v_permlane32_swap): lanes with bit S clear keep
// entry x summed with their partner's x, the others keep y summed with their
// partner's y.  Swapping x and y across the two lane groups leaves (x,
// partner's x) in the first group and (partner's y, y) in the second, so the
// sum of the two swapped registers is the step's result in every lane: no
// selects and no LDS.
template <int S>
__device__ __forceinline__ double swap_sum(double x, double y)
{
  const long long bx = __double_as_longlong(x), by = __double_as_longlong(y);
  const unsigned xl = (unsigned)(bx & 0xffffffffll), xh = (unsigned)((unsigned long long)bx >> 32);
  const unsigned yl = (unsigned)(by & 0xffffffffll), yh = (unsigned)((unsigned long long)by >> 32);
  unsigned a0, a1, b0, b1;
  if constexpr (S == 16) {
    const auto a = __builtin_amdgcn_permlane16_swap(xl, yl, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(xh, yh, false, false);
    a0 = a[0]; a1 = a[1]; b0 = b[0]; b1 = b[1];
  } else {
    static_assert(S == 32, "lane bit 4 or 5");
    const auto a = __builtin_amdgcn_permlane32_swap(xl, yl, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(xh, yh, false, false);
    a0 = a[0]; a1 = a[1]; b0 = b[0]; b1 = b[1];
  }
  const double u = __longlong_as_double((long long)(((unsigned long long)b0 << 32) | a0));
  const double v = __longlong_as_double((long long)(((unsigned long long)b1 << 32) | a1));
  return u + v;
}

// v of the lane that differs in bit 5
__device__ __forceinline__ double xor32(double v, int lane)
{
  const long long b = __double_as_longlong(v);
  const unsigned lo = (unsigned)(b & 0xffffffffll), hi = (unsigned)((unsigned long long)b >> 32);
  const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  const auto c = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  const bool up = (lane & 32) != 0;
  const unsigned rl = up ? a[0] : a[1], rh = up ? c[0] : c[1];
  return __longlong_as_double((long long)(((unsigned long long)rh << 32) | rl));
}

// halving step on lane bit LB over entries j, j + H (j < H): DPP partner
template <int LB, int H>
__device__ __forceinline__ void bfly_step_h(double *t, int lane)
{
  const bool hi = (lane & LB) != 0;
#pragma unroll
  for (int j = 0; j < H; j++) {
    const double keep = hi ? t[j + H] : t[j];
    const double send = hi ? t[j] : t[j + H];
    t[j] = keep + bfly_partner<LB>(send);
  }
}

// t[0..31] per lane -> returns, in lanes 2e and 2e + 1, the sum over the 64
// lanes of t[e]: halving steps on lane bits 5 and 4 by permlane swaps, then
// 3, 2, 1 by DPP, and the pair added.  Overwrites t.
__device__ __forceinline__ double wave_transpose_sum32x(double *t, int lane)
{
#pragma unroll
  for (int j = 0; j < 16; j++) t[j] = swap_sum<32>(t[j], t[j + 16]);
#pragma unroll
  for (int j = 0; j < 8; j++) t[j] = swap_sum<16>(t[j], t[j + 8]);
  bfly_step_h<8, 4>(t, lane);
  bfly_step_h<4, 2>(t, lane);
  bfly_step_h<2, 1>(t, lane);
  return t[0] + bfly_partner<1>(t[0]);
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
  double *part;             // SYMV partials: 128 (complex 256) per tile of the lower triangle
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
