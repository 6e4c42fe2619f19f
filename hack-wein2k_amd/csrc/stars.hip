// Interstitial potential stars on the device: STERN (stern.F:1-97) and the
// box expansion of SETWAR (setwar.f:40-74), and the band-energy gradient
// with respect to the star coefficients POTK (an extended form with no
// counterpart in the reference).
//
// A star s is the orbit of its representative G_s = (JA, JB, JC) under the
// IORD operations of the space group: member K = IMAT_i G_s, with the phase
// TAUP(K) = mean over the operations that map G_s to K of exp(2 pi i TAU_i.G_s)
// (the real build keeps the cosine, stern.F:61-63,73-75).  SETWAR writes
// WARP(K) = POTK_s conj(TAUP(K)) for every member; HAMILT reads that box.
//
// Gradient: with g(K) = dE/dRe WARP(K) + i dE/dIm WARP(K) from the warped-
// box adjoint, the chain rule through WARP = POTK conj(TAUP) gives
//   dE/dRe POTK_s + i dE/dIm POTK_s = sum over the members K of g(K) TAUP(K)
// (real build: the real part; its box entries are read through their real
// part, so dE/dIm POTK = 0 there).
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "w2k_internal.h"

namespace w2k {

// STERN for every star: one thread per star, members in the order the
// operations produce them; the box extent (largest |K_d|) by atomicMax
__global__ void k_stern(int nkk, const int *__restrict__ ja, const int *__restrict__ jb, const int *__restrict__ jc,
                        int iord, const int *__restrict__ imat, const double *__restrict__ tau, int cpx,
                        int *__restrict__ kkk, double *__restrict__ taup, int *__restrict__ nst,
                        int *__restrict__ ext)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nkk) return;
  const double tpi = 6.283185307179586476925286766559;
  const int g[3] = {ja[s], jb[s], jc[s]};
  int *K = kkk + (size_t)s * iord * 3;
  double *T = taup + (size_t)s * iord * 2;
  int cnt[64];  // operations mapped onto each member (IORD <= 48)
  int n = 0;
  for (int i = 0; i < iord; i++) {
    double tk = 0.0;
    int k[3];
    for (int j = 0; j < 3; j++) {
      tk += tau[j + 3 * i] * g[j] * tpi;
      int v = 0;
      for (int l = 0; l < 3; l++) v += imat[j + 3 * (l + 3 * i)] * g[l];
      k[j] = v;
    }
    int m = 0;
    for (; m < n; m++)
      if (K[3 * m] == k[0] && K[3 * m + 1] == k[1] && K[3 * m + 2] == k[2]) break;
    const double c = cos(tk), sn = cpx ? sin(tk) : 0.0;
    if (m < n) {
      T[2 * m] += c;
      T[2 * m + 1] += sn;
      cnt[m]++;
    } else {
      K[3 * n] = k[0]; K[3 * n + 1] = k[1]; K[3 * n + 2] = k[2];
      T[2 * n] = c;
      T[2 * n + 1] = sn;
      cnt[n] = 1;
      n++;
    }
  }
  for (int m = 0; m < n; m++) {
    T[2 * m] /= cnt[m];
    T[2 * m + 1] /= cnt[m];
    for (int d = 0; d < 3; d++) atomicMax(ext + d, abs(K[3 * m + d]));
  }
  nst[s] = n;
}

__global__ void k_zero(size_t cnt, double *__restrict__ p)
{
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = 0.0;
}

// SETWAR's expansion: WARP(K) = POTK_s conj(TAUP(K)), one thread per
// (star, operation slot); the stars' members are distinct, so no two
// threads write one entry
__global__ void k_setwar(int nkk, int iord, const int *__restrict__ kkk, const double *__restrict__ taup,
                         const int *__restrict__ nst, const double *__restrict__ potk, int w1, int w2, int w3,
                         double *__restrict__ warp)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = t / iord, m = t % iord;
  if (s >= nkk || m >= nst[s]) return;
  const int *K = kkk + ((size_t)s * iord + m) * 3;
  const double tr = taup[((size_t)s * iord + m) * 2], ti = taup[((size_t)s * iord + m) * 2 + 1];
  const double pr = potk[2 * s], pi = potk[2 * s + 1];
  const long b = (K[0] + w1) + (long)(2 * w1 + 1) * ((K[1] + w2) + (long)(2 * w2 + 1) * (K[2] + w3));
  warp[2 * b] = pr * tr + pi * ti;
  warp[2 * b + 1] = pi * tr - pr * ti;
}

// dE/dPOTK_s = sum over the members of g(K) TAUP(K), in member order
__global__ void k_star_grad(int nkk, int iord, const int *__restrict__ kkk, const double *__restrict__ taup,
                            const int *__restrict__ nst, int w1, int w2, int w3, int cpx,
                            const double *__restrict__ g, double *__restrict__ out)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nkk) return;
  double ar = 0.0, ai = 0.0;
  for (int m = 0; m < nst[s]; m++) {
    const int *K = kkk + ((size_t)s * iord + m) * 3;
    const double tr = taup[((size_t)s * iord + m) * 2], ti = taup[((size_t)s * iord + m) * 2 + 1];
    const long b = (K[0] + w1) + (long)(2 * w1 + 1) * ((K[1] + w2) + (long)(2 * w2 + 1) * (K[2] + w3));
    const double gr = g[2 * b], gi = g[2 * b + 1];
    ar += gr * tr - gi * ti;
    ai += gr * ti + gi * tr;
  }
  out[2 * s] = ar;
  out[2 * s + 1] = cpx ? ai : 0.0;
}

template <class T>
static int put(DBuf<T> &b, const T *src, size_t count, hipStream_t s)
{
  int rc = b.alloc(count > 0 ? count : 1);
  if (rc) return rc;
  if (count) W2K_HIP(hipMemcpyAsync(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice, s));
  return 0;
}

int set_stars(Ctx &c, const w2k_stars *st)
{
  RunDev &R = c.run();
  hipStream_t s = c.stream;
  if (!st || st->nkk < 1 || st->iord < 1 || st->iord > 48 || !st->ja || !st->jb || !st->jc || !st->potk ||
      !st->imat || !st->tau)
    return W2K_EINVAL;
  if (R.ins != 0) return W2K_ESTATE;  // no warped interstitial without a vns file (INS /= 0)
  const int nkk = st->nkk, iord = st->iord, cpx = R.is_complex;
  int rc;
  DBuf<int> dja, djb, djc, dimat, dext;
  DBuf<double> dtau;
  if ((rc = put(dja, st->ja, nkk, s)) || (rc = put(djb, st->jb, nkk, s)) || (rc = put(djc, st->jc, nkk, s)) ||
      (rc = put(dimat, st->imat, (size_t)9 * iord, s)) || (rc = put(dtau, st->tau, (size_t)3 * iord, s)) ||
      (rc = put(R.d_potk, st->potk, (size_t)2 * nkk, s)))
    return rc;
  if ((rc = R.d_kkk.alloc((size_t)nkk * iord * 3)) || (rc = R.d_taup.alloc((size_t)nkk * iord * 2)) ||
      (rc = R.d_nst.alloc(nkk)) || (rc = dext.alloc(3)))
    return rc;
  W2K_HIP(hipMemsetAsync(dext.p, 0, sizeof(int) * 3, s));
  hipLaunchKernelGGL(k_stern, dim3((nkk + 63) / 64), dim3(64), 0, s, nkk, dja.p, djb.p, djc.p, iord, dimat.p, dtau.p,
                     cpx, R.d_kkk.p, R.d_taup.p, R.d_nst.p, dext.p);
  W2K_STAGE(s, "k_stern");
  int ext[3];
  W2K_HIP(hipMemcpyAsync(ext, dext.p, sizeof ext, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  R.nkk = nkk;
  R.iord = iord;
  R.warp1 = ext[0];
  R.warp2 = ext[1];
  R.warp3 = ext[2];
  const size_t nbox = (size_t)(2 * ext[0] + 1) * (2 * ext[1] + 1) * (2 * ext[2] + 1);
  if ((rc = R.d_warp.alloc(2 * nbox))) return rc;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((2 * nbox + 255) / 256)), dim3(256), 0, s, 2 * nbox, R.d_warp.p);
  W2K_STAGE(s, "k_zero");
  const int nt = nkk * iord;
  hipLaunchKernelGGL(k_setwar, dim3((nt + 255) / 256), dim3(256), 0, s, nkk, iord, R.d_kkk.p, R.d_taup.p, R.d_nst.p,
                     R.d_potk.p, ext[0], ext[1], ext[2], R.d_warp.p);
  W2K_STAGE(s, "k_setwar");
  W2K_HIP(hipStreamSynchronize(s));
  return 0;
}

int adjoint_stars(Ctx &c, const double *f_host, double *e_band, double *g_potk)
{
  RunDev &R = c.run();
  if (R.nkk < 1) return W2K_ESTATE;  // no stars set on this run
  const double *box = nullptr;
  int rc = adjoint_warp_box(c, f_host, e_band, g_potk ? &box : nullptr);
  if (rc || !g_potk) return rc;
  if (!box) {
    std::fill(g_potk, g_potk + 2 * R.nkk, 0.0);
    return 0;
  }
  hipStream_t s = c.stream;
  if ((rc = c.w_v.alloc((size_t)2 * R.nkk))) return rc;
  hipLaunchKernelGGL(k_star_grad, dim3((R.nkk + 63) / 64), dim3(64), 0, s, R.nkk, R.iord, R.d_kkk.p, R.d_taup.p,
                     R.d_nst.p, R.warp1, R.warp2, R.warp3, R.is_complex, box, c.w_v.p);
  W2K_STAGE(s, "k_star_grad");
  W2K_HIP(hipMemcpyAsync(g_potk, c.w_v.p, sizeof(double) * 2 * R.nkk, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace w2k
