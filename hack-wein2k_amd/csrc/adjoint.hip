// Band-energy adjoint at one k-point (an extended form with no counterpart
// in the reference): E_band = sum_n f_n eps_n with the weights f_n held fixed.
// By Hellmann-Feynman, dE_band/dp = sum_n f_n z_n^H (dH/dp - eps_n dS/dp) z_n
// for S-normalised eigenvectors, and S does not depend on the potential, so
// only dH/dp enters.
//
// Interstitial part.  HAMILT puts the warped potential into the plane-wave
// block as H(J,I) = ... + WARP(KZZ_J - KZZ_I) for J <= I (hamilt.F:335-345,
// WARPIN = 0 outside the box, warpin.f:31-37); the real build takes its real
// part.  The solved Hermitian matrix has that value in its upper triangle
// (the stored lower triangle of HS holds it at (I,J)), so with
//   Gamma(I,J) = sum_n f_n z(I,n) conj z(J,n)
// the gradient of E_band with respect to the box entry WARP(Q), written as
// dE/dRe W + i dE/dIm W, is
//   g(Q) = sum over I > J with KZZ_J - KZZ_I = Q of 2 conj Gamma(I,J)
//          (+ sum_I Gamma(I,I) for Q = 0, real part only),
// and in the real build the real part of that (dE/dIm W = 0).  Gamma comes
// from GEMMs on the lower triangle; each box entry sums its pairs in a fixed
// order through an index grid over the basis' KZZ (no atomics).
#include <algorithm>
#include <vector>

#include "w2k_internal.h"

namespace w2k {

__global__ void k_adj_fill_int(int *__restrict__ p, int cnt, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) p[i] = v;
}

// planar copies of the first rows of Z and of Z diag(f)
__global__ void k_adj_zf(int rows, int n, int ne, int cpx, const double *__restrict__ z, const double *__restrict__ f,
                         double *__restrict__ zr, double *__restrict__ zi, double *__restrict__ fzr,
                         double *__restrict__ fzi)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= rows || k >= ne) return;
  const size_t s = (size_t)i + (size_t)n * k, d = (size_t)i + (size_t)rows * k;
  const double w = f[k];
  if (cpx) {
    const double re = z[2 * s], im = z[2 * s + 1];
    zr[d] = re;
    zi[d] = im;
    fzr[d] = w * re;
    fzi[d] = w * im;
  } else {
    const double re = z[s];
    zr[d] = re;
    fzr[d] = w * re;
  }
}

// index grid over the plane-wave KZZ: idx[(x-lo) + dx((y-lo) + dy(z-lo))] = I
__global__ void k_adj_kzz_index(int nv, const int *__restrict__ kzz, int lx, int ly, int lz, int dx, int dy,
                                int *__restrict__ idx)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nv) return;
  const int x = kzz[3 * i] - lx, y = kzz[3 * i + 1] - ly, z = kzz[3 * i + 2] - lz;
  idx[x + dx * (y + dy * z)] = i;
}

// one thread per box entry Q: its (I, J) pairs summed in order of I
__global__ void k_adj_warp_grad(int nv, const int *__restrict__ kzz, const int *__restrict__ idx, int lx, int ly,
                                int lz, int dx, int dy, int dz, int w1, int w2, int w3, int cpx,
                                const double *__restrict__ gr, const double *__restrict__ gi, double *__restrict__ g)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int n1 = 2 * w1 + 1, n2 = 2 * w2 + 1, n3 = 2 * w3 + 1;
  if (b >= n1 * n2 * n3) return;
  const int qx = b % n1 - w1, qy = (b / n1) % n2 - w2, qz = b / (n1 * n2) - w3;
  double ar = 0.0, ai = 0.0;
  for (int I = 0; I < nv; I++) {
    const int x = kzz[3 * I] + qx - lx, y = kzz[3 * I + 1] + qy - ly, z = kzz[3 * I + 2] + qz - lz;
    if (x < 0 || y < 0 || z < 0 || x >= dx || y >= dy || z >= dz) continue;
    const int J = idx[x + dx * (y + dy * z)];
    if (J < 0 || J > I) continue;
    const size_t o = (size_t)I + (size_t)nv * J;
    if (J < I) {
      ar += 2.0 * gr[o];
      if (cpx) ai -= 2.0 * gi[o];
    } else {
      ar += gr[o];  // diagonal: the solver reads its real part only
    }
  }
  g[2 * b] = ar;
  g[2 * b + 1] = ai;
}

int adjoint_gamma(Ctx &c, const double *d_f, int rows, double *Gr, double *Gi)
{
  const int n = c.n, ne = c.ne;
  const int cpx = c.run().is_complex;
  hipStream_t s = c.stream;
  int rc;
  const size_t zsz = (size_t)rows * ne;
  if ((rc = c.w_x.alloc(4 * zsz))) return rc;
  double *zr = c.w_x.p, *zi = zr + zsz, *fzr = zi + zsz, *fzi = fzr + zsz;
  hipLaunchKernelGGL(k_adj_zf, dim3((rows + 255) / 256, ne), dim3(256), 0, s, rows, n, ne, cpx, c.d_z.p, d_f, zr, zi,
                     fzr, fzi);
  W2K_STAGE(s, "k_adj_zf");
  // real:    Gamma = (Z F) Z^T
  // complex: Re = Zr F Zr^T + Zi F Zi^T,  Im = Zi F Zr^T - Zr F Zi^T
  if ((rc = dgemm(s, 'N', 'T', rows, rows, ne, 1.0, fzr, rows, zr, rows, 0.0, Gr, rows, 1))) return rc;
  if (cpx) {
    if ((rc = dgemm(s, 'N', 'T', rows, rows, ne, 1.0, fzi, rows, zi, rows, 1.0, Gr, rows, 1))) return rc;
    if ((rc = dgemm(s, 'N', 'T', rows, rows, ne, 1.0, fzi, rows, zr, rows, 0.0, Gi, rows, 1))) return rc;
    if ((rc = dgemm(s, 'N', 'T', rows, rows, ne, -1.0, fzr, rows, zi, rows, 1.0, Gi, rows, 1))) return rc;
  }
  return 0;
}

int adjoint_warp(Ctx &c, const double *f_host, double *e_band, double *g_warp)
{
  const double *box = nullptr;
  int rc = adjoint_warp_box(c, f_host, e_band, g_warp ? &box : nullptr);
  if (rc || !g_warp) return rc;
  RunDev &R = c.run();
  const size_t nbox = (size_t)(2 * R.warp1 + 1) * (2 * R.warp2 + 1) * (2 * R.warp3 + 1);
  if (!box) {
    std::fill(g_warp, g_warp + 2 * nbox, 0.0);
    return 0;
  }
  W2K_HIP(hipMemcpyAsync(g_warp, box, sizeof(double) * 2 * nbox, hipMemcpyDeviceToHost, c.stream));
  W2K_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

int adjoint_warp_box(Ctx &c, const double *f_host, double *e_band, const double **d_box)
{
  RunDev &R = c.run();
  const int n = c.n, nv = c.nv, ne = c.ne;
  const int cpx = R.is_complex;
  hipStream_t s = c.stream;
  const int w1 = R.warp1, w2 = R.warp2, w3 = R.warp3;
  const size_t nbox = (size_t)(2 * w1 + 1) * (2 * w2 + 1) * (2 * w3 + 1);
  int rc;
  // E_band from the eigenvalues on the device
  std::vector<double> ev(ne > 0 ? ne : 1);
  if (ne > 0) W2K_HIP(hipMemcpyAsync(ev.data(), c.d_eig.p, sizeof(double) * ne, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  double e = 0.0;
  for (int k = 0; k < ne; k++) e += f_host[k] * ev[k];
  if (e_band) *e_band = e;
  if (!d_box) return 0;
  *d_box = nullptr;  // no box: every gradient is zero
  if (ne == 0 || nv == 0 || R.ins != 0) return 0;
  // Gamma on the lower triangle of the plane-wave block
  const size_t gsz = (size_t)nv * nv;
  if ((rc = c.w_a.alloc(gsz)) || (cpx && (rc = c.w_b.alloc(gsz)))) return rc;
  double *Gr = c.w_a.p, *Gi = cpx ? c.w_b.p : nullptr;
  if ((rc = c.w_v.alloc((size_t)ne))) return rc;
  W2K_HIP(hipMemcpyAsync(c.w_v.p, f_host, sizeof(double) * ne, hipMemcpyHostToDevice, s));
  if ((rc = adjoint_gamma(c, c.w_v.p, nv, Gr, Gi))) return rc;
  // KZZ index grid
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    lo[d] = hi[d] = c.h_kzz[d];
    for (int i = 1; i < nv; i++) {
      lo[d] = std::min(lo[d], c.h_kzz[3 * i + d]);
      hi[d] = std::max(hi[d], c.h_kzz[3 * i + d]);
    }
  }
  const int dx = hi[0] - lo[0] + 1, dy = hi[1] - lo[1] + 1, dz = hi[2] - lo[2] + 1;
  const size_t ngrid = (size_t)dx * dy * dz;
  if ((rc = c.w_int.alloc(ngrid))) return rc;
  hipLaunchKernelGGL(k_adj_fill_int, dim3((unsigned)((ngrid + 255) / 256)), dim3(256), 0, s, c.w_int.p, (int)ngrid, -1);
  W2K_STAGE(s, "k_adj_fill_int");
  hipLaunchKernelGGL(k_adj_kzz_index, dim3((nv + 255) / 256), dim3(256), 0, s, nv, c.d_kzz.p, lo[0], lo[1], lo[2], dx,
                     dy, c.w_int.p);
  W2K_STAGE(s, "k_adj_kzz_index");
  if ((rc = c.w_c.alloc(2 * nbox))) return rc;
  hipLaunchKernelGGL(k_adj_warp_grad, dim3((unsigned)((nbox + 127) / 128)), dim3(128), 0, s, nv, c.d_kzz.p,
                     c.w_int.p, lo[0], lo[1], lo[2], dx, dy, dz, w1, w2, w3, cpx, Gr, cpx ? Gi : Gr, c.w_c.p);
  W2K_STAGE(s, "k_adj_warp_grad");
  *d_box = c.w_c.p;
  return 0;
}

}  // namespace w2k
