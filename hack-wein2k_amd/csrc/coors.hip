// The basis of one k-point on the device: COORS (coors.f:47-201), every k+G
// with |k+G| <= RKM in order of length, and LOPW (lopw.f:40-128), the plane
// waves that carry the local orbitals.
//
// COORS walks the box -NVEC < J < NVEC in (JA, JB, JC) order and inserts each
// k+G inside the sphere into a list kept sorted by length, a later vector
// going after the earlier ones of equal length; the list holds NMATMAX+1-NLO
// vectors and, once it is full, is cut back to the last gap wider than 1e-5
// in length (coors.f:149-155).  Here every box point is a thread, the points
// inside the sphere are gathered (in any order) and each one's place in the
// list is its rank under the order (length, box index): the same list.  The
// lengths are formed with the reference's sequence of roundings (no fused
// multiply-adds), so XK, YK, ZK and RK are the reference's bit for bit.
//
// The box: the reference starts from NVEC = 25 and doubles an axis while a
// box corner on that axis lies closer than the last list entry
// (coors.f:172-230).  Here the box is sized from the reciprocal lattice so
// that it holds every k+G with |k+G| <= RKM (|J_i| <= |row i of BR2^-1|
// (RKM + |k|)).  Both boxes hold the whole sphere, so the lists agree.
//
// LOPW: each (atom, l, jlo) block of MULT (2l+1) local orbitals is one
// thread that repeats the reference's greedy modified Gram-Schmidt in its
// order: the vectors SF(atom) Y_lm(rotated k+G) of successive plane waves
// are orthogonalised against the ones already taken and a plane wave is
// taken when the residual norm exceeds 1e-3.  The blocks are independent (the
// search restarts at the first plane wave for every block), so they run side
// by side; only the choice of plane waves leaves the kernel.
#include <cmath>
#include <cstdio>
#include <vector>

#include "special.h"
#include "w2k_internal.h"

namespace w2k {

struct CoorsArgs {
  double b[9];  // BR2, column-major
  double sk[3];  // k in cartesian coordinates (coors.f:47-55)
  double s[3];   // k in the reciprocal basis (SX, SY, SZ)
  double pia[3];
  double rkm;
  int n1, n2, n3;  // box: J_i in [-n_i, n_i]
  int ortho, cxz;
};

// k+G of box point (ja, jb, jc) and its length, with the roundings of
// coors.f:81-84 (left to right, each product and sum rounded on its own)
__device__ inline double kplusg(const CoorsArgs &a, int ja, int jb, int jc, double sb[3])
{
#pragma clang fp contract(off)
  for (int d = 0; d < 3; d++) {
    double x = a.sk[d] + (double)ja * a.b[d];
    x = x + (double)jb * a.b[d + 3];
    sb[d] = x + (double)jc * a.b[d + 6];
  }
  const double q = sb[0] * sb[0] + sb[1] * sb[1];
  return __dsqrt_rn(q + sb[2] * sb[2]);
}

// one thread per box point: the points inside the sphere are appended (in
// any order) with their length and box index; the box index grows with
// (JA, JB, JC) in the reference's walk order
__global__ void k_coors_sphere(CoorsArgs a, long npts, double *__restrict__ len, int *__restrict__ idx,
                               int *__restrict__ count)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const int w2 = 2 * a.n2 + 1, w3 = 2 * a.n3 + 1;
  const int jc = (int)(t % w3) - a.n3;
  const long u = t / w3;
  const int jb = (int)(u % w2) - a.n2;
  const int ja = (int)(u / w2) - a.n1;
  double sb[3];
  const double r = kplusg(a, ja, jb, jc, sb);
  if (r <= a.rkm) {
    const int p = atomicAdd(count, 1);
    len[p] = r;
    idx[p] = (int)t;
  }
}

// place of every gathered vector in the list: its rank under (length, box
// index), counted against all others through LDS tiles; order[rank] = i
constexpr int RANK_TILE = 256;
__global__ __launch_bounds__(RANK_TILE) void k_coors_rank(int m, const double *__restrict__ len,
                                                          const int *__restrict__ idx, int *__restrict__ order)
{
  __shared__ double ls[RANK_TILE];
  __shared__ int is[RANK_TILE];
  const int i = blockIdx.x * RANK_TILE + threadIdx.x;
  const double li = i < m ? len[i] : 0.0;
  const int ii = i < m ? idx[i] : 0;
  int rank = 0;
  for (int j0 = 0; j0 < m; j0 += RANK_TILE) {
    __syncthreads();
    const int j = j0 + threadIdx.x;
    if (j < m) {
      ls[threadIdx.x] = len[j];
      is[threadIdx.x] = idx[j];
    }
    __syncthreads();
    const int cnt = min(RANK_TILE, m - j0);
    for (int q = 0; q < cnt; q++) {
      const double lq = ls[q];
      rank += (lq < li) || (lq == li && is[q] < ii);
    }
  }
  if (i < m) order[rank] = i;
}

// NV from the sorted lengths (coors.f:149-155): with a full list (m >= NDM1-1)
// the list is cut back to the last gap wider than 1e-5, searching from slot
// NDM1-1; slot NDM1 holds the sentinel 1e10 while the list has fewer than
// NDM1 entries.  out[0] = NV (0 if no gap is found)
__global__ void k_coors_cut(int m, int ndm1, const double *__restrict__ len, const int *__restrict__ order,
                            int *__restrict__ out)
{
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (m < ndm1 - 1) {
    out[0] = m;
    return;
  }
  auto slot = [&](int p) { return p < m ? len[order[p]] : 1.0e10; };  // 0-based slot p < NDM1
  int nv = 0;
  for (int n = ndm1 - 1; n >= 1; n--)
    if (fabs(slot(n - 1) - slot(n)) > 1.0e-5) {
      nv = n;
      break;
    }
  out[0] = nv;
}

// rows 0..n-1 of the basis: plane waves (p < nv) from the sorted list, KZZ by
// the rules of coors.f:123-140; RK of the local-orbital rows p >= nv is what
// the reference's list leaves in those slots (list entry, the sentinel 1e10
// at slot NDM1 of a list that never filled, else 0)
__global__ void k_coors_fill(CoorsArgs a, int nv, int n, int m, int ndm1, const double *__restrict__ len,
                             const int *__restrict__ idx, const int *__restrict__ order, int *__restrict__ kzz,
                             double *__restrict__ kv)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  if (p >= nv) {
    double r = 0.0;
    if (p < m && p < ndm1) r = len[order[p]];
    else if (p == ndm1 - 1) r = 1.0e10;
    kv[4 * p + 3] = r;
    return;
  }
  const int i = order[p];
  const long t = idx[i];
  const int w2 = 2 * a.n2 + 1, w3 = 2 * a.n3 + 1;
  const int jc = (int)(t % w3) - a.n3;
  const long u = t / w3;
  const int jb = (int)(u % w2) - a.n2;
  const int ja = (int)(u / w2) - a.n1;
  double sb[3];
  const double r = kplusg(a, ja, jb, jc, sb);
  kv[4 * p] = sb[0];
  kv[4 * p + 1] = sb[1];
  kv[4 * p + 2] = sb[2];
  kv[4 * p + 3] = r;
  int *k = kzz + 3 * p;
  if (a.ortho) {
    for (int d = 0; d < 3; d++) {
      const double q = sb[d] / a.pia[d] - a.s[d];
      k[d] = (int)(q + copysign(1.0e-2, q));
    }
  } else if (a.cxz) {
    k[0] = ja + jc;
    k[1] = jb;
    k[2] = -ja + jc;
  } else {
    k[0] = ja;
    k[1] = jb;
    k[2] = jc;
  }
}

// LOPW: one thread per (atom, l, jlo) block.  blk[b] = {atom type, l, first
// row, first equivalent atom, MULT, offset of its Gram-Schmidt columns in hh
// (complex entries)}.  err[0] is set when a block runs out of plane waves.
__global__ void k_lopw(int nblk, const int *__restrict__ blk, int nv, const double *__restrict__ pos,
                       const double *__restrict__ rotij, const double *__restrict__ rotloc, int *__restrict__ kzz,
                       double *__restrict__ kv, cd *__restrict__ hh, int *__restrict__ err)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const int *q = blk + 6 * b;
  const int a = q[0], l = q[1], koff = q[2], ia1 = q[3], mult = q[4];
  cd *H = hh + q[5];
  const int nl = 2 * l + 1, nbm = mult * nl;
  const double tpi = 8.0 * atan(1.0);
  int index = 0;
  for (int nb = 0; nb < nbm; nb++) {
    const int k = koff + nb;
    cd *h = H + (size_t)nb * nbm;
    for (;;) {
      index++;
      if (index > nv) {
        atomicExch(err, 1);
        return;
      }
      const int src = index - 1;
      const int k0 = kzz[3 * src], k1 = kzz[3 * src + 1], k2 = kzz[3 * src + 2];
      const double x = kv[4 * src], y = kv[4 * src + 1], z = kv[4 * src + 2];
      kzz[3 * k] = k0;
      kzz[3 * k + 1] = k1;
      kzz[3 * k + 2] = k2;
      kv[4 * k] = x;
      kv[4 * k + 1] = y;
      kv[4 * k + 2] = z;
      if (nbm == 1) break;
      const double rkgm = sqrt(x * x + y * y + z * z);
      for (int nx = 0; nx < mult; nx++) {
        const int e = ia1 + nx;
        const double arg = k0 * pos[3 * e] + k1 * pos[3 * e + 1] + k2 * pos[3 * e + 2];
        double sn, cs;
        sincos(tpi * arg, &sn, &cs);
        cd yl[16];
        if (rkgm <= 1.0e-5) {
          for (int m = 0; m < nl; m++) yl[l * l + m] = {0.0, 0.0};
          yl[0] = {1.0, 0.0};
        } else {
          const double v[3] = {x, y, z};
          double r1[3], r2[3];
          rotate3(v, rotij + 9 * e, r1);
          rotate3(r1, rotloc + 9 * a, r2);
          ylm(r2, l, yl);
        }
        for (int m = 0; m < nl; m++) h[nx * nl + m] = cmul({cs, sn}, yl[l * l + m]);
      }
      for (int j = 0; j < nb; j++) {
        const cd *hj = H + (size_t)j * nbm;
        cd cc = {0.0, 0.0};
        for (int i = 0; i < nbm; i++) {
          const cd t = cmul(h[i], conj(hj[i]));
          cc.re += t.re;
          cc.im += t.im;
        }
        for (int i = 0; i < nbm; i++) {
          const cd t = cmul(cc, hj[i]);
          h[i].re -= t.re;
          h[i].im -= t.im;
        }
      }
      double hl = 0.0;
      for (int i = 0; i < nbm; i++) hl += h[i].re * h[i].re + h[i].im * h[i].im;
      if (hl <= 1.0e-3) continue;
      const double sc = 1.0 / sqrt(hl);
      for (int i = 0; i < nbm; i++) {
        h[i].re *= sc;
        h[i].im *= sc;
      }
      break;
    }
  }
}

template <class... A>
static int fail(int code, const char *fmt, A... args)
{
  set_last_error(fmt, args...);
  return code;
}

// 3x3 inverse (column-major), by cofactors
static bool inv3(const double *m, double *r)
{
  auto M = [&](int i, int j) { return m[i + 3 * j]; };
  const double c00 = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
  const double c01 = M(1, 2) * M(2, 0) - M(1, 0) * M(2, 2);
  const double c02 = M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0);
  const double det = M(0, 0) * c00 + M(0, 1) * c01 + M(0, 2) * c02;
  if (!(std::fabs(det) > 0.0)) return false;
  const double id = 1.0 / det;
  // r = adj(m) / det, adj(m)(i,j) = cofactor(j,i)
  r[0 + 3 * 0] = c00 * id;
  r[1 + 3 * 0] = c01 * id;
  r[2 + 3 * 0] = c02 * id;
  r[0 + 3 * 1] = (M(0, 2) * M(2, 1) - M(0, 1) * M(2, 2)) * id;
  r[1 + 3 * 1] = (M(0, 0) * M(2, 2) - M(0, 2) * M(2, 0)) * id;
  r[2 + 3 * 1] = (M(0, 1) * M(2, 0) - M(0, 0) * M(2, 1)) * id;
  r[0 + 3 * 2] = (M(0, 1) * M(1, 2) - M(0, 2) * M(1, 1)) * id;
  r[1 + 3 * 2] = (M(0, 2) * M(1, 0) - M(0, 0) * M(1, 2)) * id;
  r[2 + 3 * 2] = (M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0)) * id;
  return true;
}

int coors(Ctx &c, const w2k_lattice *L, double sx, double sy, double sz, int *nv_out, int *nlo_out)
{
  RunDev &R = c.run();
  hipStream_t s = c.stream;
  if (!L || !(L->rkm > 0.0) || L->nmatmax < 1) return W2K_EINVAL;
  const int nlo = R.nlo_cols;
  const int ndm1 = L->nmatmax + 1 - nlo;
  if (ndm1 < 2) return fail(W2K_EINVAL, "COORS: NMATMAX %d leaves no room beside %d local orbitals",
                                 L->nmatmax, nlo);
  CoorsArgs a{};
  for (int i = 0; i < 9; i++) a.b[i] = L->br2[i];
  a.s[0] = sx;
  a.s[1] = sy;
  a.s[2] = sz;
  for (int d = 0; d < 3; d++) a.pia[d] = L->pia[d];
  if (L->ortho) {
    for (int d = 0; d < 3; d++) a.sk[d] = a.s[d] * L->pia[d];
  } else {
    for (int d = 0; d < 3; d++) a.sk[d] = sx * L->br2[d] + sy * L->br2[d + 3] + sz * L->br2[d + 6];
  }
  a.rkm = L->rkm;
  a.ortho = L->ortho != 0;
  a.cxz = L->cxz != 0;
  // box: |J_i| <= |row i of BR2^-1| (RKM + |k|), one point of margin
  double binv[9];
  if (!inv3(L->br2, binv)) return fail(W2K_EINVAL, "COORS: BR2 is singular");
  const double kl = std::sqrt(a.sk[0] * a.sk[0] + a.sk[1] * a.sk[1] + a.sk[2] * a.sk[2]);
  int nb[3];
  for (int i = 0; i < 3; i++) {
    const double rn = std::sqrt(binv[i] * binv[i] + binv[i + 3] * binv[i + 3] + binv[i + 6] * binv[i + 6]);
    const double e = rn * (L->rkm + kl);
    if (!(e < 4.0e3)) return fail(W2K_EINVAL, "COORS: box of %g points per axis", e);
    nb[i] = (int)e + 1;
  }
  a.n1 = nb[0];
  a.n2 = nb[1];
  a.n3 = nb[2];
  const long npts = (long)(2 * nb[0] + 1) * (2 * nb[1] + 1) * (2 * nb[2] + 1);
  if (npts > (1L << 30)) return fail(W2K_EINVAL, "COORS: box of %ld points", npts);
  // sphere points: at most the box, and (volume bound) far fewer; the buffers
  // are sized by the box so that no count can overrun them
  DBuf<double> len;
  DBuf<int> idx, order, cnt;
  int rc;
  if ((rc = len.alloc(npts)) || (rc = idx.alloc(npts)) || (rc = cnt.alloc(2))) return rc;
  W2K_HIP(hipMemsetAsync(cnt.p, 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(k_coors_sphere, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, s, a, npts, len.p, idx.p,
                     cnt.p);
  W2K_STAGE(s, "k_coors_sphere");
  int m = 0;
  W2K_HIP(hipMemcpyAsync(&m, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  if (m < 1) return fail(W2K_EINVAL, "COORS: no k+G inside RKM = %g", L->rkm);
  if ((rc = order.alloc(m))) return rc;
  hipLaunchKernelGGL(k_coors_rank, dim3((m + RANK_TILE - 1) / RANK_TILE), dim3(RANK_TILE), 0, s, m, len.p, idx.p,
                     order.p);
  W2K_STAGE(s, "k_coors_rank");
  hipLaunchKernelGGL(k_coors_cut, dim3(1), dim3(64), 0, s, m, ndm1, len.p, order.p, cnt.p + 1);
  W2K_STAGE(s, "k_coors_cut");
  int nv = 0;
  W2K_HIP(hipMemcpyAsync(&nv, cnt.p + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  if (nv < 1) return fail(W2K_EINVAL, "COORS: the list of %d vectors has no gap wider than 1e-5", m);
  // the basis buffers of the context (as w2k_set_basis), then the rows
  if ((rc = basis_buffers(c, nv, nlo))) return rc;
  const int n = nv + nlo;
  hipLaunchKernelGGL(k_coors_fill, dim3((n + 255) / 256), dim3(256), 0, s, a, nv, n, m, ndm1, len.p, idx.p, order.p,
                     c.d_kzz.p, c.d_kvec.p);
  W2K_STAGE(s, "k_coors_fill");
  if (nlo > 0) {
    // LOPW blocks in the reference's order: atom type, l, jlo
    std::vector<int> blk;
    long hoff = 0;
    int koff = nv;
    for (int at = 0; at < R.nat; at++) {
      const int mult = R.mult[at];
      for (int l = 0; l <= W2K_LOMAX; l++) {
        const int nil = R.ilo[l + W2K_LOMAXP * at];
        for (int jlo = 0; jlo < nil; jlo++) {
          const int nbm = mult * (2 * l + 1);
          if (hoff > (1L << 30)) return fail(W2K_EINVAL, "LOPW: Gram-Schmidt work too large");
          const int q[6] = {at, l, koff, R.first_eq[at], mult, (int)hoff};
          blk.insert(blk.end(), q, q + 6);
          hoff += (long)nbm * nbm;
          koff += nbm;
        }
      }
    }
    if (koff != n) return fail(W2K_EINVAL, "LOPW: %d local-orbital rows, the run has %d", koff - nv, nlo);
    const int nblk = (int)(blk.size() / 6);
    DBuf<int> dblk;
    DBuf<double> dhh;
    if ((rc = dblk.alloc(blk.size())) || (rc = dhh.alloc(2 * (size_t)std::max(hoff, 1L)))) return rc;
    W2K_HIP(hipMemcpyAsync(dblk.p, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice, s));
    W2K_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), s));
    hipLaunchKernelGGL(k_lopw, dim3((nblk + 63) / 64), dim3(64), 0, s, nblk, dblk.p, nv, R.d_pos.p, R.d_rotij.p,
                       R.d_rotloc.p, c.d_kzz.p, c.d_kvec.p, reinterpret_cast<cd *>(dhh.p), cnt.p);
    W2K_STAGE(s, "k_lopw");
    int err = 0;
    W2K_HIP(hipMemcpyAsync(&err, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
    W2K_HIP(hipStreamSynchronize(s));
    if (err) return fail(W2K_EINVAL, "LOPW: plane waves exhausted (NV = %d)", nv);
  }
  c.h_kzz.resize((size_t)3 * n);
  W2K_HIP(hipMemcpyAsync(c.h_kzz.data(), c.d_kzz.p, sizeof(int) * 3 * n, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  c.have_basis = 1;
  if (nv_out) *nv_out = nv;
  if (nlo_out) *nlo_out = nlo;
  return 0;
}

int get_basis(Ctx &c, int *kzz, double *xk, double *yk, double *zk, double *rk)
{
  if (!c.have_basis) return W2K_ESTATE;
  const int n = c.n;
  hipStream_t s = c.stream;
  if (kzz) W2K_HIP(hipMemcpyAsync(kzz, c.d_kzz.p, sizeof(int) * 3 * n, hipMemcpyDeviceToHost, s));
  std::vector<double> kv;
  if (xk || yk || zk || rk) {
    kv.resize((size_t)4 * n);
    W2K_HIP(hipMemcpyAsync(kv.data(), c.d_kvec.p, sizeof(double) * 4 * n, hipMemcpyDeviceToHost, s));
  }
  W2K_HIP(hipStreamSynchronize(s));
  double *dst[4] = {xk, yk, zk, rk};
  for (int d = 0; d < 4; d++)
    if (dst[d])
      for (int i = 0; i < n; i++) dst[d][i] = kv[4 * (size_t)i + d];
  return 0;
}

}  // namespace w2k
