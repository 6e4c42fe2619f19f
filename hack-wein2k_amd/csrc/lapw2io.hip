// What lapw2 reads from a solved k-point (extended forms with no counterpart
// in lapw1's hot path):
//  - the vector file record of the k-point exactly as TAPEWF writes it
//    (tapewf.F:45-56, called from prtkpt.F:163-166): Fortran unformatted
//    sequential records, each framed by 4-byte length markers,
//      (SX, SY, SZ, KNAME, NV+NLO, NE, WEIGHT, IPGR)
//      (KZZ(1:3, I), I = 1, NV+NLO)
//      per eigenpair: (I, E(I)) and (Z(1:NV+NLO, I)), conjugated in the
//      complex build;
//  - the interstitial band density in reciprocal space, as lapw2's fourir.F
//    forms it from |psi_n(r)|^2 on an FFT grid, here as the pair sum
//      rho(Q) = sum_n f_n sum over plane waves I, J with KZZ_J - KZZ_I = Q
//               of conj z(I,n) z(J,n),
//    from the warped-box gradient g of the band-energy adjoint (HAMILT fills
//    one triangle with WARP(KZZ_J - KZZ_I), so g(Q) + conj g(-Q) = 2 rho(Q)
//    for Q /= 0 and g(0) = rho(0)).
#include <cstdio>
#include <cstring>
#include <vector>

#include "w2k_internal.h"

namespace w2k {

// rho on the run's box from the adjoint's g (both interleaved complex)
__global__ void k_density_from_g(int w1, int w2, int w3, int cpx, const double *__restrict__ g,
                                 double *__restrict__ rho)
{
  const int n1 = 2 * w1 + 1, n2 = 2 * w2 + 1, n3 = 2 * w3 + 1;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n1 * n2 * n3) return;
  const int qx = b % n1 - w1, qy = (b / n1) % n2 - w2, qz = b / (n1 * n2) - w3;
  const int m = (-qx + w1) + n1 * ((-qy + w2) + n2 * (-qz + w3));
  double rr, ri;
  if (qx == 0 && qy == 0 && qz == 0) {
    rr = g[2 * b];
    ri = 0.0;
  } else {
    rr = 0.5 * (g[2 * b] + g[2 * m]);
    ri = cpx ? 0.5 * (g[2 * b + 1] - g[2 * m + 1]) : 0.0;
  }
  rho[2 * b] = rr;
  rho[2 * b + 1] = ri;
}

int density_interstitial(Ctx &c, const double *f_host, double *rho_host)
{
  RunDev &R = c.run();
  const size_t nbox = (size_t)(2 * R.warp1 + 1) * (2 * R.warp2 + 1) * (2 * R.warp3 + 1);
  const double *box = nullptr;
  int rc = adjoint_warp_box(c, f_host, nullptr, &box);
  if (rc) return rc;
  if (!box) {
    std::memset(rho_host, 0, sizeof(double) * 2 * nbox);
    return 0;
  }
  hipStream_t s = c.stream;
  if ((rc = c.w_v.alloc(2 * nbox))) return rc;
  hipLaunchKernelGGL(k_density_from_g, dim3((unsigned)((nbox + 255) / 256)), dim3(256), 0, s, R.warp1, R.warp2,
                     R.warp3, R.is_complex, box, c.w_v.p);
  W2K_STAGE(s, "k_density_from_g");
  W2K_HIP(hipMemcpyAsync(rho_host, c.w_v.p, sizeof(double) * 2 * nbox, hipMemcpyDeviceToHost, s));
  W2K_HIP(hipStreamSynchronize(s));
  return 0;
}

// one Fortran unformatted sequential record: length, payload, length
static bool put_record(FILE *fp, const void *data, size_t bytes)
{
  const unsigned int len = (unsigned int)bytes;
  return fwrite(&len, 4, 1, fp) == 1 && (bytes == 0 || fwrite(data, 1, bytes, fp) == bytes) &&
         fwrite(&len, 4, 1, fp) == 1;
}

int write_vector(Ctx &c, const char *path, int append, const char *kname, double sx, double sy, double sz,
                 double weight, const char *ipgr)
{
  const int n = c.n, ne = c.ne;
  const int cpx = c.run().is_complex;
  if ((size_t)n * (cpx ? 16 : 8) > 0xffffffffull) return W2K_EINVAL;  // one record per vector
  std::vector<double> ev(ne > 0 ? ne : 1), z((size_t)(cpx ? 2 : 1) * n * (ne > 0 ? ne : 1));
  if (ne > 0) {
    W2K_HIP(hipMemcpyAsync(ev.data(), c.d_eig.p, sizeof(double) * ne, hipMemcpyDeviceToHost, c.stream));
    W2K_HIP(hipMemcpyAsync(z.data(), c.d_z.p, sizeof(double) * z.size(), hipMemcpyDeviceToHost, c.stream));
  }
  W2K_HIP(hipStreamSynchronize(c.stream));
  FILE *fp = fopen(path, append ? "ab" : "wb");
  if (!fp) {
    set_last_error("w2k_write_vector: cannot open %s", path);
    return W2K_EINVAL;
  }
  bool ok = true;
  {
    // SX, SY, SZ, KNAME(10), NV+NLO, NE, WEIGHT, IPGR(3): 3*8 + 10 + 4 + 4 + 8 + 3 bytes
    unsigned char hdr[53];
    size_t o = 0;
    const double s3[3] = {sx, sy, sz};
    std::memcpy(hdr + o, s3, 24); o += 24;
    char kn[10];
    std::memset(kn, ' ', 10);
    if (kname) std::memcpy(kn, kname, strnlen(kname, 10));
    std::memcpy(hdr + o, kn, 10); o += 10;
    const int ii[2] = {n, ne};
    std::memcpy(hdr + o, ii, 8); o += 8;
    std::memcpy(hdr + o, &weight, 8); o += 8;
    char ip[3];
    std::memset(ip, ' ', 3);
    if (ipgr) std::memcpy(ip, ipgr, strnlen(ipgr, 3));
    std::memcpy(hdr + o, ip, 3); o += 3;
    ok = ok && put_record(fp, hdr, o);
  }
  ok = ok && put_record(fp, c.h_kzz.data(), sizeof(int) * 3 * (size_t)n);
  std::vector<double> col((size_t)(cpx ? 2 : 1) * n);
  for (int i = 0; i < ne && ok; i++) {
    unsigned char rec[12];
    const int num = i + 1;
    std::memcpy(rec, &num, 4);
    std::memcpy(rec + 4, &ev[i], 8);
    ok = ok && put_record(fp, rec, 12);
    const double *src = z.data() + (size_t)(cpx ? 2 : 1) * n * i;
    for (size_t k = 0; k < col.size(); k++) col[k] = (cpx && (k & 1)) ? -src[k] : src[k];  // CONJG in the complex build
    ok = ok && put_record(fp, col.data(), sizeof(double) * col.size());
  }
  if (fclose(fp) != 0) ok = false;
  if (!ok) {
    set_last_error("w2k_write_vector: write to %s failed", path);
    return W2K_EINVAL;
  }
  return 0;
}

}  // namespace w2k
