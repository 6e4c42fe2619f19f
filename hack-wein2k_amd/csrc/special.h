// Device special functions of the lapw1 hot path, restated from the
// reference algorithms so that tables built on the GPU agree with HAMILT/HNS:
//   sphbes  -- sphbes.f:47-142 (downward recursion normalised by the sum
//              rule; power series for x <= 0.1)
//   dvbes1  -- dvbes1.f:32-49
//   ylm     -- ylm.f:160-263 (Condon-Shortley recursion)
//   rotate  -- rotate.f:22-28
#pragma once
#include <hip/hip_runtime.h>

namespace w2k {

struct cd {
  double re, im;
};
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cd conj(cd a) { return {a.re, -a.im}; }

// j_0..j_n(x) into fj[0..n] (n <= 13)
__device__ inline void sphbes(int n, double x, double *fj)
{
  const double XLIM = 0.1, HF = 0.5, ONE = 1.0, TWO = 2.0, TNHF = 10.5, FFT = 15.0;
  const double T25 = 1.0e25, TN25 = 1.0e-25, TN50 = 1.0e-50;
  if (x > XLIM) {
    double cufac = 4.2;
    if (x < (double)(n - 2)) cufac = TNHF / (n + HF - x);
    int ns = (int)(n + 5 + x * cufac);
    ns = (int)(ns + FFT / (ONE + sqrt(x)));
    if (ns + 2 > 100) ns = 98;  // x > 15 never occurs for |k+G| R <= RKmax
    double ffo = 0.0, ffn = TN25, ffp;
    int m = ns - 1;
    const double xi = ONE / x;
    double fm = (m + m) + ONE;
    double sdr = fm * TN50;
    for (;;) {
      ffp = fm * xi * ffn - ffo;
      if (fabs(ffp) >= T25) {
        sdr = sdr * TN50;
        ffp = ffp * TN25;
        ffn = ffn * TN25;
      }
      sdr = sdr + (fm - TWO) * ffp * ffp;
      ffo = ffn;
      ffn = ffp;
      if (m <= n) break;
      m = m - 1;
      fm = fm - TWO;
    }
    // downward sweep: only F(1..n+1) need to be stored (Fortran 1-based)
    double g[16];
    double gm1 = ffo, gm = ffn;  // F(M+1), F(M)
    if (m <= n + 1) g[m] = gm;
    if (m + 1 <= n + 1) g[m + 1] = gm1;
    do {
      m = m - 1;
      fm = fm - TWO;
      double gn;
      for (;;) {
        gn = fm * xi * gm - gm1;
        if (fabs(gn) >= T25) {
          // rescale F(M+1..N+1)
          for (int j = m + 1; j <= n + 1; j++) g[j] = g[j] * TN25;
          gm = gm * TN25;
          gm1 = gm1 * TN25;
          sdr = sdr * TN50;
          continue;
        }
        break;
      }
      sdr = sdr + (fm - TWO) * gn * gn;
      gm1 = gm;
      gm = gn;
      if (m <= n + 1) g[m] = gn;
    } while (m > 1);
    const double ser = ONE / sqrt(sdr);
    for (int j = 1; j <= n + 1; j++) fj[j - 1] = g[j] * ser;
  } else {
    const double hfxsq = HF * x * x;
    double xl = ONE, twm = ONE, ta, xlp;
    for (int m = 1; m <= n + 1; m++) {
      ta = xl;
      twm = twm + TWO;
      xl = xl / twm;
      ta = ta - xl * hfxsq;
      xlp = xl / (twm + TWO);
      fj[m - 1] = ta + HF * xlp * hfxsq * hfxsq;
      xl = xl * x;
    }
  }
}

__device__ inline void dvbes1(const double *fj, double *dj, double sm, double ri, int nt)
{
  const double x = sm * ri;
  if (x > 1.0e-5) {
    const double q2 = -1.0 / x;
    double q3 = q2;
    dj[0] = -fj[1];
    for (int l = 1; l < nt; l++) {
      q3 = q3 + q2;
      dj[l] = fj[l - 1] + q3 * fj[l];
    }
  } else {
    dj[0] = 0.0;
    dj[1] = 1.0 / 3.0;
    for (int l = 2; l < nt; l++) dj[l] = 0.0;
  }
}

// Y_lm, l = 0..lmax, stored y[l*l + l + m]
__device__ inline void ylm(const double *v, int lmax, cd *y)
{
  const double pi = 3.141592653589793238462643;
  double yllr = 1.0 / sqrt(4.0 * pi);
  y[0] = {yllr, 0.0};
  if (lmax <= 0) return;
  double a, b, c, ab, abc, costh, sinth, cosph, sinph;
  const double abmax = fmax(fabs(v[0]), fabs(v[1]));
  if (abmax > 0.0) {
    a = v[0] / abmax;
    b = v[1] / abmax;
    ab = sqrt(a * a + b * b);
    cosph = a / ab;
    sinph = b / ab;
  } else {
    cosph = 1.0;
    sinph = 0.0;
  }
  const double abcmax = fmax(abmax, fabs(v[2]));
  if (abcmax > 0.0) {
    a = v[0] / abcmax;
    b = v[1] / abcmax;
    c = v[2] / abcmax;
    ab = a * a + b * b;
    abc = sqrt(ab + c * c);
    costh = c / abc;
    sinth = sqrt(ab) / abc;
  } else {
    costh = 1.0;
    sinth = 0.0;
  }
  y[2] = {sqrt(3.0) * yllr * costh, 0.0};
  double temp1 = -sqrt(1.5) * yllr * sinth;
  y[3] = {temp1 * cosph, temp1 * sinph};
  y[1] = {-y[3].re, y[3].im};
  for (int l = 2; l <= lmax; l++) {
    int index = l * l;          // 0-based position of Y(l,-l)
    int index2 = index + 2 * l; // Y(l,l)
    int msign = 1 - 2 * (l % 2);
    const double yl1l1r = y[index - 1].re, yl1l1i = y[index - 1].im;
    temp1 = -sqrt((double)(2 * l + 1) / (double)(2 * l)) * sinth;
    yllr = temp1 * (cosph * yl1l1r - sinph * yl1l1i);
    double ylli = temp1 * (cosph * yl1l1i + sinph * yl1l1r);
    y[index2] = {yllr, ylli};
    y[index] = {msign * yllr, -msign * ylli};
    index2--;
    index++;
    const double temp2 = sqrt((double)(2 * l + 1)) * costh;
    y[index2] = {temp2 * yl1l1r, temp2 * yl1l1i};
    y[index] = {-msign * y[index2].re, msign * y[index2].im};
    index2--;
    index++;
    int i4l2 = index2 - 4 * l + 2;
    int i2l = index2 - 2 * l;
    const double d4ll1c = costh * sqrt((double)(4 * l * l - 1));
    const double d2l13 = -sqrt((double)(2 * l + 1) / (double)(2 * l - 3));
    for (int m = l - 2; m >= 0; m--) {
      temp1 = 1.0 / sqrt((double)((l + m) * (l - m)));
      const double t2 = d4ll1c * temp1;
      const double t3 = d2l13 * sqrt((double)((l + m - 1) * (l - m - 1))) * temp1;
      const double ylmr = t2 * y[i2l].re + t3 * y[i4l2].re;
      const double ylmi = t2 * y[i2l].im + t3 * y[i4l2].im;
      y[index2] = {ylmr, ylmi};
      y[index] = {msign * ylmr, -msign * ylmi};
      msign = -msign;
      index2--;
      index++;
      i4l2--;
      i2l--;
    }
  }
}

__device__ __forceinline__ void rotate3(const double *vec, const double *rot, double *out)
{
  for (int jc = 0; jc < 3; jc++) {
    double d = 0.0;
    for (int j = 0; j < 3; j++) d = d + vec[j] * rot[jc + 3 * j];
    out[jc] = d;
  }
}

}  // namespace w2k
