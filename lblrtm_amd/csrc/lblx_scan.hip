// FFTSCN on the device: load + boxcar, radix-2 Stockham complex FFT of the
// padded record set, the interferogram-domain filter (apodization x boxcar
// deconvolution), inverse FFT, and INTPDR's four-point interpolation.
// The algorithm's steps are those of src/fftscn_dbl.f90 (FFTSCN, Boxcar,
// Scntrn/Scnfnt, Multrn, Wrtspc) and src/postsub.f90 (INTERP).
#pragma once
#include <hip/hip_runtime.h>

namespace lblx {

// Boxcar (fftscn_dbl.f90:844-1043) of M input points per output point,
// padded with FILL to the FFT length; M = 1 is Loadsp's plain copy.
__global__ void k_scan_load(const double *in, int64_t n0, int M, int64_t ktot, int64_t N,
                            double fill, double2 *z) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  double v = fill;
  if (k < ktot) {
    if (M == 1) {
      v = in[n0 + k];
    } else {
      const double *p = in + n0 + k * (int64_t)M;
      double sum = 0.0;
      for (int mm = 0; mm < M; mm++) sum = sum + p[mm];
      v = sum / (double)M;
    }
  }
  z[k] = make_double2(v, 0.0);
}

// One radix-2 Stockham pass: sub-transforms of length ns merge into 2 ns.
// tw[m] = exp(-2 pi i m / N); sign > 0 takes the conjugate (inverse).
__global__ void k_fft_pass(const double2 *src, double2 *dst, const double2 *tw, int64_t N,
                           int64_t ns, int sign) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t h = N >> 1;
  if (i >= h) return;
  const int64_t k = i & (ns - 1);
  double2 w = tw[k * (h / ns)];
  if (sign > 0) w.y = -w.y;
  const double2 a = src[i], b0 = src[i + h];
  const double2 b = make_double2(b0.x * w.x - b0.y * w.y, b0.x * w.y + b0.y * w.x);
  const int64_t o = ((i - k) << 1) + k;
  dst[o] = make_double2(a.x + b.x, a.y + b.y);
  dst[o + ns] = make_double2(a.x - b.x, a.y - b.y);
}

// Twiddles exp(-2 pi i m / N), m < N/2, formed on the device (the
// allocation-free scan has no host table to copy).
__global__ void k_scan_twiddle(double2 *tw, int64_t N) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N / 2) return;
  double sn, cs;
  sincospi(2.0 * (double)m / (double)N, &sn, &cs);
  tw[m] = make_double2(cs, -sn);
}

// INTERP's four-point coefficients over 203 fractional positions
// (postsub.f90:2966-2974): cof[c], cof[203 + c], cof[406 + c], cof[609 + c].
__global__ void k_scan_cof(double *cof) {
  const int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip > 202) return;
  const double xnumcf = 201.;
  const double p = ((double)ip - 1.0) / (xnumcf - 1.0), pp = p * p;
  cof[ip] = -p / 2.0 * (1 - p) * (1 - p);
  cof[203 + ip] = 1.0 - pp * (3.0 - 2.0 * p) + pp / 2.0 * (1.0 - p);
  cof[406 + ip] = pp * (3.0 - 2.0 * p) + p / 2.0 * (1.0 - p) * (1.0 - p);
  cof[609 + ip] = -pp / 2 * (1.0 - p);
}

// Scnfnt interferogram-domain functions (IAPSC = 1) at OPD x.
__device__ __forceinline__ double scan_apod(int jfn, double a, double x) {
  const double pi = 3.141592653589793;
  switch (jfn) {
    case 1: {
      if (x == 0) return 1.0;
      double u = sin(pi * x * a) / (pi * x * a);
      return u * u;
    }
    case 2: {
      double u = pi * x * a;
      return exp(-fmin(2.0 * (u * u), 34.0));
    }
    case 3: return x < 1 / a ? 1.0 - x * a : 0.0;
    case 4: return x < 1.0 / a ? 1.0 : 0.0;
    case 5: {
      double u = 1.0 - (x * a) * (x * a);
      return x < 1 / a ? u * u : 0.0;
    }
    case 6: return x < 1 / a ? 0.53835685 * (1.0 + 0.857504 * cos(pi * x * a)) : 0.0;
  }
  return 0.0;
}

struct ScanFilter {
  int64_t N, k4;
  double DX, A, dvb, r4;
  int jfn, fix;
};

// Multrn with Scntrn's functions: bins k and N-k times H(k DX); the
// Nyquist bin is multiplied by the zero imaginary slot of the packed
// record, i.e. cleared.
__global__ void k_scan_filter(ScanFilter F, double2 *z) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t h = F.N >> 1;
  if (k > h) return;
  double v = 0.;
  if (k < h) {
    const double x = (double)k * F.DX;
    v = scan_apod(F.jfn, F.A, x);
    if (k == F.k4) v = F.r4;
    if (F.fix) {
      const double zz = 2.0 * 3.141592653589793 * x * (F.dvb / 2.0);
      v = v * (zz != 0 ? zz / sin(zz) : 1.0);
    }
  }
  double2 a = z[k];
  z[k] = make_double2(a.x * v, a.y * v);
  if (k > 0 && k < h) {
    double2 b = z[F.N - k];
    z[F.N - k] = make_double2(b.x * v, b.y * v);
  }
}

// Scanned spectrum point h of the extracted range (Wrtspc), with INTERP's
// end handling: T(4) = 2 T(5) - T(6) before the first point, the EOF
// extrapolation one point past the last, zeros beyond.
__device__ __forceinline__ double scan_pt(const double2 *z, int64_t off, int64_t nz, double rn,
                                          int64_t h) {
  if (h >= 0 && h < nz) return z[off + h].x * rn;
  if (h == -1) return 2.0 * (z[off].x * rn) - z[off + 1].x * rn;
  if (h == nz) return 2.0 * (z[off + nz - 1].x * rn) - z[off + nz - 2].x * rn;
  return 0.0;
}

// Output: the extracted points themselves (DVOUT = 0) or INTPDR's
// four-point interpolation with the 203-entry coefficient tables.
__global__ void k_scan_out(const double2 *z, int64_t off, int64_t nz, double rn, int64_t nout,
                           const int64_t *g, const int32_t *ip, const double *cof, double *out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nout) return;
  if (!g) {
    out[t] = z[off + t].x * rn;
    return;
  }
  const int c = ip[t];
  if (c < 0) {
    out[t] = 0.0;
    return;
  }
  const int64_t q = g[t];  // extracted-point index of T(I)
  out[t] = cof[c] * scan_pt(z, off, nz, rn, q - 1) + cof[203 + c] * scan_pt(z, off, nz, rn, q) +
           cof[406 + c] * scan_pt(z, off, nz, rn, q + 1) + cof[609 + c] * scan_pt(z, off, nz, rn, q + 2);
}

// ---------------------------------------------------------------------
// SCANFN (src/postsub.f90:25-560) in the spectral domain.

// RDSCAN's transform of one input point (postsub.f90:1465-1485): an
// optical depth becomes its transmittance (linear below 1e-4, EXPMIN past
// ARGMIN = 34); JABS takes 1 - T.
__device__ __forceinline__ double sfn_in(double si, int jtrem, int jabs) {
  double v = si;
  if (jtrem == 0) {
    v = 1.;
    if (si > 1.0e-04) v = si < 34. ? exp(-si) : exp(-34.);
    else v = 1. - si;
  }
  return jabs ? 1. - v : v;
}

// SHRKSC (postsub.f90:1326-1338): group g is the mean of input points
// o + g R .. o + g R + R - 1, added in order.
__global__ void k_sfn_shrink(const double *in, int64_t o, int R, int64_t K, int jtrem, int jabs,
                             double *s) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= K) return;
  const double *p = in + o + g * (int64_t)R;
  if (R == 1) {
    s[g] = sfn_in(p[0], jtrem, jabs);
    return;
  }
  double sum = 0.;
  for (int m = 0; m < R; m++) sum = sum + sfn_in(p[m], jtrem, jabs);
  s[g] = sum / (double)R;
}

// CONVSC (postsub.f90:1516-1598) gathered per output point.  Group g was
// convolved into R1 panel q_g (R1 moves by NLIMF - 1 = 2400 points per
// panel), with its peak at ZPEAK_g in that panel's index: it adds
// XNORM S_g XF(IT) to the points JMIN_g .. JMAX_g, IT = INT(|ZF| + 1.5),
// ZF advanced by ZSLOPE from (JMIN_g - 1 - ZPEAK_g) ZSLOPE point by point.
// ZF is formed directly, and where |ZF| + 1.5 lies within 1e-6 of an
// integer (the running sum could round across it) by CONVSC's running sum
// itself.  The groups that reach a point are added in order of g, as
// CONVSC's loop over the input adds them; output n is R1 point
// NSHIFT + 1 + n of the global numbering.
struct SfnArgs {
  double z0, ratio, zbound, zslope, xnorm;
  int64_t g_lo, g_hi;
  int64_t j0;  // global index J of output 0
  int nfmax;
};
__global__ void k_sfn_conv(SfnArgs A, const double *s, const double *zpk, const int32_t *qg, const double *xf,
                           int64_t nout, double *out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= nout) return;
  const int64_t J = A.j0 + n;
  // global peak z0 + g RATIO within [J - 1.5 - ZBOUND, J - 0.5 + ZBOUND],
  // widened by three groups for the rounding of the per-panel ZPEAK
  int64_t ga = (int64_t)floor(((double)(J - 1) - 0.5 - A.zbound - A.z0) / A.ratio) - 3;
  int64_t gb = (int64_t)floor(((double)(J - 1) + 0.5 + A.zbound - A.z0) / A.ratio) + 3;
  if (ga < A.g_lo) ga = A.g_lo;
  if (gb > A.g_hi) gb = A.g_hi;
  double r = 0.;
  for (int64_t g = ga; g <= gb; g++) {
    const int64_t k = g - A.g_lo;
    const int q = qg[k];
    if (q < 0) continue;
    const double zp = zpk[k];
    const int64_t jl = J - 2400 * (int64_t)q;  // index in the group's R1 panel
    const int64_t jmax = (int64_t)(zp + A.zbound + 1.5);
    int64_t jmin = (int64_t)(zp - A.zbound + 1.5);
    if (jmin < 1) jmin = 1;
    if (jl < jmin || jl > jmax) continue;
    const double sg = s[g];
    if (sg == 0.) continue;
    double zf = ((double)(jl - 1) - zp) * A.zslope;
    const double t = fabs(zf) + 1.5;
    if (fabs(t - rint(t)) < 1e-6) {  // near a step of the table index
      zf = ((double)(jmin - 1) - zp) * A.zslope;
      for (int64_t jf = jmin; jf < jl; jf++) zf = zf + A.zslope;
    }
    int it = (int)(fabs(zf) + 1.5);
    if (it > A.nfmax) it = A.nfmax;  // beyond the table: its zero tail
    const double si = A.xnorm * sg;
    r = r + si * xf[it - 1];
  }
  out[n] = r;
}

}  // namespace lblx
