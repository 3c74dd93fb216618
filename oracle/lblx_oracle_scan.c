/*
 * lblx_oracle_scan.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C double-precision restatement of FFTSCN (the dbl build) for the
 * checker: src/fftscn_dbl.f90 FFTSCN 25-789, Apod_Init 791-841, Loadsp
 * 1517-1817, Boxcar 844-1043, Scntrn 2491-2779 (IVX = 1), Scnfnt
 * 1934-2470 (interferogram-domain branches of JFN 1-6 and the boxcar
 * correction JFN = -1), Multrn 1818-1933, Wrtspc 2780-2962, and INTPDR
 * 2963-3174 with INTERP / RDPANL (src/postsub.f90:2768-3100, four-point
 * branch).  The record-by-record disk FFT (REALBG/BIGFFT) is replaced by
 * an in-memory radix-2 complex FFT of the same length; the transform
 * pair is taken as the identity round trip, as REALFT's is.
 *
 * Pins (tests/test_scan.py): a constant transmittance scans to H(0) times
 * the constant; a line much narrower than the instrument scans to the
 * apodized instrument function in closed form; the FFT against numpy.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lblx.h"

#define LPTSMX 65536
static const double C_[14] = {0.0, 2.0, 0.849322, 2.257609, 3.314800, 2.100669, 2.195676,
                              2.0, 2.570274, 2.367714, 2.071759, -3.3, -2.5, 3.314800};
static const double CRATIO_[14] = {0.0, 40., 10., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., -1.};
static const double CLIMIT_[14] = {0., 2., 3., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., 160.};

static int ifix(double x) { return (int)x; }

/* radix-2 complex FFT in place, sign = -1 forward, +1 backward (no scale) */
static void cfft(double *re, double *im, int64_t n, int sign)
{
  for (int64_t i = 1, j = 0; i < n; i++) {
    int64_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int64_t len = 2; len <= n; len <<= 1) {
    int64_t h = len >> 1;
    for (int64_t k = 0; k < h; k++) {
      double ang = sign * 4.0 * asin(1.0) * (double)k / (double)len;
      double wr = cos(ang), wi = sin(ang);
      for (int64_t i = k; i < n; i += len) {
        int64_t j = i + h;
        double xr = re[j] * wr - im[j] * wi, xi = re[j] * wi + im[j] * wr;
        re[j] = re[i] - xr; im[j] = im[i] - xi;
        re[i] += xr; im[i] += xi;
      }
    }
  }
}

/* interferogram-domain apodization of Scnfnt (IAPSC = 1) at OPD x */
static double apod(int jfn, double a, double x, double pi)
{
  switch (jfn) {
    case 1: return x == 0 ? 1.0 : pow(sin(pi * x * a) / (pi * x * a), 2);
    case 2: return exp(-fmin(2.0 * pow(pi * x * a, 2), 34.0));
    case 3: return x < 1 / a ? 1.0 - x * a : 0.0;
    case 4: return x < 1.0 / a ? 1.0 : 0.0;
    case 5: return x < 1 / a ? pow(1.0 - pow(x * a, 2), 2) : 0.0;
    case 6: return x < 1 / a ? 0.53835685 * (1.0 + 0.857504 * cos(pi * x * a)) : 0.0;
  }
  return 0.0;
}

/* INTPDR/INTERP four-point walk over input panels of 2400 points
 * (Wrtspc layout) onto v1..v2 step dvo; returns the number of points. */
static int64_t intpdr(const double *z, int64_t nz, double v1s, double dvi, double v1, double v2,
                      double dvo, double *out, int64_t cap)
{
  double C1[203], C2[203], C3[203], C4[203];
  const double xnumcf = 201.;
  for (int ip = 0; ip <= 202; ip++) {
    double p = ((double)ip - 1.0) / (xnumcf - 1.0), pp = p * p;
    C1[ip] = -p / 2.0 * (1 - p) * (1 - p);
    C2[ip] = 1.0 - pp * (3.0 - 2.0 * p) + pp / 2.0 * (1.0 - p);
    C3[ip] = pp * (3.0 - 2.0 * p) + p / 2.0 * (1.0 - p) * (1.0 - p);
    C4[ip] = -pp / 2 * (1.0 - p);
  }
  /* input panel headers as Wrtspc writes them */
  int npan = (int)((nz + 2399) / 2400);
  double *pv1 = malloc(npan * sizeof(double)), *pv2 = malloc(npan * sizeof(double));
  int *pn = malloc(npan * sizeof(int));
  double v1p = v1s, v2p = 0.;
  for (int k = 0; k < npan; k++) {
    int np = (int)((nz - 2400LL * k) < 2400 ? nz - 2400LL * k : 2400);
    if (k > 0) v1p = v2p + dvi;
    v2p = v1p + dvi * (double)(np - 1);
    pv1[k] = v1p; pv2[k] = v2p; pn[k] = np;
  }
  double T[2410];
  double R[2401];
  int64_t nout = 0;
  const double bound = 4.0 * dvo, vbot = v1 - bound;
  int pk = 0;
  while (pk < npan && pv2[pk] < vbot) pk++;
  for (int i = 0; i < 4; i++) T[i + 1] = 0.0;
  /* RDPANL */
  double V1I, V2I;
  int NNI;
#define LOADP(k)                                                                 \
  do {                                                                           \
    for (int q = 0; q < pn[k]; q++) T[5 + q] = z[2400LL * (k) + q];              \
    V1I = pv1[k] - 4.0 * dvi;                                                    \
    V2I = pv2[k];                                                                \
    NNI = pn[k] + 4;                                                             \
  } while (0)
  if (pk >= npan) goto done;
  LOADP(pk);
  int hit_eof = 0;
  T[4] = 2.0 * T[5] - T[6];
  double V1J = v1, DVJ = dvo, vratio = DVJ / dvi, VJ = V1J;
  int J2 = 0;
  if (V1J < V1I) goto done;  /* output starting before the data: not used */
  for (;;) {
    int I = ifix((VJ - V1I) / dvi + 1.00001);
    double VI = V1I + dvi * (double)(I - 1);
    double P = (VJ - VI) / dvi;
    int J1 = ifix((VJ - V1J) / DVJ + 1.001);
    int a = ifix((v2 - V1J) / DVJ + 1.001), b = ifix((V2I - dvi - V1J) / DVJ + 1.);
    J2 = a < b ? a : b;
    if (J2 > 2400) J2 = 2400;
    for (int J = J1; J <= J2; J++) {
      int IP = ifix(P * xnumcf + 1.00001);
      R[J] = C1[IP] * T[I - 1] + C2[IP] * T[I] + C3[IP] * T[I + 1] + C4[IP] * T[I + 2];
      P = P + vratio;
      if (P >= 1.0) {
        I = ifix((double)I + P);
        P = P - (double)ifix(P);
      }
    }
    VJ = V1J + DVJ * J2;
    if (J2 >= 2400 || VJ >= v2) {
      for (int J = 1; J <= J2 && nout < cap; J++) out[nout++] = R[J];
      double V2J = V1J + DVJ * (J2 - 1);
      V1J = V2J + DVJ;
      J2 = 0;
    }
    if (VJ >= v2) break;
    if (hit_eof) {
      /* label 80: zeros to the end of the band */
      for (;;) {
        int J1b = J2 + 1;
        int J2b = ifix((v2 - V1J) / DVJ + 1.0001);
        if (J2b > 2400) J2b = 2400;
        for (int J = J1b; J <= J2b; J++) R[J] = 0.0;
        J2 = J2b;
        VJ = V1J + DVJ * J2;
        if (J2 >= 2400 || VJ >= v2) {
          for (int J = 1; J <= J2 && nout < cap; J++) out[nout++] = R[J];
          double V2J = V1J + DVJ * (J2 - 1);
          V1J = V2J + DVJ;
          J2 = 0;
        }
        if (VJ >= v2) break;
      }
      break;
    }
    if (I >= NNI - 2) {
      for (int ii = 1; ii <= 4; ii++) T[ii] = T[ii + NNI - 4];
      if (pk + 1 < npan) {
        pk++;
        LOADP(pk);
      } else {
        hit_eof = 1;
        T[NNI + 1] = 2.0 * T[NNI] - T[NNI - 1];
        T[NNI + 2] = 0.0;
        V2I = V2I + dvi;
      }
    }
  }
#undef LOADP
done:
  free(pv1); free(pv2); free(pn);
  return nout;
}

int oracle_fftscn(const lblx_scan *S, const double *in, double v1c, double dv, int64_t n_in,
                  double *out, int64_t cap, int64_t *n_out, double *v1_out, double *dv_out)
{
  const double pi = 2.0 * asin(1.0);
  double HWHM = S->hwhm, V1 = S->v1, V2 = S->v2;
  int jemit = S->jemit, jfnin = S->jfnin, jfn = abs(jfnin);
  if (HWHM <= 0 || jemit < 0 || jemit > 1 || jfn > 6) return LBLX_E_UNSUPPORTED;
  double A = 0., PATHL = 0.;
  if (jfnin > 0) {
    A = C_[jfn] * HWHM;
    PATHL = 1.0 / A;
  } else if (jfnin < 0) {
    PATHL = HWHM;
    A = 1. / PATHL;
    HWHM = S->param == 0. ? A / fabs(C_[jfn]) : S->param;
  }
  (void)PATHL;
  if (jfn == 2 && S->param != 0.) return LBLX_E_UNSUPPORTED;
  int nobox = 0, mratio = 12;
  if (S->mratin < 0) nobox = 1;
  else if (S->mratin > 0) mratio = S->mratin;
  double V2C = v1c + dv * (double)(n_in - 1);
  if (V1 >= V2 || V1 >= V2C || V2 <= v1c) return LBLX_E_ARG;
  double V1Z = V1 - HWHM * CLIMIT_[jfn], V2Z = V2 + HWHM * CLIMIT_[jfn];
  V1Z = v1c + (double)ifix((V1Z - v1c) / dv + 1.01) * dv;
  V2Z = v1c + (double)ifix((V2Z - v1c) / dv + 1.01) * dv;
  if (V1Z < v1c) V1Z = v1c;
  if (V2Z > V2C) V2Z = V2C;
  if ((V2Z - V1Z) / HWHM < CLIMIT_[jfn]) return LBLX_E_ARG;
  if (dv > 2. * HWHM / 12) return LBLX_E_UNSUPPORTED;  /* INTPDR pre-interpolation */
  int ivx = S->ivx;
  if (ivx == 0) ivx = (V2Z - V1Z) / HWHM > CRATIO_[jfn] ? 1 : -1;
  if (jfn == 0) ivx = 0;
  if (ivx == -1) return LBLX_E_UNSUPPORTED;
  /* Loadsp */
  int64_t n0 = (int64_t)llround((V1Z - v1c) / dv);
  int64_t ltot = (int64_t)((V2Z - V1Z) / dv + 1.1);
  V2Z = V1Z + dv * (double)(ltot - 1);
  if (n0 < 0 || n0 + ltot > n_in) return LBLX_E_ARG;
  double ardata = (double)ltot / (double)LPTSMX;
  int lrdata = (ardata - (int)ardata) == 0.0 ? (int)ardata : (int)(ardata + 1);
  double power = log((double)lrdata) / log(2.0);
  int lrec = (power - (int)power) == 0 ? 1 << (int)(power + 2) : 1 << ((int)power + 3);
  double fill = jemit == 0 ? 1.0 : 0.0;
  int M = 1;
  if (jfn == 0) {
    M = (int)(2.0 * HWHM / dv);
    if (M < 2) return LBLX_E_ARG;
  } else if (nobox == 0) {
    M = (int)(2.0 * HWHM / (mratio * dv));
  }
  double *y = NULL;
  int64_t N;
  if (M >= 2 && nobox == 0) {
    V1Z = V1Z + dv * (M - 1) / 2.0;
    int64_t ktot = ltot / M;
    dv = M * dv;
    V2Z = V1Z + dv * (double)(ktot - 1);
    double ar = (double)ktot / (double)LPTSMX;
    int krdata = (ar - (int)ar) == 0.0 ? (int)ar : (int)ar + 1;
    double pw = log((double)krdata) / log(2.0);
    int krec = (pw - (int)pw) == 0 ? 1 << (int)(pw + 2) : 1 << (int)(pw + 3);
    N = (int64_t)krec * LPTSMX;
    y = malloc(N * sizeof(double));
    for (int64_t k = 0; k < ktot; k++) {
      double sum = 0.0;
      for (int mm = 0; mm < M; mm++) sum = sum + in[n0 + k * M + mm];
      y[k] = sum / (double)M;
    }
    for (int64_t k = ktot; k < N; k++) y[k] = fill;
    ltot = ktot;
  } else {
    N = (int64_t)lrec * LPTSMX;
    y = malloc(N * sizeof(double));
    for (int64_t k = 0; k < ltot; k++) y[k] = in[n0 + k];
    for (int64_t k = ltot; k < N; k++) y[k] = fill;
  }
  if (jfn != 0) {
    double *re = malloc(N * sizeof(double)), *im = calloc(N, sizeof(double));
    memcpy(re, y, N * sizeof(double));
    cfft(re, im, N, -1);
    const double DX = 1.0 / ((double)N * dv);
    const int fix = (M >= 2 && nobox == 0 && S->nofix == 0);
    /* JFN = 4: the sample straddling the end of the boxcar apodization */
    int64_t k4 = -1;
    double r4 = 0.;
    if (jfn == 4) {
      double alen = 1.0 / A;
      int64_t nn = (int64_t)(alen / DX) + 1;
      double r = alen / DX + 1 - (double)nn;
      if (r <= 0.5) { r = r + 1.0; nn = nn - 1; }
      if (nn >= 1) { k4 = nn; r4 = r - 0.5; }
    }
    for (int64_t k = 0; k <= N / 2; k++) {
      double h = 0.;
      if (k < N / 2) {
        double x = (double)k * DX;
        h = apod(jfn, A, x, pi);
        if (k == k4) h = r4;
        if (fix) {
          double z = 2.0 * pi * x * (dv / 2.0);
          h = h * (z != 0 ? z / sin(z) : 1.0);
        }
      }
      re[k] *= h; im[k] *= h;
      if (k > 0 && k < N / 2) { re[N - k] *= h; im[N - k] *= h; }
    }
    cfft(re, im, N, +1);
    for (int64_t k = 0; k < N; k++) y[k] = re[k] / (double)N;
    free(re); free(im);
  }
  /* label 95: extraction */
  int64_t N1 = (int64_t)((V1 - V1Z) / dv);
  double V1S = V1Z + (double)(N1 - 1) * dv;
  if (V1S < V1Z) { N1 = 1; V1S = V1Z; }
  int64_t N2 = (int64_t)((V2 - V1Z) / dv + 3.01);
  double V2S = V1Z + (double)(N2 - 1) * dv;
  if (V2S > V2Z) {
    N2 = (int64_t)((V2Z - V1Z) / dv + 1.01);
    V2S = V1Z + (double)(N2 - 1) * dv;
  }
  int64_t ntot = N2 - N1 + 1;
  const double *z = y + (N1 - 1);
  int64_t nres;
  if (S->dvout > 0) {
    nres = intpdr(z, ntot, V1S, dv, V1, V2, S->dvout, out, cap);
    if (v1_out) *v1_out = V1;
    if (dv_out) *dv_out = S->dvout;
  } else {
    nres = ntot < cap ? ntot : cap;
    memcpy(out, z, nres * sizeof(double));
    if (v1_out) *v1_out = V1S;
    if (dv_out) *dv_out = dv;
  }
  *n_out = nres;
  free(y);
  return LBLX_OK;
}
