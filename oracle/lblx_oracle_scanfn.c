/*
 * lblx_oracle_scanfn.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C double-precision restatement of SCANFN (the dbl build), panel by
 * panel as the reference streams it: SCANFN's driver src/postsub.f90:25-560
 * (JFN 1-4, JVAR = 0), RDSCAN 1411-1514 (panels of the input file, the
 * JTREM = 0 exponential and JABS), SHRKSC 1276-1373 (running average
 * carried across panels in SS), CONVSC 1516-1598 (the table convolution
 * into the R1 buffer, ZF advanced point by point), PANLSC 1600-1688
 * (output panels of NLIMF - 1 points and the buffer shift), and the
 * function tables SHAPET 1377-1408, SHAPEG (src/oprop.f90:1912-1941),
 * SINCSQ 2445-2484, SINC 2486-2525.  The input "file" is an array cut
 * into panels of npanel points that ends with an ENDFIL record.
 *
 * Pins (tests/test_scanfn.py): a constant spectrum scans to the constant
 * times the table's sum rule; one nonzero input point scans to the
 * scanning function itself; a linear ramp scans to the ramp at the output
 * point (the shrink's half-group shift); Gaussian band against the closed
 * form; linearity.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lblx.h"

static int ifix_(double x) { return (int)x; }

static void table(int jfn, double dxf, int nf, int nfmax, double xscale, double *xf)
{
  const double pi = 2. * asin(1.);
  for (int i = 0; i < nfmax; i++) xf[i] = 0.;
  if (jfn == 1) { /* SHAPET: XTRIAN(X) = 1 - X/2, halved */
    xf[0] = 0.5;
    for (int i = 2; i <= nf; i++) {
      double x = (double)(i - 1) * dxf;
      xf[i - 1] = 0.5 * (1. - 0.5 * x);
    }
  } else if (jfn == 2) { /* SHAPEG: sqrt(ln2/pi) exp(-ln2 x^2), last entry 0 */
    double fln2 = log(2.), recpi = 1. / pi, fgnorm = sqrt(fln2 * recpi);
    xf[0] = fgnorm * exp(-fln2 * 0.);
    for (int jj = 2; jj <= nf; jj++) {
      double x = (double)(jj - 1) * dxf;
      xf[jj - 1] = fgnorm * exp(-fln2 * (x * x));
    }
    xf[nf - 1] = 0.;
  } else { /* SINCSQ / SINC: (XSCALE/pi) (sin y / y)^(2|1), y = x XSCALE */
    double xnorm = xscale / pi;
    xf[0] = xnorm;
    for (int i = 2; i <= nf; i++) {
      double x = (double)(i - 1) * dxf;
      double y = x * xscale, sc = sin(y) / y;
      xf[i - 1] = jfn == 3 ? xnorm * (sc * sc) : xnorm * sc;
    }
  }
}

/* Returns 0, -1 for arguments the reference stops on or that give no data,
 * -2 for options outside JFN 1-4 / JVAR 0 / input spacing <= HWHM/12.
 * out receives every output panel's points in order (cap values at most). */
int oracle_scanfn(const lblx_scanfn *S_, int iemit, const double *in, double v1c, double dvin, int64_t n_in,
                  double *out, int64_t cap, int64_t *n_out, double *v1_out, double *dv_out)
{
  static const double HWJ[5] = {1., 2., 4., 54.1826, 119.332818};
  static const double DXJ[5] = {0., 0.02, 0.02, 0.02, 0.02};
  static const int NJ[5] = {0, 101, 201, 2710, 5968}, NJMX[5] = {0, 251, 251, 2760, 6018};
  static const double SMPLJ[5] = {.5, 2., 4., 4., 4.};
  static const double XSCAL[5] = {0., 0., 0., 1.391557377, 1.89549425};
  static const int JRATIO[24] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 25, 30, 32, 40, 48, 50,
                                 60, 75, 80, 100, 120};
  const double PI = 2. * asin(1.);
  const int NLIMF = 2401, NSHIFT = 32;
  double HWHM = S_->hwhm, V1 = S_->v1, V2 = S_->v2;
  int JEMIT = S_->jemit, JFN = S_->jfn, JVAR = S_->jvar;
  const int64_t LPAN = S_->npanel > 0 ? S_->npanel : 2400;
  if (HWHM <= 0.) return -1;
  int JABS = 0;
  if (JEMIT < 0) { JABS = 1; JEMIT = 0; }
  int IFN = abs(JFN);
  if (IFN > 6) return -1;
  if (IFN == 0 || IFN == 5 || IFN == 6 || JVAR != 0) return -2;
  const double XSCALE = XSCAL[IFN];
  if (JFN < 0) {
    JFN = -JFN;
    if (JFN == 3 || JFN == 4) HWHM = HWHM / (PI / XSCALE);
    else return -1;
  }
  const double HWF = HWJ[IFN], DXF = DXJ[IFN];
  const int NF = NJ[IFN], NFMAX = NJMX[IFN];
  double SAMPLE = SMPLJ[IFN];
  const double DVINT = HWHM / 12.;
  if (S_->sampl < 0.) SAMPLE = HWHM / (-S_->sampl);
  else if (S_->sampl > 0.) SAMPLE = S_->sampl;
  double *XF = calloc((size_t)NFMAX, sizeof(double));
  table(JFN, DXF, NF, NFMAX, XSCALE, XF);
  if (dvin > DVINT) { free(XF); return -2; } /* SCNINT */
  int JTREM = -1;
  if (iemit == 0 && JEMIT == 0) JTREM = 0;
  if (iemit == 1 && JEMIT == 0) JTREM = 2;
  if (iemit == 1 && JEMIT == 1) JTREM = 1;
  if (JTREM < 0) { free(XF); return -1; }
  double DVO = HWHM / SAMPLE;
  double BOUND = HWF * HWHM;
  int NBOUND = ifix_((2. * HWF) * SAMPLE + 0.01);
  BOUND = (double)NBOUND * DVO / 2.;
  const int MAXF = NLIMF + 2 * NBOUND + NSHIFT;
  int NLO = NSHIFT + 1, NHI = NLIMF + NSHIFT - 1;
  double *R1 = calloc((size_t)MAXF + 1, sizeof(double)); /* R1(1..MAXF) */
  double VFT = V1 - (double)NSHIFT * DVO;
  const double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
  /* S holds a panel plus the carried remainder; SS the remainder */
  double *S = calloc((size_t)(LPAN + 200 + 2), sizeof(double));
  double SS[200];
  int NREN = 0, INIT = 0, IRATSH = 1, IRATM1 = 0;
  double SRATIO = 1., DVSC = 0., V1SHFT = 0.;
  int ILO = 1, IHI = 0, IDATA = -1, IPANEL = -1, ISTOP = 0, ILAST = 0;
  double V1I = 0., DVI = dvin;
  int NLIM = 0;
  int64_t pnext = 0;   /* next input panel */
  int IEOFSC = 1;
  int64_t nout = 0;
  int first_out = 1;
  double ZSLOPE = 0., ZBOUND = 0., XNORM = 0.;
  int rc = 0;

  for (;;) {
    /* 40: RDSCAN */
    if (IEOFSC > 0) {
      for (;;) {
        if (pnext >= n_in) { IEOFSC = -99; break; } /* the ENDFIL record */
        int64_t p0 = pnext, np = (n_in - p0 < LPAN) ? n_in - p0 : LPAN;
        pnext += np;
        double VMIN = v1c + (double)p0 * dvin, VMAX = v1c + (double)(p0 + np - 1) * dvin;
        DVI = dvin;
        int NNI = (int)np;
        int NLOW = NREN + 1;
        if (NREN <= 0) NLOW = 1;
        VMIN = VMIN - (double)(NLOW - 1) * DVI;
        NNI = NNI + NLOW - 1;
        IDATA = 0;
        if (VMAX < VBOT) continue; /* panel skipped */
        for (int i = NLOW; i <= NNI; i++) {
          double si = in[p0 + (i - NLOW)], v = si;
          if (JTREM == 0) {
            v = 1.;
            if (si > 1.0e-04) v = si < 34. ? exp(-si) : exp(-34.);
            else v = 1. - si;
          }
          S[i] = v;
        }
        if (JABS != 0)
          for (int i = NLOW; i <= NNI; i++) S[i] = 1. - S[i];
        ILO = 1;
        IHI = NNI;
        double DIF = (VMIN - VBOT) / DVI;
        if (DIF < 0.) ILO = ifix_(-DIF + 1.5);
        if (VMAX > VTOP) {
          IHI = ifix_((VTOP - VMIN) / DVI + 1.5);
          IDATA = 1;
        }
        V1I = VMIN;
        NLIM = NNI;
        break;
      }
      if (IEOFSC > 0) {
        /* SHRKSC */
        int NLIMS = NLIM;
        if (NREN > 0)
          for (int i = 1; i <= NREN; i++) S[i] = SS[i - 1];
        NREN = 0;
        int doit = 1;
        if (INIT == 0) {
          DVSC = HWHM / 12.;
          IRATSH = ifix_(DVSC / DVI + 0.5);
          for (int i = 2; i <= 24; i++)
            if (JRATIO[i - 1] > IRATSH) { IRATSH = JRATIO[i - 2]; break; }
          if (IRATSH > JRATIO[23]) IRATSH = JRATIO[23];
          if (IRATSH <= 1) doit = 0;
          else {
            DVSC = (double)IRATSH * DVI;
            V1SHFT = (double)(IRATSH - 1) * DVI / 2.;
            SRATIO = IRATSH;
            IRATM1 = IRATSH - 1;
            INIT = 1;
          }
        }
        if (IRATSH <= 1) doit = 0;
        if (doit) {
          NREN = NLIM - (NLIM / IRATSH) * IRATSH;
          V1I = V1I + V1SHFT;
          int IMAX = NLIM - IRATM1 - NREN, K = 0;
          for (int i = 1; i <= IMAX; i += IRATSH) {
            double sumk = 0.;
            for (int j = i; j <= i + IRATM1; j++) sumk = sumk + S[j];
            K++;
            S[K] = sumk / SRATIO;
          }
          NLIM = K;
          DVI = DVSC;
          ILO = ifix_((VBOT - V1I) / DVI + 1.5);
          if (ILO < 1) ILO = 1;
          IHI = ifix_((VTOP - V1I) / DVI + 1.5);
          if (IHI > NLIM) IHI = NLIM;
          if (NREN > 0)
            for (int i = 1; i <= NREN; i++) SS[i - 1] = S[NLIMS - NREN + i];
        }
      }
    }
    int go_panel = 0;
    if (IEOFSC <= 0) go_panel = 1;
    for (;;) {
      if (!go_panel) {
        /* 50: CONVSC */
        if (ILO > IHI) {
          IPANEL = 1;
        } else {
          double RATIO = DVI / DVO, DVODX = DVO / DXF, HWBND = HWF / DVO;
          double ZINT = ((V1I - VFT) / DVO);
          int ITST = -1, full = 0;
          for (int I = ILO; I <= IHI; I++) {
            if (S[I] == 0.) continue;
            if (I >= ITST) {
              ITST = 9999;
              ZSLOPE = DVODX / HWHM;
              ZBOUND = HWBND * HWHM;
              XNORM = DVI / HWHM;
            }
            double ZPEAK = (double)(I - 1) * RATIO + ZINT;
            int JMAX = ifix_(ZPEAK + ZBOUND + 1.5);
            if (JMAX > MAXF) { ILAST = I - 1; full = 1; break; }
            int JMIN = ifix_(ZPEAK - ZBOUND + 1.5);
            if (JMIN < 1) JMIN = 1;
            double SI = XNORM * S[I];
            double ZF = ((double)(JMIN - 1) - ZPEAK) * ZSLOPE;
            for (int JF = JMIN; JF <= JMAX; JF++) {
              int IT = ifix_(fabs(ZF) + 1.5);
              if (IT > NFMAX) IT = NFMAX; /* the table's zero tail */
              R1[JF] = R1[JF] + SI * XF[IT - 1];
              ZF = ZF + ZSLOPE;
            }
          }
          if (full) IPANEL = 1;
          else { ILAST = IHI; IPANEL = IDATA; }
        }
        ILO = ILAST + 1;
        if (IPANEL == 0) break; /* GO TO 40: next input panel */
      }
      go_panel = 0;
      /* 60: PANLSC */
      ISTOP = 0;
      int NNHI = ifix_((V2 - VFT) / DVO + 1.5);
      if (NHI >= NNHI) ISTOP = 1;
      if (ISTOP == 1) NHI = NNHI;
      int NL = NHI - NLO + 1;
      double V1P = VFT + (double)(NLO - 1) * DVO;
      if (first_out) { *v1_out = V1P; first_out = 0; }
      for (int i = NLO; i <= NHI; i++) {
        if (nout < cap) out[nout] = R1[i];
        nout++;
      }
      (void)NL;
      VFT = VFT + (double)(NLIMF - 1) * DVO;
      if (ISTOP == 1) goto done;
      for (int J = NLIMF; J <= MAXF; J++) R1[J - NLIMF + 1] = R1[J];
      for (int J = MAXF - NLIMF + 2; J <= MAXF; J++) R1[J] = 0.;
      NLO = NSHIFT + 1;
      if (IEOFSC < 0) { go_panel = 1; continue; } /* GO TO 60 */
      if (IEOFSC > 0) continue;                    /* GO TO 50 */
      goto done;
    }
  }
done:
  *n_out = nout;
  *dv_out = DVO;
  if (nout > cap) rc = -3;
  free(XF);
  free(R1);
  free(S);
  return rc;
}
