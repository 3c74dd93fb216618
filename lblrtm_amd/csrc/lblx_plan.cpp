// Host planner: see lblx_plan.h.  Reference lines are cited per block.
#include "lblx_plan.h"

#include <algorithm>
#include <cmath>
#include <climits>
#include <cstring>

#include "lblx_contdata.h"
#include "lblx_refdata.h"

namespace lblx {

static const double BOLTZ = 1.3806503E-16;   // phys_consts.f90:36
static const double CLIGHT = 2.99792458E+10; // phys_consts.f90:37
static const double AVOGAD = 6.02214199E+23; // phys_consts.f90:38

// RADFN (oprop.f90:2235-2296): nu * tanh(nu / 2kT) with its three regimes.
double radfn(double vi, double xkt) {
  if (xkt > 0.0) {
    double x = vi / xkt;
    if (x <= 0.01) return 0.5 * x * vi;
    if (x <= 10.0) {
      double e = std::exp(-x);
      return vi * (1. - e) / (1. + e);
    }
  }
  return vi;
}

// RADFNI (oprop.f90:2297-2475): radiation term at vi and the next node
// (*vinew); *rdel is the per-step slope of the chord to that node.
double radfni(double vi, double dvi, double xkt, double *vinew, double *rdel,
              double *rdlast) {
  const double FACT1 = 3.0E-03;
  if (*rdlast < 0.) *rdlast = radfn(vi, xkt);
  double rad = *rdlast;
  int n = 1;
  double rdnext;
  auto step = [&](double target) {
    if (*vinew >= 0.0) {
      *vinew = target;
      n = fint((*vinew - vi) / dvi);
      if (n < 1) n = 1;
      *vinew = vi + dvi * (double)n;
    } else {
      *vinew = std::fabs(*vinew);
      n = fint((*vinew - vi) / dvi);
      if (n < 1) n = 1;
    }
  };
  if (xkt > 0.0) {
    double x = vi / xkt;
    if (x <= 0.01) {
      step(vi + FACT1 * 0.5 * vi);
      double xn = *vinew / xkt;
      rdnext = 0.5 * xn * *vinew;
    } else if (x <= 10.0) {
      double e = std::exp(-x);
      double xm = 1. - e, xp = 1. + e;
      double c = x * e;
      step(vi + FACT1 * vi / (1. + (c / xm + c / xp)));
      double en = std::exp(-(*vinew / xkt));
      rdnext = *vinew * (1. - en) / (1. + en);
    } else {
      step(vi + (FACT1 * vi));
      rdnext = *vinew;
    }
  } else {
    step(vi + (FACT1 * vi));
    rdnext = *vinew;
  }
  *rdel = (rdnext - rad) / (double)n;
  *rdlast = rdnext;
  return rad;
}

// TIPS_2003 with AtoB Lagrange interpolation (oprop.f90:5919-6295,
// 11708-11798) for molecules 1..7.
static double atob(double aa, const double *A, const double *B, int npt) {
  auto d = [](double x) { return x == 0. ? 0.0001 : x; };
  for (int i = 2; i <= npt; i++) {
    if (A[i - 1] < aa) continue;
    auto a = [&](int k) { return A[k - 1]; };
    auto b = [&](int k) { return B[k - 1]; };
    if (i < 3 || i == npt) {
      int j = i < 3 ? 3 : npt;
      double a0 = (aa - a(j - 1)) * (aa - a(j)) / (d(a(j - 2) - a(j - 1)) * d(a(j - 2) - a(j)));
      double a1 = (aa - a(j - 2)) * (aa - a(j)) / (d(a(j - 1) - a(j - 2)) * d(a(j - 1) - a(j)));
      double a2 = (aa - a(j - 2)) * (aa - a(j - 1)) / (d(a(j) - a(j - 2)) * d(a(j) - a(j - 1)));
      return a0 * b(j - 2) + a1 * b(j - 1) + a2 * b(j);
    }
    int j = i;
    double a0 = (aa - a(j - 1)) * (aa - a(j)) * (aa - a(j + 1));
    a0 = a0 / (d(a(j - 2) - a(j - 1)) * d(a(j - 2) - a(j)) * d(a(j - 2) - a(j + 1)));
    double a1 = (aa - a(j - 2)) * (aa - a(j)) * (aa - a(j + 1));
    a1 = a1 / (d(a(j - 1) - a(j - 2)) * d(a(j - 1) - a(j)) * d(a(j - 1) - a(j + 1)));
    double a2 = (aa - a(j - 2)) * (aa - a(j - 1)) * (aa - a(j + 1));
    a2 = a2 / (d(a(j) - a(j - 2)) * d(a(j) - a(j - 1)) * d(a(j) - a(j + 1)));
    double a3 = (aa - a(j - 2)) * (aa - a(j - 1)) * (aa - a(j));
    a3 = a3 / (d(a(j + 1) - a(j - 2)) * d(a(j + 1) - a(j - 1)) * d(a(j + 1) - a(j)));
    return a0 * b(j - 2) + a1 * b(j - 1) + a2 * b(j) + a3 * b(j + 1);
  }
  return 0.;
}

void tips_scor(double temp, double scor[7][10]) {
  for (int m = 0; m < 7; m++)
    for (int iso = 0; iso < 10; iso++) {
      scor[m][iso] = 0.;
      if (iso >= LBLX_ISO_MAX[m] || temp < 70. || temp > 3000.) continue;
      double q296 = atob(296., LBLX_TDAT, LBLX_QOFT[m][iso], LBLX_NT_TIPS);
      double qt = atob(temp, LBLX_TDAT, LBLX_QOFT[m][iso], LBLX_NT_TIPS);
      scor[m][iso] = q296 / qt;
    }
}

void tips_dlnscor(double temp, double h, double out[7][10]) {
  double sp[7][10], sm[7][10];
  tips_scor(temp + h, sp);
  tips_scor(temp - h, sm);
  for (int m = 0; m < 7; m++)
    for (int iso = 0; iso < 10; iso++)
      out[m][iso] = (sp[m][iso] > 0. && sm[m][iso] > 0.) ? std::log(sp[m][iso] / sm[m][iso]) / (2. * h) : 0.;
}

int plan_dv(const lblx_run *run, int nlay, int nmol, const double *pave,
            const double *tave, const double *wk, const double *wbroad,
            int iemit, double *dv_out, int32_t *itype_out) {
  if (!run || nlay <= 0 || nmol <= 0 || !pave || !tave || !wk || !wbroad || !dv_out)
    return LBLX_E_ARG;
  double DV = 0.;
  for (int l = 0; l < nlay; l++) {
    double sumwk = 0.;
    for (int m = 0; m < nmol; m++) sumwk = sumwk + wk[(size_t)l * nmol + m];
    double wtotl = sumwk + wbroad[l];
    double frh2o = wk[(size_t)l * nmol] / wtotl;
    double alfcor = (pave[l] / P0) * std::sqrt(TEMP0 / tave[l]);
    double h2oslf = (1. - frh2o + 5. * frh2o);
    double albar = run->alfal0 * alfcor * h2oslf;
    double adbar = 3.58115E-07 * (0.5 * (run->v1 + run->v2)) * std::sqrt(tave[l] / run->avmass);
    double avbar = 0.5 * (albar + std::sqrt(albar * albar + 4. * adbar * adbar));
    double olddv = DV;
    DV = avbar / run->sample;
    int itype = 99;
    if (l == 0) {  // three significant figures, even (lblrtm.f90:5444-5457)
      int iscal = fint(std::log10(DV) - 3.);
      double scal = std::pow(10., (double)iscal);
      int idv = fint((DV / scal) + 0.5);
      if (idv % 2 > 0) idv = idv + 1;
      DV = scal * (double)idv;
    } else {     // DV ratio chaining (lblrtm.f90:5460-5498)
      double type = olddv / DV;
      if (type > 2.5) return LBLX_E_DVRATIO;
      if (type >= 1.2) {
        itype = fint(1. / (type - 1.) + 0.5);
        if (itype == 3) itype = 2;
        DV = olddv * (double)itype / (double)(itype + 1);
      } else if (type >= 0.8) {
        DV = olddv;
        itype = 0;
      } else {
        DV = olddv;
        itype = 0;
        if (iemit == 0) {
          itype = fint(type / (1. - type) + 0.5);
          DV = DV * (double)(itype + 1) / (double)itype;
          itype = -itype;
        }
      }
    }
    dv_out[l] = DV;
    if (itype_out) itype_out[l] = itype;
  }
  return LBLX_OK;
}

// LBLF4 window geometry (oprop.f90:3688-3700) and, for JRAD = 1, the R4
// radiation segments (oprop.f90:3733-3758).
static void add_window(LayerPlan *P, double v1, double v2, int p_first) {
  WinPlan w{};
  w.v1r4 = v1;
  w.v2r4 = v2;
  w.nptr4 = fint((w.v2r4 - w.v1r4) / P->dvr4 + ONEPL);
  if (w.nptr4 > LIMOUT4) w.nptr4 = LIMOUT4;
  w.v2r4 = w.v1r4 + P->dvr4 * (double)(w.nptr4 - 1);
  w.p_first = p_first;
  w.seg0 = (int32_t)P->segs.size();
  w.nseg = 0;
  if (P->jrad == 1) {
    double xkt = 1. / (RADCN2 / P->tave);  // LBLF4: XKT = 1/BETA, BETA = RADCN2/TAVE
    double vitst = w.v1r4 - P->dvr4, rdlast = -1.;
    int n1 = 0, n2 = 0;
    do {
      n1 = n2 + 1;
      double vi = w.v1r4 + P->dvr4 * (double)(n1 - 1);
      double rdel;
      double radvi = radfni(vi, P->dvr4, xkt, &vitst, &rdel, &rdlast);
      n2 = fint((vitst - w.v1r4) / P->dvr4 + 0.001);
      if (n2 > w.nptr4) n2 = w.nptr4;
      if (n2 >= n1) P->segs.push_back(RadSeg{n1, n2, radvi, rdel});
      else if (n2 < n1) { /* empty segment: the Fortran loop does nothing */ }
    } while (n2 < w.nptr4);
    w.nseg = (int32_t)P->segs.size() - w.seg0;
  }
  P->wins.push_back(w);
}

// pre_xint (src/contnm.f90:994-1012)
static void pre_xint(double v1ss, double v2ss, double v1abs, double dvabs, int nptabs, int *ist,
                     int *last) {
  int nb1 = fint(2. + (v1ss - v1abs) / dvabs + 1.e-5);
  *ist = std::max(1, nb1);
  int nb2 = fint(1. + (v2ss - v1abs) / dvabs + 1.e-5);
  *last = std::min(nptabs, nb2);
}

// Coarse walk shared by FRNCO2 / o2_ver_1 / n2_ver_1 (contnm.f90:1312-1326).
static void cont_walk(const ContPlan &C, const double *hdr, ContTerm *T) {
  const double v1s = hdr[0], v2s = hdr[1], dvs = hdr[2];
  const int npts = (int)hdr[3];
  double V1C = C.v1abs - dvs, V2C = C.v2abs + dvs;
  int I1 = V1C < v1s ? -1 : fint((V1C - v1s) / dvs + 0.01);
  V1C = v1s + dvs * (double)(I1 - 1);
  int I2 = fint((V2C - v1s) / dvs + 0.01);
  int NPTC = I2 - I1 + 3;
  if (NPTC > npts) NPTC = npts + 4;
  T->on = 1;
  T->i1 = I1;
  T->nptc = NPTC;
  T->v1c = V1C;
  T->v2c = V1C + dvs * (double)(NPTC - 1);
  T->dvc = dvs;
  pre_xint(v1s, v2s, C.v1abs, C.dvabs, C.nptabs, &T->ist, &T->last);
}

// CONTNM scalars for the continua of 0..4340 cm-1 (H2O MT_CKD, CO2,
// O2 and N2 fundamentals, Rayleigh).
static int plan_cont(const lblx_run *run, const lblx_layer *Ly, int jrad, ContPlan *C, int icflg) {
  *C = ContPlan{};
  double X[7] = {1, 1, 1, 1, 1, 1, 1};
  switch (run->icntnm) {  // lblrtm.f90:705-760
    case 0: return LBLX_OK;
    case 1: break;
    case 2: X[0] = 0.; break;
    case 3: X[1] = 0.; break;
    case 4: X[0] = X[1] = 0.; break;
    case 5: X[6] = 0.; break;
    case 6: for (int k = 0; k < 7; k++) X[k] = run->cntscl[k]; break;
    default: return LBLX_E_ARG;
  }
  const double xself = X[0], xfrgn = X[1], xco2c = X[2], xo3cn = X[3], xo2cn = X[4],
               xn2cn = X[5], xrayl = X[6];
  const double V1 = Ly->v1, V2 = Ly->v2, PAVE = Ly->pave, TAVE = Ly->tave;
  if (V2 > 4340.0 && (xn2cn > 0. || xo2cn > 0. || xo3cn > 0.)) return LBLX_E_UNSUPPORTED;
  C->on = 1;
  // OPDPTH ABSRB grid (lblrtm.f90:6216-6221)
  C->dvabs = 1.;
  double a = (double)fint(V1);
  if (V1 < 0.) a = a - 1.;
  C->v1abs = a - 3. * C->dvabs;
  C->v2abs = (double)fint(V2 + 3. * C->dvabs + 0.5);
  C->nptabs = fint((C->v2abs - C->v1abs) / C->dvabs + 1.5);
  if (C->nptabs + 4 > 5050) return LBLX_E_SIZE;  // n_absrb (lblparams.f90:36)
  C->jrad = jrad;
  C->tave = TAVE;
  const double P0c = 1013., T0c = 296., XLOSMT = 2.68675E+19;
  const double RHOAVE = (PAVE / P0c) * (T0c / TAVE);
  C->xkt = TAVE / RADCN2;
  const double amagat = (PAVE / P0c) * (273. / TAVE);
  double WTOT = Ly->wbroad;
  for (int m = 0; m < Ly->nmol; m++) WTOT = WTOT + Ly->wk[m];
  const double wk1 = Ly->nmol >= 1 ? Ly->wk[0] : 0., wk2 = Ly->nmol >= 2 ? Ly->wk[1] : 0.;
  const double wk7 = Ly->nmol >= 7 ? Ly->wk[6] : 0.;
  const double x_vmr_h2o = wk1 / WTOT, x_vmr_o2 = wk7 / WTOT;
  const double x_vmr_n2 = 1. - x_vmr_h2o - x_vmr_o2;
  // an analytic-Jacobian species run has no N2 continuum (contnm.f90:226)
  const double wn2 = icflg > 0 ? 0. : x_vmr_n2 * WTOT;
  // H2O (contnm.f90:296-327, mt_ckd_h2o_module.f90:68-176)
  if (V2 > -20.0 && V1 < 20000. && (xself > 0. || xfrgn > 0.)) {
    ContTerm &T = C->h2o;
    const double *W = LBLX_MTCKD_WVN;
    const double dvc = W[1] - W[0];
    int i = 1;
    while (W[i - 1] <= C->v1abs - 2 * dvc) i++;
    int i1 = i - 1;
    if (i1 == 0) i1 = 1;
    while (W[i - 1] < C->v2abs + 2 * dvc) i++;
    T.on = 1;
    T.i1 = i1;                 // 1-based table index of coefficient 1
    T.nptc = i - i1 + 1;       // ncoeff
    T.v1c = W[i1 - 1];
    T.v2c = W[i - 1];
    T.dvc = dvc;
    int npt_l = fint((C->v2abs - C->v1abs) / C->dvabs + 1);
    pre_xint(W[0], W[LBLX_MTCKD_N - 1], C->v1abs, C->dvabs, npt_l, &T.ist, &T.last);
    T.f0 = wk1 / WTOT;                                                 // h2o_vmr
    T.f1 = (PAVE / LBLX_MTCKD_REF_P) * (LBLX_MTCKD_REF_T / TAVE);      // rho_rat
    T.f2 = LBLX_MTCKD_REF_T / TAVE;
    T.f3 = jrad != 0 ? 1. : 0.;                                        // radflag
    C->xself = xself;
    C->xfrgn = xfrgn;
    C->wk1 = wk1;
    C->forcl = run->ifrgnx == 1;
    C->aj = icflg == 1 ? 1 : 0;
  }
  // CO2 (contnm.f90:330-353)
  if (V2 > -20.0 && V1 < 10000. && xco2c > 0.) {
    cont_walk(*C, LBLX_FCO2_HDR, &C->co2);
    C->co2.f0 = wk2 * RHOAVE * 1.0E-20 * xco2c;  // WCO2
    C->co2.f1 = TAVE / 246.;                     // trat (FRNCO2 t_eff)
  }
  // O2 fundamental (contnm.f90:583-597)
  if (V2 > 1340.0 && V1 < 1850. && xo2cn > 0.) {
    cont_walk(*C, LBLX_O2F_HDR, &C->o2);
    C->o2.f0 = xo2cn * wk7 * 1.e-20 * amagat;    // tau_fac
    C->o2.f1 = (1. / 296.) - (1. / TAVE);        // xktfac
    C->o2.f2 = (1.e+20 / XLOSMT);                // factor
  }
  // N2 rotational (contnm.f90:754-790, xn2_r :2497-2567)
  if (V2 > -10.0 && V1 < 350. && xn2cn > 0.) {
    cont_walk(*C, LBLX_N2R_HDR, &C->n2r);
    C->n2r.f0 = xn2cn * (wn2 / XLOSMT) * amagat;   // tau_fac
    C->n2r.f1 = (TAVE - 296.) / (220. - 296.);     // tfac
    C->n2r.f4 = x_vmr_n2;
    C->n2r.f5 = x_vmr_o2;
    C->n2r.f6 = x_vmr_h2o;
  }
  // N2 fundamental (contnm.f90:734-753)
  if (V2 > 2001.77 && V1 < 2897.59 && xn2cn > 0.) {
    cont_walk(*C, LBLX_N2F_HDR, &C->n2);
    const double T = TAVE, T_272 = 272., T_228 = 228.;
    C->n2.f0 = xn2cn * (wn2 / XLOSMT) * amagat;  // tau_fac
    C->n2.f1 = ((1. / T) - (1. / T_272)) / ((1. / T_228) - (1. / T_272));  // xtfac
    C->n2.f2 = (T - T_272) / (T_228 - T_272);    // xt_lin
    C->n2.f3 = 1.294 - 0.4545 * T / 296.;         // a_o2
    C->n2.f4 = x_vmr_n2;
    C->n2.f5 = x_vmr_o2;
    C->n2.f6 = x_vmr_h2o;
  }
  // Rayleigh (contnm.f90:955-977)
  if (run->iaersl == 0 && V2 >= 820. && xrayl > 0.) {
    C->ray_on = 1;
    double conv = xrayl * 1.E-20 / (2.68675e-1 * 1.e5);
    C->ray_f = WTOT * conv;
  }
  return LBLX_OK;
}

int plan_cont_layer(const lblx_run *run, const lblx_layer *Ly, int jrad, ContPlan *C) {
  return plan_cont(run, Ly, jrad, C, 0);
}

// JRAD after OPDPTH's test (lblrtm.f90:6202-6204): the radiation term is
// folded into the optical depth (JRAD = 0) when V1 / XKT < 5, unless the
// caller asked for none (-1).  It becomes the next layer's JRAD.
int layer_jrad(const lblx_layer *Ly) {
  const double xkt_l = Ly->tave / RADCN2;
  return ((Ly->v1 / xkt_l) < 5. && Ly->jrad != -1) ? 0 : Ly->jrad;
}

int plan_layer(const lblx_run *run, lblx_layer *Ly, LayerPlan *P, int icflg) {
  if (!run || !Ly || !P) return LBLX_E_ARG;
  if (Ly->nmol < 1 || Ly->nmol > 7) return LBLX_E_UNSUPPORTED;
  if (run->ilblf4 < 0 || run->ilblf4 > 2) return LBLX_E_ARG;
  {  // reset, keeping the tables' storage for reuse
    std::vector<PanelPlan> pv = std::move(P->panels);
    std::vector<WinPlan> wv = std::move(P->wins);
    std::vector<RadSeg> sv = std::move(P->segs);
    *P = LayerPlan{};
    pv.clear();
    wv.clear();
    sv.clear();
    P->panels = std::move(pv);
    P->wins = std::move(wv);
    P->segs = std::move(sv);
  }
  const double V1 = Ly->v1, V2 = Ly->v2, DV = Ly->dv;
  const double SAMPLE = run->sample, ALFAL0 = run->alfal0;
  if (!(DV > 0.) || !(V2 > V1)) return LBLX_E_ARG;
  P->v1 = V1;
  P->v2 = V2;
  P->dv = DV;
  // OPDPTH (lblrtm.f90:6202-6256)
  double xkt_l = Ly->tave / RADCN2;
  Ly->jrad = layer_jrad(Ly);
  P->jrad = Ly->jrad;
  P->xkt = xkt_l;
  P->tave = Ly->tave;
  P->pave = Ly->pave;
  if (run->icntnm != 0) {
    int rc = plan_cont(run, Ly, P->jrad, &P->cont, icflg);
    if (rc) return rc;
  }
  int ipflag = (run->icntnm != 0 && Ly->pave <= 0.5) ? 1 : 0;
  double alfav = SAMPLE * DV;
  double alfav4 = 64. * alfav;
  P->dvr4 = alfav4 / SAMPLE;
  P->bound4 = 25.;
  if (run->ilblf4 == 2 && ipflag == 1) P->bound4 = 5.;
  P->v1r4st = V1 - 2. * P->dvr4;
  P->v2r4st = V2 + 2. * P->dvr4;
  P->v1l4 = P->v1r4st - P->bound4 - P->dvr4;
  P->v2l4 = P->v2r4st + P->bound4 + 2 * P->dvr4;
  // HIRAC1 (oprop.f90:223-315)
  P->dptmn_hirac = run->dptmin;
  if (P->jrad != 1) P->dptmn_hirac = run->dptmin / radfn(V2, Ly->tave / RADCN2);
  P->dptfc = run->dptfac;
  P->dptmn4 = run->dptmin / radfn(P->v2r4st, Ly->tave / RADCN2);  // LBLF4 uses its V2 argument
  P->alfmax = 4 * SAMPLE * DV * 0.04 / ALFAL0;
  P->nalfmax = fint(P->alfmax / DV);
  P->nbound = fint((2. * 64.) * P->nalfmax + 0.01);
  P->dvr2 = (0.008 / 0.002) * DV;
  P->dvr3 = (0.032 / 0.002) * DV;
  int max1 = NSHIFT + NLIM1 + NSHIFT + P->nbound / 2;
  int max2 = max1 / 4, max3 = max1 / 16;
  P->max1 = max1 + 4 * P->nalfmax;
  P->max2 = max2 + 16 * P->nalfmax / 4 + 1;
  P->max3 = max3 + 64 * P->nalfmax / 16 + 1;
  P->bound = (double)P->nbound * DV / 2.;
  P->vft0 = V1 - (double)NSHIFT * DV;
  P->vbot = V1 - P->bound;
  P->vtop = V2 + P->bound;
  // molecular quantities (MOLEC, LNCOR1 and LINF4 first-call constants)
  P->rhorat = (Ly->pave / P0) * (TEMP0 / Ly->tave);
  P->tratio = Ly->tave / TEMP0;
  P->pavp0 = Ly->pave / P0;
  P->xkt0 = TEMP0 / RADCN2;
  P->betacr_l = (1. / xkt_l) - (1. / P->xkt0);
  P->beta4 = RADCN2 / Ly->tave;
  P->betacr4 = P->beta4 - RADCN2 / TEMP0;
  tips_scor(Ly->tave, P->scor);
  double fad = std::log(2.) * 2. * AVOGAD * BOLTZ / (CLIGHT * CLIGHT);
  for (int m = 0; m < 7; m++)
    for (int i = 0; i < 10; i++)
      P->alfd1[m][i] = i < LBLX_ISO_MAX[m] ? std::sqrt(fad * Ly->tave / LBLX_SMASS[m][i]) : 0.;
  double sumwk = 0.;
  for (int m = 0; m < Ly->nmol; m++) sumwk = sumwk + Ly->wk[m];
  double wtot = sumwk + Ly->wbroad;
  for (int m = 0; m < LBLX_MXMOL; m++) P->rhoslf[m] = 0.;
  for (int m = 0; m < Ly->nmol; m++) P->rhoslf[m] = P->rhorat * Ly->wk[m] / wtot;

  // Panel walk (HIRAC1 label 80 / PANEL / LBLF4 refill).
  add_window(P, P->v1r4st, P->v2r4st, 0);
  double VFT = P->vft0;
  const int NLO = NSHIFT + 1;
  int64_t out = 0;
  for (int p = 0;; p++) {
    PanelPlan pp{};
    pp.vft = VFT;
    pp.w = (int32_t)P->wins.size() - 1;
    int NHI = NLIM1 + NSHIFT - 1;
    int istop = 0;
    if ((VFT + (NHI - 1) * DV) > V2) {
      NHI = fint((V2 - VFT) / DV + 1.);
      double v2p = VFT + (double)(NHI - 1) * DV;
      if (v2p < V2) { v2p = v2p + DV; NHI = NHI + 1; }
      pp.v2p = v2p;
      istop = 1;
    } else {
      pp.v2p = VFT + (double)(NHI - 1) * DV;
    }
    pp.nhi = NHI;
    pp.istop = istop;
    pp.out0 = out;
    out += NHI - NLO + 1;
    pp.seg0 = (int32_t)P->segs.size();
    if (P->jrad == 0) {  // PANEL radiation (oprop.f90:1492-1519)
      double v1p = VFT + (double)(NLO - 1) * DV;
      double vitst = v1p - DV, rdlast = -1.;
      int n1 = NLO - 1, n2 = NLO - 1;
      do {
        n1 = n2 + 1;
        double vi = VFT + (double)(n1 - 1) * DV;
        double rdel;
        double radvi = radfni(vi, DV, Ly->tave / RADCN2, &vitst, &rdel, &rdlast);
        n2 = fint((vitst - VFT) / DV + 0.001);
        if (n2 > NHI) n2 = NHI;
        if (n2 >= n1) P->segs.push_back(RadSeg{n1, n2, radvi, rdel});
      } while (n2 < NHI);
    }
    pp.nseg = (int32_t)P->segs.size() - pp.seg0;
    P->panels.push_back(pp);
    VFT = VFT + (double)(NLIM1 - 1) * DV;
    if (istop) break;
    // LBLF4 refill (oprop.f90:525-540); N2R3 = 155 for every panel
    const int N2R3 = NLIM3 + (NSHIFT / 16) - 1 + 3;
    double vf1 = VFT - 2. * P->dvr4;
    vf1 = V1 + std::floor((vf1 - V1) / P->dvr4) * P->dvr4;
    double vf2 = VFT + 2. * P->dvr4 + (double)(N2R3 + 4) * P->dvr3;
    const WinPlan &cw = P->wins.back();
    if (vf2 > cw.v2r4 && cw.v2r4 != P->v2r4st) add_window(P, vf1, P->v2r4st, p + 1);
    if (p > 100000000) return LBLX_E_SIZE;
  }
  int np = (int)P->panels.size();
  int pl = np - 1;
  const PanelPlan &last = P->panels.back();
  P->npts = out;
  P->g_last = (int64_t)2400 * pl + (last.nhi - 1);
  // R2 range that receives L4(R3) in the last panel (PANEL loop 10 + the
  // ISTOP extra pair, oprop.f90:1459-1477)
  {
    int limlo = (np == 1) ? 5 : 13;
    int limhi = last.nhi / 4 + 1;
    if (limlo <= limhi) {
      int jl = limlo + 4 * ((limhi - limlo) / 4);
      P->g2max = (int64_t)(jl + 5) - 1 + (int64_t)600 * pl;
    } else {
      P->g2max = (int64_t)600 * pl + 11;
    }
  }
  P->g3max = 154 + (int64_t)150 * pl;
  P->n1 = (int64_t)2400 * pl + P->max1 + 16;
  P->n2 = (int64_t)600 * pl + P->max2 + 16;
  P->n3 = (int64_t)150 * pl + P->max3 + 16;
  if (P->n1 >= (int64_t)INT32_MAX) return LBLX_E_SIZE;  // grid indices are kept in 32 bits
  return LBLX_OK;
}

}  // namespace lblx

namespace lblx {

// Line-shape sub-function tables on 2601 points (SHAPEL, SHAPEG, VERFN,
// oprop.f90:1822-1991): F1/F2/F3 = Lorentz minus the quadratic matched at
// 4, 16 and 64 half-widths (grid 0.002, 0.008, 0.032), FG = normalised
// Gaussian, XVER = the zeta = 0.3 Voigt correction.  tab holds F1, F2, F3,
// FG, XVER back to back, entry k = Fortran index k+1.
void shape_tables_host(double *tab) {
  const int NX = 2001, NF = 2601;
  double *f1 = tab, *f2 = tab + NF, *f3 = tab + 2 * NF, *fg = tab + 3 * NF, *xv = tab + 4 * NF;
  for (int i = 0; i < 5 * NF; i++) tab[i] = 0.;
  auto A = [](double z) { return (1. + 2. * z * z) / ((1. + z * z) * (1. + z * z)); };
  auto B = [](double z) { return -1. / ((1. + z * z) * (1. + z * z)); };
  const double recpi = 1. / (2. * std::asin(1.));
  const double a1 = A(4.), b1 = B(4.), a2 = A(16.), b2 = B(16.), a3 = A(64.), b3 = B(64.);
  auto lor = [](double xsq) { return 1. / (1. + xsq); };
  f1[0] = recpi * (lor(0.) - (a1 + b1 * 0.));
  for (int jj = 2; jj <= NX; jj++) {
    double x = (double)(jj - 1) * 0.002, xsq = x * x;
    f1[jj - 1] = recpi * (lor(xsq) - (a1 + b1 * xsq));
  }
  f1[NX - 1] = 0.;
  f2[0] = recpi * ((a1 + b1 * 0.) - (a2 + b2 * 0.));
  int j1lim = fint(4. / 0.008 + 1.001);
  for (int jj = 2; jj <= j1lim; jj++) {
    double x = (double)(jj - 1) * 0.008, xsq = x * x;
    f2[jj - 1] = recpi * ((a1 + b1 * xsq) - (a2 + b2 * xsq));
  }
  for (int jj = j1lim + 1; jj <= NX; jj++) {
    double x = (double)(jj - 1) * 0.008, xsq = x * x;
    f2[jj - 1] = recpi * (lor(xsq) - (a2 + b2 * xsq));
  }
  f2[NX - 1] = 0.;
  f3[0] = recpi * ((a2 + b2 * 0.) - (a3 + b3 * 0.));
  int j2lim = fint(16. / 0.032 + 1.001);
  for (int jj = 2; jj <= j2lim; jj++) {
    double x = (double)(jj - 1) * 0.032, xsq = x * x;
    f3[jj - 1] = recpi * ((a2 + b2 * xsq) - (a3 + b3 * xsq));
  }
  for (int jj = j2lim + 1; jj <= NX; jj++) {
    double x = (double)(jj - 1) * 0.032, xsq = x * x;
    f3[jj - 1] = recpi * (lor(xsq) - (a3 + b3 * xsq));
  }
  const double fln2 = std::log(2.);
  const double fgnorm = std::sqrt(fln2 * recpi);
  fg[0] = fgnorm * std::exp(-fln2 * 0.);
  for (int jj = 2; jj <= NX; jj++) {
    double x = (double)(jj - 1) * 0.002, xsq = x * x;
    fg[jj - 1] = fgnorm * std::exp(-fln2 * xsq);
  }
  fg[NX - 1] = 0.;
  const double PI = 3.1415926535898, CEXP = 0.45, CE0 = 1., CE2 = -.20737285249,
               CE4 = -.00872684335747;
  const double se0 = std::sqrt(CEXP / PI), ae2 = 2. * CEXP, ae4 = ae2 * ae2 / 3.;
  for (int i = 1; i <= NF; i++) {
    double z = 0.002 * (double)(i - 1), z2 = z * z;
    double xe0 = se0 * std::exp(-CEXP * z2);
    xv[i - 1] = (1. / (CE0 + CE2 + CE4)) * (CE0 + CE2 * ae2 * z2 + CE4 * ae4 * z2 * z2) * xe0;
  }
}

}  // namespace lblx

namespace lblx {

namespace {
const double SC_C[14] = {0.0, 2.0, 0.849322, 2.257609, 3.314800, 2.100669, 2.195676,
                         2.0, 2.570274, 2.367714, 2.071759, -3.3, -2.5, 3.314800};
const double SC_CRATIO[14] = {0.0, 40., 10., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., -1.};
const double SC_CLIMIT[14] = {0., 2., 3., 40., 160., 20., 20., 20., 40., 40., 20., 100., 10., 160.};
constexpr int64_t LPTSMX = 65536;

// Records of 65536 points that FFTSCN uses for npts data points: four
// times the data records, rounded up to a power of two (Loadsp, Boxcar).
int scan_records(int64_t npts) {
  double ar = (double)npts / (double)LPTSMX;
  int rd = (ar - (int)ar) == 0.0 ? (int)ar : (int)ar + 1;
  double pw = std::log((double)rd) / std::log(2.0);
  return (pw - (int)pw) == 0 ? 1 << (int)(pw + 2) : 1 << ((int)pw + 3);
}
}  // namespace

// SCANFN's function table (postsub.f90:74-107 DATA; SHAPET :1377-1408,
// SHAPEG oprop.f90:1912-1941, SINCSQ postsub.f90:2445-2484, SINC :2486-2525).
static void scanfn_table(int jfn, double dxf, int nf, int nfmax, double xscale, std::vector<double> &xf) {
  xf.assign(nfmax, 0.0);
  const double pi = 2. * std::asin(1.);
  if (jfn == 1) {
    xf[0] = 0.5;
    for (int i = 2; i <= nf; i++) {
      const double x = (double)(i - 1) * dxf;
      xf[i - 1] = 0.5 * (1. - 0.5 * x);
    }
  } else if (jfn == 2) {
    const double fln2 = std::log(2.), recpi = 1. / pi;
    const double fgnorm = std::sqrt(fln2 * recpi);
    xf[0] = fgnorm * std::exp(-fln2 * 0.);
    for (int jj = 2; jj <= nf; jj++) {
      const double x = (double)(jj - 1) * dxf;
      xf[jj - 1] = fgnorm * std::exp(-fln2 * (x * x));
    }
    xf[nf - 1] = 0.;
  } else {
    const double xnorm = xscale / pi;
    xf[0] = xnorm;
    for (int i = 2; i <= nf; i++) {
      const double x = (double)(i - 1) * dxf * xscale;
      const double sc = std::sin(x) / x;
      xf[i - 1] = jfn == 3 ? xnorm * (sc * sc) : xnorm * sc;
    }
  }
}

int plan_scanfn(const lblx_scanfn *S, int iemit, double v1c, double dv, int64_t n_in, ScanFnPlan *P) {
  if (!S || !P || !(dv > 0.) || n_in < 1) return LBLX_E_ARG;
  // postsub.f90:74-107: HWJ, DXJ, NJ, NJMX, SMPLJ, XSCAL for JFN 0..4
  static const double HWJ[5] = {1., 2., 4., 54.1826, 119.332818};
  static const double DXJ[5] = {0., 0.02, 0.02, 0.02, 0.02};
  static const int NJ[5] = {0, 101, 201, 2710, 5968}, NJMX[5] = {0, 251, 251, 2760, 6018};
  static const double SMPLJ[5] = {.5, 2., 4., 4., 4.};
  static const double XSCAL[5] = {0., 0., 0., 1.391557377, 1.89549425};
  static const int JRATIO[24] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 15, 16, 20, 24, 25, 30, 32, 40, 48, 50,
                                 60, 75, 80, 100, 120};
  const int NSHIFT = 32;
  double HWHM = S->hwhm;
  if (!(HWHM > 0.)) return LBLX_E_ARG;
  int jemit = S->jemit, jabs = 0;
  if (jemit < -1 || jemit > 1 || (iemit != 0 && iemit != 1)) return LBLX_E_ARG;
  if (jemit < 0) {  // :190-198
    jabs = 1;
    jemit = 0;
  }
  int jfn = S->jfn;
  const int ifn = jfn < 0 ? -jfn : jfn;
  if (ifn > 6) return LBLX_E_ARG;  // 'SCANF; JFN GT LIMIT'
  if (ifn == 0 || ifn == 5 || ifn == 6 || S->jvar != 0) return LBLX_E_UNSUPPORTED;
  const double pi = 2. * std::asin(1.);
  if (jfn < 0) {  // :222-233: HWHM holds the first zero
    jfn = -jfn;
    if (jfn != 3 && jfn != 4) return LBLX_E_ARG;
    HWHM = HWHM / (pi / XSCAL[jfn]);
  }
  const double HWF = HWJ[jfn], DXF = DXJ[jfn];
  double SAMPLE = SMPLJ[jfn];
  const double DVINT = HWHM / 12.;
  if (S->sampl < 0.) SAMPLE = HWHM / (-S->sampl);
  else if (S->sampl > 0.) SAMPLE = S->sampl;
  if (dv > DVINT) return LBLX_E_UNSUPPORTED;  // SCNINT interpolates onto DVINT first
  // JTREM (:325-335): 0 exp(-OD), 1 radiance, 2 transmittance
  int jtrem = -1;
  if (iemit == 0 && jemit == 0) jtrem = 0;
  if (iemit == 1 && jemit == 0) jtrem = 2;
  if (iemit == 1 && jemit == 1) jtrem = 1;
  if (jtrem < 0) return LBLX_E_ARG;  // ' SCANF; JTREM LT 0'
  const double V1 = S->v1, V2 = S->v2;
  if (!(V2 > V1)) return LBLX_E_ARG;
  const double DVO = HWHM / SAMPLE;
  const int NBOUND = (int)((2. * HWF) * SAMPLE + 0.01);  // :384-394
  const double BOUND = (double)NBOUND * DVO / 2.;
  const double VFT = V1 - (double)NSHIFT * DVO;
  const double VBOT = V1 - BOUND, VTOP = V2 + BOUND;
  // SHRKSC's ratio (postsub.f90:1305-1322)
  int R = 1;
  double DVSC = dv;
  {
    const double dvsc = HWHM / 12.;
    int irat = fint(dvsc / dv + 0.5);
    for (int i = 2; i <= 24; i++)
      if (JRATIO[i - 1] > irat) {
        irat = JRATIO[i - 2];
        break;
      }
    if (irat > JRATIO[23]) irat = JRATIO[23];
    if (irat > 1) {
      R = irat;
      DVSC = (double)irat * dv;
    }
  }
  // The input streams in panels of L points (RDSCAN :1439-1496); SHRKSC's
  // average runs on from the first panel read, carrying NREN points
  // across panels (:1298-1372); CONVSC takes each panel's ILO..IHI with
  // ZPEAK = (I - 1) RATIO + (V1I - VFT) / DVO, and a group whose JMAX
  // passes MAXF first has PANLSC write the R1 panel and move VFT on by
  // NLIMF - 1 points (:1562-1566, 1633-1643).  The walk records each
  // group's ZPEAK and R1 panel; it needs no data (a group of zeros adds
  // nothing and moves no later group's panel).
  const int64_t L = S->npanel > 0 ? S->npanel : 2400;
  const int NLIMF = 2401;
  const double ZBOUND = (HWF / DVO) * HWHM;
  const int MAXF = NLIMF + 2 * NBOUND + NSHIFT;
  const double V1SHFT = R > 1 ? (double)(R - 1) * dv / 2. : 0.;
  int64_t o = -1, gbase = 0;
  int nren = 0, q = 0;
  double vft = VFT;
  std::vector<double> zpk;
  std::vector<int32_t> qg;
  int64_t glo = -1, ghi = -1;
  for (int64_t p0 = 0; p0 < n_in;) {
    const int64_t np = std::min(L, n_in - p0);
    double vmin = v1c + (double)p0 * dv;
    const double vmax = v1c + (double)(p0 + np - 1) * dv;
    const int64_t pstart = p0;
    p0 += np;
    if (vmax < VBOT) continue;  // skipped before any data: nren is 0
    if (o < 0) o = pstart;
    double dvi = dv;
    const int nlow = nren > 0 ? nren + 1 : 1;
    vmin = vmin - (double)(nlow - 1) * dvi;
    const int nni = (int)np + nlow - 1;
    int ilo = 1, ihi = nni, idata = 0;
    const double dif = (vmin - VBOT) / dvi;
    if (dif < 0.) ilo = fint(-dif + 1.5);
    if (vmax > VTOP) {
      ihi = fint((VTOP - vmin) / dvi + 1.5);
      idata = 1;
    }
    double v1i = vmin;
    int nlim = nni;
    if (R > 1) {
      nren = nlim - (nlim / R) * R;
      v1i = v1i + V1SHFT;
      nlim = (nlim - nren) / R;
      dvi = DVSC;
      ilo = fint((VBOT - v1i) / dvi + 1.5);
      if (ilo < 1) ilo = 1;
      ihi = fint((VTOP - v1i) / dvi + 1.5);
      if (ihi > nlim) ihi = nlim;
    }
    const double ratio = dvi / DVO;
    double zint = ((v1i - vft) / DVO);
    for (int i = ilo; i <= ihi; i++) {
      double zpeak = (double)(i - 1) * ratio + zint;
      if (fint(zpeak + ZBOUND + 1.5) > MAXF) {  // PANLSC, then CONVSC again from I
        vft = vft + (double)(NLIMF - 1) * DVO;
        q++;
        zint = ((v1i - vft) / DVO);
        zpeak = (double)(i - 1) * ratio + zint;
        if (fint(zpeak + ZBOUND + 1.5) > MAXF) return LBLX_E_ARG;  // a group spans a panel
      }
      const int64_t g = gbase + i - 1;
      if (glo < 0) glo = g;
      for (int64_t h = ghi + 1; ghi >= 0 && h < g; h++) {  // not convolved
        zpk.push_back(0.);
        qg.push_back(-1);
      }
      zpk.push_back(zpeak);
      qg.push_back(q);
      ghi = g;
    }
    gbase += nlim;
    if (idata) break;  // CONVSC's IPANEL = IDATA: no further panel is read
  }
  if (o < 0 || glo < 0) return LBLX_E_ARG;  // 'END OF FILE ENCOUNTERED'
  const int64_t K = (n_in - o) / R;
  const double vg0 = v1c + (double)o * dv + V1SHFT;
  P->jfn = jfn;
  P->jtrem = jtrem;
  P->jabs = jabs;
  P->hwhm = HWHM;
  P->dvo = DVO;
  P->vft = VFT;
  P->v1out = VFT + (double)NSHIFT * DVO;
  P->nout = (int64_t)((V2 - VFT) / DVO + 1.5) - NSHIFT;  // PANLSC NNHI (:1627-1630)
  if (P->nout < 1) return LBLX_E_ARG;
  P->R = R;
  P->o = o;
  P->K = K;
  P->vg0 = vg0;
  P->dvsc = DVSC;
  P->g_lo = glo;
  P->g_hi = ghi;
  P->maxf = MAXF;
  P->zpk.swap(zpk);
  P->q.swap(qg);
  // CONVSC's constants (:1541-1560), JVAR = 0
  P->zbound = ZBOUND;
  P->zslope = (DVO / DXF) / HWHM;
  P->xnorm = DVSC / HWHM;
  scanfn_table(jfn, DXF, NJ[jfn], NJMX[jfn], XSCAL[jfn], P->xf);
  return LBLX_OK;
}

int plan_scan(const lblx_scan *S, double v1c, double dv, int64_t n_in, ScanPlan *P, const ScanTab *fixed) {
  if (!S || !P || !(dv > 0.) || n_in < 4) return LBLX_E_ARG;
  *P = ScanPlan{};
  double HWHM = S->hwhm;
  const double V1 = S->v1, V2 = S->v2;
  const int jfnin = S->jfnin, jfn = jfnin < 0 ? -jfnin : jfnin;
  if (HWHM <= 0 || S->jemit < 0 || S->jemit > 1 || jfn > 6) return LBLX_E_UNSUPPORTED;
  double A = 0.;
  if (jfnin > 0) {
    A = SC_C[jfn] * HWHM;
  } else if (jfnin < 0) {
    A = 1. / HWHM;  // PATHL = HWHM field
    HWHM = S->param == 0. ? A / std::fabs(SC_C[jfn]) : S->param;
  }
  if (jfn == 2 && S->param != 0.) return LBLX_E_UNSUPPORTED;
  const int nobox = S->mratin < 0;
  const int mratio = S->mratin > 0 ? S->mratin : 12;
  const double V2C = v1c + dv * (double)(n_in - 1);
  if (V1 >= V2 || V1 >= V2C || V2 <= v1c) return LBLX_E_ARG;
  // FFTSCN 442-470
  double V1Z = V1 - HWHM * SC_CLIMIT[jfn], V2Z = V2 + HWHM * SC_CLIMIT[jfn];
  V1Z = v1c + (double)fint((V1Z - v1c) / dv + 1.01) * dv;
  V2Z = v1c + (double)fint((V2Z - v1c) / dv + 1.01) * dv;
  if (V1Z < v1c) V1Z = v1c;
  if (V2Z > V2C) V2Z = V2C;
  if ((V2Z - V1Z) / HWHM < SC_CLIMIT[jfn]) return LBLX_E_ARG;
  if (dv > 2. * HWHM / 12) return LBLX_E_UNSUPPORTED;  // INTPDR onto 2 HWHM / 12 first
  int ivx = S->ivx;
  if (ivx == 0) ivx = (V2Z - V1Z) / HWHM > SC_CRATIO[jfn] ? 1 : -1;
  if (jfn == 0) ivx = 0;
  if (ivx == -1) return LBLX_E_UNSUPPORTED;  // spectral-domain instrument function
  // Loadsp
  P->n0 = (int64_t)std::llround((V1Z - v1c) / dv);
  P->ltot = (int64_t)((V2Z - V1Z) / dv + 1.1);
  V2Z = V1Z + dv * (double)(P->ltot - 1);
  if (P->n0 < 0 || P->n0 + P->ltot > n_in) return LBLX_E_ARG;
  int lrec = scan_records(P->ltot);
  P->fill = S->jemit == 0 ? 1.0 : 0.0;
  int M = 1;
  if (jfn == 0) {
    M = (int)(2.0 * HWHM / dv);
    if (M < 2) return LBLX_E_ARG;
  } else if (!nobox) {
    M = (int)(2.0 * HWHM / (mratio * dv));
  }
  if (M >= 2 && !nobox) {
    V1Z = V1Z + dv * (M - 1) / 2.0;
    P->ktot = P->ltot / M;
    dv = M * dv;
    V2Z = V1Z + dv * (double)(P->ktot - 1);
    P->N = (int64_t)scan_records(P->ktot) * LPTSMX;
    P->M = M;
  } else {
    P->M = 1;
    P->ktot = P->ltot;
    P->N = (int64_t)lrec * LPTSMX;
  }
  P->jfn = jfn;
  P->A = A;
  P->hwhm = HWHM;
  P->dvb = dv;
  P->DX = 1.0 / ((double)P->N * dv);
  P->fix = (P->M >= 2 && !nobox && S->nofix == 0) ? 1 : 0;
  P->k4 = -1;
  if (jfn == 4) {  // Scnfnt label 40: the sample straddling 1/A
    double alen = 1.0 / A;
    int64_t nn = (int64_t)(alen / P->DX) + 1;
    double r = alen / P->DX + 1 - (double)nn;
    if (r <= 0.5) {
      r = r + 1.0;
      nn = nn - 1;
    }
    if (nn >= 1) {
      P->k4 = nn;
      P->r4 = r - 0.5;
    }
  }
  // label 95: points handed to Wrtspc
  int64_t N1 = (int64_t)((V1 - V1Z) / dv);
  double V1S = V1Z + (double)(N1 - 1) * dv;
  if (V1S < V1Z) {
    N1 = 1;
    V1S = V1Z;
  }
  int64_t N2 = (int64_t)((V2 - V1Z) / dv + 3.01);
  double V2S = V1Z + (double)(N2 - 1) * dv;
  if (V2S > V2Z) N2 = (int64_t)((V2Z - V1Z) / dv + 1.01);
  P->n1 = N1;
  P->ntot = N2 - N1 + 1;
  P->v1s = V1S;
  if (!(S->dvout > 0)) {
    P->v1out = V1S;
    P->dvout = dv;
    P->nout = P->ntot;
    return LBLX_OK;
  }
  // INTPDR / INTERP four-point walk over Wrtspc's 2400-point panels.
  // Panel k holds min(2400, nz - 2400 k) points from v1p(k); v1p(k+1) =
  // v2p(k) + DV.  The walk only moves forward, so each panel's bounds are
  // formed when it is reached, in the order Wrtspc writes them.
  const int64_t nz = P->ntot;
  const double dvi = dv, dvo = S->dvout;
  const int npan = (int)((nz + 2399) / 2400);
  int pk = 0;
  double pv1 = V1S, pv2 = 0.;
  int pn = 0;
  auto panel_at = [&](int k) {  // bounds of panel k, the one after pk
    pn = (int)std::min<int64_t>(2400, nz - 2400LL * k);
    if (k > 0) pv1 = pv2 + dvi;
    pv2 = pv1 + dvi * (double)(pn - 1);
  };
  panel_at(0);
  const double xnumcf = 201.;
  while (pk < npan && pv2 < V1 - 4.0 * dvo) {
    pk++;
    if (pk < npan) panel_at(pk);
  }
  if (pk >= npan) return LBLX_E_ARG;
  double V1I = pv1 - 4.0 * dvi, V2I = pv2;
  int NNI = pn + 4;
  int64_t base = 2400LL * pk - 5;  // global index of T(0)
  bool hit_eof = false;
  double V1J = V1, DVJ = dvo, VJ = V1J;
  const double vratio = DVJ / dvi;
  if (V1J < V1I) return LBLX_E_UNSUPPORTED;
  int J2 = 0;
  // one output panel of INTPDR (2400 points) before it is flushed
  int64_t gpan[2400];
  int32_t ippan[2400];
  int64_t nout = 0;
  bool over = false;
  auto flush = [&](int n) {
    if (fixed) {
      if (nout + n > fixed->cap) {
        over = true;
        return;
      }
      std::memcpy(fixed->g + nout, gpan, (size_t)n * sizeof(int64_t));
      std::memcpy(fixed->ip + nout, ippan, (size_t)n * sizeof(int32_t));
    } else {
      P->g.insert(P->g.end(), gpan, gpan + n);
      P->ip.insert(P->ip.end(), ippan, ippan + n);
    }
    nout += n;
  };
  for (int j = 0; j < 2400; j++) {
    gpan[j] = 0;
    ippan[j] = -1;
  }
  for (;;) {
    int I = fint((VJ - V1I) / dvi + 1.00001);
    double VI = V1I + dvi * (double)(I - 1);
    double Pf = (VJ - VI) / dvi;
    int J1 = fint((VJ - V1J) / DVJ + 1.001);
    J2 = std::min(fint((V2 - V1J) / DVJ + 1.001), fint((V2I - dvi - V1J) / DVJ + 1.));
    if (J2 > 2400) J2 = 2400;
    for (int J = J1; J <= J2; J++) {
      gpan[J - 1] = base + I;
      ippan[J - 1] = fint(Pf * xnumcf + 1.00001);
      Pf = Pf + vratio;
      if (Pf >= 1.0) {
        I = fint((double)I + Pf);
        Pf = Pf - (double)fint(Pf);
      }
    }
    VJ = V1J + DVJ * J2;
    if (J2 >= 2400 || VJ >= V2) {
      flush(J2);
      V1J = V1J + DVJ * (J2 - 1) + DVJ;
      J2 = 0;
    }
    if (over) return LBLX_E_SCRATCH;
    if (VJ >= V2) break;
    if (hit_eof) {  // label 80: zeros to the end of the band
      for (;;) {
        int J1b = J2 + 1;
        int J2b = std::min(fint((V2 - V1J) / DVJ + 1.0001), 2400);
        for (int J = J1b; J <= J2b; J++) ippan[J - 1] = -1;
        J2 = J2b;
        VJ = V1J + DVJ * J2;
        if (J2 >= 2400 || VJ >= V2) {
          flush(J2);
          V1J = V1J + DVJ * (J2 - 1) + DVJ;
          J2 = 0;
        }
        if (over) return LBLX_E_SCRATCH;
        if (VJ >= V2) break;
      }
      break;
    }
    if (I >= NNI - 2) {
      if (pk + 1 < npan) {
        pk++;
        panel_at(pk);
        V1I = pv1 - 4.0 * dvi;
        V2I = pv2;
        NNI = pn + 4;
        base = 2400LL * pk - 5;
      } else {
        hit_eof = true;  // T(NNI+1) extrapolated, T(NNI+2) = 0
        V2I = V2I + dvi;
      }
    }
  }
  P->v1out = V1;
  P->dvout = dvo;
  P->nout = nout;
  return LBLX_OK;
}

// K(x, y) = Re w(x + iy) by Armstrong's three-region method, as the
// exact-Voigt build evaluates it (armstrong, src/oprop_voigt.f90:12010-12134):
// Dawson's function from Hummer's Chebyshev series (asymptotic beyond
// x = 5) plus the y power series near the axis; 10-point Gauss-Hermite
// quadrature of the Voigt integral (and of its x-antiderivative for
// moderate y) elsewhere.
static double voigt_k(double x, double y) {
  const double y2 = y * y;
  const double *W = LBLX_VOIGT_ARM_W, *T = LBLX_VOIGT_ARM_T;
  if ((y < 1.0 && x < 4.0) || y < 1.8 / (x + 1.0)) {
    const double u1 = (x * x - y2 > 70.0) ? 0.0 : std::exp(-(x * x) + y2) * std::cos(2.0 * x * y);
    double d1, d0;  // D_n and D_{n-1} of the recurrence
    if (x > 5.0) {
      const double p2 = x * x, p4 = p2 * p2, p8 = p4 * p4;
      d1 = -(.5 / p2 + .75 / p4 + 1.875 / (p4 * p2) + 6.5625 / p8 + 29.53125 / (p8 * p2) +
             162.4218 / (p8 * p4) + 1055.7421 / ((p8 * p4) * p2));
      d0 = (1.0 - d1) / (2.0 * x);
    } else {
      const double t = x / 5.0, c = 4.0 * t * t - 2.0;
      double b0 = 0., b1 = 0., b2 = 0.;
      for (int k = 33; k >= 0; k--) {
        b0 = c * b1 - b2 + LBLX_VOIGT_ARM_C[k];
        b2 = b1;
        b1 = b0;
      }
      d0 = t * (b0 - b2);
      d1 = 1.0 - 2.0 * x * d0;
    }
    double sum = y * d1;
    if (y > 1.0e-08) {
      double sgn = 1.0, yp = y;
      for (int n = 2; n <= 50; n++) {
        const double dn = (x * d1 + d0) * (-2.) / (double)n;
        d0 = d1;
        d1 = dn;
        if (n & 1) {
          sgn = -sgn;
          yp = yp * y2;
          const double term = dn * yp;
          sum = sum + sgn * term;
          if (std::fabs(term / sum) <= 1.0e-08) break;
        }
      }
    }
    return u1 - 1.12837917 * sum;
  }
  double g = 0.;
  if (y < 2.5 && x < 4.0) {
    for (int i = 0; i < 10; i++) {
      const double r = T[i] - x, q = T[i] + x;
      g = g + (4.0 * T[i] * T[i] - 2.0) *
                  (r * std::atan(r / y) + q * std::atan(q / y) - 0.5 * y * (std::log(y2 + r * r) + std::log(y2 + q * q))) *
                  W[i];
    }
    return 0.318309886 * g;
  }
  for (int i = 0; i < 10; i++)
    g = g + (1.0 / ((x - T[i]) * (x - T[i]) + y2) + 1.0 / ((x + T[i]) * (x + T[i]) + y2)) * W[i];
  return 0.318309886 * y * g;
}

// Tables of the exact-Voigt build (voigt_init, src/oprop_voigt.f90:11739-12008).
// For zeta = 0, 0.01, .., 1 (Doppler HWHM 1, Lorentz zeta/(1-zeta)) the
// Voigt profile in units of its halfwidth AV = AVC(zeta)(AL + AD) is
// split over the domains 4, 16, 64 halfwidths (2001-point grids):
//   F1 = V - q1 on [0,4);  F2 = q1 - q2 on [0,4], V - q2 on (4,16);
//   F3 = q2 - q3 on [0,16], V - q3 on (16,64),
// q_k = a_k + b_k z^2 matching V and its central-difference slope at the
// domain end.  f: 3 tables x 101 columns x 2601 (column = zeta index,
// entry = Fortran IZ - 1), then one zero column, in COMMON /FNSH/ order so
// that column 101 of a table is column 0 of the next.  coef: a_1 a_2 a_3
// b_1 b_2 b_3, 101 each.
void voigt_tables_host(double *f, double *coef) {
  const int NZ = 101, NX = 2001, NF = 2601;
  std::fill(f, f + (size_t)(3 * NZ + 1) * NF, 0.);
  const double dom[3] = {4.0, 16.0, 64.0};
  const double pi = 2.0 * std::asin(1.0);
  const double cl = std::sqrt(std::log(2.0));
  std::vector<double> zg(NX + 1), v(NX + 1);
  double *a = coef, *b = coef + 3 * NZ;  // a[k*NZ + iz], b[k*NZ + iz]
  for (int iz = 0; iz < NZ; iz++) {
    const double zeta = (double)iz / (double)(NZ - 1);
    const bool lorentz = !(zeta < 1.0);
    const double ad = lorentz ? 0. : 1., al = lorentz ? 1.0 : zeta / (1 - zeta);
    const double av = LBLX_VOIGT_AVC[iz] * (ad + al);
    const double norm = cl / std::sqrt(pi) * av;
    const double y = lorentz ? 100.0 : cl * zeta / (1. - zeta);
    for (int d = 0; d < 3; d++) {
      const double step = dom[d] / 2000.;
      const double dx = ad != 0.0 ? (cl * av) / ad * step : (cl * av) * step;
      for (int i = 0; i <= NX; i++) {
        zg[i] = (double)i * dom[d] / 2000.;
        v[i] = lorentz ? (1. / pi) * (1. / (1. + zg[i] * zg[i])) : voigt_k((double)i * dx, y) * norm;
      }
      const double slope = .5 * (v[NX] - v[NX - 2]) / step;
      b[d * NZ + iz] = slope / (2 * dom[d]);
      a[d * NZ + iz] = v[NX - 1] - b[d * NZ + iz] * (dom[d] * dom[d]);
      double *col = f + (size_t)(d * NZ + iz) * NF;
      for (int i = 0; i <= NX - 2; i++) {
        const double z2 = zg[i] * zg[i];
        const double q_lo = d > 0 ? a[(d - 1) * NZ + iz] + b[(d - 1) * NZ + iz] * z2 : 0.;
        const double q = a[d * NZ + iz] + b[d * NZ + iz] * z2;
        col[i] = (d > 0 && zg[i] <= dom[d - 1]) ? q_lo - q : v[i] - q;
      }
    }
  }
}

}  // namespace lblx
