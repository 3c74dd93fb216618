// Host-side planner for one layer: the panel sequence, far-wing window
// schedule and radiation-term node tables that HIRAC1/LBLF4 would walk
// through sequentially (src/oprop.f90:223-540, 1418-1558, 3641-3766).
// They depend only on the layer grid, never on line data, so the host
// computes them once per layer in double precision, with the same
// operation order as the Fortran, and the kernels index them.
#pragma once
#include <cstdint>
#include <vector>

#include "../../include/lblx.h"

namespace lblx {

// Physical and numerical constants of the dbl build.
constexpr double RADCN2 = 1.4387752;       // phys_consts.f90:42
constexpr double RADCN1 = 1.191042722E-12; // phys_consts.f90:41
constexpr double ONEPL = 1.001, ONEMI = 0.999;  // lblrtm.f90:631-632
constexpr double P0 = 1013.25, TEMP0 = 296.;    // lblrtm.f90:2115-2116
constexpr int NSHIFT = 32, NLIM1 = 2401, NLIM2 = 601, NLIM3 = 151;
constexpr int LIMOUT4 = 2500;   // R4 window points (LBLF4 LIMOUT)
constexpr int R4STRIDE = 2504;  // device row length of one window

inline int fint(double x) { return (int)x; }

// One segment of the piecewise-linear radiation term (RADFNI).
struct RadSeg {
  int32_t i1, i2;   // local point range (1-based, inclusive)
  double radvi;     // value at i1
  double rdel;      // increment per point
};

struct PanelPlan {
  double vft;       // VFT of this panel
  double v2p;       // V2P of the panel header
  int32_t w;        // far-wing window active at XINT time
  int32_t nhi;      // NHI (2432 except the last panel)
  int32_t istop;    // last panel
  int32_t seg0, nseg;  // radiation segments (JRAD = 0)
  int64_t out0;     // index of the panel's first point in the layer output
};

struct WinPlan {
  double v1r4, v2r4;
  int32_t nptr4;
  int32_t seg0, nseg;  // radiation segments of R4 (JRAD = 1)
  int32_t p_first;     // first panel that uses this window
};

// CONTNM set-up of one layer (src/contnm.f90:20-990, OPDPTH grid
// src/lblrtm.f90:6214-6226): the ABSRB grid and, per continuum, the
// coarse-grid walk (V1C, NPTC, first table index I1), the pre_xint clip
// [ist, last] and the layer scalars that multiply the tables.
struct ContTerm {
  int32_t on;
  int32_t i1, nptc, ist, last;  // I1 (table index of coarse point 1), NPTC, clip
  double v1c, v2c, dvc;
  double f0, f1, f2, f3, f4, f5, f6, f7;  // term scalars (see plan_cont)
};
struct ContPlan {
  int32_t on;
  double v1abs, v2abs, dvabs;
  int32_t nptabs, jrad;
  double xkt, tave;
  ContTerm h2o, co2, o2, n2, n2r;
  int32_t ray_on;
  double ray_f;      // WTOT * conv_cm2mol
  double xself, xfrgn, wk1;
  int32_t forcl;     // MT_CKD foreign closure table
  int32_t aj;        // analytic-Jacobian species run: H2O derivative term added
};

struct LayerPlan {
  // grid
  double dv, dvr2, dvr3, dvr4, bound4, alfmax, bound;
  int32_t nalfmax, nbound, max1, max2, max3;
  double v1, v2, vft0, vbot, vtop;
  double v1r4st, v2r4st, v1l4, v2l4;
  // radiation and rejection
  int32_t jrad;
  double tave, pave;
  double xkt, dptmn_hirac, dptfc, dptmn4;
  // molecular quantities
  double rhorat, tratio, pavp0, xkt0, betacr_l, beta4, betacr4;
  double scor[7][10], alfd1[7][10], rhoslf[LBLX_MXMOL];
  // plan tables
  std::vector<PanelPlan> panels;
  std::vector<WinPlan> wins;
  std::vector<RadSeg> segs;
  int64_t npts;       // output points
  int64_t g_last;     // last global fine index written
  int64_t g2max;      // last R2 index that receives L4(R3)
  int64_t g3max;      // last R3 index that receives XINT
  int64_t n1, n2, n3; // array sizes (fine, R2, R3 grids)
  ContPlan cont;      // ICNTNM != 0
};

// FFTSCN set-up (src/fftscn_dbl.f90:25-789) for an input spectrum on
// v1c + k*dv, k < n_in: the loaded range, the boxcar, the FFT length, the
// transform-domain filter parameters, the extracted range and, for
// DVOUT > 0, INTPDR's four-point walk (src/postsub.f90:2940-3100) as one
// (input index, coefficient index) pair per output point.
struct ScanPlan {
  int jfn;                // |JFNIN|
  double A;               // apodization parameter (1/OPD)
  double hwhm;
  int64_t n0, ltot;       // first loaded input point, points loaded
  int M;                  // boxcar width (1: none)
  int64_t ktot;           // points after the boxcar
  double dvb;             // grid after the boxcar
  int64_t N;              // FFT length
  double fill;            // padding (1 transmittance, 0 radiance)
  double DX;              // OPD step of the transform bins
  int fix;                // boxcar deconvolution Z / sin Z
  int64_t k4;             // JFN 4: bin holding the fractional edge value
  double r4;
  int64_t n1, ntot;       // extracted points N1 (1-based) .. N1+ntot-1
  double v1s;             // first extracted wavenumber
  double v1out, dvout;
  int64_t nout;
  std::vector<int64_t> g; // INTPDR: input index of T(I) per output point
  std::vector<int32_t> ip;  // coefficient index (0..202), -1: zero fill
};

// Fixed storage for the INTPDR tables (g, ip of ScanPlan) for callers that
// allocate nothing: plan_scan writes up to cap entries there instead of
// growing the vectors, and returns LBLX_E_SCRATCH when the output has more.
struct ScanTab {
  int64_t *g;
  int32_t *ip;
  int64_t cap;
};
int plan_scan(const lblx_scan *scan, double v1c, double dv, int64_t n_in, ScanPlan *P,
              const ScanTab *fixed = nullptr);

// SCANFN (postsub.f90:25-560) for one input spectrum: the function table,
// SHRKSC's average (ratio R from JRATIO, running from the first point of
// the first input panel reaching VBOT), the shrunk points CONVSC takes
// (g_lo..g_hi), and CONVSC's grid constants in one global output index J
// (1-based, v = vft + (J - 1) dvo), whose points NSHIFT + 1 .. NSHIFT +
// nout PANLSC writes.
struct ScanFnPlan {
  int jfn, jtrem, jabs;
  double hwhm, dvo, vft, v1out;
  int64_t nout;
  int R;                 // SHRKSC ratio (1: no average)
  int64_t o;             // first input point of the first group
  int64_t K;             // groups (the last partial group is dropped)
  double vg0, dvsc;      // wavenumber of group 0, group spacing
  int64_t g_lo, g_hi;    // groups CONVSC convolves
  double zbound, zslope, xnorm;
  int maxf;              // R1 buffer length NLIMF + 2 NBOUND + NSHIFT
  std::vector<double> xf;  // XF(1..NFMAX) at xf[0..NFMAX-1]
  // per group g_lo..g_hi: ZPEAK as CONVSC computes it, in the R1 panel q
  // current when the group is convolved (q = -1: outside ILO..IHI)
  std::vector<double> zpk;
  std::vector<int32_t> q;
};
int plan_scanfn(const lblx_scanfn *scan, int iemit, double v1c, double dv, int64_t n_in, ScanFnPlan *P);

// Build the plan; layer->jrad is updated as OPDPTH does.  Returns LBLX_OK
// or an LBLX_E_* code.
// icflg > 0: the layer is an analytic-Jacobian single-species run (CONTNM's
// /CDERIV/ ICFLG, contnm.f90:221-239, 314-322): no N2 continuum, and the
// H2O continuum carries its ln q derivative term.
int plan_layer(const lblx_run *run, lblx_layer *layer, LayerPlan *P, int icflg = 0);
// the layer's JRAD as plan_layer leaves it (the next layer's JRAD in a column)
int layer_jrad(const lblx_layer *layer);

// PATH DV selection for a column (lblrtm.f90:5336-5500).
int plan_dv(const lblx_run *run, int nlay, int nmol, const double *pave,
            const double *tave, const double *wk, const double *wbroad,
            int iemit, double *dv, int32_t *itype);

double radfn(double vi, double xkt);
double radfni(double vi, double dvi, double xkt, double *vinew, double *rdel,
              double *rdlast);
void tips_scor(double temp, double scor[7][10]);
// d ln SCOR/dT: central difference of the TIPS interpolation over +-h
void tips_dlnscor(double temp, double h, double out[7][10]);
// CONTNM's plan alone for a layer whose JRAD is already set
int plan_cont_layer(const lblx_run *run, const lblx_layer *layer, int jrad, ContPlan *C);

}  // namespace lblx
