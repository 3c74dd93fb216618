// Tangent-linear layer optical depth: the analytic derivatives of each
// layer's optical depth with respect to its temperature and the log column
// amounts of chosen molecules, taken on the decisions of the forward run
// (its kept near lines and their scatter records, its SHRINK groups, the
// far-wing lines each window accepted, its panels and windows).  They feed
// the exact tangents of the merge (k_jac_*), so a column's Jacobian costs
// one forward run plus these passes instead of one OPDPTH run per
// parameter.
//
// Near lines (LNCOR1, src/oprop.f90:995-1193; CNVFNV, :1262-1320):
//   S' = S W [RADFN(v, T)] SCOR(T) exp(-E'' (c2/T - c2/T0)) (1 + exp(-c2 v'/T))
//   aL = (T/T0)^n (alfa0 (rho - rho_s) + hwhm rho_s),   rho, rho_s ~ 1/T
//   aD = v' ALFD1,  ALFD1 ~ sqrt(T),  v' = v + rho shift
//   zeta = aL/(aL + aD),   aV = AVRAT(zeta) (aL + aD)  (no derivative where
//   ALFV is clamped to [DV, ALFMAX])
// and the table weights w_k = (S'/aV) CF_k(zeta).  A line adds w F(|z|) at
// z = (x - v')/(aV DXF_G), so
//   d(w F) = dw F(|z|) - w F'(|z|) (sign(z) dv'/(aV DXF_G) + |z| daV/aV),
// F' being the slope of the tabulated shape per table step (central
// differences of the table).  For ln q of molecule m, a line of molecule
// m' has d ln S'/d ln q = [m' = m] and, since rho_s(m') = rho W_m'/W_tot,
//   d rho_s(m')/d ln q_m = rho_s(m') ([m' = m] - W_m/W_tot):
// its own molecule's lines widen with their self-broadened share, and every
// line narrows through the total amount (O2, at 21 %, carries most of
// that).  The second term is the same for all channels up to the factor
// W_m/W_tot, so it runs as one internal channel U (every line's tangent for
// d rho_s = rho_s, strength fixed) and channel m takes
//   d tau/d ln q_m = own_m - (W_m/W_tot) U
// at the output.
//
// Far wing (LINF4 + SHRINK, :3093-3235, 3508-3640; CONVF4, :3878-4039):
// each TAPE9 line is the strength-weighted mean of a group's members, so
// its derivative follows from the members' exactly:
//   dS = sum dS_k,  d<a> = (sum dS_k (a_k - <a>) + S_k da_k)/S
// for a = centre, Lorentz and Doppler width.  CONVF4's value is then
// differentiated through SIL, SIV, F4BND, ZVSQ and the centre.
//
// Radiation: d ln RADFN/dT = -x/(T sinh x), x = v/XKT, with RADFN's
// branches, enters the strengths (JRAD = 1), the far-wing window
// (JRAD = 1) and the optical depth (JRAD = 0).  The continuum channels are
// central differences of CONTNM's closed forms on host-perturbed plans.
#pragma once
#include <hip/hip_runtime.h>

#include "lblx_kernels.h"

namespace lblx {

constexpr int TAN_NQ = 9;  // internal channels: temperature, up to seven molecules, U

// Per-job constants of the tangent.
struct TanLayer {
  double tave;
  double dlnscor[7][10];  // d ln SCOR/dT (slope of the TIPS interpolation)
  double wq[TAN_NQ];      // W_m / W_tot of channel q's molecule (0 for the temperature)
  int32_t chan[7];        // channel of molecule m+1 (0: none, 1..nspec)
  int32_t nchan;          // output channels: 1 + nspec
  int32_t nint;           // internal channels: nchan, + U (index nchan) when nspec > 0
};

// Derivative of one near line's contribution for one parameter: weights of
// the F1, FG, XVER, F2, F3 tables, the centre term dv' recalf / DXF1 and
// the width term daV/aV.
struct TanPar {
  float d1, dd, dv, d2, d3, a, b;
};
struct TanLine {
  TanPar t, s, u;  // temperature; the line's own molecule; U
  int32_t q;       // channel of that molecule (0: none)
  int32_t pad;
};

// d ln RADFN(v, XKT)/dT with RADFN's branches (oprop.f90:2235-2296):
// 0.5 x v ~ 1/T up to x = 0.01, v tanh(x/2) up to x = 10, v beyond
__device__ __forceinline__ double d_dlnradfn_dT(double v, double xkt, double t) {
  const double x = v / xkt;
  if (x <= 0.01) return -1. / t;
  if (x <= 10.) return -x / (t * sinh(x));
  return 0.;
}

// d ln(strength)/dT of the temperature correction that LNCOR1 and LINF4
// share: SCOR, the lower-state Boltzmann factor and the stimulated-emission
// term 1 + exp(-c2 v'/T), with dv'/dT = dnu.
__device__ __forceinline__ double d_dlns_dT(double dlnscor, double epp, double vnu, double dnu, double t) {
  const double c2 = 1.4387752;
  const double e = exp(-vnu * c2 / t);
  return dlnscor + epp * c2 / (t * t) + (c2 * vnu / (t * t) - c2 * dnu / t) * e / (1. + e);
}

// ---------------------------------------------------------------------
// Near lines: LNCOR1's correction of each line again, with its
// derivatives, as the weights of the tangent tiles.  For the kept lines
// only, stored like their scatter records (at the kept-line index kidx).
__global__ __launch_bounds__(256) void k_near_tan(DevLines L, const DevLayer *Ys, const TanLayer *Ts,
                                                  Strides S, const int *keep, const int32_t *kidx,
                                                  TanLine *out) {
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  const TanLayer &TL = Ts[job];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Y.nn || !keep[job * S.line + t]) return;
  const int64_t slot = job * S.line + kidx[job * S.line + t];
  TanLine o{};
  const int i = Y.nlo + t;
  const int mc = L.mol[i];
  const int m = mc % 100, iso = (mc % 1000) / 100;
  if (m < 1 || m > 7 || m > Y.nmol || iso < 1 || iso > c_isomax[m - 1]) {
    out[slot] = o;
    return;
  }
  const double T = TL.tave;
  double sui = (double)L.sp[i] * Y.wk[m - 1];
  double dlns = 0.;
  if (Y.jrad == 1) {
    sui = sui * d_radfn(L.vnu[i], Y.xkt);
    dlns = d_dlnradfn_dT(L.vnu[i], Y.xkt, T);
  }
  double alfa0 = L.alfa[i], hwhm = L.hwhm[i];
  if (m == 7) alfa0 = (alfa0 - 0.21 * hwhm) / (1.0 - 0.21);
  const double vnu = L.vnu[i] + Y.rhorat * (double)L.pshift[i];
  const double dnu_t = -Y.rhorat * (double)L.pshift[i] / T;
  const double n = L.tmpalf[i];
  const double tmpcor = pow(Y.tratio, n);
  const double alfa0i = alfa0 * tmpcor, hwhmsi = hwhm * tmpcor;
  const double rs = Y.rhoslf[m - 1];
  const double alfl = alfa0i * (Y.rhorat - rs) + hwhmsi * rs;
  const double alfd1 = Y.alfd1[m - 1][iso - 1];
  const double alfad = vnu * alfd1;
  const double sum = alfl + alfad;
  const double zeta = alfl / sum;
  int iz = fint(100. * zeta + ONEPL_D);
  iz = iz < 1 ? 1 : (iz > 101 ? 101 : iz);
  const double zetdif = 100. * zeta - (double)(iz - 1);
  const double avr = c_avrat[iz - 1] + zetdif * (c_avrat[iz] - c_avrat[iz - 1]);
  const double avs = 100. * (c_avrat[iz] - c_avrat[iz - 1]);  // dAVRAT/dzeta
  double alfv = avr * sum;
  const bool clamped = alfv < Y.dv || alfv > Y.alfmax;
  if (alfv < Y.dv) alfv = Y.dv;
  if (alfv > Y.alfmax) alfv = Y.alfmax;
  const double recalf = 1. / alfv;
  double epp = L.epp[i];
  if (epp <= -1.001) epp = fabs(epp);
  if (epp > -0.999) {
    sui = sui * Y.scor[m - 1][iso - 1] * exp(-epp * Y.betacr_l) * (1. + exp(-vnu / Y.xkt));
    dlns += d_dlns_dT(TL.dlnscor[m - 1][iso - 1], epp, vnu, dnu_t, T);
  }
  const double depth = sui * recalf;
  const double *ct[5] = {c_cf1, c_cgauss, c_cer, c_cf2, c_cf3};
  double cw[5], cs[5];
  for (int k = 0; k < 5; k++) {
    cw[k] = ct[k][iz - 1] + zetdif * (ct[k][iz] - ct[k][iz - 1]);
    cs[k] = 100. * (ct[k][iz] - ct[k][iz - 1]);
  }
  auto fill = [&](TanPar &p, double dln_s, double dal, double dad, double dnu) {
    const double dz = (dal * alfad - alfl * dad) / (sum * sum);
    const double dav = clamped ? 0. : avs * dz * sum + avr * (dal + dad);
    const double dlnd = dln_s - dav * recalf;
    double w[5];
    for (int k = 0; k < 5; k++) w[k] = depth * (cw[k] * dlnd + cs[k] * dz);
    p.d1 = (float)w[0];
    p.dd = (float)w[1];
    p.dv = (float)w[2];
    p.d2 = (float)w[3];
    p.d3 = (float)w[4];
    p.a = (float)(dnu * recalf / 0.002);
    p.b = (float)(dav * recalf);
  };
  // temperature: widths ~ T^(n-1) (rho ~ 1/T), Doppler ~ v' sqrt(T)
  fill(o.t, dlns, alfl * (n - 1.) / T, alfad / (2. * T) + alfd1 * dnu_t, dnu_t);
  const int q = TL.chan[m - 1];
  o.q = q;
  const double dals = (hwhmsi - alfa0i) * rs;  // daL for d rho_s = rho_s
  if (q > 0) fill(o.s, 1., dals, 0., 0.);
  if (TL.nint > TL.nchan) fill(o.u, 0., dals, 0., 0.);
  out[slot] = o;
}

// ---------------------------------------------------------------------
// Tangent tiles of grid G: the channels' near-line scatter, then PANEL's
// combine per channel as k_tile does it:
//   G = 3: R3' = lines + XINT(R4') + XINT(ABSRB')
//   G = 2: R2' = lines + Lagrange4(R3')
//   G = 1: OD' = (lines + Lagrange4(R2')) R  [+ OD d ln R/dT for the
//          temperature when JRAD = 0]
// tabs: F1 FG XVER F2 F3, then their slopes per table step (10 x NFMX
// floats).  Channel q of the R2/R3 rows is at rin/rout + q * sin/sout;
// its optical-depth derivative goes to odc + odcoff[job * nchan + q].
//
// No LDS accumulators (k_tan_seg): each of the 16 waves
// owns a contiguous segment of the tile, 64 NS points, lane i the points
// i + 64 s, and keeps their channels in registers; it walks the lines whose
// extent may reach its segment (the tile's line range narrowed by a binary
// search on the line centres) and adds each line's contributions to its own
// points.  Lines reaching two segments are walked by both waves; nothing is
// shared between waves, so no atomics and no accumulator in LDS (on gfx950
// ds_add_f32 runs near lane-serial and ds_add_f64 needs twice the LDS).
// NSP: species channels held
// (>= nspec); NS points per lane, sized to keep the registers of a
// 1024-thread workgroup (128 per lane) free of spills: 6 (4 with seven
// species) for R1, 4 for R2, 2 for R3 (tiles of 64 NS 16 points).
constexpr int TSEG_T = 1024;
template <int G, int NSP>
struct TSeg {
  static constexpr int NS = G == 1 ? (NSP <= 3 ? 6 : 4) : (G == 2 ? 4 : 2);
  static constexpr int TILE = 64 * NS * (TSEG_T / 64);
};

template <int G, typename OUT, int NSP>
__global__ __launch_bounds__(TSEG_T) void k_tan_seg(Recs recs, const TanLine *__restrict__ tl, TileMap M,
                                                     const double *vraw, const int32_t *kidx,
                                                     const double *margin, const float *tabs, const DevLayer *Ys,
                                                     const TanLayer *Ts, Strides S, const DevPanel *pans,
                                                     const DevWin *wins, const double *r4c, const double *absc,
                                                     int abs_stride, const DevSeg *segs, const double *rin, int64_t sin,
                                                     double *rout, int64_t sout, const OUT *od, OUT *odc,
                                                     const int64_t *odcoff) {
  constexpr int NS = TSeg<G, NSP>::NS;
  constexpr int TILE = TSeg<G, NSP>::TILE;
  constexpr int NT = G == 1 ? 3 : 1;  // tables of the grid
  extern __shared__ double sm_t[];
  float2 *tf = reinterpret_cast<float2 *>(sm_t);  // each table entry with its slope
  const int tid = threadIdx.x;
  const int job = M.job[blockIdx.x];
  const DevLayer &Y = Ys[job];
  const TanLayer &TL = Ts[job];
  const int NO = TL.nchan, nsp = NO - 1;
  const bool hasu = TL.nint > TL.nchan;
  const int t0 = G == 1 ? 0 : (G == 2 ? 3 : 4);
  for (int k = tid; k < NT * NFMX; k += TSEG_T) {
    const int a = k / NFMX, r = k - a * NFMX;
    tf[k] = make_float2(tabs[(size_t)(t0 + a) * NFMX + r], tabs[(size_t)(5 + t0 + a) * NFMX + r]);
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int64_t T0 = (int64_t)M.loc[blockIdx.x] * TILE;
  const int64_t s_lo = T0 + (int64_t)wave * 64 * NS;  // this wave's segment [s_lo, s_lo + 64 NS)
  const int64_t s_hi = s_lo + 64 * NS - 1;
  // lines whose extent may reach the segment: centres within the margin
  const double dvg = G == 1 ? Y.dv : (G == 2 ? Y.dvr2 : Y.dvr3);
  const double mg = margin[job * 3 + G - 1];
  const double va = Y.vft0 + (double)s_lo * dvg - mg, vb = Y.vft0 + (double)(s_hi + 1) * dvg + mg;
  const double *v = vraw + Y.nlo;
  int l0 = M.lo[blockIdx.x], h0 = M.hi[blockIdx.x];
  {
    int l = l0, h = h0;
    while (l < h) {
      const int m = (l + h) >> 1;
      if (v[m] < va) l = m + 1;
      else h = m;
    }
    int l2 = l, h2 = h0;
    while (l2 < h2) {
      const int m = (l2 + h2) >> 1;
      if (v[m] <= vb) l2 = m + 1;
      else h2 = m;
    }
    // record range -> kept-line range (the records and weights are stored
    // for the kept lines only)
    l0 = Y.nn > 0 ? kidx[job * S.line + l] : 0;
    h0 = Y.nn > 0 ? kidx[job * S.line + l2] : 0;
  }
  float accT[NS], accU[NS], accS[NS][NSP > 0 ? NSP : 1];
#pragma unroll
  for (int sl = 0; sl < NS; sl++) {
    accT[sl] = accU[sl] = 0.f;
#pragma unroll
    for (int c = 0; c < (NSP > 0 ? NSP : 1); c++) accS[sl][c] = 0.f;
  }
  const int64_t rbase = job * S.line;
  const float ascale = G == 1 ? 1.f : (G == 2 ? 0.25f : 0.0625f);  // DXF1 / DXF_G
  for (int c0 = l0; c0 < h0; c0 += 64) {
    // lane i holds line c0 + i; the wave walks the lines that reach its
    // segment, each line's fields broadcast with readlane
    const int li = c0 + lane;
    int npts = 0;
    int64_t glo = 0;
    float zfl = 0.f, zslope = 0.f;
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    TanLine tn{};
    if (li < h0) {
      const int64_t r = rbase + li;
      if (G == 1) {
        const SRec1 q = recs.r1[r];
        npts = q.npts;
        glo = q.glo;
        zfl = (float)q.zfl;
        zslope = (float)q.zslope;
        s0 = (float)q.str;
        s1 = (float)q.strd;
        s2 = (float)q.strver;
      } else {
        const SRecA q = (G == 2 ? recs.r2 : recs.r3)[r];
        npts = q.npts;
        glo = q.glo;
        zfl = (float)q.zfl;
        zslope = (float)q.zslope;
        s0 = (float)q.str;
      }
      // lines that miss the segment are not read further
      if (npts > 0 && (glo + npts - 1 < s_lo || glo > s_hi)) npts = 0;
      if (npts > 0) tn = tl[r];
    }
    uint64_t todo = __ballot(npts > 0);
    while (todo) {
      const int j = __builtin_ctzll(todo);
      todo &= todo - 1;
      const int np = __builtin_amdgcn_readlane(npts, j);
      const int64_t g0 = rl_i64(glo, j);
      const float zf0 = rl_f32(zfl, j), zs = rl_f32(zslope, j);
      const float w0 = rl_f32(s0, j);
      const float w1 = G == 1 ? rl_f32(s1, j) : 0.f, w2 = G == 1 ? rl_f32(s2, j) : 0.f;
      const int q = __builtin_amdgcn_readlane(tn.q, j);
      const float ta = rl_f32(tn.t.a, j) * ascale, tb = rl_f32(tn.t.b, j);
      const float sb = rl_f32(tn.s.b, j), ub = rl_f32(tn.u.b, j);
      float td0, td1 = 0.f, td2 = 0.f, sd0, sd1 = 0.f, sd2 = 0.f, ud0, ud1 = 0.f, ud2 = 0.f;
      if (G == 1) {
        td0 = rl_f32(tn.t.d1, j);
        td1 = rl_f32(tn.t.dd, j);
        td2 = rl_f32(tn.t.dv, j);
        sd0 = rl_f32(tn.s.d1, j);
        sd1 = rl_f32(tn.s.dd, j);
        sd2 = rl_f32(tn.s.dv, j);
        ud0 = rl_f32(tn.u.d1, j);
        ud1 = rl_f32(tn.u.dd, j);
        ud2 = rl_f32(tn.u.dv, j);
      } else {
        td0 = rl_f32(G == 2 ? tn.t.d2 : tn.t.d3, j);
        sd0 = rl_f32(G == 2 ? tn.s.d2 : tn.s.d3, j);
        ud0 = rl_f32(G == 2 ? tn.u.d2 : tn.u.d3, j);
      }
      // the line's points on the segment, segment-local: [pa, pb]; point
      // p of the segment is the line's point p + d0
      const int64_t ga = g0 > s_lo ? g0 : s_lo;
      const int64_t gb = g0 + np - 1 < s_hi ? g0 + np - 1 : s_hi;
      const int pa = (int)(ga - s_lo), pb = (int)(gb - s_lo), d0 = (int)(s_lo - g0) + 1;
#pragma unroll
      for (int sl = 0; sl < NS; sl++) {
        if (64 * sl + 63 < pa || 64 * sl > pb) continue;  // wave-uniform
        const int p = 64 * sl + lane;
        if (p < pa || p > pb) continue;
        // table position in FP32: the tangent needs the shape's slope, not
        // the forward's exact choice of entry at half-way points
        const float zf = fmaf((float)(p + d0), zs, zf0);
        const float u = fabsf(zf);
        int iz = (int)(u + 1.5f) - 1;
        iz = iz > NFMX - 1 ? NFMX - 1 : iz;
        const float2 fe = tf[iz];
        const float f = fe.x;
        float sl2 = w0 * fe.y;  // the weights times the slopes
        float tv = td0 * f, sv = sd0 * f, uv = ud0 * f;
        if (G == 1) {
          const float2 ge = tf[NFMX + iz], xe = tf[2 * NFMX + iz];
          sl2 = fmaf(w1, ge.y, fmaf(w2, xe.y, sl2));
          tv = fmaf(td1, ge.x, fmaf(td2, xe.x, tv));
          sv = fmaf(sd1, ge.x, fmaf(sd2, xe.x, sv));
          uv = fmaf(ud1, ge.x, fmaf(ud2, xe.x, uv));
        }
        const float sgn = zf < 0.f ? -1.f : 1.f;
        accT[sl] += fmaf(-sl2, fmaf(ta, sgn, tb * u), tv);
        if (hasu) accU[sl] += fmaf(-sl2, ub * u, uv);
#pragma unroll
        for (int c = 0; c < NSP; c++)
          if (c + 1 == q) accS[sl][c] += fmaf(-sl2, sb * u, sv);
      }
    }
  }
  // the combine, point by point
  const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
  const double X10 = -1. / 16., X11 = 9. / 16.;
  const DevPanel *pan = pans + Y.pan0;
  const int P = TL.nint;
  const int PU = hasu ? NO : -1;
#pragma unroll
  for (int sl = 0; sl < NS; sl++) {
    const int64_t g = s_lo + 64 * sl + lane;
    double vq[TAN_NQ];
    vq[0] = (double)accT[sl];
#pragma unroll
    for (int c = 1; c < TAN_NQ; c++) {
      double x = 0.;
      if (c - 1 < NSP && c <= nsp) x = (double)accS[sl][(c - 1 < NSP ? c - 1 : 0)];
      if (c == PU) x = (double)accU[sl];
      vq[c] = x;
    }
    if (G == 3) {
      if (g >= Y.n3) continue;
      int p = 0, I = 0, n1r3 = 1, w = 0;
      double vft = 0.;
      const bool inx = g <= Y.g3max;
      if (inx) {
        p = g < 155 ? 0 : (int)((g - 5) / 150);
        if (p > Y.npanel - 1) p = Y.npanel - 1;
        I = (int)(g - (int64_t)150 * p) + 1;
        n1r3 = p == 0 ? 1 : 6;
        w = Y.win0 + pan[p].w;
        vft = pan[p].vft;
      }
#pragma unroll
      for (int qc = 0; qc < TAN_NQ; qc++) {
        if (qc >= P) break;
        double x = vq[qc];
        if (inx) {
          const DevWin W = wins[w];
          x = x + xint_at(W.v1r4, W.v2r4, Y.dvr4, r4c + ((size_t)w * P + qc) * R4ROW, vft, Y.dvr3, I, n1r3, 155);
          if (Y.icntnm >= 1 && qc < NO)
            x = x + xint_at(Y.v1abs, Y.v2abs, Y.dvabs, absc + ((size_t)job * NO + qc) * abs_stride, vft, Y.dvr3, I,
                            n1r3, 155);
        }
        rout[qc * sout + Y.r3o + g] = x;
      }
    } else if (G == 2) {
      if (g >= Y.n2) continue;
      const bool inl = g >= 4 && g <= Y.g2max;
      const int64_t q4 = g >> 2;
      const int ph = (int)(g & 3);
#pragma unroll
      for (int qc = 0; qc < TAN_NQ; qc++) {
        if (qc >= P) break;
        double x = vq[qc];
        if (inl) {
          const double *r3 = rin + qc * sin + Y.r3o;
          const double a0 = r3[q4 - 1], a1 = r3[q4], a2 = r3[q4 + 1], a3 = r3[q4 + 2];
          if (ph == 0) x = x + a1;
          else if (ph == 1) x = x + X00 * a0 + X01 * a1 + X02 * a2 + X03 * a3;
          else if (ph == 2) x = x + X10 * (a0 + a3) + X11 * (a1 + a2);
          else x = x + X03 * a0 + X02 * a1 + X01 * a2 + X00 * a3;
        }
        rout[qc * sout + Y.r2o + g] = x;
      }
    } else {
      if (g < 32 || g > Y.g_last) continue;
      const int64_t q4 = g >> 2;
      const int ph = (int)(g & 3);
#pragma unroll
      for (int qc = 0; qc < TAN_NQ; qc++) {
        if (qc >= P) break;
        const double *r2 = rin + qc * sin + Y.r2o;
        const double a0 = r2[q4 - 1], a1 = r2[q4], a2 = r2[q4 + 1], a3 = r2[q4 + 2];
        double x = vq[qc];
        if (ph == 0) x = x + a1;
        else if (ph == 1) x = x + X00 * a0 + X01 * a1 + X02 * a2 + X03 * a3;
        else if (ph == 2) x = x + X10 * (a0 + a3) + X11 * (a1 + a2);
        else x = x + X03 * a0 + X02 * a1 + X01 * a2 + X00 * a3;
        vq[qc] = x;
      }
      double rad = 1., dlr = 0.;
      if (Y.jrad == 0) {
        int p = (int)((g - 32) / 2400);
        if (p > Y.npanel - 1) p = Y.npanel - 1;
        const DevPanel pp = pan[p];
        const int I = (int)(g - (int64_t)2400 * p) + 1;
        int lo2 = 0, hi2 = pp.nseg - 1;
        while (lo2 < hi2) {
          const int mid = (lo2 + hi2 + 1) >> 1;
          if (segs[Y.seg0 + pp.seg0 + mid].i1 <= I) lo2 = mid;
          else hi2 = mid - 1;
        }
        const DevSeg sg = segs[Y.seg0 + pp.seg0 + lo2];
        rad = sg.radvi + (double)(I - sg.i1) * sg.rdel;
        dlr = d_dlnradfn_dT(pp.vft + (double)(I - 1) * Y.dv, Y.xkt, TL.tave);
      }
#pragma unroll
      for (int qc = 0; qc < TAN_NQ; qc++) {
        if (qc >= NO) break;
        double x = vq[qc];
        if (qc > 0 && PU > 0) x = x - TL.wq[qc] * vq[PU];
        x = x * rad;
        if (qc == 0 && Y.jrad == 0) x = x + (double)od[Y.od0 + g - 32] * dlr;
        odc[odcoff[job * NO + qc] + g - 32] = (OUT)x;
      }
    }
  }
}

// ---------------------------------------------------------------------
// Far-wing lines.  Each TAPE9 line is one class of a SHRINK group; its
// derivatives per channel (t9tan[(TAPE9 line) 4P + 4q + {0: dS, 1: daL,
// 2: daD, 3: dv}]) are the strength-weighted sums over its members of
//   dS_k,  dS_k (a_k - <a>) + S_k da_k  (a = centre, aL, aD),
// divided by S (the group means are the TAPE9 line itself).  A member adds
// to the temperature, to its own molecule's channel and to U.
//
// k_shr_mem_tan, one thread per compacted LINF4 line (W.craw: its TAPE3
// record): the member's temperature terms dS_T, S dv_T, S daL_T, S daD_T,
// its width change S daL_u for d rho_s = rho_s ((T/T0)^n = aL / (alfa0
// (rho - rho_s) + hwhm rho_s)), and the channel of its molecule.
struct MemTan {
  double ds, sdv, sdal, sdad, sdalu;
};

__global__ void k_shr_mem_tan(DevLines L, const DevLayer *Ys, const TanLayer *Ts, Strides S, ShrinkWork W,
                              MemTan *mt, int8_t *mq) {
  const int job = blockIdx.y;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= W.npass[job]) return;
  const int64_t o = job * S.line;
  const DevLayer &Y = Ys[job];
  const TanLayer &TL = Ts[job];
  const double T = TL.tave;
  const double sp = W.csp[o + k];
  const int i = W.craw[o + k];
  const int mc = L.mol[i];
  const int m = mc % 100, iso = (mc % 1000) / 100;
  const double x = W.cx[o + k], al = W.cal[o + k], ad = W.cad[o + k];
  const double dnu = -Y.rhorat * (double)L.pshift[i] / T;
  double epp = L.epp[i];
  if (epp <= -1.001) epp = fabs(epp);
  MemTan r;
  r.ds = epp > -0.999 ? sp * d_dlns_dT(TL.dlnscor[m - 1][iso - 1], epp, x, dnu, T) : 0.;
  r.sdv = sp * dnu;
  r.sdal = sp * al * ((double)L.tmpalf[i] - 1.) / T;
  r.sdad = sp * (ad / (2. * T) + Y.alfd1[m - 1][iso - 1] * dnu);
  const int qm = TL.chan[m - 1];
  r.sdalu = 0.;
  if (qm > 0 || TL.nint > TL.nchan) {
    double alfa0 = L.alfa[i], hwhm = L.hwhm[i];
    if (m == 7) alfa0 = (alfa0 - 0.21 * hwhm) / (1.0 - 0.21);
    const double rs = Y.rhoslf[m - 1];
    const double den = alfa0 * (Y.rhorat - rs) + hwhm * rs;
    r.sdalu = den > 0. ? sp * al * (hwhm - alfa0) * rs / den : 0.;
  }
  mt[o + k] = r;
  mq[o + k] = (int8_t)qm;
}

// per output slot: the sums over the group's members, class by class as
// k_shr_emit writes the TAPE9 lines
__global__ void k_shr_emit_tan(const TanLayer *Ts, Strides S, ShrinkWork W, Tape9 T9, const MemTan *mt,
                               const int8_t *mq, float *t9tan) {
  const int job = blockIdx.y;
  const int qs = blockIdx.x * blockDim.x + threadIdx.x;
  if (qs >= T9.n[job]) return;
  const int64_t o = job * S.line;
  const int gs = W.grp[o + qs];
  if (gs < 0) return;
  const int ge = W.gout[o + qs];
  const int P = Ts[job].nint;
  const int PU = Ts[job].nint > Ts[job].nchan ? Ts[job].nchan : 0;  // channel of U (0: none)
  int64_t slot = o + qs;
  for (int c = 0; c < 2; c++) {  // non-CO2 first, then CO2
    // [q][dS, dS x + S dv, dS aL + S daL, dS aD + S daD]
    double acc[TAN_NQ][4];
#pragma unroll
    for (int q = 0; q < TAN_NQ; q++)
#pragma unroll
      for (int e = 0; e < 4; e++) acc[q][e] = 0.;
    bool any = false;
    for (int k = gs; k <= ge; k++) {
      if ((W.cfl[o + k] & 1) != c) continue;
      const double sp = W.csp[o + k];
      if (sp == 0.) continue;
      any = true;
      const double x = W.cx[o + k], al = W.cal[o + k], ad = W.cad[o + k];
      const MemTan r = mt[o + k];
      acc[0][0] += r.ds;
      acc[0][1] += r.ds * x + r.sdv;
      acc[0][2] += r.ds * al + r.sdal;
      acc[0][3] += r.ds * ad + r.sdad;
      const int qm = mq[o + k];
#pragma unroll
      for (int q = 1; q < TAN_NQ; q++) {
        if (q == qm) {
          acc[q][0] += sp;
          acc[q][1] += sp * x;
          acc[q][2] += sp * al + r.sdalu;
          acc[q][3] += sp * ad;
        }
        if (q == PU) acc[q][2] += r.sdalu;
      }
    }
    if (!any) continue;
    const double rs = 1. / T9.sp[slot], vg = T9.vnu[slot], alg = T9.alfa[slot], adg = T9.epp[slot];
    float *dst = t9tan + slot * 4 * P;
#pragma unroll
    for (int q = 0; q < TAN_NQ; q++)
      if (q < P) {
        dst[4 * q] = (float)acc[q][0];
        dst[4 * q + 1] = (float)((acc[q][2] - alg * acc[q][0]) * rs);
        dst[4 * q + 2] = (float)((acc[q][3] - adg * acc[q][0]) * rs);
        dst[4 * q + 3] = (float)((acc[q][1] - vg * acc[q][0]) * rs);
      }
    slot++;
  }
}

// ---------------------------------------------------------------------
// Far-wing tangent.  CONVF4's value f(XMSQ) (k_convf4) is linear in the
// strength, so with r = 1/(aL^2 + XMSQ) and xm = XJJ - XNUI, per channel
//   df = dS f/S + daL df/daL + daV df/daV - dv 2 xm df/dXMSQ,
//   near (XMSQ <= ZSQBND aV^2): f = SIV (A3 + B3 XMSQ/aV^2) - SUB,
//   far: f = SIL r - SUB,   SUB = F4BND (k0 - k1 XMSQ),
// SIV ~ S aL/aV^2, SIL ~ S aL, F4BND = SIL/(aL^2 + BOUND4^2) (k0 = 1,
// k1 = 0; CO2: k0 = 2, k1 = 1/BOUND4^2).
//
// k_f4_tan_lines: the window-independent constants of every TAPE9 line
// (a line is read by every window within BOUND4 of it) and each channel's
// (dS, daL, daV, dv), daV from daL, daD through AVRAT(zeta).
struct F4TConst {
  float alfli2, thr;      // aL^2, ZSQBND aV^2
  float rs;               // 1/S
  float siv, sivl, sivv;  // SIV, SIV/aL, -2 SIV/aV
  float b3v, silv, sil2;  // B3/aV^2, SIL/aL, 2 SIL aL
  float sil, c1, c0;      // SIL, k1 F4BND, k0 F4BND
  float fl0, fl1;         // dSUB/daL = dF4BND/daL (k0 - k1 XMSQ): its k0, k1 parts
};

__global__ void k_f4_tan_lines(Tape9 T9, const float *t9tan, const DevLayer *Ys, const TanLayer *Ts, Strides S,
                               F4TConst *fc, float4 *fd) {
  const int job = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T9.n[job]) return;
  const int P = Ts[job].nint;
  const int64_t o = job * S.line;
  const double bound4 = Ys[job].bound4;
  const double recpi = 1. / (2. * asin(1.));
  const double zsqbnd = 64. * 64.;
  const double A3 = (1. + 2. * zsqbnd) / ((1. + zsqbnd) * (1. + zsqbnd));
  const double B3 = -1. / ((1. + zsqbnd) * (1. + zsqbnd));
  (void)A3;
  const double bndsq = bound4 * bound4;
  const double sp = T9.sp[o + i];
  const double alfadi = T9.epp[o + i], alfali = T9.alfa[o + i];
  const double sum = alfali + alfadi;
  const double zetai = alfali / sum;
  int iz = fint(100. * zetai + ONEPL_D);
  iz = iz < 1 ? 1 : (iz > 101 ? 101 : iz);
  const double zetdif = 100. * zetai - (double)(iz - 1);
  const double avr = c_avrat[iz - 1] + zetdif * (c_avrat[iz] - c_avrat[iz - 1]);
  const double avs = 100. * (c_avrat[iz] - c_avrat[iz - 1]);
  const double alfavi = avr * sum;
  const double sil = sp * recpi * alfali;
  const double siv = (alfali / alfavi) * sp * recpi / alfavi;
  const double alfli2 = alfali * alfali, alfvi2 = alfavi * alfavi;
  const double den = alfli2 + bndsq;
  const double f4b = sil / den;
  const bool co2 = T9.mol[o + i] == 2;
  const double k0 = co2 ? 2. : 1., k1 = co2 ? 1. / bndsq : 0.;
  const double df4l = f4b * (1. / alfali - 2. * alfali / den);  // dF4BND/daL
  F4TConst p;
  p.alfli2 = (float)alfli2;
  p.thr = (float)(zsqbnd * alfvi2);
  p.rs = (float)(1. / sp);
  p.siv = (float)siv;
  p.sivl = (float)(siv / alfali);
  p.sivv = (float)(-2. * siv / alfavi);
  p.b3v = (float)(B3 / alfvi2);
  p.silv = (float)(sil / alfali);
  p.sil2 = (float)(2. * sil * alfali);
  p.sil = (float)sil;
  p.c1 = (float)(k1 * f4b);
  p.c0 = (float)(k0 * f4b);
  p.fl0 = (float)(k0 * df4l);
  p.fl1 = (float)(k1 * df4l);
  fc[o + i] = p;
  const float *tq = t9tan + (o + i) * 4 * P;
  for (int q = 0; q < P; q++) {
    const double dL = tq[4 * q + 1], dD = tq[4 * q + 2];
    const double dz = (dL * alfadi - alfali * dD) / (sum * sum);
    const double dV = avs * dz * sum + avr * (dL + dD);
    fd[(o + i) * P + q] = make_float4(tq[4 * q], (float)dL, (float)dV, tq[4 * q + 3]);
  }
}

// k_convf4_tan: every line the forward run accepted with points in a
// window (amask: one 64-bit word per 64-line chunk from the window's first
// line; wspan: that line and the chunk count) adds its derivative to each
// channel's R4 row.  The decisions are already taken, so a heavy window's
// chunks are shared out over several workgroups (F4TPart: window, part,
// parts, and the row of partial sums the part writes); k_r4_tan_sum adds a
// window's parts in a fixed order.  The accepted lines of CFT_T lines at a
// time are compacted into LDS with their constants; every thread holds the
// channels of its points j = tid + 256 s in registers, so a point's
// additions happen in line order without atomics.
constexpr int CFT_T = 256;
constexpr int CFT_S = (R4ROW + CFT_T - 1) / CFT_T;  // points per thread

struct F4TPart {
  int32_t wi, part, nparts, dst;
};

template <int P>
__global__ __launch_bounds__(CFT_T) void k_convf4_tan(Tape9 T9, const F4TConst *fc, const float4 *fd,
                                                      const DevLayer *Ys, Strides S, const DevWin *wins,
                                                      const F4TPart *parts, const uint64_t *amask, int64_t mwords,
                                                      const int32_t *wspan, float *r4part) {
  __shared__ F4TConst par[CFT_T];
  __shared__ double pxn[CFT_T];
  __shared__ int2 pj[CFT_T];
  __shared__ float4 cof[CFT_T * P];
  __shared__ int s_wcnt[CFT_T / 64];
  const F4TPart pt = parts[blockIdx.x];
  const int wi = pt.wi;
  const DevWin W = wins[wi];
  const int job = W.job;
  const DevLayer &Y = Ys[job];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int64_t o = job * S.line;
  const double dvr4 = Y.dvr4, bound4 = Y.bound4;
  const double zsqbnd = 64. * 64.;
  const float a3 = (float)((1. + 2. * zsqbnd) / ((1. + zsqbnd) * (1. + zsqbnd)));
  const int nptr4 = W.nptr4;
  float acc[CFT_S][P];
#pragma unroll
  for (int sl = 0; sl < CFT_S; sl++)
#pragma unroll
    for (int c = 0; c < P; c++) acc[sl][c] = 0.f;
  const int pos = wspan[2 * wi], nch = wspan[2 * wi + 1];
  // this part's chunks, rounded to CFT_T-line steps
  const int nstep = (nch + CFT_T / 64 - 1) / (CFT_T / 64);
  const int s0 = (int)((int64_t)nstep * pt.part / pt.nparts), s1 = (int)((int64_t)nstep * (pt.part + 1) / pt.nparts);
  const int end = pos + 64 * nch;
  const uint64_t *mk = amask + (size_t)wi * mwords;
  for (int c0 = pos + CFT_T * s0; c0 < pos + CFT_T * s1 && c0 < end; c0 += CFT_T) {
    // lines c0 .. c0 + 255: chunk words (c0 - pos)/64 + wave
    const int i = c0 + tid;
    const int w = (c0 - pos) / 64 + wave;
    const uint64_t word = (i < end) ? mk[w] : 0ull;  // wave-uniform
    const bool acc_me = (word >> lane) & 1ull;
    const int below = lane == 0 ? 0 : __builtin_popcountll(word & ((1ull << lane) - 1));
    if (lane == 0) s_wcnt[wave] = __builtin_popcountll(word);
    __syncthreads();  // counts ready; the previous step's constants no longer read
    int base = 0, nacc = 0;
#pragma unroll
    for (int v = 0; v < CFT_T / 64; v++) {
      if (v < wave) base += s_wcnt[v];
      nacc += s_wcnt[v];
    }
    if (acc_me) {
      const int k = base + below;
      const double xnui = T9.vnu[o + i] - W.v1r4;
      const int jmin = fint((xnui - bound4) / dvr4 + 2.);
      const int jmax = fint((xnui + bound4) / dvr4 + 1.);
      int2 jr;
      jr.x = jmin < 1 ? 1 : jmin;
      jr.y = jmax > nptr4 ? nptr4 : jmax;
      if (jmin > nptr4) jr.y = 0;
      pxn[k] = xnui;
      pj[k] = jr;
      par[k] = fc[o + i];
#pragma unroll
      for (int c = 0; c < P; c++) cof[k * P + c] = fd[(o + i) * P + c];
    }
    __syncthreads();
    for (int a = 0; a < nacc; a++) {
      // the line's point range is the same for every lane
      const int2 jl = pj[a];
      const int jx = __builtin_amdgcn_readfirstlane(jl.x), jy = __builtin_amdgcn_readfirstlane(jl.y);
      if (jy < jx) continue;
      const F4TConst q = par[a];
      const double xnui = pxn[a];
      float4 k[P];
#pragma unroll
      for (int c = 0; c < P; c++) k[c] = cof[a * P + c];
#pragma unroll
      for (int sl = 0; sl < CFT_S; sl++) {
        // this wave's points of slot sl: j = 256 sl + 64 wave + lane
        const int wlo = CFT_T * sl + 64 * wave;
        if (wlo + 63 < jx || wlo > jy) continue;  // wave-uniform
        const int j = wlo + lane;
        if (j < jx || j > jy) continue;
        const float xm = (float)((double)(j - 1) * dvr4 - xnui);
        const float xmsq = xm * xm;
        const float sub = fmaf(-xmsq, q.c1, q.c0);      // SUB
        const float dsubl = fmaf(-xmsq, q.fl1, q.fl0);  // dSUB/daL
        // both branches, then the one the forward took (a branch would
        // diverge at the ZSQBND boundary inside most waves)
        const float sh = fmaf(xmsq, q.b3v, a3);         // A3 + B3 XMSQ/aV^2
        const float rr = __builtin_amdgcn_rcpf(q.alfli2 + xmsq);
        const float rr2 = rr * rr;
        const bool nearb = xmsq <= q.thr;
        // f, df/daL, df/daV, df/dXMSQ
        const float f = (nearb ? q.siv * sh : q.sil * rr) - sub;
        const float fl = (nearb ? q.sivl * sh : fmaf(q.silv, rr, -q.sil2 * rr2)) - dsubl;
        const float fv = nearb ? q.sivv * fmaf(2.f * xmsq, q.b3v, a3) : 0.f;
        const float fx = (nearb ? q.siv * q.b3v : -q.sil * rr2) + q.c1;
        const float fs = f * q.rs;
        const float fn = -2.f * xm * fx;  // df/dv: dXMSQ/dv = -2 xm
#pragma unroll
        for (int c = 0; c < P; c++)
          acc[sl][c] = fmaf(k[c].x, fs, fmaf(k[c].y, fl, fmaf(k[c].z, fv, fmaf(k[c].w, fn, acc[sl][c]))));
      }
    }
  }
  float *row = r4part + (size_t)pt.dst * P * R4ROW;
#pragma unroll
  for (int sl = 0; sl < CFT_S; sl++) {
    const int j = CFT_T * sl + tid;
    if (j >= R4ROW) continue;
#pragma unroll
    for (int c = 0; c < P; c++) row[(size_t)c * R4ROW + j] = acc[sl][c];
  }
}

// A window's R4 channel rows: the sum of its parts (wfirst[w] .. +
// nparts - 1, in order), the window's piecewise-linear radiation factor
// (JRAD = 1) on every channel, and for the temperature that factor's
// derivative on the forward R4.  One thread per (channel, point).
__global__ void k_r4_tan_sum(const DevLayer *Ys, const TanLayer *Ts, const DevWin *wins, const DevSeg *segs,
                             const int32_t *wfirst, const int32_t *wnparts, const float *r4part, int P,
                             const double *r4fwd, double *r4c) {
  const int wi = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= P * R4ROW) return;
  const int c = t / R4ROW, j = t - c * R4ROW;
  const DevWin W = wins[wi];
  const DevLayer &Y = Ys[W.job];
  double *rows = r4c + (size_t)wi * P * R4ROW;
  const bool in = j >= 1 && j <= W.nptr4;
  if (!in) {
    rows[t] = 0.;
    return;
  }
  double v = 0.;
  const int f0 = wfirst[wi], np = wnparts[wi];
  for (int s = 0; s < np; s++) v += (double)r4part[(size_t)(f0 + s) * P * R4ROW + t];
  if (W.nseg > 0) {
    int lo2 = 0, hi2 = W.nseg - 1;
    while (lo2 < hi2) {
      const int mid = (lo2 + hi2 + 1) >> 1;
      if (segs[W.seg0 + mid].i1 <= j) lo2 = mid;
      else hi2 = mid - 1;
    }
    const DevSeg sg = segs[W.seg0 + lo2];
    if (j >= sg.i1 && j <= sg.i2) {
      v = v * (sg.radvi + (double)(j - sg.i1) * sg.rdel);
      if (c == 0)
        v += r4fwd[(size_t)wi * R4ROW + j] * d_dlnradfn_dT(W.v1r4 + (double)(j - 1) * Y.dvr4, Y.xkt, Ts[W.job].tave);
    }
  }
  rows[t] = v;
}

// Central differences of the continuum rows: absc[(job P + q) stride + k]
// = (plus - minus) rh[q] over the perturbed plans' ABSRB rows (plan
// 2 (job P + q) holds the plus step, the next one the minus step).
__global__ void k_cont_diff(const double *ab, int stride, int njobs, int P, const double *rh, double *absc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)njobs * P * stride;
  if (t >= n) return;
  const int64_t row = t / stride;
  const int k = (int)(t - row * stride);
  const int q = (int)(row % P);
  const double *plus = ab + 2 * row * (int64_t)stride;
  absc[t] = (plus[k] - plus[stride + k]) * rh[q];
}

}  // namespace lblx
