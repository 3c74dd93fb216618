// HIP kernels for the layer optical-depth path on MI355X (gfx950).
//
// The reference computes one layer panel by panel on one core (HIRAC1,
// src/oprop.f90:25-668).  Here the whole band of a layer is resident in
// HBM and each stage is one launch:
//   k_linf4       LINF4 line correction            oprop.f90:3093-3235
//   k_shrink      LINF4 batching + SHRINK          oprop.f90:3236-3290, 3508-3640
//   k_convf4      LBLF4/CONVF4, one window per WG  oprop.f90:3641-4039
//   k_lncor1      LNCOR1 line correction           oprop.f90:995-1193
//   k_reject*     LNCOR1 rejection + block-read panel fixpoint
//   k_scatter_rec CNVFNV per-line set-up           oprop.f90:1262-1305
//   k_tile<G>     CNVFNV scatter into R3/R2/R1 tiles in LDS, with the
//                 PANEL/XINT combine fused into the epilogue
//                 oprop.f90:1307-1320, 1459-1519, 2201-2234
// All order-dependent decisions of the reference (line rejection against
// the running far-wing array, SHRINK batch boundaries, far-wing windows)
// are reproduced exactly; see the kernel comments.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lblx_kernels.h"

namespace lblx {

__constant__ double c_avrat[102], c_cf1[102], c_cf2[102], c_cf3[102], c_cgauss[102], c_cer[102];
__constant__ int c_isomax[7];

constexpr double ONEPL_D = 1.001, ONEMI_D = 0.999;
constexpr int NFMX = 2601;

__device__ __forceinline__ int fint(double x) { return (int)x; }

__device__ __forceinline__ double d_radfn(double vi, double xkt) {
  if (xkt > 0.0) {
    double x = vi / xkt;
    if (x <= 0.01) return 0.5 * x * vi;
    if (x <= 10.0) {
      double e = exp(-x);
      return vi * (1. - e) / (1. + e);
    }
  }
  return vi;
}

// ---------------------------------------------------------------------
// Block-wide exclusive scan (sum) for 256 threads.
template <int NT>
__device__ int block_scan_excl(int v, int *buf, int *total) {
  int tid = threadIdx.x;
  buf[tid] = v;
  __syncthreads();
  for (int off = 1; off < NT; off <<= 1) {
    int t = (tid >= off) ? buf[tid - off] : 0;
    __syncthreads();
    buf[tid] += t;
    __syncthreads();
  }
  int incl = buf[tid];
  int tot = buf[NT - 1];
  __syncthreads();
  *total = tot;
  return incl - v;
}

// ---------------------------------------------------------------------
// LINF4 correction of every record: shifted centre, temperature-corrected
// strength, Lorentz and Doppler widths (oprop.f90:3093-3235).  Records the
// first passing line above V2L4, where LINF4 stops reading (IEOF = 1).
__global__ void k_linf4(DevLines L, DevLayer Y, int first, F4Line *out, int *stop_idx) {
  int i = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L.n) return;
  F4Line o;
  o.pass = 0;
  o.x = 0.;
  o.sp = 0.;
  o.al = 0.;
  o.ad = 0.;
  o.mol = 0;
  int b = L.blk[i];
  int mc = L.mol[i];
  int m = mc % 100, iso = (mc % 1000) / 100;
  double vraw = L.vnu[i];
  if (L.bvmax[b] >= Y.v1l4 && m >= 1 && m <= Y.nmol && iso >= 1 && iso <= c_isomax[m - 1]) {
    double sui = (double)L.sp[i] * Y.wk[m - 1];
    if (sui != 0. && vraw >= Y.v1l4) {
      double alfa = L.alfa[i], hwhm = L.hwhm[i];
      if (m == 7) alfa = (alfa - 0.21 * hwhm) / (1.0 - 0.21);  // RDLNFL O2 (oprop.f90:3456-3462)
      double x = vraw + Y.rhorat * (double)L.pshift[i];
      double epp = L.epp[i];
      if (epp <= -1.001) epp = fabs(epp);
      if (epp > -0.999)
        sui = sui * Y.scor[m - 1][iso - 1] * exp(-epp * Y.betacr4) * (1. + exp(-x * Y.beta4));
      double tmpcor = pow(Y.tratio, (double)L.tmpalf[i]);
      double alfa0i = alfa * tmpcor;
      double hwhmsi = hwhm * tmpcor;
      o.al = alfa0i * (Y.rhorat - Y.rhoslf[m - 1]) + hwhmsi * Y.rhoslf[m - 1];
      o.ad = x * Y.alfd1[m - 1][iso - 1];
      o.x = x;
      o.sp = sui;
      o.mol = m;
      o.pass = 1;
      if (x > Y.v2l4) atomicMin(stop_idx, i);
    }
  }
  out[i] = o;
}

// ---------------------------------------------------------------------
// LINF4 batching and SHRINK for one layer in one workgroup.
//
// LINF4 reads 250-line TAPE3 blocks until the buffer holds >= 1000 lines
// (shrunk lines carried from earlier batches plus new raw lines), then
// SHRINK merges consecutive lines into groups: a group ends before the
// first line whose centre exceeds VLMT, and VLMT is reset to (centre of
// the group's last line) + DVR4/2.  Group boundaries therefore form a
// chain e_{k+1} = f(e_k) with f(e) = (first k >= e+2, x_k > x_e + DVR4/2)
// - 1; the chain of each batch is marked with pointer doubling in LDS.
// Each group emits one non-CO2 and one CO2 effective line (strength-
// weighted centre and widths) if it holds any positive strength of that
// class.  TAPE9 blocks close when they hold >= 1000 lines.
constexpr int SHR_T = 256;
constexpr int SHR_MAX = 1280;
constexpr int SHR_K = 11;

__global__ __launch_bounds__(SHR_T) void k_shrink(const F4Line *f4, DevLines L, DevLayer Y,
                                                  const int *stop_idx, Tape9 T9) {
  __shared__ double bx[SHR_MAX], bsp[SHR_MAX], bal[SHR_MAX], bad[SHR_MAX];
  __shared__ int8_t bco2[SHR_MAX];
  __shared__ int16_t jmp[SHR_K][SHR_MAX + 1];
  __shared__ int16_t gst[SHR_MAX];
  __shared__ int scnt[SHR_MAX];
  __shared__ int sbuf[SHR_T];
  __shared__ int s_g;
  __shared__ double s_lastx;
  const int tid = threadIdx.x;
  const double dv2 = Y.dvr4 / 2.;
  const int stop_i = *stop_idx;
  int J = 0, out_n = 0, t9s = 0, t9nb = 0, nb = 0;
  bool stop = false;

  auto flush = [&]() {
    if (J > 0 && tid == 0) {
      T9.bstart[t9nb] = t9s;
      T9.bcount[t9nb] = J;
      T9.bvhi[t9nb] = s_lastx;
    }
    if (J > 0) t9nb++;
    t9s = out_n;
    J = 0;
  };

  auto process = [&](int n) -> int {
    if (n == 0) return 0;
    for (int i = tid; i < n; i += SHR_T)
      if (bsp[i] < 1.0e-35) bsp[i] = 0.0;
    for (int i = tid; i < n; i += SHR_T) {
      double thr = bx[i] + dv2;
      int k = i + 2;
      while (k < n && !(bx[k] > thr)) k++;
      int f = k - 1;
      if (f > n - 1) f = n - 1;
      jmp[0][i] = (int16_t)((i == n - 1) ? n : f);
    }
    if (tid == 0) {
      jmp[0][n] = (int16_t)n;
      double thr = bx[0] + dv2;
      int k = 1;
      while (k < n && !(bx[k] > thr)) k++;
      s_g = k - 1;
    }
    __syncthreads();
    int nlev = 1;
    while ((1 << nlev) <= n && nlev < SHR_K) nlev++;
    for (int lev = 1; lev < nlev; lev++) {
      for (int i = tid; i <= n; i += SHR_T) jmp[lev][i] = jmp[lev - 1][jmp[lev - 1][i]];
      __syncthreads();
    }
    const int g = s_g;
    for (int i = tid; i < n; i += SHR_T) {
      int ge, gs;
      if (i <= g) {
        ge = g;
        gs = 0;
      } else {
        int cur = g;
        for (int lev = nlev - 1; lev >= 0; lev--) {
          int nx = jmp[lev][cur];
          if (nx <= i - 1) cur = nx;
        }
        ge = jmp[0][cur];
        gs = cur + 1;
      }
      gst[i] = (int16_t)((ge == i) ? gs : -1);
    }
    __syncthreads();
    // emitted lines per group end
    for (int i = tid; i < n; i += SHR_T) {
      int c = 0;
      int gs = gst[i];
      if (gs >= 0) {
        int a = 0, b = 0;
        for (int k = gs; k <= i; k++) {
          if (bsp[k] > 0.) {
            if (bco2[k]) b = 1;
            else a = 1;
          }
        }
        c = a + b;
      }
      scnt[i] = c;
    }
    __syncthreads();
    // exclusive scan of scnt over n entries (5 per thread)
    constexpr int PER = (SHR_MAX + SHR_T - 1) / SHR_T;
    int loc[PER];
    int s = 0;
    for (int q = 0; q < PER; q++) {
      int i = tid * PER + q;
      loc[q] = (i < n) ? scnt[i] : 0;
      s += loc[q];
    }
    int total;
    int base = block_scan_excl<SHR_T>(s, sbuf, &total);
    __syncthreads();
    for (int q = 0; q < PER; q++) {
      int i = tid * PER + q;
      if (i < n) scnt[i] = base;
      base += loc[q];
    }
    __syncthreads();
    for (int i = tid; i < n; i += SHR_T) {
      int gs = gst[i];
      if (gs < 0) continue;
      double sumv = 0., sums = 0., sumal = 0., sumad = 0.;
      double sumv2 = 0., sums2 = 0., sumal2 = 0., sumad2 = 0.;
      for (int k = gs; k <= i; k++) {
        double sp = bsp[k];
        if (!bco2[k]) {
          sumv = sumv + bx[k] * sp;
          sums = sums + sp;
          sumal = sumal + sp * bal[k];
          sumad = sumad + sp * bad[k];
        } else {
          sumv2 = sumv2 + bx[k] * sp;
          sums2 = sums2 + sp;
          sumal2 = sumal2 + sp * bal[k];
          sumad2 = sumad2 + sp * bad[k];
        }
      }
      int o = out_n + scnt[i];
      if (sums > 0.) {
        T9.vnu[o] = sumv / sums;
        T9.sp[o] = sums;
        T9.alfa[o] = sumal / sums;
        T9.epp[o] = sumad / sums;
        T9.mol[o] = 0;
        if (o == out_n + total - 1) s_lastx = sumv / sums;
        o++;
      }
      if (sums2 > 0.) {
        T9.vnu[o] = sumv2 / sums2;
        T9.sp[o] = sums2;
        T9.alfa[o] = sumal2 / sums2;
        T9.epp[o] = sumad2 / sums2;
        T9.mol[o] = 2;
        if (o == out_n + total - 1) s_lastx = sumv2 / sums2;
      }
    }
    __syncthreads();
    return total;
  };

  for (int b = 0; b < L.nblk && !stop; b++) {
    if (L.bvmax[b] < Y.v1l4) continue;  // RDLNFL skips the block
    int s = L.bstart[b], e = L.bstart[b + 1];
    int i = s + tid;
    int fl = 0;
    if (i < e && i <= stop_i) fl = f4[i].pass;
    int total;
    int pos = block_scan_excl<SHR_T>(fl, sbuf, &total);
    if (fl) {
      int k = nb + pos;
      bx[k] = f4[i].x;
      bsp[k] = f4[i].sp;
      bal[k] = f4[i].al;
      bad[k] = f4[i].ad;
      bco2[k] = (int8_t)(f4[i].mol == 2);
    }
    if (stop_i >= s && stop_i < e) stop = true;
    nb += total;
    __syncthreads();
    if (!stop && J + nb < 1000) continue;
    int em = process(nb);
    J += em;
    out_n += em;
    nb = 0;
    if (J >= 1000 || stop) flush();
  }
  if (!stop) {  // end of file: shrink what was read and close the block
    int em = process(nb);
    J += em;
    out_n += em;
    flush();
  }
  if (tid == 0) {
    *T9.n = out_n;
    *T9.nblk = t9nb;
  }
}

// ---------------------------------------------------------------------
// CONVF4 for one far-wing window per workgroup (oprop.f90:3878-4039).
//
// The reference rejects a line when its peak (A3 |SIV|) is at most
// DPTMN + DPTFC * R4(JJ), where R4 holds the contributions of the lines
// accepted before it in this window.  Lines are therefore taken in TAPE9
// order in chunks of 64: wave 0 resolves the chunk's decisions in order
// (each lane's running R4(JJ) is the LDS value plus the in-chunk
// contributions of earlier accepted lanes), then all threads add the
// accepted lines into the window held in LDS.  Each thread owns a fixed
// residue class of points, so additions to a point happen in line order
// without atomics.
constexpr int CF4_T = 256;

struct F4Par {
  double xnui, sil, siv_a3, siv_b3, f4bnd, rec_alfvi2, alfli2;
  int jmin, jmax, co2;
};

__device__ __forceinline__ double f4_at(const F4Par &p, double xjj, double r_bndsq, double zsqbnd) {
  double xm = xjj - p.xnui;
  double xmsq = xm * xm;
  double zvsq = xmsq * p.rec_alfvi2;
  double sub = p.co2 ? (2. - (xmsq * r_bndsq)) * p.f4bnd : p.f4bnd;
  if (zvsq <= zsqbnd) return (p.siv_a3 + zvsq * p.siv_b3) - sub;
  return p.sil / (p.alfli2 + xmsq) - sub;
}

__global__ __launch_bounds__(CF4_T) void k_convf4(Tape9 T9, DevLayer Y, const DevWin *wins,
                                                  const DevSeg *segs, double *r4out,
                                                  unsigned long long *k3upd) {
  __shared__ double r4[2504];
  __shared__ F4Par par[64];
  __shared__ int s_nacc, s_done, s_b;
  const int tid = threadIdx.x;
  const DevWin W = wins[blockIdx.x];
  const double dvr4 = Y.dvr4, bound4 = Y.bound4;
  const double recpi = 1. / (2. * asin(1.));
  const double zsqbnd = 64. * 64.;
  const double A3 = (1. + 2. * zsqbnd) / ((1. + zsqbnd) * (1. + zsqbnd));
  const double B3 = -1. / ((1. + zsqbnd) * (1. + zsqbnd));
  const double bndsq = bound4 * bound4;
  const double r_bndsq = 1. / bndsq;
  const double vnulst = W.v2r4 + bound4;
  const double vlo = W.v1r4 - bound4;
  const int nptr4 = W.nptr4;
  const int nl = *T9.n;
  const int nb9 = *T9.nblk;
  for (int j = tid; j < 2504; j += CF4_T) r4[j] = 0.;
  if (tid == 0) {
    s_b = nb9;
    s_done = 0;
  }
  __syncthreads();
  for (int b = tid; b < nb9; b += CF4_T)
    if (T9.bvhi[b] >= vlo) atomicMin(&s_b, b);
  __syncthreads();
  int pos = (s_b < nb9) ? T9.bstart[s_b] : nl;
  while (pos < nl) {
    if (tid < 64) {
      const int lane = tid;
      int i = pos + lane;
      bool valid = i < nl;
      F4Par p;
      p.co2 = 0;
      p.jmin = 1;
      p.jmax = 0;
      double speak = 0., partial = 0., vnui = 0.;
      bool nz = false;
      int jj = 1;
      if (valid) {
        double sp = T9.sp[i];
        nz = sp != 0.;
        double alfadi = T9.epp[i], alfali = T9.alfa[i];
        double zetai = alfali / (alfali + alfadi);
        int iz = fint(100. * zetai + ONEPL_D);
        iz = iz < 1 ? 1 : (iz > 101 ? 101 : iz);  // zeta in [0,1] keeps iz in range
        double zetdif = 100. * zetai - (double)(iz - 1);
        double alfavi = (c_avrat[iz - 1] + zetdif * (c_avrat[iz] - c_avrat[iz - 1])) * (alfali + alfadi);
        double ralfvi = 1. / alfavi;
        p.sil = sp * recpi * alfali;
        double siv = (alfali * ralfvi) * sp * recpi * ralfvi;
        speak = A3 * fabs(siv);
        jj = fint((T9.vnu[i] - W.v1r4) / dvr4 + 1.);
        jj = jj < 1 ? 1 : (jj > nptr4 ? nptr4 : jj);
        vnui = T9.vnu[i];
        p.xnui = vnui - W.v1r4;
        int jmin = fint((p.xnui - bound4) / dvr4 + 2.);
        int jmax = fint((p.xnui + bound4) / dvr4 + 1.);
        p.jmin = jmin < 1 ? 1 : jmin;
        p.jmax = jmax > nptr4 ? nptr4 : jmax;
        if (jmin > nptr4) p.jmax = 0;  // accepted but nothing to add (GO TO 60)
        p.alfli2 = alfali * alfali;
        p.f4bnd = p.sil / (p.alfli2 + bndsq);
        p.rec_alfvi2 = 1. / (alfavi * alfavi);
        p.siv_a3 = siv * A3;
        p.siv_b3 = siv * B3;
        p.co2 = T9.mol[i] == 2;
        partial = r4[jj];
      }
      const double xjj_me = (double)(jj - 1) * dvr4;
      int nacc = 0;
      int done = 0;
      for (int j = 0; j < 64; j++) {
        bool acc = valid && nz && !(speak <= (Y.dptmn4 + Y.dptfc * partial));
        bool stp = acc && (vnui >= vnulst);
        int accj = __shfl(acc ? 1 : 0, j);
        int stpj = __shfl(stp ? 1 : 0, j);
        if (stpj) {
          done = 1;
          break;
        }
        if (accj) {
          F4Par q;
          q.xnui = __shfl(p.xnui, j);
          q.sil = __shfl(p.sil, j);
          q.siv_a3 = __shfl(p.siv_a3, j);
          q.siv_b3 = __shfl(p.siv_b3, j);
          q.f4bnd = __shfl(p.f4bnd, j);
          q.rec_alfvi2 = __shfl(p.rec_alfvi2, j);
          q.alfli2 = __shfl(p.alfli2, j);
          q.jmin = __shfl(p.jmin, j);
          q.jmax = __shfl(p.jmax, j);
          q.co2 = __shfl(p.co2, j);
          if (lane > j && jj >= q.jmin && jj <= q.jmax) partial += f4_at(q, xjj_me, r_bndsq, zsqbnd);
          if (lane == j && p.jmax >= p.jmin) par[nacc] = p;
          if (q.jmax >= q.jmin) nacc++;
        }
        // lanes at or below j are final; a rejected lane j simply adds nothing
      }
      if (lane == 0) {
        s_nacc = nacc;
        s_done = done;
      }
    }
    __syncthreads();
    const int nacc = s_nacc;
    if (tid == 0) {
      unsigned long long u = 0;
      for (int a = 0; a < nacc; a++) u += (unsigned long long)(par[a].jmax - par[a].jmin + 1);
      atomicAdd(k3upd, u);
    }
    for (int a = 0; a < nacc; a++) {
      F4Par q = par[a];
      // owned points j = 1 + tid + k*CF4_T
      int j0 = q.jmin - 1;
      int r = j0 % CF4_T;
      int first = j0 - r + tid + (tid < r ? CF4_T : 0) + 1;
      for (int j = first; j <= q.jmax; j += CF4_T) {
        double xjj = (double)(j - 1) * dvr4;
        r4[j] = r4[j] + f4_at(q, xjj, r_bndsq, zsqbnd);
      }
    }
    __syncthreads();
    if (s_done) break;
    pos += 64;
  }
  __syncthreads();
  // JRAD = 1: radiation term as a piecewise-linear factor (oprop.f90:3733-3758)
  if (W.nseg > 0) {
    for (int j = tid + 1; j <= nptr4; j += CF4_T) {
      int lo = 0, hi = W.nseg - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (segs[W.seg0 + mid].i1 <= j) lo = mid;
        else hi = mid - 1;
      }
      DevSeg sg = segs[W.seg0 + lo];
      if (j >= sg.i1 && j <= sg.i2) r4[j] = r4[j] * (sg.radvi + (double)(j - sg.i1) * sg.rdel);
    }
    __syncthreads();
  }
  double *row = r4out + (size_t)blockIdx.x * 2504;
  for (int j = tid; j < 2504; j += CF4_T) row[j] = (j >= 1 && j <= nptr4) ? r4[j] : 0.;
}

// ---------------------------------------------------------------------
// LNCOR1 for the near lines (oprop.f90:995-1193) except the two tests
// that depend on the panel at which the block is read (the VFT test and
// the R4 rejection); those are applied in k_reject.  Also computes the
// first panel whose R1 buffer can hold the line (CNVFNV JMAX1 <= MAX1).
__global__ void k_lncor1(DevLines L, DevLayer Y, const DevPanel *pan, int nlo, int nn,
                         NearLine *out) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  int i = nlo + t;
  NearLine o;
  o.vnu = L.vnu[i];
  o.recalf = 0.;
  o.sp = 0.;
  o.speak = 0.;
  o.zetai = 0.;
  o.izeta = 1;
  o.ff = 0;
  o.keep = 0;
  o.proc = -1;
  int mc = L.mol[i];
  int m = mc % 100, iso = (mc % 1000) / 100;
  if (m >= 1 && m <= 7 && iso >= 1 && iso <= c_isomax[m - 1]) {
    double sui = (m <= Y.nmol) ? (double)L.sp[i] * Y.wk[m - 1] : 0.;
    if (Y.jrad == 1) sui = sui * d_radfn(L.vnu[i], Y.xkt);
    if (sui != 0.) {
      double alfa0 = L.alfa[i], hwhm = L.hwhm[i];
      if (m == 7) alfa0 = (alfa0 - 0.21 * hwhm) / (1.0 - 0.21);  // RDLIN (oprop.f90:818-827)
      double vnu = L.vnu[i] + Y.rhorat * (double)L.pshift[i];
      double tmpcor = pow(Y.tratio, (double)L.tmpalf[i]);
      double alfa0i = alfa0 * tmpcor;
      double hwhmsi = hwhm * tmpcor;
      double alfl = alfa0i * (Y.rhorat - Y.rhoslf[m - 1]) + hwhmsi * Y.rhoslf[m - 1];
      double alfad = vnu * Y.alfd1[m - 1][iso - 1];
      double zeta = alfl / (alfl + alfad);
      double fzeta = 100. * zeta;
      int iz = fint(fzeta + ONEPL_D);
      iz = iz < 1 ? 1 : (iz > 101 ? 101 : iz);  // zeta in [0,1] keeps iz in range
      double zetdif = fzeta - (double)(iz - 1);
      double alfv = (c_avrat[iz - 1] + zetdif * (c_avrat[iz] - c_avrat[iz - 1])) * (alfl + alfad);
      if (alfv < Y.dv) alfv = Y.dv;
      if (alfv > Y.alfmax) alfv = Y.alfmax;
      double recalf = 1. / alfv;
      double epp = L.epp[i];
      if (epp <= -1.001) epp = fabs(epp);
      if (epp > -0.999) sui = sui * Y.scor[m - 1][iso - 1] * exp(-epp * Y.betacr_l) * (1. + exp(-vnu / Y.xkt));
      o.vnu = vnu;
      o.recalf = recalf;
      o.sp = sui;
      o.speak = sui * recalf;
      o.zetai = zeta;
      o.izeta = iz;
      o.keep = 1;
      // first fit: INT(ZINT_p + BHWDXF + 1.5) <= MAX1
      double zslope = recalf * (Y.dv / 0.002);
      double bhwdxf = (4. / 0.002) / zslope;
      double z0 = (vnu - Y.vft0) / Y.dv;
      int p = (int)floor((z0 + bhwdxf + 0.5 - (double)Y.max1) / 2400.) + 1;
      if (p < 0) p = 0;
      if (p > Y.npanel) p = Y.npanel;
      auto fits = [&](int q) {
        double zint = (vnu - pan[q].vft) / Y.dv;
        return fint(zint + bhwdxf + 1.5) <= Y.max1;
      };
      while (p > 0 && fits(p - 1)) p--;
      while (p < Y.npanel && !fits(p)) p++;
      o.ff = p;  // == npanel: never processed before the last panel
    }
  }
  out[t] = o;
}

// One step of the block-read fixpoint: with the current guess of the panel
// at which each block is read, apply LNCOR1's VFT test and R4 rejection
// (oprop.f90:1120, 1153-1177) and reduce the first-fit panel of the kept
// lines per block.
__global__ void k_reject(const NearLine *nl, int nn, int nlo, const int32_t *blk, int blk0,
                         const int *pread, const DevPanel *pan, const DevWin *wins,
                         const double *r4, DevLayer Y, int *keep_out, int *blkmax) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  NearLine o = nl[t];
  int keep = o.keep;
  int k = blk[nlo + t] - blk0;
  if (keep) {
    int pr = pread[k];
    int pq = pr < Y.npanel ? pr : Y.npanel - 1;
    double vft = pan[pq].vft;
    double alfv = 1. / o.recalf;
    if (64. * alfv + o.vnu < vft) keep = 0;
    if (keep) {
      DevWin W = wins[pan[pq].w];
      int jj = fint((o.vnu - W.v1r4) / Y.dvr4 + 1.);
      jj = jj < 1 ? 1 : (jj > W.nptr4 ? W.nptr4 : jj);
      double r = r4[(size_t)pan[pq].w * 2504 + jj];
      if (o.speak <= (Y.dptmn_hirac + Y.dptfc * r)) keep = 0;
    }
  }
  keep_out[t] = keep;
  if (keep) atomicMax(&blkmax[k], o.ff);
}

// Exclusive prefix max over blocks: panel at which each block is read.
// Sets *changed when the guess moved; resets blkmax for the next pass.
__global__ __launch_bounds__(256) void k_pread_scan(int *blkmax, int nbk, int *pread, int *changed) {
  __shared__ int buf[256];
  __shared__ int carry;
  int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nbk; base += 256) {
    int k = base + tid;
    int v = (k < nbk) ? blkmax[k] : 0;
    buf[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      int t = tid >= off ? buf[tid - off] : 0;
      __syncthreads();
      buf[tid] = max(buf[tid], t);
      __syncthreads();
    }
    int excl = max(carry, tid > 0 ? buf[tid - 1] : 0);
    if (k < nbk) {
      if (pread[k] != excl) *changed = 1;
      pread[k] = excl;
      blkmax[k] = 0;
    }
    __syncthreads();
    if (tid == 0) carry = max(carry, buf[255]);
    __syncthreads();
  }
}

// Processing panel of each kept line: the panel at which its block is read,
// raised by the first-fit panels of kept lines before it in the block
// (CNVFNV stops at the first line that does not fit, oprop.f90:1287-1292).
// Then the scatter record of CNVFNV (oprop.f90:1262-1305).
__global__ __launch_bounds__(256) void k_scatter_rec(const NearLine *nl, const int *keep, int nlo,
                                                     DevLines L, int blk0, int nbk, const int *pread,
                                                     const DevPanel *pan, DevLayer Y, int nhi_t,
                                                     ScatterRec *rec) {
  __shared__ int buf[256];
  int k = blockIdx.x;  // near block
  int b = blk0 + k;
  int s = L.bstart[b] - nlo, e = L.bstart[b + 1] - nlo;
  if (s < 0) s = 0;
  if (e > nhi_t) e = nhi_t;
  int tid = threadIdx.x;
  int t = s + tid;
  bool valid = t < e;
  int kp = valid ? keep[t] : 0;
  NearLine o;
  if (valid) o = nl[t];
  int v = (valid && kp) ? o.ff : 0;
  buf[tid] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int q = tid >= off ? buf[tid - off] : 0;
    __syncthreads();
    buf[tid] = max(buf[tid], q);
    __syncthreads();
  }
  if (!valid) return;
  ScatterRec r;
  r.npts = 0;
  r.pad = 0;
  r.g1lo = 0;
  r.s2 = r.s3 = 0;
  r.zfl1 = r.zfl2 = r.zfl3 = r.zslope = 0.;
  r.strf1 = r.strf2 = r.strf3 = r.strd = r.strver = 0.;
  int P = max(pread[k], buf[tid]);
  if (kp && P < Y.npanel) {
    double vft = pan[P].vft;
    double depthi = o.sp * o.recalf;
    int izm = o.izeta;
    double zetdif = 100. * o.zetai - (double)(izm - 1);
    r.strf1 = depthi * (c_cf1[izm - 1] + zetdif * (c_cf1[izm] - c_cf1[izm - 1]));
    r.strf2 = depthi * (c_cf2[izm - 1] + zetdif * (c_cf2[izm] - c_cf2[izm - 1]));
    r.strf3 = depthi * (c_cf3[izm - 1] + zetdif * (c_cf3[izm] - c_cf3[izm - 1]));
    r.strd = depthi * (c_cgauss[izm - 1] + zetdif * (c_cgauss[izm] - c_cgauss[izm - 1]));
    r.strver = depthi * (c_cer[izm - 1] + zetdif * (c_cer[izm] - c_cer[izm - 1]));
    double zslope = o.recalf * (Y.dv / 0.002);
    double zint = (o.vnu - vft) / Y.dv;
    double bhwdxf = (4. / 0.002) / zslope;
    int jmax1 = fint(zint + bhwdxf + 1.5);
    int jmin1 = fint(zint - bhwdxf + 1.5);
    double conf2 = Y.dv / Y.dvr2, conf3 = Y.dv / Y.dvr3;
    double rshft = zint < 0.0 ? -0.5 : 0.5;
    int j2shft = fint(zint * (1. - conf2) + rshft);
    int j3shft = fint(zint * (1. - conf3) + rshft);
    int jmin2 = jmin1 - j2shft, jmin3 = jmin1 - j3shft;
    r.zfl1 = ((double)(jmin1 - 2) - zint) * zslope;
    r.zfl2 = ((double)(jmin2 - 2) - zint * conf2) * zslope;
    r.zfl3 = ((double)(jmin3 - 2) - zint * conf3) * zslope;
    r.zslope = zslope;
    r.g1lo = (int64_t)jmin1 - 1 + (int64_t)2400 * P;
    r.s2 = (int64_t)j2shft + (int64_t)1800 * P;
    r.s3 = (int64_t)j3shft + (int64_t)2250 * P;
    r.npts = jmax1 - jmin1 + 1;
  }
  rec[t] = r;
}

__global__ void k_fill_i32(int *p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Near-wing point updates of one grid (sum of JMAX1-JMIN1+1 over kept lines).
__global__ void k_sum_npts(const ScatterRec *rec, int n, unsigned long long *acc) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = (i < n && rec[i].npts > 0) ? (unsigned long long)rec[i].npts : 0ull;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(acc, v);
}

// Line range [lo, hi) whose scatter extent may touch each tile of grid G.
__global__ void k_tile_ranges(const double *vraw, int nn, double vft0, double dvg, int tile,
                              int ntiles, double margin, int *lo, int *hi) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  double a = vft0 + (double)t * tile * dvg - margin;
  double b = vft0 + (double)(t + 1) * tile * dvg + margin;
  int l = 0, h = nn;
  while (l < h) {
    int m = (l + h) >> 1;
    if (vraw[m] < a) l = m + 1;
    else h = m;
  }
  int l2 = l, h2 = nn;
  while (l2 < h2) {
    int m = (l2 + h2) >> 1;
    if (vraw[m] <= b) l2 = m + 1;
    else h2 = m;
  }
  lo[t] = l;
  hi[t] = l2;
}

// XINT of a coarse array into R3 point G3 (oprop.f90:2219-2230) for the
// panel that owns G3.
__device__ __forceinline__ double xint_at(double v1a, double v2a, double dva, const double *A,
                                          double vft, double dvr3, int I, int n1r3, int n2r3) {
  double recdva = 1. / dva;
  int ilo = fint((v1a + dva - vft) / dvr3 + 1. + ONEMI_D);
  if (ilo < n1r3) ilo = n1r3;
  int ihi = fint((v2a - dva - vft) / dvr3 + ONEMI_D);
  if (ihi > n2r3) ihi = n2r3;
  if (I < ilo || I > ihi) return 0.;
  double vi = vft + dvr3 * (double)(I - 1);
  int j = fint((vi - v1a) * recdva + ONEPL_D);
  double vj = v1a + dva * (double)(j - 1);
  double p = recdva * (vi - vj);
  double c = (3. - 2. * p) * p * p;
  double b = 0.5 * p * (1. - p);
  double b1 = b * (1. - p);
  double b2 = b * p;
  return -A[j - 1] * b1 + A[j] * (1. - c + b2) + A[j + 1] * (c + b1) - A[j + 2] * b2;
}

// Near-line scatter into one tile of grid G (3: DVR3, 2: DVR2, 1: DV) with
// the PANEL combine in the epilogue.
//   G = 3: R3 = lines(F3) + XINT(R4) + XINT(ABSRB)           -> r3 (global)
//   G = 2: R2 = lines(F2) + Lagrange4(R3)                      -> r2 (global)
//   G = 1: OD = (lines(F1,FG,XVER) + Lagrange4(R2)) * radiation -> od
template <int G, typename OUT>
__global__ __launch_bounds__(256) void k_tile(const ScatterRec *rec, const int *tlo, const int *thi,
                                              int tile, const double *tabs, DevLayer Y,
                                              const DevPanel *pan, const DevWin *wins,
                                              const double *r4, const double *absrb1,
                                              const DevSeg *segs, const double *rin, double *rout,
                                              OUT *od) {
  extern __shared__ double sm[];
  double *acc = sm;                 // [tile]
  double *t1 = sm + tile;           // F_G table
  double *tg = t1 + NFMX;           // FG   (G = 1)
  double *tv = tg + NFMX;           // XVER (G = 1)
  const int tid = threadIdx.x;
  for (int k = tid; k < tile; k += 256) acc[k] = 0.;
  const double *src = tabs + (size_t)(G == 1 ? 0 : (G == 2 ? 1 : 2)) * NFMX;
  for (int k = tid; k < NFMX; k += 256) t1[k] = src[k];
  if (G == 1) {
    for (int k = tid; k < NFMX; k += 256) {
      tg[k] = tabs[3 * NFMX + k];
      tv[k] = tabs[4 * NFMX + k];
    }
  }
  __syncthreads();
  const int64_t T0 = (int64_t)blockIdx.x * tile;
  const int lo = tlo[blockIdx.x], hi = thi[blockIdx.x];
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  for (int li = lo + wave; li < hi; li += 4) {
    const ScatterRec r = rec[li];
    if (r.npts <= 0) continue;
    int64_t glo = G == 1 ? r.g1lo : (G == 2 ? r.g1lo - r.s2 : r.g1lo - r.s3);
    int64_t a = glo > T0 ? glo : T0;
    int64_t b = glo + r.npts - 1;
    if (b > T0 + tile - 1) b = T0 + tile - 1;
    const double zfl = G == 1 ? r.zfl1 : (G == 2 ? r.zfl2 : r.zfl3);
    const double str = G == 1 ? r.strf1 : (G == 2 ? r.strf2 : r.strf3);
    for (int64_t k = a + lane; k <= b; k += 64) {
      int off = (int)(k - glo);
      double zf = zfl + (double)(off + 1) * r.zslope;
      int iz = fint(fabs(zf) + 1.5) - 1;
      iz = iz > NFMX - 1 ? NFMX - 1 : iz;  // |zf| <= 2500 by construction
      double v;
      if (G == 1) v = str * t1[iz] + r.strd * tg[iz] + r.strver * tv[iz];
      else v = str * t1[iz];
      atomicAdd(&acc[k - T0], v);
    }
  }
  __syncthreads();
  const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
  const double X10 = -1. / 16., X11 = 9. / 16.;
  for (int k = tid; k < tile; k += 256) {
    int64_t g = T0 + k;
    if (G == 3) {
      if (g >= Y.n3) continue;
      double v = acc[k];
      if (g <= Y.g3max) {
        int p = g < 155 ? 0 : (int)((g - 5) / 150);
        if (p > Y.npanel - 1) p = Y.npanel - 1;
        int I = (int)(g - (int64_t)150 * p) + 1;
        int n1r3 = p == 0 ? 1 : 6;
        DevPanel pp = pan[p];
        DevWin W = wins[pp.w];
        v = v + xint_at(W.v1r4, W.v2r4, Y.dvr4, r4 + (size_t)pp.w * 2504, pp.vft, Y.dvr3, I, n1r3,
                        155) * 1.0;
        if (Y.icntnm >= 1)
          v = v + xint_at(Y.v1abs, Y.v2abs, Y.dvabs, absrb1, pp.vft, Y.dvr3, I, n1r3, 155) * 1.;
      }
      rout[g] = v;
    } else if (G == 2) {
      if (g >= Y.n2) continue;
      double v = acc[k];
      if (g >= 4 && g <= Y.g2max) {
        int64_t q = g >> 2;
        int ph = (int)(g & 3);
        double a0 = rin[q - 1], a1 = rin[q], a2 = rin[q + 1], a3 = rin[q + 2];
        if (ph == 0) v = v + a1;
        else if (ph == 1) v = v + X00 * a0 + X01 * a1 + X02 * a2 + X03 * a3;
        else if (ph == 2) v = v + X10 * (a0 + a3) + X11 * (a1 + a2);
        else v = v + X03 * a0 + X02 * a1 + X01 * a2 + X00 * a3;
      }
      rout[g] = v;
    } else {
      if (g < 32 || g > Y.g_last) continue;
      double v = acc[k];
      int64_t q = g >> 2;
      int ph = (int)(g & 3);
      double a0 = rin[q - 1], a1 = rin[q], a2 = rin[q + 1], a3 = rin[q + 2];
      if (ph == 0) v = v + a1;
      else if (ph == 1) v = v + X00 * a0 + X01 * a1 + X02 * a2 + X03 * a3;
      else if (ph == 2) v = v + X10 * (a0 + a3) + X11 * (a1 + a2);
      else v = v + X03 * a0 + X02 * a1 + X01 * a2 + X00 * a3;
      if (Y.jrad == 0) {
        int p = (int)((g - 32) / 2400);
        if (p > Y.npanel - 1) p = Y.npanel - 1;
        DevPanel pp = pan[p];
        int I = (int)(g - (int64_t)2400 * p) + 1;
        int lo2 = 0, hi2 = pp.nseg - 1;
        while (lo2 < hi2) {
          int mid = (lo2 + hi2 + 1) >> 1;
          if (segs[pp.seg0 + mid].i1 <= I) lo2 = mid;
          else hi2 = mid - 1;
        }
        DevSeg sg = segs[pp.seg0 + lo2];
        v = v * (sg.radvi + (double)(I - sg.i1) * sg.rdel);
      }
      od[g - 32] = (OUT)v;
    }
  }
}

}  // namespace lblx
