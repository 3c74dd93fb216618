// HIP kernels for the layer optical-depth path on MI355X (gfx950).
//
// The reference computes one layer panel by panel on one core (HIRAC1,
// src/oprop.f90:25-668).  Here every layer ("job") of a batch is resident
// in HBM over its whole band and each stage is one launch for all jobs:
//   LINF4 + SHRINK   k_shr_count, k_shr_scan, k_shr_compact,
//                    k_shr_ewalk, k_shr_t9sim/t9chain/t9replay, k_shr_bwalk,
//                    k_shr_emit
//                    (oprop.f90:3093-3290, 3508-3640)
//   LBLF4/CONVF4     k_convf4, one far-wing window per workgroup
//                    (oprop.f90:3641-4039)
//   LNCOR1           k_lncor1 + the block-read fixpoint k_reject/k_pread_scan
//                    (oprop.f90:734-885, 995-1193)
//   CNVFNV set-up    k_scatter_rec (oprop.f90:1262-1305)
//   CNVFNV + PANEL   k_tile<G>: scatter into R3/R2/R1 tiles in LDS with the
//                    XINT / Lagrange / radiation combine in the epilogue
//                    (oprop.f90:1307-1320, 1459-1519, 2201-2234)
// Every order-dependent decision of the reference (line rejection against
// the running far-wing array, SHRINK batch boundaries, far-wing windows,
// the panel at which a line block is read) is reproduced exactly.
#include <hip/hip_runtime.h>
#include <type_traits>

#include <climits>
#include <cstdint>

#include "lblx_kernels.h"

namespace lblx {

__constant__ double c_avrat[102], c_cf1[102], c_cf2[102], c_cf3[102], c_cgauss[102], c_cer[102];
__constant__ int c_isomax[7];
// exact-Voigt build: a_3, b_3 of voigt_init over zeta index 0..100, and the
// value CONVF4 reads one past the end (a_3(101) = b_1(0) in COMMON
// /voigt_cf/; b_3(101) taken as b_3(100))
__constant__ double c_va3[102], c_vb3[102];
constexpr int NZETA = 101;

constexpr double ONEPL_D = 1.001, ONEMI_D = 0.999;
constexpr int NFMX = 2601;
constexpr int LIMIN4 = 1000;  // LINF4 batch size

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

// Block-wide exclusive scan (sum) for NT threads.
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

__global__ void k_fill_i32(int *p, int64_t n, int v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------
// Line coupling of record i at the layer's temperature: Y and G (IFLG 1)
// or GAMMA1 and GAMMA2 (IFLG 3) interpolated linearly between the two
// TEMPLC nodes of interval w (0: LNCOR1, oprop.f90:1039-1061; 1: LINF4,
// :3118-3135).  flag: 0 for IFLG 0, 2 for another IFLG > 0.
struct LcCoef {
  double yi, gi, g1, g2;
  int flag;
};
__device__ __forceinline__ LcCoef lc_coef(const DevLines &L, const DevLayer &Y, int i, int w) {
  LcCoef r{0., 0., 0., 0., 0};
  if (!L.cidx) return r;
  const int k = L.cidx[i];
  if (k == -1) return r;
  if (k < 0) {
    r.flag = 2;
    return r;
  }
  const double *a = L.cco + (size_t)(k >> 1) * 8, *b = a + 4;
  const int lo = Y.lc_lo[w], an = Y.lc_an[w];
  const double va = a[an] + (a[lo + 1] - a[lo]) * Y.lc_rect[w] * Y.lc_dif[w];
  const double vb = b[an] + (b[lo + 1] - b[lo]) * Y.lc_rect[w] * Y.lc_dif[w];
  if (k & 1) {
    r.flag = 3;
    r.g1 = va;
    r.g2 = vb;
  } else {
    r.flag = 1;
    r.yi = va;
    r.gi = vb;
  }
  return r;
}

// ---------------------------------------------------------------------
// LINF4 correction of one record (oprop.f90:3093-3235).  Returns whether
// the record enters SHRINK and fills the corrected centre, strength and
// widths.  RDLNFL's O2 air->foreign width conversion (oprop.f90:3456-3462)
// is applied first.
struct F4Line {
  double x, sp, al, ad, spp;
  int co2, m;
};

// LINF4's tests alone: whether record i enters SHRINK, and its shifted
// centre.  linf4_line applies the same tests.
__device__ __forceinline__ bool linf4_pass(const DevLines &L, const DevLayer &Y, int i, double &x) {
  const int b = L.blk[i];
  const int mc = L.mol[i];
  const int m = mc % 100, iso = (mc % 1000) / 100;
  const double vraw = L.vnu[i];
  if (L.bvmax[b] < Y.v1l4 || m < 1 || m > Y.nmol || iso < 1 || iso > c_isomax[m - 1]) return false;
  const double sui = (double)L.sp[i] * Y.wk[m - 1];
  if (sui == 0. || vraw < Y.v1l4) return false;
  x = vraw + Y.rhorat * (double)L.pshift[i];
  return true;
}

__device__ __forceinline__ bool linf4_line(const DevLines &L, const DevLayer &Y, int i, F4Line &o) {
  double x;
  if (!linf4_pass(L, Y, i, x)) return false;
  const int mc = L.mol[i];
  const int m = mc % 100, iso = (mc % 1000) / 100;
  double sui = (double)L.sp[i] * Y.wk[m - 1];
  double alfa = L.alfa[i], hwhm = L.hwhm[i];
  if (m == 7) alfa = (alfa - 0.21 * hwhm) / (1.0 - 0.21);
  if (x == 0.) sui = 2. * sui;  // a line at 0 cm-1 counts twice (oprop.f90:3172)
  double epp = L.epp[i];
  if (epp <= -1.001) epp = fabs(epp);
  if (epp > -0.999) sui = sui * Y.scor[m - 1][iso - 1] * exp(-epp * Y.betacr4) * (1. + exp(-x * Y.beta4));
  double tmpcor = pow(Y.tratio, (double)L.tmpalf[i]);
  double alfa0i = alfa * tmpcor;
  double hwhmsi = hwhm * tmpcor;
  const int ci = L.cidx ? L.cidx[i] : -1;
  o.al = alfa0i * (Y.rhorat - Y.rhoslf[m - 1]) + hwhmsi * Y.rhoslf[m - 1];
  o.ad = x * Y.alfd1[m - 1][iso - 1];
  o.x = x;
  o.spp = 0.;
  if (ci < 0) {
    o.sp = sui < 1.0e-35 ? 0.0 : sui;  // SHRINK zeroes S < 1e-35 (oprop.f90:3546-3549)
  } else {
    // IFLG 3: reduced width; IFLG 1: SP = SUI (1 + G P^2), SPP = SUI Y P;
    // SHRINK zeroes both where SP < 1e-35 (oprop.f90:3192-3225, 3546-3549)
    const LcCoef lc = lc_coef(L, Y, i, 1);
    if (lc.flag == 3) o.al = o.al * (1.0 - lc.g1 * Y.pavp0 - lc.g2 * Y.pavp2);
    const double sp = sui * (1. + lc.gi * Y.pavp2);
    const bool z = sp < 1.0e-35;
    o.sp = z ? 0.0 : sp;
    o.spp = z ? 0.0 : sui * lc.yi * Y.pavp0;
  }
  o.co2 = m == 2;
  o.m = m;
  return true;
}
// SHRINK's flags of a passing line (bit 0 CO2, bit 1 positive strength,
// bit 2 coupled, bits 3-5 molecule)
__device__ __forceinline__ int8_t f4_flags(const F4Line &o) {
  return (int8_t)((o.co2 ? 1 : 0) | (o.sp > 0. ? 2 : 0) | (o.spp != 0. ? 4 : 0) | (o.m << 3));
}

// ---------------------------------------------------------------------
// SHRINK, in parallel.
//
// LINF4 reads 250-line TAPE3 blocks until its buffer holds >= 1000 lines
// (shrunk lines kept from earlier batches of the current TAPE9 block plus
// new raw lines) and then SHRINKs the new lines as one batch.  Inside a
// batch a group ends before the first line whose centre exceeds VLMT;
// VLMT starts at x_s + DVR4/2 and is reset to (centre of the group's last
// line) + DVR4/2.  Group ends therefore form the chain
//     e_1 = g(s),  e_{k+1} = f(e_k),
//     f(e) = (first k >= e+2 with x_k > x_e + DVR4/2) - 1,
//     g(s) = (first k >= s+1 with x_k > x_s + DVR4/2) - 1,
// cut at the batch end.  A group emits one non-CO2 and one CO2 effective
// line if it holds positive strength of that class.  g, f and the count
// each step emits (jmp, wsum) are tabled per line as the passing lines are
// compacted; a wave per TAPE3 block walks the chain from the block's first
// line through the next EBLK blocks and tables the count a batch of those
// blocks would emit; one wave per job then walks LINF4's batch boundaries
// over those tables, and one wave per batch walks its chain once more to
// mark each group's last line with the group's first line and output slot.
constexpr int SB_T = 256;

// Per passing line, cfl: bit 0 CO2, bit 1 positive strength.  The lines a
// group [a, b] emits: one per class with a positive line.
// A coupled line (bit 2) is emitted on its own and enters neither sum.
__device__ __forceinline__ int em_bits(int f) { return (f & 4) ? 0 : ((f & 2) ? 1 << (f & 1) : 0); }
__device__ __forceinline__ int em_of(int m) { return (m & 1) + (m >> 1); }
// the same over [a, b] by a whole wave (a, b wave-uniform), coupled lines
// counted one each
__device__ __forceinline__ int em_cnt(const int8_t *cfl, int a, int b) {
  int m = 0, c = 0;
  for (int k0 = a; k0 <= b; k0 += 64) {
    const int k = k0 + (int)(threadIdx.x & 63);
    const int f = k <= b ? cfl[k] : 0;
    m |= em_bits(f);
    c += __builtin_popcountll(__ballot(f & 4));
  }
  return (__ballot(m & 1) ? 1 : 0) + (__ballot(m & 2) ? 1 : 0) + c;
}

// per TAPE3 block: passing lines, and the block-local position of the
// first passing line above V2L4 (LINF4 stops there with IEOF = 1,
// oprop.f90:3228-3231)
__global__ __launch_bounds__(SB_T) void k_shr_count(DevLines L, const DevLayer *Ys, Strides S,
                                                    ShrinkWork W) {
  __shared__ int buf[SB_T];
  __shared__ int s_stop;
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  const int bl = blockIdx.x;  // block index relative to Y.b4
  if (bl >= Y.nb4) return;
  const int b = Y.b4 + bl;
  int i = L.bstart[b] + threadIdx.x;
  int pass = 0, stp = 0;
  if (i < L.bstart[b + 1]) {
    double x;
    if (linf4_pass(L, Y, i, x)) {
      pass = 1;
      stp = x > Y.v2l4;
    }
  }
  if (threadIdx.x == 0) s_stop = INT_MAX;
  int tot;
  int pos = block_scan_excl<SB_T>(pass, buf, &tot);
  if (stp) atomicMin(&s_stop, pos);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t o = job * S.blk + bl;
    W.bcnt[o] = tot;
    W.bstop[o] = s_stop == INT_MAX ? -1 : s_stop;
  }
}

// per job: exclusive prefix of the passing lines over its blocks; passing
// index of the stop line (or -1); the chain's closing entry at n
__global__ __launch_bounds__(256) void k_shr_scan(const DevLayer *Ys, Strides S, ShrinkWork W) {
  __shared__ int buf[256];
  __shared__ int carry;
  __shared__ int s_stop;
  const int job = blockIdx.x;
  const int nblk = Ys[job].nb4;
  const int64_t o = job * S.blk;
  if (threadIdx.x == 0) {
    carry = 0;
    s_stop = INT_MAX;
  }
  __syncthreads();
  for (int base = 0; base < nblk; base += 256) {
    const int b = base + threadIdx.x;
    const int v = b < nblk ? W.bcnt[o + b] : 0;
    int tot;
    const int ex = block_scan_excl<256>(v, buf, &tot);
    if (b < nblk) {
      W.bp[o + b] = carry + ex;
      if (W.bstop[o + b] >= 0) atomicMin(&s_stop, carry + ex + W.bstop[o + b]);
    }
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const int n = carry;
    W.bp[o + nblk] = n;
    W.npass[job] = n;
    W.stop[job] = s_stop == INT_MAX ? -1 : s_stop;
    W.jmp[job * S.line + n] = n;
    W.wsum[job * S.line + n] = 0;
  }
}

// Passing lines after the current TAPE3 block, read from the raw records
// (for the chain entries of a block's last lines): x and cfl of passing
// index k = k0, k0 + 1, ... up to n (the job's passing lines).
struct RawCursor {
  const DevLines *L;
  const DevLayer *Y;
  int i, iend, k, n;
  // the next passing line: false past the job's last one
  __device__ bool next(double &x, int &f) {
    while (k < n && i < iend) {
      F4Line o;
      const int r = i++;
      if (linf4_line(*L, *Y, r, o)) {
        k++;
        x = o.x;
        f = f4_flags(o);
        return true;
      }
    }
    return false;
  }
};

// per TAPE3 block: write the passing lines compacted in file order, then
// each one's chain entries g(e), f(e) and the count the group after e
// emits.  The block's lines are in LDS; a line whose search runs past the
// block continues over the raw records that follow.
constexpr int SB_HALO = 64;  // raw records after the block read into LDS
__global__ __launch_bounds__(SB_T) void k_shr_compact(DevLines L, const DevLayer *Ys, Strides S,
                                                      ShrinkWork W) {
  __shared__ int buf[SB_T];
  __shared__ double sx[SB_T + SB_HALO];
  __shared__ int8_t sf[SB_T + SB_HALO];
  __shared__ int s_hn;
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  const int bl = blockIdx.x;
  if (bl >= Y.nb4) return;
  const int b = Y.b4 + bl;
  const int i = L.bstart[b] + threadIdx.x;
  F4Line o{};
  int pass = 0;
  if (i < L.bstart[b + 1] && linf4_line(L, Y, i, o)) pass = 1;
  int tot;
  const int pos = block_scan_excl<SB_T>(pass, buf, &tot);
  const int64_t ob = job * S.blk + bl;
  const int64_t oj = job * S.line;
  const int k0 = W.bp[ob];  // passing index of the block's first passing line
  const int8_t fl = f4_flags(o);
  if (pass) {
    const int64_t k = oj + k0 + pos;
    W.cx[k] = o.x;
    W.csp[k] = o.sp;
    if (W.cspp) W.cspp[k] = o.spp;
    W.cal[k] = o.al;
    W.cad[k] = o.ad;
    W.cfl[k] = fl;
    if (W.craw) W.craw[k] = i;
    sx[pos] = o.x;
    sf[pos] = fl;
  }
  // the passing lines among the SB_HALO raw records after the block (the
  // searches of the block's last lines run into them), after its own
  const int iend = L.bstart[Y.b4 + Y.nb4];
  const int ih = L.bstart[b + 1] + (int)threadIdx.x;
  if (threadIdx.x < SB_HALO) {
    F4Line oh{};
    const bool ph = ih < iend && linf4_line(L, Y, ih, oh);
    const uint64_t mh = __ballot(ph);
    const int ph_pos = __builtin_popcountll(mh & ((1ull << threadIdx.x) - 1ull));
    if (ph) {
      sx[tot + ph_pos] = oh.x;
      sf[tot + ph_pos] = f4_flags(oh);
    }
    if (threadIdx.x == 0) s_hn = __builtin_popcountll(mh);
  }
  __syncthreads();
  if (threadIdx.x >= tot) return;
  const int n = W.npass[job];
  const int nl = tot + s_hn;  // lines in LDS
  const int el = threadIdx.x;  // the block's el-th passing line
  const int e = k0 + el;
  const double thr = sx[el] + Y.dvr4 / 2.;
  RawCursor rc{&L, &Y, L.bstart[b + 1] + SB_HALO, iend, k0 + nl, n};
  // passing line k > e: from LDS within the block and its halo, else from
  // the cursor (k increases by one per call)
  auto line_at = [&](int k, double &x, int &f) -> bool {
    if (k >= n) return false;
    if (k - k0 < nl) {
      x = sx[k - k0];
      f = sf[k - k0];
      return true;
    }
    return rc.next(x, f);
  };
  // g: first k >= e+1 with x_k > thr; f from e+2 on; the emitted classes
  // of [e+1, f]
  int k = e + 1, m = 0, g = -1, nc = 0;
  double x = 0.;
  int f = 0;
  for (;; k++) {
    if (!line_at(k, x, f)) break;  // k = n
    if (x > thr) {
      if (g < 0) g = k - 1;
      if (k >= e + 2) break;
      m |= em_bits(f);  // line e+1 is past VLMT but opens the group after e
      nc += (f >> 2) & 1;
      continue;
    }
    m |= em_bits(f);
    nc += (f >> 2) & 1;
  }
  if (g < 0) g = k - 1;
  W.g[oj + e] = g;
  if (e == n - 1) {
    W.jmp[oj + e] = n;
    W.wsum[oj + e] = 0;
    return;
  }
  W.jmp[oj + e] = k - 1;
  W.wsum[oj + e] = (int16_t)(em_of(m) + nc);
}

// Lines SHRINK emits for the batch [s, t] of passing indices: the first
// group ends at g(s) (or is cut at t), then the chain e -> f(e) while it
// stays below t, and the last group is cut at t.
// One wave walks the chain: the lanes hold f and the emitted count of 64
// consecutive lines from the window base, so a step is a lane read and the
// line table is fetched once per 64 lines.  Every value but the window is
// wave-uniform; all lanes of the wave call it together.
struct ChainWalk {
  const int8_t *cfl;
  const int32_t *jmp;
  const int16_t *wsum;
  int lim, s, e1, cur, acc, base, wj, ww;
  __device__ void load(int b) {
    base = b;
    const int i = b + (int)(threadIdx.x & 63);
    wj = i < lim ? jmp[i] : INT_MAX;
    ww = i < lim ? (int)wsum[i] : 0;
  }
  __device__ ChainWalk(const ShrinkWork &W, int64_t o, int n, int s_) {
    cfl = W.cfl + o;
    jmp = W.jmp + o;
    wsum = W.wsum + o;
    lim = n + 1;
    s = s_;
    e1 = __builtin_amdgcn_readfirstlane(W.g[o + s]);
    cur = e1;
    acc = em_cnt(cfl, s, e1);
    load(cur);
  }
  // f(cur) and the count the group after cur emits
  __device__ int next(int *w) {
    if (cur >= base + 64) load(cur);
    const int k = __builtin_amdgcn_readfirstlane(cur - base);
    *w = __builtin_amdgcn_readlane(ww, k);
    return __builtin_amdgcn_readlane(wj, k);
  }
  // t never decreases from one call to the next
  __device__ int count_to(int t) {
    if (e1 >= t) return em_cnt(cfl, s, t);
    for (;;) {
      int w;
      const int nx = next(&w);
      if (nx >= t) break;
      acc += w;
      cur = nx;
    }
    return acc + em_cnt(cfl, cur + 1, t);
  }
};

// per (job, block b): lines emitted by a batch of blocks b..b+k for k <
// EBLK (block indices relative to the job's first far-wing block), and by
// the batch from b to the stop line when that lies within those blocks
// (-1 otherwise).  A batch ends at the first block that brings its passing
// lines to LIMIN4 or more, so the table stops there (-1 beyond).
__global__ __launch_bounds__(64) void k_shr_ewalk(const DevLayer *Ys, Strides S, ShrinkWork W) {
  const int job = blockIdx.y;
  const int b = blockIdx.x;  // one wave per block
  const int lane = threadIdx.x;
  const int nblk = Ys[job].nb4;
  if (b >= nblk) return;
  const int64_t ob = job * S.blk;
  const int s = W.bp[ob + b];
  const int stop_p = W.stop[job];
  int32_t *et = W.etab + (ob + b) * EBLK;
  if (W.bp[ob + nblk] <= s) {  // no passing line from here on
    for (int k = lane; k < EBLK; k += 64) et[k] = 0;
    if (lane == 0) W.estop[ob + b] = -1;
    return;
  }
  ChainWalk cw(W, job * S.line, W.npass[job], s);
  int es = -1, mine = -1;  // lane k keeps entry k
  int k = 0;
  for (; k < EBLK; k++) {
    if (b + k + 1 > nblk) {
      if (lane == k) mine = 0;
      continue;
    }
    const int t = W.bp[ob + b + k + 1] - 1;
    if (es < 0 && stop_p >= s && stop_p <= t) es = cw.count_to(stop_p);
    const int v = t >= s ? cw.count_to(t) : 0;
    if (lane == k) mine = v;
    if (t + 1 - s >= LIMIN4) break;
  }
  if (lane < EBLK) et[lane] = mine;
  if (lane == 0) W.estop[ob + b] = es;
}

// LINF4's read loop over blocks (oprop.f90:3130-3290) decides the batch
// boundaries and TAPE9 blocks in order.  A batch reads blocks until it
// holds the stop line or J + (passing lines read) >= LIMIN4, J being the
// shrunk lines kept from the TAPE9 block's earlier batches; both tests are
// monotone in the block prefix bp.  Blocks run from the job's first
// far-wing block; its last one holds the stop line unless the file ends
// first.  A TAPE9 block starts at a block boundary with J = 0, so the walk
// runs in three launches:
//   k_shr_t9sim    every block as a possible TAPE9-block start, one lane
//                  each: where the next start is, the lines the block
//                  emits and its batch count;
//   k_shr_t9chain  one wave per job follows the starts from block 0;
//   k_shr_t9replay one wave per start on the chain writes its batches and
//                  its TAPE9 block entry.
struct Batch {
  int b, s, t;  // next block, passing-line range [s, t]
  bool stop, eof;
};

// one batch from block bs, one lane alone
__device__ __forceinline__ Batch batch_lane(const int32_t *bp, int nblk, int stop_p, int J, int bs) {
  Batch B;
  B.s = bp[bs];
  B.stop = B.eof = false;
  int pe = 0;  // bp after the batch's last block
  if (bs >= nblk) {
    B.eof = true;
    B.b = bs;
  } else {
    const int lim = LIMIN4 - J + B.s;
    int k = bs;
    for (;;) {
      const int pn = bp[k + 1];
      if (stop_p >= 0 && pn > stop_p) {
        B.stop = true;
        pe = pn;
        break;
      }
      if (pn >= lim) {
        pe = pn;
        break;
      }
      if (++k >= nblk) {
        B.eof = true;
        pe = pn;
        break;
      }
    }
    B.b = B.eof ? nblk : k + 1;
  }
  B.t = B.stop ? stop_p : pe - 1;
  return B;
}

// the same batch, the wave testing 64 blocks at once
__device__ __forceinline__ Batch batch_wave(const int32_t *bp, int nblk, int stop_p, int J, int bs) {
  const int lane = threadIdx.x & 63;
  Batch B;
  B.s = bp[bs];
  B.stop = B.eof = false;
  int pe = 0;
  B.b = bs;
  if (bs >= nblk) {
    B.eof = true;
  } else {
    const int lim = LIMIN4 - J + B.s;
    for (int base = bs;; base += 64) {
      const int k = base + lane;
      const int pn = k < nblk ? bp[k + 1] : 0;
      const bool hs = k < nblk && stop_p >= 0 && pn > stop_p;
      const bool hit = hs || (k < nblk && pn >= lim);
      const uint64_t m = __ballot(hit);
      if (m) {
        const int f = __ffsll((unsigned long long)m) - 1;
        B.b = base + f + 1;
        pe = __builtin_amdgcn_readlane(pn, f);
        B.stop = (__ballot(hs) >> f) & 1;
        break;
      }
      if (base + 64 >= nblk) {
        B.b = nblk;
        B.eof = true;
        pe = bp[nblk];
        break;
      }
    }
  }
  B.t = B.stop ? stop_p : pe - 1;
  return B;
}

// Lines a batch emits from the block tables, or -1 when they do not cover it
__device__ __forceinline__ int batch_em(const ShrinkWork &W, int64_t ob, const Batch &B, int bs) {
  if (B.t < B.s) return 0;
  const int nb = B.b - bs;
  if (B.stop && nb <= EBLK && W.estop[ob + bs] >= 0) return W.estop[ob + bs];
  return (!B.stop && nb >= 1 && nb <= EBLK) ? W.etab[(ob + bs) * EBLK + nb - 1] : -1;
}

// One TAPE9 block from block b0, the whole wave: its batches (a batch longer
// than the tables takes the chain walk), written when WRITE with the block's
// output offset out0, batch offset bat0 and TAPE9 block index k9.
template <bool WRITE>
__device__ void t9_block_wave(const ShrinkWork &W, const Tape9 &T9, int64_t ob, int64_t o, int64_t ot, int stop_p,
                              int nblk, int npass, int b0, int out0, int bat0, int k9, int *nxt, int *nem,
                              int *nbt) {
  const bool lead = (threadIdx.x & 63) == 0;
  int b = b0, J = 0, nbat = 0;
  for (;;) {
    const int bs = b;
    const Batch B = batch_wave(W.bp + ob, nblk, stop_p, J, bs);
    b = B.b;
    int em = 0;
    if (B.t >= B.s) {
      em = batch_em(W, ob, B, bs);
      if (em < 0) em = ChainWalk(W, o, npass, B.s).count_to(B.t);
      if (WRITE && lead) {
        W.bat_s[o + bat0 + nbat] = B.s;
        W.bat_t[o + bat0 + nbat] = B.t;
        W.bat_off[o + bat0 + nbat] = out0 + J;
      }
      nbat++;
    }
    J += em;
    if (J >= LIMIN4 || B.stop || B.eof) {
      if (WRITE && lead && J > 0) {
        T9.bstart[ot + k9] = out0;
        T9.bcount[ot + k9] = J;
      }
      *nxt = (B.stop || B.eof) ? -1 : b;
      *nem = J;
      *nbt = nbat;
      return;
    }
  }
}

// all_wave: leave every start to the chain wave (a test of that path)
__global__ __launch_bounds__(64) void k_shr_t9sim(const DevLayer *Ys, Strides S, ShrinkWork W, int all_wave) {
  const int job = blockIdx.y;
  const int nblk = Ys[job].nb4;
  const int b0 = blockIdx.x * 64 + (int)threadIdx.x;
  if (b0 >= nblk) return;
  const int64_t ob = job * S.blk;
  if (all_wave) {
    W.t9nxt[ob + b0] = -2;
    return;
  }
  const int stop_p = W.stop[job];
  const int32_t *bp = W.bp + ob;
  int b = b0, J = 0, nbat = 0, nxt = -1;
  for (;;) {
    const int bs = b;
    const Batch B = batch_lane(bp, nblk, stop_p, J, bs);
    b = B.b;
    if (B.t >= B.s) {
      const int em = batch_em(W, ob, B, bs);
      if (em < 0) {  // left to the chain wave
        nxt = -2;
        break;
      }
      J += em;
      nbat++;
    }
    if (J >= LIMIN4 || B.stop || B.eof) {
      nxt = (B.stop || B.eof) ? -1 : b;
      break;
    }
  }
  W.t9nxt[ob + b0] = nxt;
  W.t9nem[ob + b0] = J;
  W.t9nbt[ob + b0] = nbat;
}

// One wave per job; the start tables in LDS when the job has at most lcap
// blocks.  t9k must hold -1 for every block on entry.
__global__ __launch_bounds__(64) void k_shr_t9chain(const DevLayer *Ys, Strides S, ShrinkWork W, Tape9 T9,
                                                    int lcap) {
  extern __shared__ int32_t sm_c[];
  const int job = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const int nblk = Ys[job].nb4;
  const int64_t ob = job * S.blk, o = job * S.line, ot = job * S.t9blk;
  const int stop_p = W.stop[job], npass = W.npass[job];
  const bool inl = nblk <= lcap;
  int32_t *l_nxt = sm_c, *l_nem = sm_c + lcap, *l_nbt = sm_c + 2 * lcap;
  if (inl)
    for (int i = lane; i < nblk; i += 64) {
      l_nxt[i] = W.t9nxt[ob + i];
      l_nem[i] = W.t9nem[ob + i];
      l_nbt[i] = W.t9nbt[ob + i];
    }
  __syncthreads();
  int b = 0, out = 0, nbat = 0, k = 0;
  while (b < nblk) {
    int nx = inl ? l_nxt[b] : W.t9nxt[ob + b];
    int J, nb;
    if (nx == -2) {
      t9_block_wave<false>(W, T9, ob, o, ot, stop_p, nblk, npass, b, 0, 0, 0, &nx, &J, &nb);
    } else {
      J = inl ? l_nem[b] : W.t9nem[ob + b];
      nb = inl ? l_nbt[b] : W.t9nbt[ob + b];
    }
    if (lane == 0) {
      W.t9k[ob + b] = k;
      W.t9out[ob + b] = out;
      W.t9bat[ob + b] = nbat;
    }
    out += J;
    nbat += nb;
    if (J > 0) k++;
    if (nx < 0) break;
    b = nx;
  }
  if (lane == 0) {
    W.nbat[job] = nbat;
    T9.n[job] = out;
    T9.nblk[job] = k;
  }
}

__global__ __launch_bounds__(64) void k_shr_t9replay(const DevLayer *Ys, Strides S, ShrinkWork W, Tape9 T9) {
  const int job = blockIdx.y;
  const int b0 = blockIdx.x;
  const int nblk = Ys[job].nb4;
  if (b0 >= nblk) return;
  const int64_t ob = job * S.blk;
  const int k = W.t9k[ob + b0];
  if (k < 0) return;
  int nx, J, nb;
  t9_block_wave<true>(W, T9, ob, job * S.line, job * S.t9blk, W.stop[job], nblk, W.npass[job], b0,
                      W.t9out[ob + b0], W.t9bat[ob + b0], k, &nx, &J, &nb);
}

// per (job, batch), one wave each: the batch's chain.  Each group that
// emits lines is recorded at its first output slot: its first line (grp)
// and last line (gout); the second slot of a group that emits both a
// non-CO2 and a CO2 line is marked -1.  Every output slot is written.
__global__ __launch_bounds__(64) void k_shr_bwalk(Strides S, ShrinkWork W) {
  const int job = blockIdx.y;
  const int q = blockIdx.x;
  if (q >= W.nbat[job]) return;
  const bool lead = threadIdx.x == 0;
  const int64_t o = job * S.line;
  const int s = W.bat_s[o + q], t = W.bat_t[o + q], off = W.bat_off[o + q];
  ChainWalk cw(W, o, W.npass[job], s);
  const int lane = threadIdx.x & 63;
  auto mark = [&](int gs, int ge, int slot, int em) {
    if (em <= 0) return;
    if (lead) {
      W.grp[o + slot] = gs;
      W.gout[o + slot] = ge;
    }
    for (int k = 1 + lane; k < em; k += 64) W.grp[o + slot + k] = -1;
  };
  if (cw.e1 >= t) {
    mark(s, t, off, em_cnt(cw.cfl, s, t));
    return;
  }
  mark(s, cw.e1, off, cw.acc);
  for (;;) {
    int w;
    const int nx = cw.next(&w);
    if (nx >= t) break;
    mark(cw.cur + 1, nx, off + cw.acc, w);
    cw.acc += w;
    cw.cur = nx;
  }
  mark(cw.cur + 1, t, off + cw.acc, em_cnt(cw.cfl, cw.cur + 1, t));
}

// per output slot: the group recorded there writes its effective lines
// (strength-weighted centre and widths, sums in line order as SHRINK does,
// oprop.f90:3560-3631)
__global__ void k_shr_emit(const DevLayer *Ys, Strides S, ShrinkWork W, Tape9 T9) {
  const int job = blockIdx.y;
  const int qs = blockIdx.x * blockDim.x + threadIdx.x;
  if (qs >= T9.n[job]) return;
  const int64_t o = job * S.line;
  const int gs = W.grp[o + qs];
  if (gs < 0) return;
  const int ge = W.gout[o + qs];
  double sumv = 0., sums = 0., sumal = 0., sumad = 0.;
  double sumv2 = 0., sums2 = 0., sumal2 = 0., sumad2 = 0.;
  int64_t q = o + qs;
  if (W.cspp) {  // the group's coupled lines first, each as it stands, in line order
    for (int k = gs; k <= ge; k++) {
      const int fl = W.cfl[o + k];
      if (!(fl & 4)) continue;
      T9.vnu[q] = W.cx[o + k];
      T9.sp[q] = W.csp[o + k];
      T9.alfa[q] = W.cal[o + k];
      T9.epp[q] = W.cad[o + k];
      T9.spp[q] = W.cspp[o + k];
      T9.mol[q] = fl >> 3;
      q++;
    }
  }
  // the sums over the other lines (built twice so that files without
  // coupled lines run the plain loop)
  auto sum_lines = [&](auto coupled) {
    for (int k = gs; k <= ge; k++) {
      const int fl = W.cfl[o + k];
      double sp = W.csp[o + k];
      if constexpr (decltype(coupled)::value) sp = (fl & 4) ? 0. : sp;  // adds nothing
      if (!(fl & 1)) {
        sumv = sumv + W.cx[o + k] * sp;
        sums = sums + sp;
        sumal = sumal + sp * W.cal[o + k];
        sumad = sumad + sp * W.cad[o + k];
      } else {
        sumv2 = sumv2 + W.cx[o + k] * sp;
        sums2 = sums2 + sp;
        sumal2 = sumal2 + sp * W.cal[o + k];
        sumad2 = sumad2 + sp * W.cad[o + k];
      }
    }
  };
  if (W.cspp) sum_lines(std::true_type{});
  else sum_lines(std::false_type{});
  if (sums > 0.) {
    T9.vnu[q] = sumv / sums;
    T9.sp[q] = sums;
    T9.alfa[q] = sumal / sums;
    T9.epp[q] = sumad / sums;
    if (T9.spp) T9.spp[q] = 0.;
    T9.mol[q] = 0;
    q++;
  }
  if (sums2 > 0.) {
    T9.vnu[q] = sumv2 / sums2;
    T9.sp[q] = sums2;
    T9.alfa[q] = sumal2 / sums2;
    T9.epp[q] = sumad2 / sums2;
    if (T9.spp) T9.spp[q] = 0.;
    T9.mol[q] = 2;
  }
}

// VMAX of each TAPE9 block header: centre of its last line (oprop.f90:3272-3277)
__global__ void k_t9_vhi(Strides S, Tape9 T9) {
  const int job = blockIdx.y;
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= T9.nblk[job]) return;
  const int64_t ot = job * S.t9blk;
  T9.bvhi[ot + b] = T9.vnu[job * S.line + T9.bstart[ot + b] + T9.bcount[ot + b] - 1];
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
constexpr int CF4_S = CF4_T - 64;  // threads that scatter (waves 1-3)

// Broadcast lane j's value to the wave (j wave-uniform).
__device__ __forceinline__ double rl_f64(double v, int j) {
  int2 w = __builtin_bit_cast(int2, v);
  w.x = __builtin_amdgcn_readlane(w.x, j);
  w.y = __builtin_amdgcn_readlane(w.y, j);
  return __builtin_bit_cast(double, w);
}
__device__ __forceinline__ float rl_f32(float v, int j) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}
__device__ __forceinline__ int64_t rl_i64(int64_t v, int j) {
  int2 w = __builtin_bit_cast(int2, v);
  w.x = __builtin_amdgcn_readlane(w.x, j);
  w.y = __builtin_amdgcn_readlane(w.y, j);
  return __builtin_bit_cast(int64_t, w);
}
__device__ __forceinline__ double rl_t(double v, int j) { return rl_f64(v, j); }
__device__ __forceinline__ float rl_t(float v, int j) { return rl_f32(v, j); }

// T = double: the dbl build.  T = float (FP32 split): the far-wing shape is
// evaluated and accumulated in FP32; the line centre relative to the R4
// grid (xnui) and the grid positions stay FP64.
//
// The reference's value at distance xm from the centre (oprop.f90:3989-4006)
//   (ZVSQ <= ZSQBND ? SIV*(A3 + ZVSQ*B3) : SIL/(ALFLI2 + XMSQ)) - SUBTR,
//   ZVSQ = XMSQ/ALFVI2, SUBTR = F4BND (CO2: (2 - XMSQ/BNDSQ)*F4BND)
// is kept per line as the coefficients of one polynomial in XMSQ per branch:
//   near = nA + XMSQ*nB,  far = SIL/(ALFLI2 + XMSQ) + (XMSQ*c1 - c0),
// with c0 = SUBTR's constant, c1 its XMSQ slope, nA = SIV*A3 - c0,
// nB = SIV*B3/ALFVI2 + c1, and the branch test XMSQ <= thr = ZSQBND*ALFVI2.
template <typename T>
struct F4Par {
  double xnui;
  T sil, alfli2, nA, nB, c0, c1, thr;
  int jmin, jmax;
};

// 1/d for d > 0 in the normal range: the hardware estimate (relative error
// below 4.6e-8 on this chip) refined by one Newton step, which squares the
// error: below 2.2e-15 measured over 1e-8..1e8, far inside the far wing's
// share of the optical depth tolerance.
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp_nr(float d) { return __builtin_amdgcn_rcpf(d); }

// Value at distance xmd = XJJ - XNUI from the line centre.
template <typename T>
__device__ __forceinline__ T f4_at(const F4Par<T> &p, double xmd) {
  T xm = (T)xmd;
  T xmsq = xm * xm;
  // both branches evaluated and selected: the inner region covers a few
  // points per line, so a divergent branch would only add exec-mask work
  T near = fma(xmsq, p.nB, p.nA);
  T far = fma(p.sil, rcp_nr(p.alfli2 + xmsq), fma(xmsq, p.c1, -p.c0));
  return xmsq <= p.thr ? near : far;
}
// With the dispersive term of a coupled line: SPP != 0 adds XM DPTRAT
// F4FN, DPTRAT = SPP / (SP ALFAVI) (oprop.f90:3976-4004); cd = DPTRAT, 0
// for the other lines.  CPL: the line file has coupled lines.
template <bool CPL, typename T>
__device__ __forceinline__ T f4_val(const F4Par<T> &p, T cd, double xmd) {
  const T v = f4_at(p, xmd);
  if constexpr (!CPL) return v;
  return cd != (T)0. ? v + ((T)xmd * cd) * v : v;
}

// Five waves per SIMD (registers held to 96): with the 29 KB of LDS of the
// FP64 build that keeps five windows per compute unit resident, so the
// 1230 windows of a 50-layer column run in one round.
// CPL: the line file holds coupled lines; MIR: some window reaches below
// 50 cm-1, where lines at 0 < VNU <= 25 add their Van Vleck - Weisskopf
// mirror at -VNU (oprop.f90:4014-4024), a second entry after the line's
// own.  Separate builds of the kernel, so that other runs keep the plain
// code.
template <typename T, bool CPL, bool MIR>
__global__ __launch_bounds__(CF4_T) __attribute__((amdgpu_waves_per_eu(5)))
void k_convf4(Tape9 T9, const DevLayer *Ys, Strides S, const DevWin *wins, const DevSeg *segs,
              const int32_t *worder, double *r4out, unsigned long long *k3upd, uint64_t *amask,
              int64_t mwords, int32_t *wspan) {
  __shared__ T r4[R4ROW];
  constexpr int NPAR = MIR ? 128 : 64;  // entries a chunk of 64 lines can add
  __shared__ F4Par<T> par[2][NPAR];
  __shared__ T pcd[CPL ? 2 : 1][CPL ? NPAR : 1];  // DPTRAT of the entries in par
  __shared__ int s_nacc2[2], s_done2[2], s_b;
  const int tid = threadIdx.x;
  const int wi = worder[blockIdx.x];
  const DevWin W = wins[wi];
  const int job = W.job;
  const DevLayer &Y = Ys[job];
  const int64_t o = job * S.line;
  const int64_t ot = job * S.t9blk;
  const double dvr4 = Y.dvr4, bound4 = Y.bound4;
  const double recpi = 1. / (2. * asin(1.));
  const double zsqbnd = 64. * 64.;
  const double A3 = (1. + 2. * zsqbnd) / ((1. + zsqbnd) * (1. + zsqbnd));
  const double B3 = -1. / ((1. + zsqbnd) * (1. + zsqbnd));
  const double bndsq = bound4 * bound4;
  const double vnulst = W.v2r4 + bound4;
  const double vlo = W.v1r4 - bound4;
  const int nptr4 = W.nptr4;
  const int nl = T9.n[job];
  const int nb9 = T9.nblk[job];
  for (int j = tid; j < R4ROW; j += CF4_T) r4[j] = (T)0.;
  if (tid == 0) {
    s_b = nb9;
    s_nacc2[0] = 0;  // an empty chunk before the first
    s_done2[0] = 0;
  }
  __syncthreads();
  for (int b = tid; b < nb9; b += CF4_T)
    if (T9.bvhi[ot + b] >= vlo) atomicMin(&s_b, b);
  __syncthreads();
  int pos = (s_b < nb9) ? T9.bstart[ot + s_b] : nl;
  unsigned long long nupd = 0;
  // raw TAPE9 fields of the chunk wave 0 resolves next (prefetched)
  double f_sp = 0., f_al = 0., f_ad = 0., f_v = 0., f_spp = 0.;
  int f_mol = 0;
  if (tid < 64 && pos + tid < nl) {
    f_sp = T9.sp[o + pos + tid];
    if (CPL) f_spp = T9.spp[o + pos + tid];
    f_al = T9.alfa[o + pos + tid];
    f_ad = T9.epp[o + pos + tid];
    f_v = T9.vnu[o + pos + tid];
    f_mol = T9.mol[o + pos + tid];
  }
  // margin that sends lines within rounding of the rejection bound to the
  // exact sequential test
  const T lb_margin = sizeof(T) == 4 ? (T)1e-5 : (T)1e-12;
  // Decisions and scatter overlap: while waves 1-3 add the chunk decided
  // last (par[cur]) into R4, wave 0 decides the next chunk against R4 as it
  // stood before that scatter plus the last chunk's accepted lines
  // evaluated at its own line centres (in the order the scatter adds them).
  // Wave 0 reads R4 before the barrier that starts the scatter.
  int cur = 0;
  int npos = pos;
  for (;;) {
    const bool more = !s_done2[cur] && npos < nl;
    const int lane = tid & 63;
    const int i = npos + lane;
    const bool valid = more && i < nl;
    // wave 0's lines of the chunk it decides (unused by waves 1-3)
    F4Par<T> p, pm;  // the line, and its mirror (MIR)
    p.jmin = 1;
    p.jmax = 0;
    p.xnui = 0.;
    pm.jmin = 1;
    pm.jmax = 0;
    pm.xnui = 0.;
    bool has_m = false;
    T p_cd = (T)0.;
    T speak = (T)0., partial = (T)0., pkb = (T)0., f4bnd = (T)0.;
    double vnui = 0.;
    bool nz = false, bnd_ok = false, cpl = false;
    int jj = 1;
    if (tid < 64 && more) {
      // the in-order decisions are the window's critical path: wave 0 is
      // issued ahead of the scattering waves on its SIMD while it decides
      __builtin_amdgcn_s_setprio(3);
      if (valid) {
        double sp = f_sp;
        cpl = CPL && f_spp != 0.;  // a coupled line is never rejected
        nz = sp != 0. || cpl;
        double alfadi = f_ad, alfali = f_al;
        double zetai = alfali / (alfali + alfadi);
        int iz = fint(100. * zetai + ONEPL_D);
        iz = iz < 1 ? 1 : (iz > 101 ? 101 : iz);  // zeta in [0,1] keeps iz in range
        double zetdif = 100. * zetai - (double)(iz - 1);
        double alfavi = (c_avrat[iz - 1] + zetdif * (c_avrat[iz] - c_avrat[iz - 1])) * (alfali + alfadi);
        double ralfvi = 1. / alfavi;
        double sil = sp * recpi * alfali;
        double siv = (alfali * ralfvi) * sp * recpi * ralfvi;
        speak = (T)(A3 * fabs(siv));
        vnui = f_v;
        jj = fint((vnui - W.v1r4) / dvr4 + 1.);
        jj = jj < 1 ? 1 : (jj > nptr4 ? nptr4 : jj);
        p.xnui = vnui - W.v1r4;
        int jmin = fint((p.xnui - bound4) / dvr4 + 2.);
        int jmax = fint((p.xnui + bound4) / dvr4 + 1.);
        p.jmin = jmin < 1 ? 1 : jmin;
        p.jmax = jmax > nptr4 ? nptr4 : jmax;
        if (jmin > nptr4) p.jmax = 0;  // accepted but nothing to add (GO TO 60)
        double alfli2 = alfali * alfali;
        p.alfli2 = (T)alfli2;
        double sil4 = sil, f4b = sil / (alfli2 + bndsq), siva3 = siv * A3, sivb3 = siv * B3;
        bool co2 = f_mol == 2;
        if (Y.voigt) {  // oprop_voigt.f90:3875-3892, 3929-3931: zeta-dependent
                        // a_3, b_3; CO2 only through CHI (= 1)
          const int izx = iz - 1;  // 0..100
          const double a3x = c_va3[izx] + zetdif * (c_va3[izx + 1] - c_va3[izx]);
          const double b3x = c_vb3[izx] + zetdif * (c_vb3[izx + 1] - c_vb3[izx]);
          const double f4_64x = (a3x + b3x * zsqbnd);
          const double sivv = sp * ralfvi;
          speak = (T)(a3x * fabs(sivv));
          sil4 = sivv * f4_64x * (alfli2 + (alfavi * alfavi) * zsqbnd);
          f4b = sil4 / (alfli2 + bndsq);
          siva3 = sivv * a3x;
          sivb3 = sivv * b3x;
          co2 = false;
        }
        const double alfvi2 = alfavi * alfavi;
        const double c0 = co2 ? 2. * f4b : f4b, c1 = co2 ? f4b / bndsq : 0.;
        p.sil = (T)sil4;
        p.c0 = (T)c0;
        p.c1 = (T)c1;
        p.nA = (T)(siva3 - c0);
        p.nB = (T)(sivb3 / alfvi2 + c1);
        p.thr = (T)(zsqbnd * alfvi2);
        f4bnd = (T)f4b;
        // peak bound of the shape: near branch <= SIV*A3 (B3 < 0), far
        // branch <= SIL / (ALFLI2 + ZSQBND ALFVI2) (it applies beyond
        // ZSQBND halfwidths squared); margins cover rounding in T
        bnd_ok = sp > 0.;
        const double pkn = siva3;
        const double pkf = sil4 / (alfli2 + zsqbnd * alfvi2 * (1. - 1e-6));
        double pkd = (pkn > pkf ? pkn : pkf) * (1. + 1e-5);
        if (CPL && cpl) {
          // |XM| <= BOUND4: the factor 1 + XM DPTRAT lies within
          // 1 +- BOUND4 |DPTRAT|, and either sign of the shape may reach
          // max(peak, 2 F4BND) in magnitude
          const double dptrat = f_spp / (sp * alfavi);
          p_cd = (T)dptrat;
          const double mg = (pkd > 2. * f4b ? pkd : 2. * f4b) * (1. + bound4 * fabs(dptrat)) * (1. + 1e-5);
          pkd = mg;
          f4bnd = (T)(0.5 * mg);  // the lower bound takes -2 f4bnd
        }
        pkb = (T)pkd;
        partial = r4[jj];
        if (MIR && vnui > 0. && vnui <= 25. && jmin <= nptr4) {
          // label 50: the same shape at -VNU, SPP negated (not after GO TO 60)
          has_m = true;
          pm = p;
          pm.xnui = -vnui - W.v1r4;
          const int mn = fint((pm.xnui - bound4) / dvr4 + 2.);
          const int mx = fint((pm.xnui + bound4) / dvr4 + 1.);
          pm.jmin = mn < 1 ? 1 : mn;
          pm.jmax = mx > nptr4 ? nptr4 : mx;
          if (mn > nptr4) pm.jmax = 0;
        }
      }
    }
    __syncthreads();
    if (tid >= 64) {  // waves 1-3: the last decided chunk into R4
      const int ts = tid - 64;
      const int nacc = s_nacc2[cur];
      const double step = CF4_S * dvr4;
      // a line's parameters are read from LDS one line ahead, so the read
      // overlaps the previous line's points
      F4Par<T> qn = par[cur][0];
      for (int a = 0; a < nacc; a++) {
        const F4Par<T> q = qn;
        qn = par[cur][a + 1 < nacc ? a + 1 : a];
        const T qcd = CPL ? pcd[cur][a] : (T)0.;
        if (ts == 0) nupd += (unsigned long long)(q.jmax - q.jmin + 1);
        int j0 = q.jmin - 1;
        int r = j0 % CF4_S;
        int first = j0 - r + ts + (ts < r ? CF4_S : 0) + 1;
        // XJJ - XNUI advanced by a running sum: a line spans a few strides,
        // so the sum's rounding stays within a few ulp of XJJ
        double xmd = (double)(first - 1) * dvr4 - q.xnui;
        int j = first;
        if (CPL && qcd != (T)0.) {  // a coupled line
          for (; j <= q.jmax; j += CF4_S, xmd += step) r4[j] = r4[j] + f4_val<CPL>(q, qcd, xmd);
          continue;
        }
        if constexpr (sizeof(T) == 4) {
          // FP32 shapes: two points per step, two independent evaluations in
          // flight per thread (a thread's points are distinct, so the order
          // of additions to each point is unchanged).  The FP64 build is
          // bound by FP64 issue, where this only adds instructions.
          for (; j + CF4_S <= q.jmax; j += 2 * CF4_S) {
            const double x1 = xmd + step;
            const T v0 = f4_at(q, xmd), v1 = f4_at(q, x1);
            xmd = x1 + step;
            r4[j] = r4[j] + v0;
            r4[j + CF4_S] = r4[j + CF4_S] + v1;
          }
        }
        for (; j <= q.jmax; j += CF4_S, xmd += step) r4[j] = r4[j] + f4_at(q, xmd);
      }
    } else if (more) {
      const int nb = cur ^ 1;
      const double xjj_me = (double)(jj - 1) * dvr4;
      // the last chunk's accepted lines at this chunk's line centres
      const int nprev = s_nacc2[cur];
      F4Par<T> qn = par[cur][0];
      for (int a = 0; a < nprev; a++) {
        const F4Par<T> q = qn;
        qn = par[cur][a + 1 < nprev ? a + 1 : a];
        const T qcd = CPL ? pcd[cur][a] : (T)0.;
        if (valid && jj >= q.jmin && jj <= q.jmax) partial += f4_val<CPL>(q, qcd, xjj_me - q.xnui);
      }
      const bool cand = valid && nz;
      // Bounds of the running R4 at this line over the undecided earlier
      // lines of the chunk: each adds at least -2 F4BND (shape >= 0,
      // subtraction <= 2 F4BND) and at most its peak bound pk.  A line
      // rejected against the lower bound is rejected, and one accepted
      // against the upper bound accepted, whatever the earlier lines
      // decide.  Only lines up to the last undecided one are taken in order
      // (their exact running R4 is needed); the rest are accepted in bulk.
      T neg = cand ? (T)2. * f4bnd : (T)0.;
      T pk = cand ? pkb : (T)0.;
      if (MIR && has_m) {  // the mirror adds as much again
        neg = neg + neg;
        pk = pk + pk;
      }
      T pre = neg, preu = pk;
      for (int d = 1; d < 64; d <<= 1) {
        T t = __shfl_up(pre, d);
        T u = __shfl_up(preu, d);
        if (lane >= d) {
          pre += t;
          preu += u;
        }
      }
      const T below = pre - neg, above = preu - pk;
      const T slack = (T)Y.dptfc * (fabs(partial) + below + above) * lb_margin + fabs((T)Y.dptmn4) * lb_margin;
      const T thr_lb = (T)Y.dptmn4 + (T)Y.dptfc * (partial - below);
      const T thr_ub = (T)Y.dptmn4 + (T)Y.dptfc * (partial + above);
      const bool surely_rej = !(CPL && cpl) && speak < thr_lb - slack;
      const bool surely_acc = cand && ((CPL && cpl) || (bnd_ok && speak > thr_ub + slack));
      const uint64_t undm = __ballot(cand && !surely_rej && !surely_acc);
      const int lastund = undm ? 63 - __builtin_clzll(undm) : -1;
      const uint64_t head = lastund < 0 ? 0ull : (lastund == 63 ? ~0ull : ((1ull << (lastund + 1)) - 1));
      uint64_t todo = __ballot(cand && !surely_rej) & head;
      // a surely accepted line's decision is known without the running R4,
      // so its step does not wait on the previous line's contribution
      const uint64_t sacc_m = __ballot(surely_acc);
      const uint64_t past_m = __ballot(vnui >= vnulst);
      int nacc = 0;
      int done = 0;
      uint64_t am = 0ull;  // accepted lines with points in the window
      while (todo) {
        const int j = __builtin_ctzll(todo);
        todo &= todo - 1;
        int accj;
        if ((sacc_m >> j) & 1ull) {
          accj = 1;
        } else {
          bool acc = cand && ((CPL && cpl) || !(speak <= ((T)Y.dptmn4 + (T)Y.dptfc * partial)));
          accj = __builtin_amdgcn_readlane(acc ? 1 : 0, j);
        }
        const int stpj = accj && ((past_m >> j) & 1ull);
        if (stpj) {
          done = 1;
          break;
        }
        if (accj) {
          F4Par<T> q;
          q.xnui = rl_f64(p.xnui, j);
          q.sil = rl_t(p.sil, j);
          q.alfli2 = rl_t(p.alfli2, j);
          q.nA = rl_t(p.nA, j);
          q.nB = rl_t(p.nB, j);
          q.c0 = rl_t(p.c0, j);
          q.c1 = rl_t(p.c1, j);
          q.thr = rl_t(p.thr, j);
          const T qcd = CPL ? rl_t(p_cd, j) : (T)0.;
          q.jmin = __builtin_amdgcn_readlane(p.jmin, j);
          q.jmax = __builtin_amdgcn_readlane(p.jmax, j);
          if (lane > j && jj >= q.jmin && jj <= q.jmax) partial += f4_val<CPL>(q, qcd, xjj_me - q.xnui);
          if (lane == j && p.jmax >= p.jmin) {
            par[nb][nacc] = p;
            if (CPL) pcd[nb][nacc] = p_cd;
          }
          if (q.jmax >= q.jmin) {
            nacc++;
            am |= 1ull << j;
          }
          if (MIR && __builtin_amdgcn_readlane(has_m ? 1 : 0, j)) {
            q.xnui = rl_f64(pm.xnui, j);
            q.jmin = __builtin_amdgcn_readlane(pm.jmin, j);
            q.jmax = __builtin_amdgcn_readlane(pm.jmax, j);
            if (lane > j && jj >= q.jmin && jj <= q.jmax) partial += f4_val<CPL>(q, -qcd, xjj_me - q.xnui);
            if (q.jmax >= q.jmin) {
              if (lane == j) {
                par[nb][nacc] = pm;
                if (CPL) pcd[nb][nacc] = -p_cd;
              }
              nacc++;
            }
          }
        }
      }
      if (!done) {  // lanes after the last undecided one: accepted in bulk
        uint64_t rest = __ballot(surely_acc) & ~head;
        const uint64_t stpm = __ballot(surely_acc && vnui >= vnulst) & rest;
        if (stpm) {  // the first accepted line past VNULST ends the window
          const int js = __builtin_ctzll(stpm);
          rest &= (1ull << js) - 1;
          done = 1;
        }
        const uint64_t addm = rest & __ballot(p.jmax >= p.jmin);
        const uint64_t mirm = MIR ? rest & __ballot(has_m && pm.jmax >= pm.jmin) : 0ull;
        const uint64_t below = lane == 0 ? 0ull : ((1ull << lane) - 1);
        const int slot = nacc + __builtin_popcountll(addm & below) + __builtin_popcountll(mirm & below);
        if ((addm >> lane) & 1) {
          par[nb][slot] = p;
          if (CPL) pcd[nb][slot] = p_cd;
        }
        if (MIR && ((mirm >> lane) & 1)) {  // after the line's own entry
          const int sm = slot + (int)((addm >> lane) & 1);
          par[nb][sm] = pm;
          if (CPL) pcd[nb][sm] = -p_cd;
        }
        nacc += __builtin_popcountll(addm) + __builtin_popcountll(mirm);
        am |= addm;
      }
      if (lane == 0) {
        s_nacc2[nb] = nacc;
        s_done2[nb] = done;
        // the tangent pass reads the accepted lines back (chunk index < nl / 64)
        if (amask) amask[(size_t)wi * mwords + (npos - pos) / 64] = am;
      }
      // prefetch the next chunk while the group scatters this one
      const int in = npos + 64 + lane;
      if (!done && in < nl) {
        f_sp = T9.sp[o + in];
        if (CPL) f_spp = T9.spp[o + in];
        f_al = T9.alfa[o + in];
        f_ad = T9.epp[o + in];
        f_v = T9.vnu[o + in];
        f_mol = T9.mol[o + in];
      }
      __builtin_amdgcn_s_setprio(0);
    }
    __syncthreads();
    if (!more) break;
    cur ^= 1;
    npos += 64;
  }
  if (tid == 64 && nupd) atomicAdd(&k3upd[job], nupd);
  if (wspan && tid == 0) {
    wspan[2 * wi] = pos;
    wspan[2 * wi + 1] = (npos - pos) / 64;
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
      if (j >= sg.i1 && j <= sg.i2) r4[j] = (T)((double)r4[j] * (sg.radvi + (double)(j - sg.i1) * sg.rdel));
    }
    __syncthreads();
  }
  double *row = r4out + (size_t)wi * R4ROW;
  for (int j = tid; j < R4ROW; j += CF4_T) row[j] = (j >= 1 && j <= nptr4) ? (double)r4[j] : 0.;
}

// ---------------------------------------------------------------------
// LNCOR1's corrections of TAPE3 record i (oprop.f90:995-1193) except the
// two tests that depend on the panel at which its block is read: shifted
// centre, 1/ALFV, corrected strength and the zeta column.  False for a
// record LNCOR1 skips.
struct NearCore {
  double vnu, recalf, sui, sp, zetai, sppsp;
  int izeta;
  bool norej;  // IFLG != 0: LNCOR1 never rejects the line
};
__device__ __forceinline__ bool lncor1_core(const DevLines &L, const DevLayer &Y, int i, NearCore &c) {
  const int mc = L.mol[i];
  const int m = mc % 100, iso = (mc % 1000) / 100;
  if (m < 1 || m > 7 || iso < 1 || iso > c_isomax[m - 1]) return false;
  double sui = (m <= Y.nmol) ? (double)L.sp[i] * Y.wk[m - 1] : 0.;
  if (Y.jrad == 1) sui = sui * d_radfn(L.vnu[i], Y.xkt);
  if (sui == 0.) return false;
  const int ci = L.cidx ? L.cidx[i] : -1;
  double alfa0 = L.alfa[i], hwhm = L.hwhm[i];
  if (m == 7) alfa0 = (alfa0 - 0.21 * hwhm) / (1.0 - 0.21);  // RDLIN (oprop.f90:818-827)
  const double vnu = L.vnu[i] + Y.rhorat * (double)L.pshift[i];
  const double tmpcor = pow(Y.tratio, (double)L.tmpalf[i]);
  const double alfa0i = alfa0 * tmpcor;
  const double hwhmsi = hwhm * tmpcor;
  double alfl = alfa0i * (Y.rhorat - Y.rhoslf[m - 1]) + hwhmsi * Y.rhoslf[m - 1];
  if (ci >= 0 && (ci & 1)) {  // IFLG 3
    const LcCoef lc = lc_coef(L, Y, i, 0);
    alfl = alfl * (1.0 - lc.g1 * Y.pavp0 - lc.g2 * Y.pavp2);
  }
  const double alfad = vnu * Y.alfd1[m - 1][iso - 1];
  const double zeta = alfl / (alfl + alfad);
  const double fzeta = 100. * zeta;
  int iz = fint(fzeta + ONEPL_D);
  iz = iz < 1 ? 1 : (iz > 101 ? 101 : iz);  // zeta in [0,1] keeps iz in range
  const double zetdif = fzeta - (double)(iz - 1);
  double alfv = (c_avrat[iz - 1] + zetdif * (c_avrat[iz] - c_avrat[iz - 1])) * (alfl + alfad);
  if (alfv < Y.dv) alfv = Y.dv;
  if (alfv > Y.alfmax) alfv = Y.alfmax;
  double epp = L.epp[i];
  if (epp <= -1.001) epp = fabs(epp);
  if (epp > -0.999) sui = sui * Y.scor[m - 1][iso - 1] * exp(-epp * Y.betacr_l) * (1. + exp(-vnu / Y.xkt));
  c.vnu = vnu;
  c.recalf = 1. / alfv;
  c.sui = sui;
  double yi = 0., gi = 0.;
  if (ci >= 0 && !(ci & 1)) {  // IFLG 1
    const LcCoef lc = lc_coef(L, Y, i, 0);
    yi = lc.yi;
    gi = lc.gi;
  }
  c.sp = sui * (1. + gi * Y.pavp2);   // oprop.f90:1147-1149
  c.sppsp = (sui * yi * Y.pavp0) / c.sp;
  c.norej = ci != -1;
  c.zetai = zeta;
  c.izeta = iz;
  return true;
}

// LNCOR1 for the near lines: the fields every rejection pass reads, the
// first panel whose R1 buffer can hold the line (CNVFNV JMAX1 <= MAX1), and
// the strength and zeta column the scatter records of the surviving lines
// read.
__global__ void k_lncor1(DevLines L, const DevLayer *Ys, Strides S, const DevPanel *pans,
                         NearHot *out, NearCold *cold) {
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= Y.nn) return;
  const DevPanel *pan = pans + Y.pan0;
  const int i = Y.nlo + t;
  NearHot o{L.vnu[i], 0., 0., 0, 0};
  NearCore c;
  if (lncor1_core(L, Y, i, c)) {
    o.vnu = c.vnu;
    o.recalf = c.recalf;
    // SPEAK = SUI * RECALF; lines with IFLG != 0 pass every rejection
    o.speak = c.norej ? HUGE_VAL : c.sui * c.recalf;
    o.keep = 1;
    // first fit: INT(ZINT_p + BHWDXF + 1.5) <= MAX1
    const double zslope = c.recalf * (Y.dv / 0.002);
    const double bhwdxf = (4. / 0.002) / zslope;
    const double z0 = (c.vnu - Y.vft0) / Y.dv;
    int p = (int)floor((z0 + bhwdxf + 0.5 - (double)Y.max1) / 2400.) + 1;
    if (p < 0) p = 0;
    if (p > Y.npanel) p = Y.npanel;
    auto fits = [&](int q) {
      double zint = (c.vnu - pan[q].vft) / Y.dv;
      return fint(zint + bhwdxf + 1.5) <= Y.max1;
    };
    while (p > 0 && fits(p - 1)) p--;
    while (p < Y.npanel && !fits(p)) p++;
    o.ff = p;  // == npanel: never processed before the last panel
    cold[job * S.line + t] = NearCold{c.sp, c.zetai, c.sppsp, c.izeta, 0};
  }
  out[job * S.line + t] = o;
}

// LNCOR1's VFT test and R4 rejection (oprop.f90:1120, 1153-1177) of one
// near line when its block is read at panel pr.
__device__ __forceinline__ int reject_one(const NearHot &o, int pr, const DevLayer &Y, const DevPanel *pans,
                                          const DevWin *wins, const double *r4) {
  int keep = o.keep;
  if (!keep) return 0;
  int pq = pr < Y.npanel ? pr : Y.npanel - 1;
  const DevPanel &pp = pans[Y.pan0 + pq];
  double alfv = 1. / o.recalf;
  if (64. * alfv + o.vnu < pp.vft) return 0;
  int w = Y.win0 + pp.w;
  DevWin W = wins[w];
  int jj = fint((o.vnu - W.v1r4) / Y.dvr4 + 1.);
  jj = jj < 1 ? 1 : (jj > W.nptr4 ? W.nptr4 : jj);
  double r = r4[(size_t)w * R4ROW + jj];
  if (o.speak <= (Y.dptmn_hirac + Y.dptfc * r)) return 0;
  return 1;
}

// One step of the block-read fixpoint: with the current guess of the panel
// at which each block is read, apply LNCOR1's VFT test and R4 rejection
// and reduce the first-fit panel of the kept lines per block.  Only blocks
// whose read panel changed in the last step (dirty) are redone; the others
// keep their flags and block maxima.  The per-block maximum is reduced
// across the wave before the atomic.
__global__ __launch_bounds__(256) void k_reject(const NearHot *nl, DevLines L, const DevLayer *Ys, Strides S,
                                                const int *pread, const DevPanel *pans, const DevWin *wins,
                                                const double *r4, int *keep_out, int *blkmax, const int *dirty,
                                                const int *skip) {
  if (skip && *skip == 0) return;  // the previous pass moved no read panel: nothing is dirty
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = t < Y.nn;
  const int k = in ? L.blk[Y.nlo + t] - Y.blk0 : -1;
  const bool act = in && dirty[job * S.blk + k] != 0;
  int val = -1;
  if (act) {
    const NearHot o = nl[job * S.line + t];
    const int keep = reject_one(o, pread[job * S.blk + k], Y, pans, wins, r4);
    keep_out[job * S.line + t] = keep;
    if (keep) val = o.ff;
  }
  const int lane = threadIdx.x & 63;
  uint64_t pend = __ballot(val >= 0);
  while (pend) {
    const int leader = __builtin_ctzll(pend);
    const int kl = __builtin_amdgcn_readlane(k, leader);
    const bool mine = val >= 0 && k == kl;
    int m = mine ? val : -1;
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if (lane == leader) atomicMax(&blkmax[job * S.blk + kl], m);
    pend &= ~__ballot(mine);
  }
}

// The same decisions in RDLIN's order: block k is read at the exclusive
// prefix maximum of the kept first-fit panels of blocks 0..k-1, and its
// lines are kept or rejected at that panel before block k+1 is considered.
// One wave per job walks its blocks in file order (at most 250 lines a
// block, four per lane).  This is the recurrence the fixpoint solves; it
// runs when the fixpoint has not settled within its pass budget, and
// leaves pread, blkmax and keep exactly as a settled fixpoint would.
__global__ __launch_bounds__(64) void k_reject_seq(const NearHot *nl, DevLines L, const DevLayer *Ys, Strides S,
                                                   int *pread, const DevPanel *pans, const DevWin *wins,
                                                   const double *r4, int *keep_out, int *blkmax) {
  const int job = blockIdx.x;
  const DevLayer &Y = Ys[job];
  const int lane = threadIdx.x;
  int run = 0;
  for (int k = 0; k < Y.nbk; k++) {
    const int b = Y.blk0 + k;
    const int t0 = max(L.bstart[b] - Y.nlo, 0);
    const int t1 = min(L.bstart[b + 1] - Y.nlo, Y.nn);
    int m = 0;
    for (int t = t0 + lane; t < t1; t += 64) {
      const NearHot o = nl[job * S.line + t];
      const int keep = reject_one(o, run, Y, pans, wins, r4);
      keep_out[job * S.line + t] = keep;
      if (keep) m = max(m, o.ff);
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off));
    if (lane == 0) {
      pread[job * S.blk + k] = run;
      blkmax[job * S.blk + k] = m;
    }
    run = max(run, m);
  }
}

// Exclusive prefix max over blocks: panel at which each block is read.
// Raises *changed when the guess moved and marks the blocks whose read
// panel moved dirty (their block maximum is reset for the next pass).
__global__ __launch_bounds__(256) void k_pread_scan(const DevLayer *Ys, Strides S, int *blkmax,
                                                    int *pread, int *changed, int *dirty, const int *skip) {
  if (skip && *skip == 0) return;  // as in k_reject: the pass is a no-op
  __shared__ int buf[256];
  __shared__ int carry;
  const int job = blockIdx.x;
  const int nbk = Ys[job].nbk;
  const int64_t ob = job * S.blk;
  int tid = threadIdx.x;
  if (tid == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nbk; base += 256) {
    int k = base + tid;
    int v = (k < nbk) ? blkmax[ob + k] : 0;
    buf[tid] = v;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
      int t = tid >= off ? buf[tid - off] : 0;
      __syncthreads();
      buf[tid] = max(buf[tid], t);
      __syncthreads();
    }
    int excl = max(carry, tid > 0 ? buf[tid - 1] : 0);
    __syncthreads();
    if (k < nbk) {
      const int moved = pread[ob + k] != excl;
      if (moved) {
        *changed = 1;
        blkmax[ob + k] = 0;
      }
      dirty[ob + k] = moved;
      pread[ob + k] = excl;
    }
    __syncthreads();
    if (tid == 0) carry = max(carry, buf[255]);
    __syncthreads();
  }
}

// Lines that survive the rejection tests per near block, and their
// exclusive prefix per job over the blocks (bkp) with the total at kidx[nn].
__global__ __launch_bounds__(256) void k_keep_blk(const DevLayer *Ys, DevLines L, Strides S, const int *keep,
                                                  int32_t *bkc) {
  __shared__ int buf[256];
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  const int k = blockIdx.x;
  if (k >= Y.nbk) return;
  const int b = Y.blk0 + k;
  int s = L.bstart[b] - Y.nlo, e = L.bstart[b + 1] - Y.nlo;
  if (s < 0) s = 0;
  if (e > Y.nn) e = Y.nn;
  const int t = s + (int)threadIdx.x;
  int tot;
  block_scan_excl<256>(t < e ? keep[job * S.line + t] : 0, buf, &tot);
  if (threadIdx.x == 0) bkc[job * S.blk + k] = tot;
}
__global__ __launch_bounds__(256) void k_keep_scan(const DevLayer *Ys, Strides S, const int32_t *bkc, int32_t *bkp,
                                                   int32_t *kidx) {
  __shared__ int buf[256];
  __shared__ int carry;
  const int job = blockIdx.x;
  const int nbk = Ys[job].nbk;
  const int64_t o = job * S.blk;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nbk; base += 256) {
    const int k = base + (int)threadIdx.x;
    int tot;
    const int ex = block_scan_excl<256>(k < nbk ? bkc[o + k] : 0, buf, &tot);
    if (k < nbk) bkp[o + k] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) kidx[job * S.line + Ys[job].nn] = carry;
}

// Processing panel of each kept line: the panel at which its block is read,
// raised by the first-fit panels of kept lines before it in the block
// (CNVFNV stops at the first line that does not fit, oprop.f90:1287-1292).
// Then the scatter record of CNVFNV (oprop.f90:1262-1305).  The records of
// the kept lines are stored contiguously in file order, at their kept-line
// index (kidx, written for every near record).  With k1acc, the
// near-wing point updates and convolved lines per job (sum of JMAX1-JMIN1+1
// and count over kept lines; LINCNT of the TAPE6 table) are added up too.
__global__ __launch_bounds__(256) void k_scatter_rec(const NearHot *nh, const NearCold *nc, const int *keep, DevLines L,
                                                     const DevLayer *Ys, Strides S, const int *pread,
                                                     const DevPanel *pans, const int32_t *bkp, int32_t *kidx,
                                                     Recs recs, unsigned long long *k1acc,
                                                     unsigned long long *lines) {
  __shared__ int buf[256];
  const int job = blockIdx.y;
  const DevLayer &Y = Ys[job];
  const int k = blockIdx.x;  // near block
  if (k >= Y.nbk) return;
  const int b = Y.blk0 + k;
  int s = L.bstart[b] - Y.nlo, e = L.bstart[b + 1] - Y.nlo;
  if (s < 0) s = 0;
  if (e > Y.nn) e = Y.nn;
  const int tid = threadIdx.x;
  const int t = s + tid;
  const bool valid = t < e;
  const int64_t o = job * S.line;
  const int kp = valid ? keep[o + t] : 0;
  NearHot h{};
  if (kp) h = nh[o + t];
  int nkb;
  const int64_t ci = o + bkp[job * S.blk + k] + block_scan_excl<256>(kp, buf, &nkb);  // record slot
  if (valid) kidx[o + t] = (int32_t)(ci - o);
  buf[tid] = kp ? h.ff : 0;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int q = tid >= off ? buf[tid - off] : 0;
    __syncthreads();
    buf[tid] = max(buf[tid], q);
    __syncthreads();
  }
  int np = 0;
  if (kp) {
    const int P = max(pread[job * S.blk + k], buf[tid]);
    SRec1 r1{};
    SRecA r2{}, r3{};
    SRecV rv{};
    rv.izm = 1;
    if (P < Y.npanel) {
      const NearCold n = nc[o + t];
      double vft = pans[Y.pan0 + P].vft;
      double depthi = n.sp * h.recalf;
      int izm = n.izeta;
      double zetdif = 100. * n.zetai - (double)(izm - 1);
      if (Y.voigt) {  // oprop_voigt.f90:1273-1276: DEPTHI and the zeta column
        r1.str = r2.str = r3.str = depthi;
        rv.zdif = zetdif;
        rv.izm = izm < 1 ? 1 : (izm > NZETA ? NZETA : izm);
      } else {
        r1.str = depthi * (c_cf1[izm - 1] + zetdif * (c_cf1[izm] - c_cf1[izm - 1]));
        r2.str = depthi * (c_cf2[izm - 1] + zetdif * (c_cf2[izm] - c_cf2[izm - 1]));
        r3.str = depthi * (c_cf3[izm - 1] + zetdif * (c_cf3[izm] - c_cf3[izm - 1]));
        r1.strd = depthi * (c_cgauss[izm - 1] + zetdif * (c_cgauss[izm] - c_cgauss[izm - 1]));
        r1.strver = depthi * (c_cer[izm - 1] + zetdif * (c_cer[izm] - c_cer[izm - 1]));
      }
      double zslope = h.recalf * (Y.dv / 0.002);
      double zint = (h.vnu - vft) / Y.dv;
      double bhwdxf = (4. / 0.002) / zslope;
      int jmax1 = fint(zint + bhwdxf + 1.5);
      int jmin1 = fint(zint - bhwdxf + 1.5);
      double conf2 = Y.dv / Y.dvr2, conf3 = Y.dv / Y.dvr3;
      double rshft = zint < 0.0 ? -0.5 : 0.5;
      int j2shft = fint(zint * (1. - conf2) + rshft);
      int j3shft = fint(zint * (1. - conf3) + rshft);
      int jmin2 = jmin1 - j2shft, jmin3 = jmin1 - j3shft;
      r1.zfl = ((double)(jmin1 - 2) - zint) * zslope;
      r2.zfl = ((double)(jmin2 - 2) - zint * conf2) * zslope;
      r3.zfl = ((double)(jmin3 - 2) - zint * conf3) * zslope;
      r1.zslope = r2.zslope = r3.zslope = zslope;
      // global fine index of JMIN1 and its R2/R3 counterparts (G1 < 2^31 for
      // any layer the planner accepts)
      const int32_t g1lo = jmin1 - 1 + 2400 * P;
      r1.glo = g1lo;
      r2.glo = g1lo - (j2shft + 1800 * P);
      r3.glo = g1lo - (j3shft + 2250 * P);
      r1.npts = r2.npts = r3.npts = np = jmax1 - jmin1 + 1;
      if (n.sppsp != 0. && np > 0) {  // line coupling: the dispersive term
        r1.npts = r2.npts = r3.npts = np | NPTS_CPL;
        recs.cpl[ci] = n.sppsp;
      }
    }
    recs.r1[ci] = r1;
    recs.r2[ci] = r2;
    recs.r3[ci] = r3;
    if (Y.voigt) recs.v[ci] = rv;
  }
  if (k1acc) {  // TAPE6 counters (statistics passes only)
    unsigned long long v = np > 0 ? (unsigned long long)np : 0ull, c = np > 0 ? 1ull : 0ull;
    for (int off = 32; off > 0; off >>= 1) {
      v += __shfl_down(v, off);
      c += __shfl_down(c, off);
    }
    if ((tid & 63) == 0 && v) {
      atomicAdd(&k1acc[job], v);
      atomicAdd(&lines[job], c);
    }
  }
}

// Plain device copy (the calibration kernel for HBM traffic counters).
__global__ void k_copy_f64(double *dst, const double *src, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// Tile table of one grid level: tile -> (job, local tile).
struct TileMap {
  const int32_t *job, *loc;
  int32_t *lo, *hi;  // line range per tile
};

// Line range [lo, hi) whose scatter extent may touch each tile of grid G.
// The range is found over the near records (raw centres, in file order)
// and handed out in kept-line indices (kidx: kept lines before each
// record), which is how the scatter records are stored; without kidx, in
// record indices.
__global__ void k_tile_ranges(const double *vraw, const DevLayer *Ys, TileMap M, int ntiles,
                              int tile, int G, const double *margin, const int32_t *kidx, Strides S) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntiles) return;
  const int job = M.job[t];
  const DevLayer &Y = Ys[job];
  const double dvg = G == 1 ? Y.dv : (G == 2 ? Y.dvr2 : Y.dvr3);
  const double *v = vraw + Y.nlo;
  const int nn = Y.nn;
  double a = Y.vft0 + (double)M.loc[t] * tile * dvg - margin[job * 3 + G - 1];
  double b = Y.vft0 + (double)(M.loc[t] + 1) * tile * dvg + margin[job * 3 + G - 1];
  int l = 0, h = nn;
  while (l < h) {
    int m = (l + h) >> 1;
    if (v[m] < a) l = m + 1;
    else h = m;
  }
  int l2 = l, h2 = nn;
  while (l2 < h2) {
    int m = (l2 + h2) >> 1;
    if (v[m] <= b) l2 = m + 1;
    else h2 = m;
  }
  if (!kidx) {
    M.lo[t] = l;
    M.hi[t] = l2;
    return;
  }
  // (no near records: kidx is not written)
  M.lo[t] = nn > 0 ? kidx[job * S.line + l] : 0;
  M.hi[t] = nn > 0 ? kidx[job * S.line + l2] : 0;
}

// XINT of a coarse array into R3 point I of the panel that owns it
// (oprop.f90:2219-2230); A is the 1-based coarse array.
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


// The fields of a scatter record that grid G needs, one line per lane.
template <int G, bool V>
struct TileLine {
  int64_t glo;
  double zfl, zslope, str, strd, strver;
  int npts, izm;
  __device__ __forceinline__ void load(const Recs &R, int64_t base, int li, int hi) {
    strd = strver = 0.;
    izm = 1;
    if (li < hi) {
      const int64_t i = base + li;
      if (G == 1) {
        const SRec1 r = R.r1[i];
        npts = r.npts;
        glo = r.glo;
        zfl = r.zfl;
        zslope = r.zslope;
        str = r.str;
        if (!V) {
          strd = r.strd;
          strver = r.strver;
        }
      } else {
        const SRecA r = (G == 2 ? R.r2 : R.r3)[i];
        npts = r.npts;
        glo = r.glo;
        zfl = r.zfl;
        zslope = r.zslope;
        str = r.str;
      }
      if (V) {  // ZETDIF and IZM of the exact-Voigt build
        const SRecV v = R.v[i];
        strd = v.zdif;
        izm = v.izm;
      }
    } else {
      npts = 0;
      glo = 0;
      zfl = zslope = str = 0.;
    }
  }
};

// Threads per tile workgroup: the tables and the accumulator in LDS are
// shared by TILE_T / 64 waves.
#ifndef LBLX_TILE_T
#define LBLX_TILE_T 1024
#endif
constexpr int TILE_T = LBLX_TILE_T;

// Near-line scatter into one tile of grid G (3: DVR3, 2: DVR2, 1: DV) with
// the PANEL combine in the epilogue.
//   G = 3: R3 = lines(F3) + XINT(R4) + XINT(ABSRB)            -> r3
//   G = 2: R2 = lines(F2) + Lagrange4(R3)                       -> r2
//   G = 1: OD = (lines(F1,FG,XVER) + Lagrange4(R2)) * radiation -> od
//   V (exact-Voigt build, oprop_voigt.f90:1305-1336): the line shape is
//   DEPTHI ((1 - A) F_G(IZ, IZM) + A F_G(IZ, IZM + 1)) from the zeta tables
//   in vtab (read through L2: 101 columns per table do not fit LDS), with
//   no FG or XVER term.
//   LACC: type of the LDS accumulator (ACC, or double under FP32 shapes
//   where LDS float atomics stall).
//   CPL: the line file has coupled lines (their records carry NPTS_CPL).
template <int G, typename OUT, typename ACC, bool V, typename LACC, bool CPL>
__global__ __launch_bounds__(TILE_T) void k_tile(Recs recs, TileMap M, int tile, int ch,
                                              const double *tabs, const ACC *vtab, const DevLayer *Ys,
                                              Strides S, const DevPanel *pans, const DevWin *wins,
                                              const double *r4, const double *absrb,
                                              const DevSeg *segs, const double *rin, double *rout,
                                              OUT *od) {
  extern __shared__ double sm_d[];
  // [-1, tile): the slot before the tile holds the grid point below VFT
  // for RSYM on the first tile (R(0) of the reference's RR arrays)
  LACC *acc = reinterpret_cast<LACC *>(sm_d) + 1;
  ACC *t1 = reinterpret_cast<ACC *>(acc + tile);     // F_G table
  ACC *tg = t1 + NFMX;              // FG   (G = 1)
  ACC *tv = tg + NFMX;              // XVER (G = 1)
  const int tid = threadIdx.x;
  const int job = M.job[blockIdx.x];
  const DevLayer &Y = Ys[job];
  const int64_t rbase = job * S.line;
  for (int k = tid - 1; k < tile; k += TILE_T) acc[k] = (LACC)0.;
  if (!V) {
    const double *src = tabs + (size_t)(G == 1 ? 0 : (G == 2 ? 1 : 2)) * NFMX;
    for (int k = tid; k < NFMX; k += TILE_T) t1[k] = (ACC)src[k];
    if (G == 1) {
      for (int k = tid; k < NFMX; k += TILE_T) {
        tg[k] = (ACC)tabs[3 * NFMX + k];
        tv[k] = (ACC)tabs[4 * NFMX + k];
      }
    }
  }
  const ACC *vt = vtab + (size_t)(G == 1 ? 0 : (G == 2 ? 1 : 2)) * NZETA * NFMX;
  __syncthreads();
  const int64_t T0 = (int64_t)M.loc[blockIdx.x] * tile;
  const bool fold = T0 == 0 && Y.vft0 <= 0.;  // RSYM reads the point below VFT too
  const int64_t tlo = fold ? -1 : T0;
  const int lo = M.lo[blockIdx.x], hi = M.hi[blockIdx.x];
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  // Each wave takes chunks of CH consecutive lines: lane i < CH loads the
  // record of line c0 + i with vector loads, then the wave walks the chunk
  // line by line, broadcasting each line's fields with readlane.  The next
  // chunk's records are loaded before the current chunk is scattered.
  // The chunk length ch (1..64 lines; a workgroup step takes TILE_T / 64
  // chunks) is a tuning parameter: chunks of 16-32 spread a tile's lines
  // over its waves more evenly (16 measured best for the FP64 R1 tile, 32
  // for the others).
  const int CH = ch, CSTEP = TILE_T / 64 * CH;
  TileLine<G, V> cur, nxt;
  int c0 = lo + wave * CH;
  if (c0 < hi) cur.load(recs, rbase, c0 + lane, lane < CH ? hi : c0);
  for (; c0 < hi; c0 += CSTEP) {
    const int c1 = c0 + CSTEP;
    if (c1 < hi) nxt.load(recs, rbase, c1 + lane, lane < CH ? hi : c1);
    const int nj = hi - c0 < CH ? hi - c0 : CH;
    for (int j = 0; j < nj; j++) {
      const int nraw = __builtin_amdgcn_readlane(cur.npts, j);
      const int npts = CPL ? nraw & ~NPTS_CPL : nraw;
      if (npts <= 0) continue;
      const bool cp = CPL && (nraw & NPTS_CPL) != 0;
      const int64_t glo = rl_i64(cur.glo, j);
      const double zfl = rl_f64(cur.zfl, j), zslope = rl_f64(cur.zslope, j);
      const ACC str = (ACC)rl_f64(cur.str, j);
      const ACC strd = (G == 1 && !V) ? (ACC)rl_f64(cur.strd, j) : (ACC)0.;
      const ACC strver = (G == 1 && !V) ? (ACC)rl_f64(cur.strver, j) : (ACC)0.;
      const ACC za = V ? (ACC)rl_f64(cur.strd, j) : (ACC)0.;
      const ACC *ca = vt;
      if (V) ca = vt + (size_t)(__builtin_amdgcn_readlane(cur.izm, j) - 1) * NFMX;
      int64_t a = glo > tlo ? glo : tlo;
      int64_t b = glo + npts - 1;
      if (b > T0 + tile - 1) b = T0 + tile - 1;
      // tile-local point index kl; the line's point offset is kl + d0
      // (|d0| is at most the tile width or the line's point count)
      const int kb = (int)(b - T0), d0 = (int)(T0 - glo);
      // the point loop, built twice so that uncoupled lines run it plain
      auto points = [&](auto coupled) {
        for (int kl = (int)(a - T0) + lane; kl <= kb; kl += 64) {
          const int off = kl + d0;
          double zf = zfl + (double)(off + 1) * zslope;
          int iz = fint(fabs(zf) + 1.5) - 1;
          iz = iz > NFMX - 1 ? NFMX - 1 : iz;  // |zf| <= 2500 by construction
          ACC v;
          if (V) v = str * ((((ACC)1. - za) * ca[iz]) + (za * ca[iz + NFMX]));
          else if (G == 1) v = fma(str, t1[iz], fma(strd, tg[iz], strver * tv[iz]));
          else v = str * t1[iz];
          if constexpr (decltype(coupled)::value) {
            // line coupling: CLC * SPPSP * ZF, CLC = 4, 16, 64 halfwidths
            // over 2000 table steps for R1, R2, R3 (oprop.f90:1259-1261)
            const double cz = (G == 1 ? 4. : (G == 2 ? 16. : 64.)) / 2000. * recs.cpl[rbase + c0 + j];
            v = v + v * (ACC)(cz * zf);
          }
          atomicAdd(&acc[kl], (LACC)v);
        }
      };
      if (cp) points(std::true_type{});
      else points(std::false_type{});
    }
    cur = nxt;
  }
  __syncthreads();
  if (fold) {
    // RSYM (oprop.f90:501-505, 2145-2200): panel 0 starts at or below 0
    // cm-1, so the lines' values at negative wavenumbers and at their
    // mirror images are folded onto each other before the coarser grid and
    // the far wing are added.  Point R(j) = acc[j - 1] (j >= 0); IP is the first
    // point at or above 0, pair n = (I, K) = (IP + n, IP - 1 - n) for
    // I <= 2 IP - 1 (at most 34 pairs: VFT >= -32 DV), and each update
    // reads the values before the fold, so wave 0 reads, then writes.
    const double dvg = G == 1 ? Y.dv : (G == 2 ? Y.dvr2 : Y.dvr3);
    const int ip = fint(-Y.vft0 / dvg + 1. - .000001) + 1;
    double pp = ((double)(ip - 1) + Y.vft0 / dvg) * 2.;
    const bool low = pp <= 1.;
    if (!low) pp = pp - 1.;
    const double wn1 = -pp * (pp - 1.) * (pp - 2.) / 6., w0 = (pp * pp - 1.) * (pp - 2.) / 2.;
    const double w1 = -pp * (pp + 1.) * (pp - 2.) / 2., w2 = pp * (pp * pp - 1.) / 6.;
    const int i = ip + tid, k = ip - 1 - tid;
    double ri = 0., rk = 0.;
    bool wi = false, wk = false;
    auto R = [&](int j) { return (double)acc[j - 1]; };
    if (tid < 64 && i <= ip + ip - 1 && ip + ip + 1 <= tile) {
      if (low) {  // the walk stops at the first K <= 1
        if (k > 1) {
          ri = R(i) + wn1 * R(k + 2) + w0 * R(k + 1) + w1 * R(k) + w2 * R(k - 1);
          rk = R(k) + wn1 * R(i + 2) + w0 * R(i + 1) + w1 * R(i) + w2 * R(i - 1);
          wi = wk = true;
        }
      } else if (k >= 1) {  // stops at the first K < 1; I untouched for K <= 2
        rk = R(k) + wn1 * R(i + 1) + w0 * R(i) + w1 * R(i - 1) + w2 * R(i - 2);
        wk = true;
        if (k > 2) {
          ri = R(i) + wn1 * R(k + 1) + w0 * R(k) + w1 * R(k - 1) + w2 * R(k - 2);
          wi = true;
        }
      }
    }
    __syncthreads();
    if (wi) acc[i - 1] = (LACC)ri;
    if (wk) acc[k - 1] = (LACC)rk;
    __syncthreads();
  }
  const double X00 = -7. / 128., X01 = 105. / 128., X02 = 35. / 128., X03 = -5. / 128.;
  const double X10 = -1. / 16., X11 = 9. / 16.;
  const DevPanel *pan = pans + Y.pan0;
  for (int k = tid; k < tile; k += TILE_T) {
    int64_t g = T0 + k;
    if (G == 3) {
      if (g >= Y.n3) continue;
      double v = (double)acc[k];
      if (g <= Y.g3max) {
        int p = g < 155 ? 0 : (int)((g - 5) / 150);
        if (p > Y.npanel - 1) p = Y.npanel - 1;
        int I = (int)(g - (int64_t)150 * p) + 1;
        int n1r3 = p == 0 ? 1 : 6;
        DevPanel pp = pan[p];
        int w = Y.win0 + pp.w;
        DevWin W = wins[w];
        v = v + xint_at(W.v1r4, W.v2r4, Y.dvr4, r4 + (size_t)w * R4ROW, pp.vft, Y.dvr3, I, n1r3,
                        155) * 1.0;
        if (Y.icntnm >= 1)
          v = v + xint_at(Y.v1abs, Y.v2abs, Y.dvabs, absrb + Y.abs0, pp.vft, Y.dvr3, I, n1r3, 155) * 1.;
      }
      rout[Y.r3o + g] = v;
    } else if (G == 2) {
      if (g >= Y.n2) continue;
      double v = (double)acc[k];
      if (g >= 4 && g <= Y.g2max) {
        const double *r3 = rin + Y.r3o;
        int64_t q = g >> 2;
        int ph = (int)(g & 3);
        double a0 = r3[q - 1], a1 = r3[q], a2 = r3[q + 1], a3 = r3[q + 2];
        if (ph == 0) v = v + a1;
        else if (ph == 1) v = v + X00 * a0 + X01 * a1 + X02 * a2 + X03 * a3;
        else if (ph == 2) v = v + X10 * (a0 + a3) + X11 * (a1 + a2);
        else v = v + X03 * a0 + X02 * a1 + X01 * a2 + X00 * a3;
      }
      rout[Y.r2o + g] = v;
    } else {
      if (g < 32 || g > Y.g_last) continue;
      const double *r2 = rin + Y.r2o;
      double v = (double)acc[k];
      int64_t q = g >> 2;
      int ph = (int)(g & 3);
      double a0 = r2[q - 1], a1 = r2[q], a2 = r2[q + 1], a3 = r2[q + 2];
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
          if (segs[Y.seg0 + pp.seg0 + mid].i1 <= I) lo2 = mid;
          else hi2 = mid - 1;
        }
        DevSeg sg = segs[Y.seg0 + pp.seg0 + lo2];
        v = v * (sg.radvi + (double)(I - sg.i1) * sg.rdel);
      }
      od[Y.od0 + g - 32] = (OUT)v;
    }
  }
}

// ---------------------------------------------------------------------
// CONTNM: ABSRB(1..NPTABS) of every job on its 1 cm-1 grid, one thread
// per point.  Each continuum is XINT (Keys cubic convolution) of its
// coarse array, whose four needed values are evaluated from the tables
// in place (FRNCO2 / o2_ver_1 / n2_ver_1 / mt_ckd_h2o_absco); the terms
// are added in CONTNM's order: H2O self and foreign, CO2, O2, N2,
// Rayleigh (src/contnm.f90:296-977).
struct ContTabs {
  const double *wvn, *self, *forr, *forcl, *texp;  // MT_CKD, LBLX_MTCKD_N
  const double *fco2, *xfac, *tdep;                // 5003, 500, 25
  const double *o2x, *o2t;                         // 103
  const double *n2a, *n2b, *n2h;                   // 228
  const double *n2r296, *n2rsf296, *n2r220, *n2rsf220;  // 73
};

__device__ __forceinline__ double d_mt_radfn(double vi, double xkt) {  // myradfn
  double x = vi / xkt;
  if (x <= 0.01) return 0.5 * x * vi;
  if (x <= 10.) {
    double e = exp(-x);
    return vi * (1. - e) / (1. + e);
  }
  return vi;
}

// XINT / myxint at ABSRB point I with the coarse values from A(j).
template <class F>
__device__ __forceinline__ double xint_fn(double v1a, double v2a, double dva, F A, double vft,
                                          double dvr3, int I, int n1r3, int n2r3) {
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
  return -A(j - 1) * b1 + A(j) * (1. - c + b2) + A(j + 1) * (c + b1) - A(j + 2) * b2;
}

__global__ __launch_bounds__(256) void k_contnm(const ContPlan *cps, ContTabs TB, int stride,
                                                int njobs, double *absrb) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)stride * njobs) return;
  const int job = (int)(t / stride);
  const int i = (int)(t - (int64_t)job * stride);
  const ContPlan &C = cps[job];
  double ab = 0.;
  if (!C.on || i < 1 || i > C.nptabs) {
    absrb[t] = 0.;
    return;
  }
  const double xkt = C.xkt;
  const int jrad = C.jrad;
  if (C.h2o.on) {
    const ContTerm &H = C.h2o;
    const double *fortab = C.forcl ? TB.forcl : TB.forr;
    auto self = [&](int k) -> double {
      if (k < 1 || k > H.nptc) return 0.;
      int q = H.i1 + k - 2;
      double v = TB.self[q] * pow(H.f2, TB.texp[q]);
      v = v * H.f0 * H.f1;
      if (H.f3 != 0.) v = v * d_mt_radfn(TB.wvn[q], xkt);
      return v;
    };
    auto forn = [&](int k) -> double {
      if (k < 1 || k > H.nptc) return 0.;
      int q = H.i1 + k - 2;
      double v = fortab[q] * (1 - H.f0) * H.f1;
      if (H.f3 != 0.) v = v * d_mt_radfn(TB.wvn[q], xkt);
      return v;
    };
    double sh = 0. + xint_fn(H.v1c, H.v2c, H.dvc, self, C.v1abs, C.dvabs, i, H.ist, H.last) * 1.0;
    double fh = 0. + xint_fn(H.v1c, H.v2c, H.dvc, forn, C.v1abs, C.dvabs, i, H.ist, H.last) * 1.0;
    double cself = C.wk1 * sh * C.xself;
    double cfor = C.wk1 * fh * C.xfrgn;
    ab = ab + cself + cfor;
    if (C.aj) {  // ln q derivative of the H2O continuum (contnm.f90:314-322)
      const double h = H.f0;
      const double cfaj = h >= 1.0 ? 0. : C.wk1 * fh * h / (1.0 - h);
      ab = ab + (cself - cfaj);
    }
  }
  if (C.co2.on) {
    const ContTerm &K = C.co2;
    auto co2 = [&](int J) -> double {
      if (J < 1 || J > K.nptc) return 0.;
      int I = K.i1 + (J - 1);
      double f = 0.;
      if (I >= 1 && I <= 5003) {
        double tcor = 1.;
        if (I >= 1196 && I <= 1220) tcor = pow(K.f1, TB.tdep[I - 1196]);
        f = tcor * TB.fco2[I - 1];
      }
      double VJ = K.v1c + K.dvc * (double)(J - 1);
      double cfac = 1.;
      if (VJ >= 2000. && VJ <= 2998.) cfac = TB.xfac[fint((VJ - 1998.) / 2. + 0.00001) - 1];
      f = cfac * f;
      double c = f * K.f0;
      if (jrad == 1) c = c * d_radfn(VJ, xkt);
      return c;
    };
    ab = ab + xint_fn(K.v1c, K.v2c, K.dvc, co2, C.v1abs, C.dvabs, i, K.ist, K.last) * 1.0;
  }
  if (C.o2.on) {
    const ContTerm &K = C.o2;
    auto o2 = [&](int J) -> double {
      if (J < 1 || J > K.nptc) return 0.;
      int I = K.i1 + (J - 1);
      double VJ = K.v1c + K.dvc * (double)(J - 1);
      double c0 = 0.;
      if (I >= 1 && I <= 103) c0 = K.f2 * TB.o2x[I - 1] * exp(TB.o2t[I - 1] * K.f1) / VJ;
      double c = K.f0 * c0;
      if (jrad == 1) c = c * d_radfn(VJ, xkt);
      return c;
    };
    ab = ab + xint_fn(K.v1c, K.v2c, K.dvc, o2, C.v1abs, C.dvabs, i, K.ist, K.last) * 1.0;
  }
  if (C.n2r.on) {  // N2 rotational (xn2_r): 296 K / 220 K tables, geometric in T
    const ContTerm &K = C.n2r;
    auto n2r = [&](int J) -> double {
      if (J < 1 || J > K.nptc) return 0.;
      int I = K.i1 + (J - 1);
      double VJ = K.v1c + K.dvc * (double)(J - 1);
      double c0 = 0., c1 = 0.;
      if (I >= 1 && I <= 73) {
        const double a = TB.n2r296[I - 1], sa = TB.n2rsf296[I - 1];
        c0 = a * pow(TB.n2r220[I - 1] / a, K.f1);
        const double sf_T = sa * pow(TB.n2rsf220[I - 1] / sa, K.f1);
        c1 = (sf_T - 1.) * (0.79) / (0.21);
      }
      double c = K.f0 * c0 * (K.f4 + c1 * K.f5 + 1. * K.f6);
      if (jrad == 1) c = c * d_radfn(VJ, xkt);
      return c;
    };
    ab = ab + xint_fn(K.v1c, K.v2c, K.dvc, n2r, C.v1abs, C.dvabs, i, K.ist, K.last) * 1.0;
  }
  if (C.n2.on) {
    const ContTerm &K = C.n2;
    auto n2 = [&](int J) -> double {
      if (J < 1 || J > K.nptc) return 0.;
      int I = K.i1 + (J - 1);
      double VJ = K.v1c + K.dvc * (double)(J - 1);
      double c0 = 0., c1 = 0., c2 = 0.;
      if (I >= 1 && I <= 228) {
        double a = TB.n2a[I - 1], b = TB.n2b[I - 1];
        if (a > 0. && b > 0.) c0 = a * pow(b / a, K.f1);
        else c0 = a + (b - a) * K.f2;
        c0 = c0 / VJ;
        c1 = K.f3 * c0;
        c2 = (9. / 7.) * TB.n2h[I - 1] * c0;
      }
      double c = K.f0 * (K.f4 * c0 + K.f5 * c1 + K.f6 * c2);
      if (jrad == 1) c = c * d_radfn(VJ, xkt);
      return c;
    };
    ab = ab + xint_fn(K.v1c, K.v2c, K.dvc, n2, C.v1abs, C.dvabs, i, K.ist, K.last) * 1.0;
  }
  if (C.ray_on) {
    double vr = C.v1abs + (double)(i - 1) * C.dvabs;
    double x = vr / 1.e4;
    double ray = (x * x * x / (9.38076E2 - 10.8426 * (x * x))) * C.ray_f;
    if (jrad == 1) ray = ray * x;
    else if (jrad == 0) ray = ray * x / d_radfn(vr, xkt);
    ab = ab + ray;
  }
  absrb[t] = ab;
}

// ---------------------------------------------------------------------
// XMERGE for upwelling radiance (IEMIT=1, IPATHL=1, LTE, no aerosols).
// EMIN (xmerge.f90:1850-1879): T = exp(-tau), E = (1-T)(B(Tave) + (B(Tzu)
// - B(Tave)) F(tau)) with F = tau/6 for tau <= 0.06.  First layer (EMINIT,
// xmerge.f90:2470-2480): R = E + T eps B(Tbnd).  Later layers (RADMRG/RADNN,
// xmerge.f90:2991-2995, 3118-3130): R = T L4(R_old) + E, T = T L4(T_old).
// RADNN walks each new panel in LL = ITYPE+1 phases: local point
// i = m*LL + (NL-1) reads carry-over slot IPL = 1 + m*max(ITYPE,1) (NL = 1)
// or the four slots IPL-1..IPL+2 around IPL = NL-1 + m*ITYPE with the
// Lagrange weights of P = 1 - (NL-1)/LL.  Slot j of new panel p holds old
// point base_p + j - 1 (base_p from the host's replay of RADMRG's panel
// reads, in DevPanel::out0); RADMRG duplicates the first old point below
// slot 1 and the last one past the end, which the index clamp reproduces.
// down (IPATHL = 3, downwelling radiance at the observer, layers taken
// outward from it): EMIN's boundary temperature is the layer's side toward
// the observer (TZL), EMINIT adds no boundary, and RADNN's IPATH_FLG = 3
// branch (xmerge.f90:3022-3026, 3151-3159) adds the new layer behind the
// old ones: R = L4(R_old) + E L4(T_old), T = T L4(T_old).  xktend > 0:
// EMBND (xmerge.f90:3652-3733) after the merge, R += T eps B(Tbnd), for
// the last layer of a downwelling path.
struct MergeArgs {
  double dv, xkt, xkta, xktbnd, emis;
  int64_t n, nold;
  int32_t first, itype;
  double a1[12], a2[12], a3[12], a4[12];
  int32_t down, pad;
  double xktend;
};

// The quotients use the reciprocal of rcp_nr (within 2.2e-15 relative).
__device__ __forceinline__ double d_planck(double vi, double xkt) {
  double v = 0.;
  if (xkt > 0.0) {
    double x = vi * rcp_nr(xkt);
    if (x < 1e-2) v = (vi * vi) * xkt * rcp_nr(1. + 0.5 * x);
    else if (x <= 80.) {
      double e = exp(-x);
      v = (vi * vi * vi) * e * rcp_nr(1. - e);
    }
  }
  return v * 1.191042722E-12;
}

// The four old points and weights that RADNN combines for local point k
// of a new panel whose carry-over slot 1 holds old point base1.  Phase
// r = 0 reads one point (weights 1,0,0,0).
struct MergeTap {
  int64_t i0, i1, i2, i3;
  double w0, w1, w2, w3;
};
__device__ __forceinline__ MergeTap merge_tap(const MergeArgs &A, int64_t base1, int k) {
  const int LL = A.itype + 1;
  const int64_t base = base1 - 1;  // old index of carry-over slot 0
  const int m = k / LL;
  const int r = k - m * LL;
  const int64_t last = A.nold - 1;
  MergeTap t;
  if (r == 0) {
    int64_t q = base + 1 + (int64_t)m * (A.itype > 0 ? A.itype : 1);
    q = q < 0 ? 0 : (q > last ? last : q);
    t.i0 = t.i1 = t.i2 = t.i3 = q;
    t.w0 = 1.;
    t.w1 = t.w2 = t.w3 = 0.;
  } else {
    int64_t q = base + r + (int64_t)m * A.itype;  // slot IPL = r + m*ITYPE
    t.i0 = q - 1 < 0 ? 0 : (q - 1 > last ? last : q - 1);
    t.i1 = q < 0 ? 0 : (q > last ? last : q);
    t.i2 = q + 1 > last ? last : q + 1;
    t.i3 = q + 2 > last ? last : q + 2;
    const int nl = r + 1;
    t.w0 = A.a1[nl];
    t.w1 = A.a2[nl];
    t.w2 = A.a3[nl];
    t.w3 = A.a4[nl];
  }
  return t;
}

// EMIN's emission of a layer at wavenumber vi (and its transmittance).
__device__ __forceinline__ double merge_em(double vi, double odvi, double xkt, double xkta, double *tr_out) {
  const double od_lo = 0.06;
  double bb = d_planck(vi, xkt);
  double bb_dif = xkta > 0. ? d_planck(vi, xkta) - bb : 0.;
  if (odvi < -od_lo) odvi = -od_lo;
  double tr = exp(-odvi);
  double f = odvi <= od_lo ? (1. / 6.) * odvi : 1. - 2. * (tr * rcp_nr(tr - 1.) + rcp_nr(odvi));
  *tr_out = tr;
  return (1. - tr) * (bb + bb_dif * f);
}

template <typename IN>
__global__ void k_merge(MergeArgs A, const DevPanel *pan, const IN *od, const double *rold,
                        const double *told, double *rnew, double *tnew) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n) return;
  const int p = (int)(g / 2400);
  const int k = (int)(g - (int64_t)2400 * p);
  const double vi = (pan[p].vft + 32. * A.dv) + (double)k * A.dv;
  double tr;
  double em = merge_em(vi, (double)od[g], A.xkt, A.xkta, &tr);
  double r, t;
  if (A.first) {
    if (!A.down && A.xktbnd > 0.) em = em + tr * d_planck(vi, A.xktbnd) * A.emis;
    r = em;
    t = tr;
  } else {
    MergeTap q = merge_tap(A, pan[p].out0, k);
    double ro = q.w0 * rold[q.i0] + q.w1 * rold[q.i1] + q.w2 * rold[q.i2] + q.w3 * rold[q.i3];
    double to = q.w0 * told[q.i0] + q.w1 * told[q.i1] + q.w2 * told[q.i2] + q.w3 * told[q.i3];
    r = A.down ? ro + em * to : tr * ro + em;
    t = tr * to;
  }
  if (A.xktend > 0.) r = r + t * A.emis * d_planck(vi, A.xktend);
  rnew[g] = r;
  tnew[g] = t;
}

// A run of layers on one grid whose merges after the first are the
// identity map of RADNN (ITYPE = 0, new panel p reading old point 2400 p
// on): the first layer merges as k_merge does, the rest update R and T in
// registers point by point (R = T_l R + E_l, T = T_l T; the identity taps
// add exact zeros, so the values equal the layer-by-layer merge), and only
// the last layer's R and T are written.
struct RunLayer {
  double xkt, xkta;
  int64_t od;    // element offset of the layer's optical depth
  int64_t rout;  // element offset of the layer's radiance in rall (-1: not kept)
};

template <typename IN>
__global__ void k_merge_run(MergeArgs A, const RunLayer *rl, int nrun, const DevPanel *pan,
                            const IN *odb, const double *rold, const double *told, double *rnew,
                            double *tnew, double *rall) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n) return;
  const int p = (int)(g / 2400);
  const int k = (int)(g - (int64_t)2400 * p);
  const double vi = (pan[p].vft + 32. * A.dv) + (double)k * A.dv;
  double tr;
  double em = merge_em(vi, (double)odb[rl[0].od + g], A.xkt, A.xkta, &tr);
  double r, t;
  if (A.first) {
    if (!A.down && A.xktbnd > 0.) em = em + tr * d_planck(vi, A.xktbnd) * A.emis;
    r = em;
    t = tr;
  } else {
    MergeTap q = merge_tap(A, pan[p].out0, k);
    double ro = q.w0 * rold[q.i0] + q.w1 * rold[q.i1] + q.w2 * rold[q.i2] + q.w3 * rold[q.i3];
    double to = q.w0 * told[q.i0] + q.w1 * told[q.i1] + q.w2 * told[q.i2] + q.w3 * told[q.i3];
    r = A.down ? ro + em * to : tr * ro + em;
    t = tr * to;
  }
  if (rl[0].rout >= 0) rall[rl[0].rout + g] = r;
  for (int j = 1; j < nrun; j++) {
    const RunLayer L = rl[j];
    em = merge_em(vi, (double)odb[L.od + g], L.xkt, L.xkta, &tr);
    r = A.down ? r + em * t : tr * r + em;
    t = tr * t;
    if (L.rout >= 0) rall[L.rout + g] = r;
  }
  if (A.xktend > 0.) r = r + t * A.emis * d_planck(vi, A.xktend);
  rnew[g] = r;
  tnew[g] = t;
}


// ---------------------------------------------------------------------
// Jacobians of the TOA radiance.  With R_l = T_l L4(R_{l-1}) + E_l the
// TOA radiance is a composition of per-layer maps, so a parameter x of
// layer l enters through dR_l/dx on grid l and reaches the top as the
// tangent J <- T_k L4(J) through every layer k above (the same carry-over
// interpolation as the radiance itself).
//   dR_l/dx = D_l dtau_l/dx [+ (1 - T_l)(1 - F) dB(Tave)/dT for x = Tave]
//   D_l = dR_l/dtau_l = -T_l L4(R_{l-1}) + T_l (B + dB F) + (1 - T_l) dB F'
// (L4(R_{l-1}) -> emis B(Tbnd) at the surface layer), F as in EMIN.
struct JacSrc {
  MergeArgs A;
  double tave;
  int64_t stride;        // between parameter planes of the tangent buffer
  int32_t nparam;
  double rh[8];          // 1 / step of each parameter's optical-depth difference
  int64_t odp[8];        // element offset of each perturbed optical depth
  int32_t direct[8];     // the perturbed run's optical depth is the derivative itself
};

// The source terms of layer point g: D = dR_l/dtau_l and the
// temperature's Planck term (1 - T_l)(1 - F) dB/dT, with T_l.
template <typename IN>
__device__ __forceinline__ void jac_src_terms(const JacSrc &S, const DevPanel *pan, const IN *od, const double *rold,
                                              int64_t g, double *D_out, double *bt_out, double *tr_out) {
  const MergeArgs &A = S.A;
  const int p = (int)(g / 2400);
  const int k = (int)(g - (int64_t)2400 * p);
  const double vi = (pan[p].vft + 32. * A.dv) + (double)k * A.dv;
  const double od_lo = 0.06;
  // PLANCK at Tave with its exponential kept for dB/dT
  const double x = vi * rcp_nr(A.xkt);
  double bb = 0., ex = 0.;
  if (x < 1e-2) bb = (vi * vi) * A.xkt * rcp_nr(1. + 0.5 * x) * 1.191042722E-12;
  else if (x <= 80.) {
    ex = exp(-x);
    bb = (vi * vi * vi) * ex * rcp_nr(1. - ex) * 1.191042722E-12;
  }
  const double bb_dif = A.xkta > 0. ? d_planck(vi, A.xkta) - bb : 0.;
  double odvi = (double)od[g];
  const bool clamped = odvi < -od_lo;
  if (clamped) odvi = -od_lo;
  const double tr = exp(-odvi);
  double f, fp;
  if (odvi <= od_lo) {
    f = (1. / 6.) * odvi;
    fp = 1. / 6.;
  } else {  // the forward merge's F (merge_em) and its derivative
    const double rq = rcp_nr(tr - 1.), ro = rcp_nr(odvi);
    f = 1. - 2. * (tr * rq + ro);
    fp = -2. * (tr * rq * rq - ro * ro);
  }
  double below;
  if (A.first) {
    below = A.xktbnd > 0. ? d_planck(vi, A.xktbnd) * A.emis : 0.;
  } else {
    MergeTap t = merge_tap(A, pan[p].out0, k);
    below = t.w0 * rold[t.i0] + t.w1 * rold[t.i1] + t.w2 * rold[t.i2] + t.w3 * rold[t.i3];
  }
  double D = -tr * below + tr * (bb + bb_dif * f) + (1. - tr) * bb_dif * fp;
  if (clamped) D = 0.;
  // dB/dT of PLANCK at Tave (zero where PLANCK is cut off)
  const double dbdt = (x >= 1e-2 && x <= 80.) ? bb * (x * rcp_nr(S.tave)) * rcp_nr(1. - ex) : 0.;
  *D_out = D;
  *bt_out = (1. - tr) * (1. - f) * dbdt;
  *tr_out = tr;
}

// Tangent planes are stored in FP32 (JP = float) between merges and in the
// caller's FP64 (JP = double) where they are written out; the arithmetic is
// FP64 throughout.
template <typename IN, typename JP>
__global__ void k_jac_src(JacSrc S, const DevPanel *pan, const IN *od, const IN *odall,
                          const double *rold, JP *jout, const double *scale) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= S.A.n) return;
  double D, bt, tr;
  jac_src_terms(S, pan, od, rold, g, &D, &bt, &tr);
  for (int q = 0; q < S.nparam; q++) {
    double dtau = S.direct[q] ? (double)odall[S.odp[q] + g] * S.rh[q]
                              : ((double)odall[S.odp[q] + g] - (double)od[g]) * S.rh[q];
    double v = D * dtau;
    if (q == 0) v = v + bt;
    if (scale) v = v * scale[g];
    jout[(int64_t)q * S.stride + g] = (JP)v;
  }
}

// Reverse mode (vector-Jacobian products).  With R_l = T_l L4(R_{l-1}) +
// E_l and lam_l = dL/dR_l, layer l's parameter x gets sum_g lam_l(g)
// dR_l(g)/dx (the source terms above), and lam_{l-1} = L4^T (T_l lam_l),
// the transpose of the carry-over interpolation.
//
// k_adj_src: each block's partial sums of lam (D dtau_q + [q = 0] bt) over
// its 256 points, in a fixed order: part[block P + q].
constexpr int ADJ_T = 256;
template <typename IN>
__global__ __launch_bounds__(ADJ_T) void k_adj_src(JacSrc S, const DevPanel *pan, const IN *od, const IN *odall,
                                                   const double *rold, const double *lam, double *part) {
  __shared__ double red[ADJ_T];
  const int64_t g = (int64_t)blockIdx.x * ADJ_T + threadIdx.x;
  double D = 0., bt = 0., tr = 0., l = 0.;
  if (g < S.A.n) {
    jac_src_terms(S, pan, od, rold, g, &D, &bt, &tr);
    l = lam[g];
  }
  for (int q = 0; q < S.nparam; q++) {
    double v = 0.;
    if (g < S.A.n) {
      const double dtau = (double)odall[S.odp[q] + g] * S.rh[q];
      v = l * (D * dtau + (q == 0 ? bt : 0.));
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = ADJ_T / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * S.nparam + q] = red[0];
    __syncthreads();
  }
}

// grad[q gstride] (+)= the sum of the blocks' partials of parameter q
// (blockIdx.x = q), in a fixed order.
__global__ __launch_bounds__(ADJ_T) void k_adj_sum(const double *part, int nblk, int P, double *grad, int64_t gstride) {
  __shared__ double red[ADJ_T];
  const int q = blockIdx.x;
  double v = 0.;
  for (int b = threadIdx.x; b < nblk; b += ADJ_T) v += part[(int64_t)b * P + q];
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = ADJ_T / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[(int64_t)q * gstride] = red[0];
}

// lam_old += L4^T (T_l lam): every point of the new grid adds its four
// weighted taps (lam_old cleared first).
template <typename IN>
__global__ void k_adj_prop(MergeArgs A, const DevPanel *pan, const IN *od, const double *lam, double *lam_old) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n) return;
  const int p = (int)(g / 2400);
  const int k = (int)(g - (int64_t)2400 * p);
  double odvi = (double)od[g];
  if (odvi < -0.06) odvi = -0.06;
  const double x = exp(-odvi) * lam[g];
  if (x == 0.) return;
  const MergeTap t = merge_tap(A, pan[p].out0, k);
  atomicAdd(&lam_old[t.i0], t.w0 * x);
  if (t.w1 != 0.) atomicAdd(&lam_old[t.i1], t.w1 * x);
  if (t.w2 != 0.) atomicAdd(&lam_old[t.i2], t.w2 * x);
  if (t.w3 != 0.) atomicAdd(&lam_old[t.i3], t.w3 * x);
}

// One layer of tangent propagation for nv vectors (plane stride istr on
// the old grid): J_new = T_k L4(S J_old), where S (on the old grid) is the
// pending transmittance product of the identity merges just below, for the
// first nsc vectors only (S null: none).  Vector v = l P + q is written at
// jout + l lstr + q qstr, times oscale (on the new grid) when given.
template <typename JP>
struct JacOut {
  JP *jout;
  int64_t lstr, qstr;
  int32_t P;
  const double *oscale;
};

template <typename IN, typename JP>
__global__ void k_jac_prop(MergeArgs A, const DevPanel *pan, const IN *od, const float *jin,
                           int64_t istr, int nv, const double *st, int nsc, JacOut<JP> O) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n) return;
  const int p = (int)(g / 2400);
  const int k = (int)(g - (int64_t)2400 * p);
  double odvi = (double)od[g];
  if (odvi < -0.06) odvi = -0.06;
  double tr = exp(-odvi);
  if (O.oscale) tr = tr * O.oscale[g];
  MergeTap t = merge_tap(A, pan[p].out0, k);
  double s0 = 1., s1 = 1., s2 = 1., s3 = 1.;
  if (st) {
    s0 = st[t.i0];
    s1 = st[t.i1];
    s2 = st[t.i2];
    s3 = st[t.i3];
  }
  for (int v = 0; v < nv; v++) {
    const float *a = jin + (int64_t)v * istr;
    const double a0 = a[t.i0], a1 = a[t.i1], a2 = a[t.i2], a3 = a[t.i3];
    double x;
    if (v < nsc)
      x = t.w0 * (s0 * a0) + t.w1 * (s1 * a1) + t.w2 * (s2 * a2) + t.w3 * (s3 * a3);
    else
      x = t.w0 * a0 + t.w1 * a1 + t.w2 * a2 + t.w3 * a3;
    const int l = v / O.P, q = v - l * O.P;
    O.jout[(int64_t)l * O.lstr + (int64_t)q * O.qstr + g] = (JP)(tr * x);
  }
}




// One layer merge of many profiles in one launch (blockIdx.y = profile).
// One run of same-grid layers of one profile (k_merge_run's work), for the
// batched merges: every launch takes the next run of each profile.
struct MergeJob {
  MergeArgs A;           // the run's first layer
  const DevPanel *pan;
  const void *od;        // optical depth base (double, or float for FP32 split)
  const RunLayer *rl;    // the run's layers: temperatures and optical-depth offsets
  int32_t nrun, pad;
  const double *rold, *told;
  double *rnew, *tnew;
};

template <typename IN>
__global__ void k_merge_batch(const MergeJob *jobs) {
  const MergeJob &J = jobs[blockIdx.y];
  const MergeArgs &A = J.A;
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.n) return;
  const int p = (int)(g / 2400);
  const int k = (int)(g - (int64_t)2400 * p);
  const double vi = (J.pan[p].vft + 32. * A.dv) + (double)k * A.dv;
  const IN *odb = (const IN *)J.od;
  double tr;
  double em = merge_em(vi, (double)odb[J.rl[0].od + g], A.xkt, A.xkta, &tr);
  double r, t;
  if (A.first) {
    if (A.xktbnd > 0.) em = em + tr * d_planck(vi, A.xktbnd) * A.emis;
    r = em;
    t = tr;
  } else {
    MergeTap q = merge_tap(A, J.pan[p].out0, k);
    double ro = q.w0 * J.rold[q.i0] + q.w1 * J.rold[q.i1] + q.w2 * J.rold[q.i2] + q.w3 * J.rold[q.i3];
    double to = q.w0 * J.told[q.i0] + q.w1 * J.told[q.i1] + q.w2 * J.told[q.i2] + q.w3 * J.told[q.i3];
    r = tr * ro + em;
    t = tr * to;
  }
  for (int j = 1; j < J.nrun; j++) {
    const RunLayer L = J.rl[j];
    em = merge_em(vi, (double)odb[L.od + g], L.xkt, L.xkta, &tr);
    r = tr * r + em;
    t = tr * t;
  }
  J.rnew[g] = r;
  J.tnew[g] = t;
}

__global__ void k_fill_f64(double *p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Running transmittance product of the layers above: tt *= exp(-tau).
template <typename IN>
__global__ void k_jac_trans(const IN *od, int64_t n, double *tt) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  double odvi = (double)od[g];
  if (odvi < -0.06) odvi = -0.06;
  tt[g] = tt[g] * exp(-odvi);
}

// dst plane v = tt * src plane v (nv planes, strides istr / ostr).
// FP32 planes (stride istr) out to the caller's FP64 planes: vector v =
// l P + q at dst + l lstr + q qstr, times s[g] for v < nsc (s null: none)
__global__ void k_jac_out(const float *src, int64_t istr, int nv, int P, const double *s, int nsc,
                          int64_t n, double *dst, int64_t lstr, int64_t qstr) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const double f = s ? s[g] : 1.;
  for (int v = 0; v < nv; v++) {
    const int l = v / P, q = v - l * P;
    double x = (double)src[(int64_t)v * istr + g];
    if (v < nsc) x = x * f;
    dst[(int64_t)l * lstr + (int64_t)q * qstr + g] = x;
  }
}


// ---------------------------------------------------------------------
// IEMIT = 0 merge of layer optical depths: ABINIT / ABSMRG
// (xmerge.f90:270-685), IPATHL = 1.  One thread per point of the new
// layer's grid.  ABSMRG's walk, after its alignment at new point g0 / old
// point NVS0 (1-based), is periodic: new point g0 + t, t = c LL + j
// (LL = ITYPE + 1), is the old point NVS0 + c ITYPE when j = 0, and for
// j = 1..ITYPE the 4-point Lagrange value A(j) at taps NVS - 1 .. NVS + 2,
// NVS = NVS0 + c ITYPE + j - 1 (the tap before the first old point is the
// first point).  Past the walk's end of the old file (t > t_eof) the rest
// of that new panel takes the mean of the old file's last two points.
struct AbsArgs {
  int64_t n, n_old, g0, nvs0, t_eof;
  int32_t itype, first;
  double a1[12], a2[12], a3[12], a4[12];
};

template <typename IN>
__global__ __launch_bounds__(256) void k_absmrg(AbsArgs A, const IN *od, const double *old, double *out) {
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= A.n) return;
  const double nw = (double)od[g];
  double v;
  if (A.first) {
    v = nw;  // ABINIT
  } else if (A.itype == 0) {
    v = nw + old[g];
  } else {
    const int64_t t = g - A.g0;
    if (t < 0) {
      v = nw;
    } else if (t == 0) {
      v = 1.0 * nw + old[A.nvs0 - 1];
    } else if (A.t_eof >= 0 && t > A.t_eof) {
      v = 1.0 * nw + (old[A.n_old - 1] + old[A.n_old - 2]) * 0.5;
    } else {
      const int LL = A.itype + 1;
      const int64_t c = t / LL;
      const int j = (int)(t - c * LL);
      if (j == 0) {
        v = nw + old[A.nvs0 + c * A.itype - 1];
      } else {
        const int64_t nvs = A.nvs0 + c * A.itype + j - 1;
        const double y = nvs == 1 ? old[0] : old[nvs - 2];
        const double oi = A.a1[j] * y + A.a2[j] * old[nvs - 1] + A.a3[j] * old[nvs] + A.a4[j] * old[nvs + 1];
        v = nw + oi;
      }
    }
  }
  out[g] = v;
}


// ---------------------------------------------------------------------
// PNLINT (oprop.f90:1564-1806): the layer OD regridded onto DVOUT.  The
// reference interpolates each DV panel as PANEL emits it; the host replays
// that control flow and records, per run of output points computed from
// one input panel, FJ1DIF = (V1PO - V1P)/DVP + 1 + 2 and the input panel's
// first global point.  One thread per output point: FJJ = FJ1DIF + RATDV
// (II - 1), JJ = INT(FJJ) - 2, JP = INT((FJJ - JJ) 100 - 199.5), value =
// 4-point Lagrange with weights tabulated at phase JP/100; taps before the
// first point and past the last repeat the end points (PNLINT's pads).
struct DvoSeg {
  int64_t o0;     // global output index of this run's first point
  int64_t obase;  // global output index of the output panel's II = 1
  int64_t gq;     // global index of the input panel's first point
  double fj1;     // FJ1DIF
};
struct DvoArgs {
  int64_t n_out, n_in;
  int32_t nseg, pass;
  double ratdv;
  double a1[101], a2[101], a3[101], a4[101];
};

template <typename IN>
__global__ __launch_bounds__(256) void k_dvout(DvoArgs A, const DvoSeg *segs, const IN *od, double *out) {
  const int64_t o = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (o >= A.n_out) return;
  if (A.pass) {  // DVOUT = DV: the panels themselves
    out[o] = (double)od[o];
    return;
  }
  int lo = 0, hi = A.nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].o0 <= o) lo = mid;
    else hi = mid - 1;
  }
  const DvoSeg S = segs[lo];
  const int64_t ii = o - S.obase + 1;
  const double fjj = S.fj1 + A.ratdv * (double)(ii - 1);
  const int jj = fint(fjj) - 2;
  const int jp = fint((fjj - (double)jj) * 100. - 199.5);
  double t[4];
  for (int k = 0; k < 4; k++) {
    int64_t g = S.gq + jj - 2 + k;
    g = g < 0 ? 0 : (g > A.n_in - 1 ? A.n_in - 1 : g);
    t[k] = (double)od[g];
  }
  out[o] = A.a1[jp] * t[0] + A.a2[jp] * t[1] + A.a3[jp] * t[2] + A.a4[jp] * t[3];
}

}  // namespace lblx
