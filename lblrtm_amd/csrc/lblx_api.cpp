// C ABI of the MI355X layer optical-depth path (include/lblx.h).
//
// One context owns the device, one HIP stream, the resident line table,
// the line-shape tables and grow-only work buffers.  lblx_od_layer* runs
// the OPDPTH/HIRAC1 replacement for one layer as a fixed sequence of
// launches on that stream (see lblx_kernels.hip for each stage).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lblx.h"
#include "lblx_kernels.h"
#include "lblx_plan.h"
#include "lblx_refdata.h"

// The kernels are compiled in this translation unit (one HIP module).
#include "lblx_kernels.hip"

namespace lblx {
void shape_tables_host(double *tab /* 5 x 2601: F1 F2 F3 FG XVER */);
}

using namespace lblx;

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t e_ = (x);                       \
    if (e_ != hipSuccess) {                    \
      set_err(std::string(#x) + ": " + hipGetErrorString(e_)); \
      return LBLX_E_HIP;                       \
    }                                          \
  } while (0)

static thread_local std::string g_err;
static void set_err(const std::string &s) { g_err = s; }

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes) return 0;
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t want = std::max(n, (size_t)256);
    if (hipMalloc(&p, want) != hipSuccess) {
      p = nullptr;
      return -1;
    }
    bytes = want;
    return 0;
  }
  template <typename T>
  T *as() const { return (T *)p; }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

}  // namespace

struct lblx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // host copy of the line table (flat, file order)
  std::vector<double> h_vnu, h_bvmax;
  std::vector<int32_t> h_bstart, h_blk;
  double max_pshift = 0.;
  int nlines = 0, nblk = 0;
  bool sorted = true;
  // device line table
  DevBuf d_vnu, d_sp, d_alfa, d_epp, d_hwhm, d_tmpalf, d_pshift, d_mol, d_blk, d_bstart, d_bvmax;
  DevBuf d_tabs;  // F1 F2 F3 FG XVER
  // work buffers
  DevBuf w_f4, w_stop, w_t9vnu, w_t9sp, w_t9alfa, w_t9epp, w_t9mol, w_t9n, w_t9bstart,
      w_t9bcount, w_t9nblk, w_t9bvhi;
  DevBuf w_pan, w_win, w_seg, w_r4, w_near, w_keep, w_blkmax, w_pread, w_changed, w_rec;
  DevBuf w_tlo, w_thi, w_r3, w_r2, w_cnt, w_od, w_absrb;
  DevLines lines() const {
    DevLines L;
    L.n = nlines;
    L.nblk = nblk;
    L.bstart = d_bstart.as<int32_t>();
    L.bvmax = d_bvmax.as<double>();
    L.vnu = d_vnu.as<double>();
    L.sp = d_sp.as<float>();
    L.alfa = d_alfa.as<float>();
    L.epp = d_epp.as<float>();
    L.hwhm = d_hwhm.as<float>();
    L.tmpalf = d_tmpalf.as<float>();
    L.pshift = d_pshift.as<float>();
    L.mol = d_mol.as<int32_t>();
    L.blk = d_blk.as<int32_t>();
    return L;
  }
};

extern "C" {

const char *lblx_strerror(int code) {
  switch (code) {
    case LBLX_OK: return "ok";
    case LBLX_E_ARG: return "invalid argument";
    case LBLX_E_NODEV: return "no HIP device";
    case LBLX_E_HIP: return g_err.empty() ? "HIP runtime error" : g_err.c_str();
    case LBLX_E_NREC: return "RDLIN; NREC GT LIMIN";
    case LBLX_E_UNSUPPORTED: return "input outside the supported subset";
    case LBLX_E_SIZE: return "output buffer too small";
    case LBLX_E_NOLINES: return "no line table loaded";
    case LBLX_E_DVRATIO: return "layer DV ratio above 2.5";
    case LBLX_E_COARSER: return "RADMRG; ITYPE LT 0";
    default: return "unknown error";
  }
}

int lblx_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int lblx_create(lblx_ctx **out, int device) {
  if (!out) return LBLX_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return LBLX_E_NODEV;
  HIPCHK(hipSetDevice(device));
  lblx_ctx *c = new lblx_ctx();
  c->device = device;
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return LBLX_E_HIP;
  }
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_avrat), LBLX_AVRAT, sizeof(LBLX_AVRAT)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_cf1), LBLX_CF1, sizeof(LBLX_CF1)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_cf2), LBLX_CF2, sizeof(LBLX_CF2)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_cf3), LBLX_CF3, sizeof(LBLX_CF3)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_cgauss), LBLX_CGAUSS, sizeof(LBLX_CGAUSS)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_cer), LBLX_CER, sizeof(LBLX_CER)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_isomax), LBLX_ISO_MAX, sizeof(LBLX_ISO_MAX)));
  std::vector<double> tabs(5 * 2601);
  shape_tables_host(tabs.data());
  if (c->d_tabs.ensure(tabs.size() * sizeof(double))) {
    delete c;
    return LBLX_E_HIP;
  }
  HIPCHK(hipMemcpy(c->d_tabs.p, tabs.data(), tabs.size() * sizeof(double), hipMemcpyHostToDevice));
  *out = c;
  return LBLX_OK;
}

void lblx_destroy(lblx_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  DevBuf *bufs[] = {&c->d_vnu, &c->d_sp, &c->d_alfa, &c->d_epp, &c->d_hwhm, &c->d_tmpalf,
                    &c->d_pshift, &c->d_mol, &c->d_blk, &c->d_bstart, &c->d_bvmax, &c->d_tabs,
                    &c->w_f4, &c->w_stop, &c->w_t9vnu, &c->w_t9sp, &c->w_t9alfa, &c->w_t9epp,
                    &c->w_t9mol, &c->w_t9n, &c->w_t9bstart, &c->w_t9bcount, &c->w_t9nblk,
                    &c->w_t9bvhi, &c->w_pan, &c->w_win, &c->w_seg, &c->w_r4, &c->w_near,
                    &c->w_keep, &c->w_blkmax, &c->w_pread, &c->w_changed, &c->w_rec, &c->w_tlo,
                    &c->w_thi, &c->w_r3, &c->w_r2, &c->w_cnt, &c->w_od, &c->w_absrb};
  for (DevBuf *b : bufs) b->release();
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int lblx_load_lines(lblx_ctx *c, const lblx_lines *L) {
  if (!c || !L || L->nblk < 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  // flatten the records in file order; coupling records are not supported
  std::vector<int32_t> bstart(L->nblk + 1), blk, mol;
  std::vector<double> vnu, bvmax(L->nblk);
  std::vector<float> sp, alfa, epp, hwhm, tmpalf, pshift;
  int n = 0;
  for (int b = 0; b < L->nblk; b++) {
    if (L->nrec[b] > LBLX_LINREC || L->nrec[b] < 0) return LBLX_E_NREC;
    bstart[b] = n;
    bvmax[b] = L->vmax[b];
    for (int i = 0; i < L->nrec[b]; i++) {
      size_t k = (size_t)b * LBLX_LINREC + i;
      if (L->iflg[k] != 0) return LBLX_E_UNSUPPORTED;
      vnu.push_back(L->vnu[k]);
      sp.push_back(L->sp[k]);
      alfa.push_back(L->alfa[k]);
      epp.push_back(L->epp[k]);
      hwhm.push_back(L->hwhm[k]);
      tmpalf.push_back(L->tmpalf[k]);
      pshift.push_back(L->pshift[k]);
      mol.push_back(L->mol[k]);
      blk.push_back(b);
      n++;
    }
  }
  bstart[L->nblk] = n;
  bool sorted = true;
  for (int i = 1; i < n; i++)
    if (vnu[i] < vnu[i - 1]) sorted = false;
  for (int b = 1; b < L->nblk; b++)
    if (bvmax[b] < bvmax[b - 1]) sorted = false;
  if (!sorted) return LBLX_E_UNSUPPORTED;
  double mp = 0.;
  for (float p : pshift) mp = std::max(mp, (double)std::fabs(p));
  auto up = [&](DevBuf &d, const void *h, size_t bytes) -> int {
    if (d.ensure(bytes)) return -1;
    if (bytes && hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return -1;
    return 0;
  };
  if (up(c->d_vnu, vnu.data(), n * 8) || up(c->d_sp, sp.data(), n * 4) ||
      up(c->d_alfa, alfa.data(), n * 4) || up(c->d_epp, epp.data(), n * 4) ||
      up(c->d_hwhm, hwhm.data(), n * 4) || up(c->d_tmpalf, tmpalf.data(), n * 4) ||
      up(c->d_pshift, pshift.data(), n * 4) || up(c->d_mol, mol.data(), n * 4) ||
      up(c->d_blk, blk.data(), n * 4) || up(c->d_bstart, bstart.data(), (L->nblk + 1) * 4) ||
      up(c->d_bvmax, bvmax.data(), L->nblk * 8)) {
    set_err("line table upload failed");
    return LBLX_E_HIP;
  }
  c->h_vnu = std::move(vnu);
  c->h_bvmax = std::move(bvmax);
  c->h_bstart = std::move(bstart);
  c->h_blk = std::move(blk);
  c->nlines = n;
  c->nblk = L->nblk;
  c->max_pshift = mp;
  c->sorted = sorted;
  return LBLX_OK;
}

int lblx_plan_dv(const lblx_run *run, int nlay, int nmol, const double *pave, const double *tave,
                 const double *wk, const double *wbroad, int iemit, double *dv, int32_t *itype) {
  return plan_dv(run, nlay, nmol, pave, tave, wk, wbroad, iemit, dv, itype);
}

int lblx_od_size(const lblx_run *run, const lblx_layer *layer, int64_t *npts) {
  if (!run || !layer || !npts) return LBLX_E_ARG;
  lblx_layer tmp = *layer;
  LayerPlan P;
  int rc = plan_layer(run, &tmp, &P);
  if (rc) return rc;
  *npts = P.npts;
  return LBLX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
static int run_layer(lblx_ctx *c, const lblx_run *run, lblx_layer *layer, void *d_od,
                     int64_t n_out, bool f32out, double *v1p, lblx_stats *st) {
  if (c->nlines <= 0) return LBLX_E_NOLINES;
  LayerPlan P;
  int rc = plan_layer(run, layer, &P);
  if (rc) return rc;
  if (run->icntnm != 0) return LBLX_E_UNSUPPORTED;  // continuum: separate entry
  if (n_out < P.npts) return LBLX_E_SIZE;
  hipStream_t s = c->stream;
  DevLayer Y{};
  Y.dv = P.dv;
  Y.dvr2 = P.dvr2;
  Y.dvr3 = P.dvr3;
  Y.dvr4 = P.dvr4;
  Y.bound4 = P.bound4;
  Y.alfmax = P.alfmax;
  Y.vft0 = P.vft0;
  Y.vbot = P.vbot;
  Y.vtop = P.vtop;
  Y.v1l4 = P.v1l4;
  Y.v2l4 = P.v2l4;
  Y.rhorat = P.rhorat;
  Y.tratio = P.tratio;
  Y.xkt = P.xkt;
  Y.betacr_l = P.betacr_l;
  Y.beta4 = P.beta4;
  Y.betacr4 = P.betacr4;
  Y.dptmn_hirac = P.dptmn_hirac;
  Y.dptfc = P.dptfc;
  Y.dptmn4 = P.dptmn4;
  std::memcpy(Y.scor, P.scor, sizeof(Y.scor));
  std::memcpy(Y.alfd1, P.alfd1, sizeof(Y.alfd1));
  for (int m = 0; m < 8; m++) {
    Y.rhoslf[m] = P.rhoslf[m];
    Y.wk[m] = m < layer->nmol ? layer->wk[m] : 0.;
  }
  Y.nmol = layer->nmol;
  Y.jrad = P.jrad;
  Y.max1 = P.max1;
  Y.npanel = (int)P.panels.size();
  Y.nwin = (int)P.wins.size();
  Y.g_last = P.g_last;
  Y.g2max = P.g2max;
  Y.g3max = P.g3max;
  Y.n1 = P.n1;
  Y.n2 = P.n2;
  Y.n3 = P.n3;
  Y.icntnm = 0;

  // plan tables
  std::vector<DevPanel> pan(P.panels.size());
  for (size_t i = 0; i < P.panels.size(); i++) {
    const PanelPlan &q = P.panels[i];
    pan[i] = DevPanel{q.vft, q.w, q.nhi, q.istop, q.seg0, q.nseg, 0, q.out0};
  }
  std::vector<DevWin> win(P.wins.size());
  for (size_t i = 0; i < P.wins.size(); i++) {
    const WinPlan &w = P.wins[i];
    win[i] = DevWin{w.v1r4, w.v2r4, w.nptr4, w.seg0, w.nseg, w.p_first};
  }
  std::vector<DevSeg> seg(std::max<size_t>(P.segs.size(), 1));
  for (size_t i = 0; i < P.segs.size(); i++)
    seg[i] = DevSeg{P.segs[i].i1, P.segs[i].i2, P.segs[i].radvi, P.segs[i].rdel};

  const int n = c->nlines;
  // far-wing line range: first block with VMAX >= V1L4 (RDLNFL skip)
  int b4 = (int)(std::lower_bound(c->h_bvmax.begin(), c->h_bvmax.end(), P.v1l4) - c->h_bvmax.begin());
  int first4 = b4 < c->nblk ? c->h_bstart[b4] : n;
  // near range (RDLIN): first record >= VBOT in the first block read, up
  // to and including the first record above VTOP
  int bn = (int)(std::lower_bound(c->h_bvmax.begin(), c->h_bvmax.end(), P.vbot) - c->h_bvmax.begin());
  int nlo = n, nhi = n - 1;
  if (bn < c->nblk) {
    int s0 = c->h_bstart[bn], e0 = c->h_bstart[bn + 1];
    nlo = (int)(std::lower_bound(c->h_vnu.begin() + s0, c->h_vnu.begin() + e0, P.vbot) -
                c->h_vnu.begin());
    if (nlo >= e0) nlo = e0 - 1;
    int up = (int)(std::upper_bound(c->h_vnu.begin() + nlo, c->h_vnu.end(), P.vtop) - c->h_vnu.begin());
    nhi = up < n ? up : n - 1;
  }
  const int nn = nhi >= nlo ? nhi - nlo + 1 : 0;
  const int blk0 = nn > 0 ? c->h_blk[nlo] : 0;
  const int nbk = nn > 0 ? c->h_blk[nhi] - blk0 + 1 : 0;

  // buffers
  const int nwin = Y.nwin;
  if (c->w_f4.ensure((size_t)n * sizeof(F4Line)) || c->w_stop.ensure(16) ||
      c->w_t9vnu.ensure((size_t)n * 8 + 64) || c->w_t9sp.ensure((size_t)n * 8 + 64) ||
      c->w_t9alfa.ensure((size_t)n * 8 + 64) || c->w_t9epp.ensure((size_t)n * 8 + 64) ||
      c->w_t9mol.ensure((size_t)n * 4 + 64) || c->w_t9n.ensure(16) ||
      c->w_t9bstart.ensure((size_t)(n / 1000 + 16) * 4) ||
      c->w_t9bcount.ensure((size_t)(n / 1000 + 16) * 4) || c->w_t9nblk.ensure(16) ||
      c->w_t9bvhi.ensure((size_t)(n / 1000 + 16) * 8) ||
      c->w_pan.ensure(pan.size() * sizeof(DevPanel)) || c->w_win.ensure(win.size() * sizeof(DevWin)) ||
      c->w_seg.ensure(seg.size() * sizeof(DevSeg)) || c->w_r4.ensure((size_t)nwin * 2504 * 8) ||
      c->w_near.ensure((size_t)std::max(nn, 1) * sizeof(NearLine)) ||
      c->w_keep.ensure((size_t)std::max(nn, 1) * 4) || c->w_blkmax.ensure((size_t)(nbk + 1) * 4) ||
      c->w_pread.ensure((size_t)(nbk + 1) * 4) || c->w_changed.ensure(16) ||
      c->w_rec.ensure((size_t)std::max(nn, 1) * sizeof(ScatterRec)) ||
      c->w_r3.ensure((size_t)P.n3 * 8) || c->w_r2.ensure((size_t)P.n2 * 8) || c->w_cnt.ensure(64)) {
    set_err("work buffer allocation failed");
    return LBLX_E_HIP;
  }
  HIPCHK(hipMemcpyAsync(c->w_pan.p, pan.data(), pan.size() * sizeof(DevPanel), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->w_win.p, win.data(), win.size() * sizeof(DevWin), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->w_seg.p, seg.data(), seg.size() * sizeof(DevSeg), hipMemcpyHostToDevice, s));
  DevLines L = c->lines();
  Tape9 T9;
  T9.vnu = c->w_t9vnu.as<double>();
  T9.sp = c->w_t9sp.as<double>();
  T9.alfa = c->w_t9alfa.as<double>();
  T9.epp = c->w_t9epp.as<double>();
  T9.mol = c->w_t9mol.as<int32_t>();
  T9.n = c->w_t9n.as<int32_t>();
  T9.bstart = c->w_t9bstart.as<int32_t>();
  T9.bcount = c->w_t9bcount.as<int32_t>();
  T9.nblk = c->w_t9nblk.as<int32_t>();
  T9.bvhi = c->w_t9bvhi.as<double>();
  unsigned long long *cnt = c->w_cnt.as<unsigned long long>();
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, c->w_stop.as<int>(), 1, INT_MAX);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int *)cnt, 8, 0);
  // 1. LINF4 correction
  if (first4 < n) {
    int cnt4 = n - first4;
    hipLaunchKernelGGL(k_linf4, dim3((cnt4 + 255) / 256), dim3(256), 0, s, L, Y, first4,
                       c->w_f4.as<F4Line>(), c->w_stop.as<int>());
  }
  // 2. SHRINK into TAPE9
  hipLaunchKernelGGL(k_shrink, dim3(1), dim3(256), 0, s, c->w_f4.as<F4Line>(), L, Y,
                     c->w_stop.as<int>(), T9);
  // 3. far-wing windows
  hipLaunchKernelGGL(k_convf4, dim3(nwin), dim3(256), 0, s, T9, Y, c->w_win.as<DevWin>(),
                     c->w_seg.as<DevSeg>(), c->w_r4.as<double>(), cnt + 1);
  if (nn > 0) {
    // 4. LNCOR1 + first fit
    hipLaunchKernelGGL(k_lncor1, dim3((nn + 255) / 256), dim3(256), 0, s, L, Y,
                       c->w_pan.as<DevPanel>(), nlo, nn, c->w_near.as<NearLine>());
    // 5. block-read panel fixpoint
    hipLaunchKernelGGL(k_fill_i32, dim3((nbk + 255) / 256), dim3(256), 0, s, c->w_pread.as<int>(), nbk, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3((nbk + 255) / 256), dim3(256), 0, s, c->w_blkmax.as<int>(), nbk, 0);
    int it = 0;
    for (; it < 64; it++) {
      hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, c->w_changed.as<int>(), 1, 0);
      hipLaunchKernelGGL(k_reject, dim3((nn + 255) / 256), dim3(256), 0, s, c->w_near.as<NearLine>(), nn,
                         nlo, c->d_blk.as<int32_t>(), blk0, c->w_pread.as<int>(),
                         c->w_pan.as<DevPanel>(), c->w_win.as<DevWin>(), c->w_r4.as<double>(), Y,
                         c->w_keep.as<int>(), c->w_blkmax.as<int>());
      hipLaunchKernelGGL(k_pread_scan, dim3(1), dim3(256), 0, s, c->w_blkmax.as<int>(), nbk,
                         c->w_pread.as<int>(), c->w_changed.as<int>());
      if (it >= 1) {
        int changed = 1;
        HIPCHK(hipMemcpyAsync(&changed, c->w_changed.p, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (!changed) break;
      }
    }
    // the final pass left keep/pread consistent; rerun k_reject so that keep
    // matches the converged pread exactly
    hipLaunchKernelGGL(k_reject, dim3((nn + 255) / 256), dim3(256), 0, s, c->w_near.as<NearLine>(), nn,
                       nlo, c->d_blk.as<int32_t>(), blk0, c->w_pread.as<int>(),
                       c->w_pan.as<DevPanel>(), c->w_win.as<DevWin>(), c->w_r4.as<double>(), Y,
                       c->w_keep.as<int>(), c->w_blkmax.as<int>());
    // 6. CNVFNV scatter records
    hipLaunchKernelGGL(k_scatter_rec, dim3(nbk), dim3(256), 0, s, c->w_near.as<NearLine>(),
                       c->w_keep.as<int>(), nlo, L, blk0, nbk, c->w_pread.as<int>(),
                       c->w_pan.as<DevPanel>(), Y, nn, c->w_rec.as<ScatterRec>());
    hipLaunchKernelGGL(k_sum_npts, dim3((nn + 255) / 256), dim3(256), 0, s, c->w_rec.as<ScatterRec>(),
                       nn, cnt);
  }
  // 7. tiles: R3, R2, R1 (+ PANEL combine)
  const int tiles[3] = {2048, 1024, 512};  // grid 1, 2, 3
  const int64_t ng[3] = {P.g_last + 1, P.n2, P.n3};
  const double dvg[3] = {P.dv, P.dvr2, P.dvr3};
  double shiftm = c->max_pshift * P.rhorat + 2. * P.dv;
  for (int G = 3; G >= 1; G--) {
    int T = tiles[G - 1];
    int nt = (int)((ng[G - 1] + T - 1) / T);
    if (c->w_tlo.ensure((size_t)nt * 4) || c->w_thi.ensure((size_t)nt * 4)) return LBLX_E_HIP;
    double margin = (double)(4 * (P.nalfmax + 1) + 8) * dvg[G - 1] + shiftm;
    if (nn > 0)
      hipLaunchKernelGGL(k_tile_ranges, dim3((nt + 255) / 256), dim3(256), 0, s,
                         c->d_vnu.as<double>() + nlo, nn, P.vft0, dvg[G - 1], T, nt, margin,
                         c->w_tlo.as<int>(), c->w_thi.as<int>());
    else {
      hipLaunchKernelGGL(k_fill_i32, dim3((nt + 255) / 256), dim3(256), 0, s, c->w_tlo.as<int>(), nt, 0);
      hipLaunchKernelGGL(k_fill_i32, dim3((nt + 255) / 256), dim3(256), 0, s, c->w_thi.as<int>(), nt, 0);
    }
    size_t shm = (size_t)(T + (G == 1 ? 3 : 1) * 2601) * sizeof(double);
    const double *tabs = c->d_tabs.as<double>();
    if (G == 3) {
      HIPCHK(hipFuncSetAttribute((const void *)k_tile<3, double>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      hipLaunchKernelGGL((k_tile<3, double>), dim3(nt), dim3(256), shm, s, c->w_rec.as<ScatterRec>(),
                         c->w_tlo.as<int>(), c->w_thi.as<int>(), T, tabs, Y, c->w_pan.as<DevPanel>(),
                         c->w_win.as<DevWin>(), c->w_r4.as<double>(), (const double *)nullptr,
                         c->w_seg.as<DevSeg>(), (const double *)nullptr, c->w_r3.as<double>(),
                         (double *)nullptr);
    } else if (G == 2) {
      HIPCHK(hipFuncSetAttribute((const void *)k_tile<2, double>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      hipLaunchKernelGGL((k_tile<2, double>), dim3(nt), dim3(256), shm, s, c->w_rec.as<ScatterRec>(),
                         c->w_tlo.as<int>(), c->w_thi.as<int>(), T, tabs, Y, c->w_pan.as<DevPanel>(),
                         c->w_win.as<DevWin>(), c->w_r4.as<double>(), (const double *)nullptr,
                         c->w_seg.as<DevSeg>(), c->w_r3.as<double>(), c->w_r2.as<double>(),
                         (double *)nullptr);
    } else if (!f32out) {
      HIPCHK(hipFuncSetAttribute((const void *)k_tile<1, double>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      hipLaunchKernelGGL((k_tile<1, double>), dim3(nt), dim3(256), shm, s, c->w_rec.as<ScatterRec>(),
                         c->w_tlo.as<int>(), c->w_thi.as<int>(), T, tabs, Y, c->w_pan.as<DevPanel>(),
                         c->w_win.as<DevWin>(), c->w_r4.as<double>(), (const double *)nullptr,
                         c->w_seg.as<DevSeg>(), c->w_r2.as<double>(), (double *)nullptr,
                         (double *)d_od);
    } else {
      HIPCHK(hipFuncSetAttribute((const void *)k_tile<1, float>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
      hipLaunchKernelGGL((k_tile<1, float>), dim3(nt), dim3(256), shm, s, c->w_rec.as<ScatterRec>(),
                         c->w_tlo.as<int>(), c->w_thi.as<int>(), T, tabs, Y, c->w_pan.as<DevPanel>(),
                         c->w_win.as<DevWin>(), c->w_r4.as<double>(), (const double *)nullptr,
                         c->w_seg.as<DevSeg>(), c->w_r2.as<double>(), (double *)nullptr,
                         (float *)d_od);
    }
  }
  HIPCHK(hipGetLastError());
  if (v1p) *v1p = P.panels[0].vft + (double)(NSHIFT) * P.dv;
  if (st) {
    unsigned long long h[2] = {0, 0};
    int32_t t9n = 0;
    HIPCHK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&t9n, c->w_t9n.p, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::memset(st, 0, sizeof(*st));
    st->k1_updates = (int64_t)h[0];
    st->k3_updates = (int64_t)h[1];
    st->nlins4 = t9n;
    st->npanel = (int64_t)P.panels.size();
    st->npts = P.npts;
  }
  return LBLX_OK;
}

extern "C" {

int lblx_od_layer_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layer, void *d_od,
                         int64_t n_out, double *v1p, lblx_stats *st) {
  if (!c || !run || !layer || !d_od) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  int rc = run_layer(c, run, layer, d_od, n_out, run->precision == LBLX_FP32_SPLIT, v1p, st);
  if (rc) return rc;
  return LBLX_OK;
}

int lblx_od_layer(lblx_ctx *c, const lblx_run *run, lblx_layer *layer, void *od, int64_t n_out,
                  double *v1p, lblx_stats *st) {
  if (!c || !run || !layer || !od) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  int64_t npts = 0;
  int rc = lblx_od_size(run, layer, &npts);
  if (rc) return rc;
  if (n_out < npts) return LBLX_E_SIZE;
  size_t esz = run->precision == LBLX_FP32_SPLIT ? 4 : 8;
  if (c->w_od.ensure((size_t)npts * esz)) return LBLX_E_HIP;
  rc = run_layer(c, run, layer, c->w_od.p, npts, esz == 4, v1p, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(od, c->w_od.p, (size_t)npts * esz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBLX_OK;
}

int lblx_sync(lblx_ctx *c) {
  if (!c) return LBLX_E_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBLX_OK;
}

void *lblx_stream(lblx_ctx *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"

// ---------------------------------------------------------------------
// TAPE3 reader (layout documented in lblrtm_amd/tape.py; RDLIN,
// oprop.f90:734-776; struct_types.f90:28-44).
extern "C" int lblx_load_tape3(lblx_ctx *c, const char *path) {
  if (!c || !path) return LBLX_E_ARG;
  FILE *f = std::fopen(path, "rb");
  if (!f) return LBLX_E_ARG;
  std::vector<char> data;
  {
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data.resize(sz > 0 ? sz : 0);
    if (sz > 0 && std::fread(data.data(), 1, sz, f) != (size_t)sz) {
      std::fclose(f);
      return LBLX_E_ARG;
    }
    std::fclose(f);
  }
  size_t pos = 0;
  auto rec = [&](const char **body, int32_t *len) -> bool {
    if (pos + 4 > data.size()) return false;
    int32_t n;
    std::memcpy(&n, data.data() + pos, 4);
    if (n < 0 || pos + 8 + (size_t)n > data.size()) return false;
    int32_t n2;
    std::memcpy(&n2, data.data() + pos + 4 + n, 4);
    if (n2 != n) return false;
    *body = data.data() + pos + 4;
    *len = n;
    pos += 8 + (size_t)n;
    return true;
  };
  const char *b;
  int32_t len;
  if (!rec(&b, &len)) return LBLX_E_ARG;   // file header
  if (len >= 56 && b[55] == '^') {         // negative-EPP record (oprop.f90:746-750)
    if (!rec(&b, &len)) return LBLX_E_ARG;
  }
  std::vector<int32_t> nrec, mol, iflg;
  std::vector<double> vmin, vmax, vnu;
  std::vector<float> sp, alfa, epp, hwhm, tmpalf, pshift;
  const int R = LBLX_LINREC;
  while (rec(&b, &len)) {
    if (len < 24) return LBLX_E_ARG;
    double v0, v1;
    int32_t n, nw;
    std::memcpy(&v0, b, 8);
    std::memcpy(&v1, b + 8, 8);
    std::memcpy(&n, b + 16, 4);
    std::memcpy(&nw, b + 20, 4);
    if (n > R) return LBLX_E_NREC;
    if (!rec(&b, &len) || nw != 9750 || len != 9750 * 4) return LBLX_E_UNSUPPORTED;
    nrec.push_back(n);
    vmin.push_back(v0);
    vmax.push_back(v1);
    size_t o = 0;
    auto take = [&](auto &vec, size_t esz) {
      size_t base = vec.size();
      vec.resize(base + R);
      std::memcpy(vec.data() + base, b + o, R * esz);
      o += R * esz;
    };
    take(vnu, 8);
    take(sp, 4);
    take(alfa, 4);
    take(epp, 4);
    take(mol, 4);
    take(hwhm, 4);
    take(tmpalf, 4);
    take(pshift, 4);
    take(iflg, 4);
  }
  lblx_lines L;
  L.nblk = (int32_t)nrec.size();
  L.nrec = nrec.data();
  L.vmin = vmin.data();
  L.vmax = vmax.data();
  L.vnu = vnu.data();
  L.sp = sp.data();
  L.alfa = alfa.data();
  L.epp = epp.data();
  L.mol = mol.data();
  L.hwhm = hwhm.data();
  L.tmpalf = tmpalf.data();
  L.pshift = pshift.data();
  L.iflg = iflg.data();
  return lblx_load_lines(c, &L);
}
