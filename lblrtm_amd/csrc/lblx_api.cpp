// C ABI of the MI355X layer optical-depth path (include/lblx.h).
//
// One context owns the device, one HIP stream, the resident line table,
// the line-shape tables and grow-only work buffers.  lblx_od_layer* runs
// the OPDPTH/HIRAC1 replacement for one layer as a fixed sequence of
// launches on that stream (see lblx_kernels.hip for each stage).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <memory>
#include <condition_variable>
#include <functional>
#include <type_traits>

#include "../../include/lblx.h"
#include "lblx_kernels.h"

#ifndef LBLX_FP32_LDS64
// FP32 shapes accumulate R2/R3 in FP64 LDS: ds_add_f32 alone runs about 23x
// slower than ds_add_f64 on gfx950 (203 vs 4650 G atomics/s); in the R1
// tile, where the adds interleave with three table reads and the FP64
// index arithmetic, FP32 accumulation is still the faster (2.39 vs 2.69 ms)
#define LBLX_FP32_LDS64 0xC  // bits G = 2, 3
#endif
#include "lblx_plan.h"
#include "lblx_contdata.h"
#include "lblx_refdata.h"

// The kernels are compiled in this translation unit (one HIP module).
#include "lblx_kernels.hip"
#include "lblx_scan.hip"
#include "lblx_tangent.hip"

namespace lblx {
void shape_tables_host(double *tab /* 5 x 2601: F1 F2 F3 FG XVER */);
void voigt_tables_host(double *f /* (3 x 101 + 1) x 2601 */, double *coef /* 6 x 101 */);
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

// Work memory of the allocation-free FFTSCN entry (lblx_fftscn_noalloc_device):
// the two FFT planes, the twiddles and INTPDR's index and coefficient
// tables, sized for FFTs of up to 2^19 points and 2^18 output points
// (23 MiB).  Each device holds its own copy until the process ends.  In
// host code `scratch` is only a handle: its device address is looked up
// once per device (scratch_lookup, from lblx_create) and that address is
// what the entry passes to kernels and copies.
constexpr int64_t NA_NMAX = (int64_t)1 << 19, NA_OUTMAX = (int64_t)1 << 18;
constexpr size_t NA_OFF_Z1 = (size_t)NA_NMAX * sizeof(double2);
constexpr size_t NA_OFF_TW = 2 * NA_OFF_Z1;
constexpr size_t NA_OFF_G = NA_OFF_TW + (size_t)(NA_NMAX / 2) * sizeof(double2);
constexpr size_t NA_OFF_IP = NA_OFF_G + (size_t)NA_OUTMAX * sizeof(int64_t);
constexpr size_t NA_OFF_COF = NA_OFF_IP + (size_t)NA_OUTMAX * sizeof(int32_t);
constexpr size_t NA_BYTES = NA_OFF_COF + 8192;
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];  // device address of `scratch` per device

// With `device` current: the device address of `scratch` on it.
static int scratch_lookup(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= 64) return LBLX_E_ARG;
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_err(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return LBLX_E_HIP;
    }
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return LBLX_OK;
}

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t n) {
    if (n <= bytes) return 0;
    if (p) {  // work queued on any stream may still use the old buffer
      (void)hipDeviceSynchronize();
      hipFree(p);
    }
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

// Stage timers (the analogue of the TAPE6 TIME/CONVOLUTION/PANEL columns,
// oprop.f90:547-620): HIP events recorded on the context stream around each
// stage, summed when collected.
enum Stage { ST_LINF4, ST_SHRINK, ST_CONVF4, ST_LNCOR1, ST_REJECT, ST_SCATTER, ST_TILE3, ST_TILE2,
             ST_TILE1, ST_MERGE, ST_CONTNM, ST_SCAN, ST_HOST, ST_TANGENT, ST_N };
static const char *kStageNames[ST_N] = {"linf4", "shrink", "convf4", "lncor1", "reject",
                                        "scatter_rec", "tile_r3", "tile_r2", "tile_r1",
                                        "merge", "contnm", "scan", "host", "tangent"};

struct lblx_ctx {
  int device = 0;
  int ncu = 256;  // compute units of the device
  // LBLX_SHRINK_WAVE_WALK=1 at creation: every TAPE9 block of the batch
  // walk is simulated by the chain wave (tests compare it with the tables)
  int t9_all_wave = 0;
  // block-read fixpoint: pass budget before the sequential walk takes over
  // (LBLX_FIX_MAXPASS at creation, 1..64), and what the last OD batch used
  int fix_maxpass = 16;
  int fix_passes = 0, fix_seq = 0;
  void *scratch = nullptr;  // device address of `scratch` on this device
  hipStream_t stream = nullptr;
  bool timing = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> evs;
  std::vector<hipEvent_t> evpool;
  double stage_ms[ST_N] = {0};
  long stage_n[ST_N] = {0};
  hipEvent_t ev_get() {
    if (!evpool.empty()) {
      hipEvent_t e = evpool.back();
      evpool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }
  // second stream (default priority) for work that overlaps the main chain;
  // every fork is joined back into `stream` within the same call
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int tbegin(int st, hipStream_t on = nullptr) {
    if (!timing) return -1;
    hipEvent_t a = ev_get(), b = ev_get();
    (void)hipEventRecord(a, on ? on : stream);
    evs.push_back({st, {a, b}});
    return (int)evs.size() - 1;
  }
  void tend(int h, hipStream_t on = nullptr) {
    if (h < 0) return;
    (void)hipEventRecord(evs[h].second.second, on ? on : stream);
  }
  // Pinned staging of the per-call tables copied to the device.  A copy
  // from pageable memory returns only once the stream has reached it, which
  // would hold the host until the queued kernels finish; staged copies let
  // the host return and plan the next call while they run.  Two arenas are
  // used by alternate generations (one table upload each); an arena is
  // refilled only after the copies issued from it have completed.
  struct Stage {
    char *p = nullptr;
    size_t cap = 0, used = 0;
    hipEvent_t ev = nullptr;
    bool live = false;
  };
  Stage stg[2];
  int scur = 0;
  void stage_begin() {
    scur ^= 1;
    Stage &S = stg[scur];
    if (S.live) (void)hipEventSynchronize(S.ev);
    else if (S.used > 0) (void)hipStreamSynchronize(stream);  // a generation that never ended
    S.live = false;
    S.used = 0;
  }
  void stage_end(hipStream_t s) {
    Stage &S = stg[scur];
    if (!S.ev && hipEventCreateWithFlags(&S.ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipStreamSynchronize(s);  // no event: the copies are simply waited for
      return;
    }
    (void)hipEventRecord(S.ev, s);
    S.live = true;
  }
  hipError_t h2d(void *dst, const void *src, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    Stage &S = stg[scur];
    const size_t rn = (n + 255) & ~(size_t)255;
    if (S.used + rn > S.cap) {
      // the copies already issued from this arena finish before it is replaced
      hipError_t e = hipStreamSynchronize(s);
      if (e != hipSuccess) return e;
      const size_t cap = std::max(2 * (S.used + rn), (size_t)16 << 20);
      char *np = nullptr;
      if ((e = hipHostMalloc((void **)&np, cap, hipHostMallocDefault)) != hipSuccess) return e;
      if (S.p) (void)hipHostFree(S.p);
      S.p = np;
      S.cap = cap;
      S.used = 0;
    }
    char *q = S.p + S.used;
    std::memcpy(q, src, n);
    S.used += rn;
    return hipMemcpyAsync(dst, q, n, hipMemcpyHostToDevice, s);
  }
  void stage_release() {
    for (Stage &S : stg) {
      if (S.ev) (void)hipEventDestroy(S.ev);
      if (S.p) (void)hipHostFree(S.p);
      S = Stage{};
    }
  }
  // host wall time (ST_HOST), summed only while timing is on
  std::chrono::steady_clock::time_point h0;
  void hbegin() {
    if (timing) h0 = std::chrono::steady_clock::now();
  }
  void hend() {
    if (!timing) return;
    stage_ms[ST_HOST] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    stage_n[ST_HOST]++;
  }
  void tcollect() {
    (void)hipStreamSynchronize(stream);
    for (auto &e : evs) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) {
        stage_ms[e.first] += ms;
        stage_n[e.first] += 1;
      }
      evpool.push_back(e.second.first);
      evpool.push_back(e.second.second);
    }
    evs.clear();
  }
  // host copy of the line table (flat, file order)
  std::vector<double> h_vnu, h_bvmax;
  std::vector<int32_t> h_bstart, h_blk;
  std::vector<float> h_sp;
  std::vector<int32_t> h_mol;
  std::vector<int32_t> h_molidx[7];  // lines of molecule m+1 with a valid isotope and S != 0
  int64_t last_line_stride = 0;      // per-job line stride of the last OD run
  double max_pshift = 0.;
  int nlines = 0, nblk = 0;
  int64_t nbrd = 0;  // TAPE3 lines carrying broadener flags (IBRD data)
  bool sorted = true;
  // device line table
  DevBuf d_vnu, d_sp, d_alfa, d_epp, d_hwhm, d_tmpalf, d_pshift, d_mol, d_blk, d_bstart, d_bvmax;
  DevBuf d_cidx, d_cco;  // line coupling (DevLines::cidx, cco)
  int64_t ncpl = 0;      // lines with IFLG > 0
  DevBuf d_tabs;  // F1 F2 F3 FG XVER
  DevBuf d_vtab, d_vtabf;  // exact-Voigt build tables, FP64 and FP32 (built on first use)
  bool voigt_ready = false;
  DevBuf d_cont;  // continuum tables (ContTabs)
  ContTabs cont_tabs{};
  DevBuf w_cps;   // ContPlan per job
  // work buffers
  DevBuf w_dlay, w_pan, w_win, w_word, w_seg, w_r4, w_margin;
  DevBuf w_cx, w_csp, w_cal, w_cad, w_cfl, w_g, w_jmp, w_wsum, w_bcnt, w_etab, w_jobi,
      w_bat;
  DevBuf w_t9vnu, w_t9sp, w_t9alfa, w_t9epp, w_t9mol, w_t9n, w_t9blk, w_t9bvhi;
  DevBuf w_cspp, w_t9spp, w_cpl;  // line coupling: SPP of the passing and TAPE9 lines, SPPSP per record
  DevBuf w_dirty;  // blocks whose read panel moved in the last fixpoint step
  DevBuf w_t9s;    // TAPE9-block start tables of the batch walk (six ints per block)
  DevBuf w_near, w_keep, w_kidx, w_blkmax, w_pread, w_changed, w_rec, w_r3, w_r2, w_cnt, w_od, w_absrb;
  DevBuf w_tmap[3];
  DevBuf w_mpan, w_rt, w_mrun;
  std::shared_ptr<void> jprep;            // host job tables of run_jobs (JobsPrep)
  std::vector<LayerPlan> jplans;          // the Jacobian's job plans, storage reused
  DevBuf s_z[2], s_tw, s_g, s_ip, s_cof;  // FFTSCN
  DevBuf f_s, f_xf, f_zp, f_q;            // SCANFN: shrunk input, function table, ZPEAK and R1 panel per group
  DevBuf j_rall, j_a, j_b;                // Jacobians
  // tangent-linear optical depth: compacted-line records, CONVF4's accepted
  // lines, per-line derivative weights, channel rows, continuum channels
  DevBuf t_craw, t_amask, t_wspan, t_tl, t_t9, t_r4c, t_tlay, t_cps, t_absab, t_absc, t_rh, t_r3c, t_r2c,
      t_odcoff, t_tmap[3], t_tabs, t_mem, t_memq, t_f4c, t_f4d, t_parts, t_r4p;
  int64_t s_twN = 0;
  int tiles[3] = {8192, 4096, 2048};      // R1, R2, R3 tile widths (autotuned)
  int lbatch[3] = {0, 0, 0};              // lines per wave chunk in the tiles (0: 16 FP64 R1, else 32)
  DevBuf b_jobs, b_rt, b_rl;              // batched merges
  DevBuf w_dvo;                           // PNLINT runs
  struct GraphEntry {
    std::vector<int64_t> key;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<GraphEntry> graphs;         // instantiated merge chains
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
    L.cidx = ncpl > 0 ? d_cidx.as<int32_t>() : nullptr;
    L.cco = ncpl > 0 ? d_cco.as<double>() : nullptr;
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
    case LBLX_E_SCRATCH: return g_err.empty() ? "request larger than the fixed scratch" : g_err.c_str();
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
  {
    const char *e = std::getenv("LBLX_SHRINK_WAVE_WALK");
    c->t9_all_wave = (e && std::atoi(e) != 0) ? 1 : 0;
    const char *f = std::getenv("LBLX_FIX_MAXPASS");
    if (f) c->fix_maxpass = std::min(64, std::max(1, std::atoi(f)));
  }
  {
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
      c->ncu = ncu;
  }
  if (hipStreamCreate(&c->stream) != hipSuccess || hipStreamCreate(&c->side) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return LBLX_E_HIP;
  }
  {
    int rs = scratch_lookup(device, &c->scratch);
    if (rs) {
      (void)hipEventDestroy(c->ev_fork);
      (void)hipEventDestroy(c->ev_join);
      (void)hipStreamDestroy(c->side);
      (void)hipStreamDestroy(c->stream);
      delete c;
      return rs;
    }
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
  // continuum tables, concatenated in ContTabs order
  {
    const std::pair<const double *, size_t> parts[] = {
        {LBLX_MTCKD_WVN, LBLX_MTCKD_N}, {LBLX_MTCKD_SELF, LBLX_MTCKD_N},
        {LBLX_MTCKD_FOR, LBLX_MTCKD_N}, {LBLX_MTCKD_FORCL, LBLX_MTCKD_N},
        {LBLX_MTCKD_TEXP, LBLX_MTCKD_N}, {LBLX_FCO2, 5003}, {LBLX_XFACCO2, 500},
        {LBLX_CO2_TDEP, 25}, {LBLX_O2F_XO2, 103}, {LBLX_O2F_XO2T, 103},
        {LBLX_N2F_272, 228}, {LBLX_N2F_228, 228}, {LBLX_N2F_AH2O, 228},
        {LBLX_N2R_C296, 73}, {LBLX_N2R_SF296, 73}, {LBLX_N2R_C220, 73}, {LBLX_N2R_SF220, 73}};
    std::vector<double> all;
    std::vector<size_t> off;
    for (auto &q : parts) {
      off.push_back(all.size());
      all.insert(all.end(), q.first, q.first + q.second);
    }
    if (c->d_cont.ensure(all.size() * sizeof(double))) {
      delete c;
      return LBLX_E_HIP;
    }
    HIPCHK(hipMemcpy(c->d_cont.p, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
    const double *b = c->d_cont.as<double>();
    const double **dst[] = {&c->cont_tabs.wvn, &c->cont_tabs.self, &c->cont_tabs.forr,
                            &c->cont_tabs.forcl, &c->cont_tabs.texp, &c->cont_tabs.fco2,
                            &c->cont_tabs.xfac, &c->cont_tabs.tdep, &c->cont_tabs.o2x,
                            &c->cont_tabs.o2t, &c->cont_tabs.n2a, &c->cont_tabs.n2b,
                            &c->cont_tabs.n2h, &c->cont_tabs.n2r296, &c->cont_tabs.n2rsf296,
                            &c->cont_tabs.n2r220, &c->cont_tabs.n2rsf220};
    for (size_t k = 0; k < off.size(); k++) *dst[k] = b + off[k];
  }
  *out = c;
  return LBLX_OK;
}

void lblx_destroy(lblx_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  for (auto &g : c->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  DevBuf *bufs[] = {&c->d_vnu, &c->d_sp, &c->d_alfa, &c->d_epp, &c->d_hwhm, &c->d_tmpalf,
                    &c->d_pshift, &c->d_mol, &c->d_blk, &c->d_bstart, &c->d_bvmax, &c->d_tabs, &c->d_cont, &c->w_cps,
                    &c->d_cidx, &c->d_cco, &c->w_cspp, &c->w_t9spp, &c->w_cpl,
                    &c->w_dlay, &c->w_pan, &c->w_win, &c->w_word, &c->w_seg, &c->w_r4, &c->w_margin,
                    &c->w_cx, &c->w_csp, &c->w_cal, &c->w_cad, &c->w_cfl,
                    &c->w_g, &c->w_jmp, &c->w_wsum, &c->w_bcnt, &c->w_etab, &c->w_t9s, &c->w_jobi, &c->w_bat,
                    &c->w_t9vnu, &c->w_t9sp, &c->w_t9alfa, &c->w_t9epp, &c->w_t9mol, &c->w_t9n,
                    &c->w_t9blk, &c->w_t9bvhi, &c->w_near, &c->w_keep, &c->w_kidx, &c->w_blkmax, &c->w_pread,
                    &c->w_changed, &c->w_rec, &c->w_r3, &c->w_r2, &c->w_cnt, &c->w_od,
                    &c->w_absrb, &c->w_tmap[0], &c->w_tmap[1], &c->w_tmap[2], &c->w_mpan, &c->w_rt, &c->w_mrun, &c->s_z[0], &c->s_z[1], &c->s_tw, &c->s_g, &c->s_ip, &c->s_cof, &c->j_rall, &c->j_a, &c->j_b, &c->b_jobs, &c->b_rt, &c->b_rl,
                    &c->w_dirty, &c->d_vtab, &c->d_vtabf, &c->w_dvo,
                    &c->t_craw, &c->t_amask, &c->t_wspan, &c->t_tl, &c->t_t9, &c->t_r4c, &c->t_tlay, &c->t_cps,
                    &c->t_absab, &c->t_absc, &c->t_rh, &c->t_r3c, &c->t_r2c, &c->t_odcoff, &c->t_tmap[0],
                    &c->t_tmap[1], &c->t_tmap[2], &c->t_tabs, &c->t_mem, &c->t_memq, &c->t_f4c, &c->t_f4d,
                    &c->t_parts, &c->t_r4p};
  for (DevBuf *b : bufs) b->release();
  c->tcollect();
  c->stage_release();
  for (hipEvent_t e : c->evpool) (void)hipEventDestroy(e);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int lblx_load_lines(lblx_ctx *c, const lblx_lines *L) {
  if (!c || !L || L->nblk < 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  // Flatten the lines in file order.  A coupling record (IFLG < 0) is not
  // a line (RDLIN's IOUT skips it, oprop.f90:844-850): it is dropped, and
  // the coefficients it holds for the line before it (IFLG 1 or 3) go to
  // the coupling table (oprop.f90:789-795, 1039-1048).
  std::vector<int32_t> bstart(L->nblk + 1), blk, mol, cidx;
  std::vector<double> vnu, bvmax(L->nblk), cco;
  std::vector<float> sp, alfa, epp, hwhm, tmpalf, pshift;
  int n = 0;
  int64_t ncpl = 0;
  for (int b = 0; b < L->nblk; b++) {
    if (L->nrec[b] > LBLX_LINREC || L->nrec[b] < 0) return LBLX_E_NREC;
    bstart[b] = n;
    bvmax[b] = L->vmax[b];
    for (int i = 0; i < L->nrec[b]; i++) {
      size_t k = (size_t)b * LBLX_LINREC + i;
      const int f = L->iflg[k];
      if (f < 0) continue;
      int ci = -1;
      if (f == 1 || f == 3) {
        // the coefficient record must follow in the same block
        if (i + 1 >= L->nrec[b] || L->iflg[k + 1] >= 0) {
          set_err("a coupled line (IFLG 1 or 3) lacks its coefficient record");
          return LBLX_E_UNSUPPORTED;
        }
        float am;
        std::memcpy(&am, &L->mol[k + 1], sizeof am);  // MOL holds a REAL*4
        ci = 2 * (int)(cco.size() / 8) + (f == 3 ? 1 : 0);
        const double c8[8] = {L->vnu[k + 1], L->alfa[k + 1], (double)am, L->tmpalf[k + 1],
                              L->sp[k + 1], L->epp[k + 1], L->hwhm[k + 1], L->pshift[k + 1]};
        cco.insert(cco.end(), c8, c8 + 8);
      } else if (f > 0) {
        ci = -2;
      }
      if (f > 0) ncpl++;
      cidx.push_back(ci);
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
      up(c->d_bvmax, bvmax.data(), L->nblk * 8) ||
      (ncpl > 0 && (up(c->d_cidx, cidx.data(), n * 4) || up(c->d_cco, cco.data(), cco.size() * 8)))) {
    set_err("line table upload failed");
    return LBLX_E_HIP;
  }
  for (int m = 0; m < 7; m++) c->h_molidx[m].clear();
  for (int i = 0; i < n; i++) {
    int m = mol[i] % 100, iso = (mol[i] % 1000) / 100;
    if (m >= 1 && m <= 7 && iso >= 1 && iso <= LBLX_ISO_MAX[m - 1] && sp[i] != 0.f)
      c->h_molidx[m - 1].push_back(i);
  }
  c->h_sp = std::move(sp);
  c->h_mol = std::move(mol);
  c->h_vnu = std::move(vnu);
  c->h_bvmax = std::move(bvmax);
  c->h_bstart = std::move(bstart);
  c->h_blk = std::move(blk);
  c->nlines = n;
  c->nblk = L->nblk;
  c->max_pshift = mp;
  c->sorted = sorted;
  c->nbrd = 0;
  c->ncpl = ncpl;
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

int lblx_od_panels(const lblx_run *run, const lblx_layer *layer, int cap, double *v1p, double *v2p,
                   int32_t *nlim, int *npanel) {
  if (!run || !layer || !npanel) return LBLX_E_ARG;
  lblx_layer tmp = *layer;
  LayerPlan P;
  int rc = plan_layer(run, &tmp, &P);
  if (rc) return rc;
  const int np = (int)P.panels.size();
  *npanel = np;
  if (cap < np) return (v1p || v2p || nlim) ? LBLX_E_SIZE : LBLX_OK;
  for (int k = 0; k < np; k++) {
    const PanelPlan &q = P.panels[k];
    // PANEL: V1P = VFT + (NLO-1) DV with NLO = NSHIFT+1; NLIM = NHI-NLO+1
    if (v1p) v1p[k] = q.vft + (double)NSHIFT * P.dv;
    if (v2p) v2p[k] = q.v2p;
    if (nlim) nlim[k] = q.nhi - NSHIFT;
  }
  return LBLX_OK;
}

int lblx_write_panels(const char *path, int append, int word, const void *filhdr, int64_t filhdr_bytes,
                      int npanel, const double *v1p, const double *v2p, double dvp, const int32_t *nlim,
                      const double *a, const double *b, int endfil) {
  if (!path || (word != 4 && word != 8) || npanel < 0 || (npanel > 0 && (!v1p || !v2p || !nlim || !a)) ||
      (filhdr && filhdr_bytes <= 0) || filhdr_bytes > INT32_MAX)
    return LBLX_E_ARG;
  FILE *f = std::fopen(path, append ? "ab" : "wb");
  if (!f) return LBLX_E_ARG;
  bool ok = true;
  // one sequential unformatted record: 4-byte length, body, 4-byte length
  auto record = [&](const void *body, int64_t n) {
    const int32_t m = (int32_t)n;
    ok = ok && std::fwrite(&m, 4, 1, f) == 1 && (n == 0 || std::fwrite(body, 1, (size_t)n, f) == (size_t)n) &&
         std::fwrite(&m, 4, 1, f) == 1;
  };
  if (filhdr) record(filhdr, filhdr_bytes);
  std::vector<unsigned char> buf;
  int64_t o = 0;
  for (int k = 0; k < npanel && ok; k++) {
    const int32_t n = nlim[k];
    if (n < 0 || (int64_t)n * word > INT32_MAX) {
      ok = false;
      break;
    }
    // /BUFPNL/ V1P, V2P (REAL*8 under IMPLICIT REAL*8 (V)), DVP (REAL),
    // NLIM (INTEGER): default REAL and INTEGER are `word` bytes
    unsigned char h[32];
    size_t hn = 0;
    std::memcpy(h + hn, &v1p[k], 8);
    hn += 8;
    std::memcpy(h + hn, &v2p[k], 8);
    hn += 8;
    if (word == 8) {
      const int64_t nl = n;
      std::memcpy(h + hn, &dvp, 8);
      std::memcpy(h + hn + 8, &nl, 8);
      hn += 16;
    } else {
      const float d4 = (float)dvp;
      std::memcpy(h + hn, &d4, 4);
      std::memcpy(h + hn + 4, &n, 4);
      hn += 8;
    }
    record(h, (int64_t)hn);
    for (const double *src : {a, b}) {
      if (!src) continue;
      buf.resize((size_t)n * word);
      if (word == 8) {
        std::memcpy(buf.data(), src + o, (size_t)n * 8);
      } else {
        for (int32_t i = 0; i < n; i++) {
          const float x = (float)src[o + i];
          std::memcpy(buf.data() + 4 * (size_t)i, &x, 4);
        }
      }
      record(buf.data(), (int64_t)n * word);
    }
    o += n;
  }
  if (endfil && ok) {  // ENDFIL: six default integers of -99 (lblrtm.f90:1368-1377)
    unsigned char e[48];
    for (int i = 0; i < 6; i++) {
      if (word == 8) {
        const int64_t v = -99;
        std::memcpy(e + 8 * i, &v, 8);
      } else {
        const int32_t v = -99;
        std::memcpy(e + 4 * i, &v, 4);
      }
    }
    record(e, 6 * word);
  }
  if (std::fclose(f) != 0) ok = false;
  return ok ? LBLX_OK : LBLX_E_ARG;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Line ranges of one layer in the flat table.
//   far wing (LINF4): blocks [b4, b4 + nb4): from the first block with
//     VMAX >= V1L4 to the block of the first line that passes LINF4 and
//     lies above V2L4 even after the largest pressure shift, so the stop
//     line (the first passing line above V2L4) is inside (oprop.f90:3228);
//   near lines (RDLIN): records [nlo, nlo + nn).
// Returns the number of records the job's per-line work arrays must hold.
struct JobLines {
  int b4, nb4, nlo, nn;
};
static int64_t job_lines(const lblx_ctx *c, const LayerPlan &P, const lblx_layer *ly, JobLines *J) {
  const int n = c->nlines, nblk = c->nblk;
  J->b4 = (int)(std::lower_bound(c->h_bvmax.begin(), c->h_bvmax.end(), P.v1l4) - c->h_bvmax.begin());
  const double vlim = P.v2l4 + c->max_pshift * P.rhorat + 1e-6 * (1. + std::fabs(P.v2l4));
  int i = (int)(std::upper_bound(c->h_vnu.begin(), c->h_vnu.end(), vlim) - c->h_vnu.begin());
  int bend = nblk;
  for (;;) {
    int best = n, bm = -1;
    for (int m = 0; m < 7 && m < ly->nmol; m++) {
      if (ly->wk[m] == 0.) continue;
      const std::vector<int32_t> &ix = c->h_molidx[m];
      auto it = std::lower_bound(ix.begin(), ix.end(), i);
      if (it != ix.end() && *it < best) {
        best = *it;
        bm = m;
      }
    }
    if (best >= n) break;
    if ((double)c->h_sp[best] * ly->wk[bm] != 0.) {
      bend = c->h_blk[best] + 1;
      break;
    }
    i = best + 1;
  }
  J->nb4 = std::max(0, bend - J->b4);
  // near range (RDLIN): first record >= VBOT in the first block read, up
  // to and including the first record above VTOP
  int bn = (int)(std::lower_bound(c->h_bvmax.begin(), c->h_bvmax.end(), P.vbot) - c->h_bvmax.begin());
  int nlo = n, nhi = n - 1;
  if (bn < nblk) {
    int s0 = c->h_bstart[bn], e0 = c->h_bstart[bn + 1];
    nlo = (int)(std::lower_bound(c->h_vnu.begin() + s0, c->h_vnu.begin() + e0, P.vbot) - c->h_vnu.begin());
    if (nlo >= e0) nlo = e0 - 1;
    int up = (int)(std::upper_bound(c->h_vnu.begin() + nlo, c->h_vnu.end(), P.vtop) - c->h_vnu.begin());
    nhi = up < n ? up : n - 1;
  }
  J->nlo = nlo;
  J->nn = nhi >= nlo ? nhi - nlo + 1 : 0;
  const int64_t n4 = J->nb4 > 0 ? c->h_bstart[J->b4 + J->nb4] - c->h_bstart[J->b4] : 0;
  return std::max<int64_t>(n4, J->nn);
}

// Tables of the exact-Voigt build, computed once per process on the host
// and uploaded to the context's device on first use.
static const std::vector<double> &voigt_host(std::vector<double> **coef_out) {
  static std::vector<double> f, coef;
  static std::once_flag once;
  std::call_once(once, []() {
    f.assign((size_t)(3 * 101 + 1) * 2601, 0.);
    coef.assign(6 * 101, 0.);
    voigt_tables_host(f.data(), coef.data());
  });
  if (coef_out) *coef_out = &coef;
  return f;
}

static int ensure_voigt(lblx_ctx *c) {
  if (c->voigt_ready) return LBLX_OK;
  std::vector<double> *coef = nullptr;
  const std::vector<double> &f = voigt_host(&coef);
  std::vector<float> ff(f.begin(), f.end());
  double a3[102], b3[102];
  for (int i = 0; i < 101; i++) {
    a3[i] = (*coef)[2 * 101 + i];
    b3[i] = (*coef)[5 * 101 + i];
  }
  a3[101] = (*coef)[3 * 101];  // a_3(101) is b_1(0) in COMMON /voigt_cf/
  b3[101] = b3[100];
  if (c->d_vtab.ensure(f.size() * 8) || c->d_vtabf.ensure(ff.size() * 4)) return LBLX_E_HIP;
  HIPCHK(hipMemcpy(c->d_vtab.p, f.data(), f.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_vtabf.p, ff.data(), ff.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_va3), a3, sizeof(a3)));
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(lblx::c_vb3), b3, sizeof(b3)));
  c->voigt_ready = true;
  return LBLX_OK;
}

// Host side of an OPDPTH batch: the jobs' descriptors and plan tables, tile
// maps and the strides of the per-job work arrays.
// Host worker threads kept for the process (up to 15 beside the caller).
// One task runs at a time; a task is a counter of indices that workers and
// the caller claim until none is left, and it is finished when every index
// has run.  A worker that wakes late holds the task it saw and finds no
// index left, so it never runs a finished task's function.
class HostPool {
 public:
  static HostPool &get() {
    static HostPool pool;
    return pool;
  }
  // false when another task is running (the caller then does it otherwise)
  bool run(int n, const std::function<void(int)> &f) {
    std::unique_lock<std::mutex> busy(run_m_, std::try_to_lock);
    if (!busy.owns_lock() || th_.empty()) return false;
    auto t = std::make_shared<Task>();
    t->f = f;
    t->n = n;
    {
      std::lock_guard<std::mutex> lk(m_);
      cur_ = t;
      gen_++;
    }
    cv_.notify_all();
    work(*t);
    while (t->done.load() < n) std::this_thread::yield();
    std::lock_guard<std::mutex> lk(m_);
    if (cur_ == t) cur_.reset();
    return true;
  }
  ~HostPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      quit_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }

 private:
  struct Task {
    std::function<void(int)> f;
    int n = 0;
    std::atomic<int> next{0}, done{0};
  };
  HostPool() {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned k = std::min(15u, hw > 1 ? hw - 1 : 0u);
    for (unsigned i = 0; i < k; i++) th_.emplace_back([this]() { loop(); });
  }
  static void work(Task &t) {
    for (int i = t.next++; i < t.n; i = t.next++) {
      t.f(i);
      t.done++;
    }
  }
  void loop() {
    long seen = 0;
    for (;;) {
      std::shared_ptr<Task> t;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&]() { return quit_ || (cur_ && gen_ != seen); });
        if (quit_) return;
        seen = gen_;
        t = cur_;
      }
      work(*t);
    }
  }
  std::mutex run_m_, m_;
  std::condition_variable cv_;
  std::vector<std::thread> th_;
  std::shared_ptr<Task> cur_;
  long gen_ = 0;
  bool quit_ = false;
};

// f(i) for i in [0, n) on the host worker threads (the calling thread among
// them); f must not depend on the order.
template <class F>
static void parallel_for(int n, F f) {
  if (n <= 0) return;
  if (n < 8) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  if (HostPool::get().run(n, std::function<void(int)>(f))) return;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nth = (int)std::min<unsigned>(std::min(16u, hw), (unsigned)std::max(1, n / 8));
  if (nth <= 1) {
    for (int i = 0; i < n; i++) f(i);
    return;
  }
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int i = next++; i < n; i = next++) f(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nth; t++) pool.emplace_back(work);
  work();
  for (auto &t : pool) t.join();
}

struct JobsPrep {
  int njobs = 0;
  std::vector<LayerPlan> own;
  std::vector<const LayerPlan *> plans;
  std::vector<DevLayer> ys;
  std::vector<DevPanel> pan;
  std::vector<DevWin> win;
  std::vector<int32_t> worder;  // far-wing workgroup -> window
  std::vector<DevSeg> seg;
  std::vector<double> margins;
  std::vector<ContPlan> cps;
  std::vector<int32_t> tjob[3], tloc[3];
  int tiles[3] = {0, 0, 0};
  int64_t r2tot = 0, r3tot = 0, maxline = 0;
  int maxnn = 0, maxnbk = 0, abs_stride = 0, maxnb4 = 0;
  Strides S{};
};

// Far-wing windows are all resident at once (one workgroup each), and a
// window's time follows its line-centre density: windows of the same band
// in neighbouring layers are equally heavy, and in index order they land on
// the same compute units.  The windows are dealt out heaviest first, one per
// compute unit per tier, in alternating direction, with the raw lines within
// BOUND4 of the window times its points as the weight.
static void window_costs(const lblx_ctx *c, const JobsPrep &J, std::vector<double> &cost) {
  const int nwin = (int)J.win.size();
  cost.resize(nwin);
  for (int i = 0; i < nwin; i++) {
    const DevWin &w = J.win[i];
    const double b4 = J.ys[w.job].bound4;
    const auto lo = std::lower_bound(c->h_vnu.begin(), c->h_vnu.end(), w.v1r4 - b4);
    const auto hi = std::upper_bound(lo, c->h_vnu.end(), w.v2r4 + b4);
    cost[i] = (double)(hi - lo) * w.nptr4;
  }
}

static void far_wing_order(const lblx_ctx *c, JobsPrep &J) {
  const int nwin = (int)J.win.size();
  std::vector<double> cost;
  window_costs(c, J, cost);
  std::vector<int32_t> by(nwin);
  for (int i = 0; i < nwin; i++) by[i] = i;
  std::stable_sort(by.begin(), by.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  auto &ord = J.worder;
  ord.resize(nwin);
  const int T = std::max(1, c->ncu);
  for (int i = 0; i < nwin; i++) {
    const int tier = i / T, slot = i % T, m = std::min(T, nwin - tier * T);
    ord[(size_t)tier * T + ((tier & 1) ? m - 1 - slot : slot)] = by[i];
  }
}

// Temperature interval of the line-coupling coefficients (TEMPLC = 200,
// 250, 296, 340 K).  LNCOR1 takes the first interval [ILC, ILC+1] with
// TAVE < TEMPLC(ILC+1), ILC <= 3, anchored at ILC (oprop.f90:981-990);
// LINF4 the first [ILC-1, ILC] with TAVE <= TEMPLC(ILC), 2 <= ILC <= 4,
// anchored at ILC (:3077-3082).
static void lc_interval(double tave, DevLayer &Y) {
  static const double T[4] = {200.0, 250.0, 296.0, 340.0};
  int c = 0;  // 0-based ILC - 1
  for (int il = 0; il < 3; il++) {
    c = il;
    if (tave < T[c + 1]) break;
  }
  Y.lc_lo[0] = c;
  Y.lc_an[0] = c;
  Y.lc_rect[0] = 1.0 / (T[c + 1] - T[c]);
  Y.lc_dif[0] = tave - T[c];
  int f = 0;  // 0-based ILC
  while (f < 4 && !(tave <= T[f])) f++;
  if (f == 0) f = 1;
  if (f == 4) f = 3;
  Y.lc_lo[1] = f - 1;
  Y.lc_an[1] = f;
  Y.lc_rect[1] = 1.0 / (T[f] - T[f - 1]);
  Y.lc_dif[1] = tave - T[f];
}

// Plan a batch of jobs (layers) on the host.  Reads only the context's line
// table and settings, so it may run beside launches of another batch.
// Layer plans are computed here, or handed in (pre, with the layers' JRAD
// already chained).
static int prep_jobs(const lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int njobs, bool chain_jrad,
                     const int64_t *od_off, const LayerPlan *pre, JobsPrep &J) {
  if (c->nlines <= 0) return LBLX_E_NOLINES;
  if (njobs <= 0) return LBLX_E_ARG;
  if (run->ibrd > 0 && c->nbrd > 0) return LBLX_E_UNSUPPORTED;  // broadener widths and shifts
  const int n = c->nlines, nblk = c->nblk;
  J.njobs = njobs;
  if (!pre && (int)J.own.size() < njobs) J.own.resize(njobs);  // plans keep their storage
  J.plans.assign(njobs, nullptr);
  J.ys.assign(njobs, DevLayer{});
  J.margins.assign(3 * njobs, 0.);
  J.cps.assign(njobs, ContPlan{});
  auto &own = J.own;
  auto &plans = J.plans;
  auto &ys = J.ys;
  auto &pan = J.pan;
  auto &win = J.win;
  auto &seg = J.seg;
  auto &margins = J.margins;
  auto &cps = J.cps;
  // 1. layer plans: handed in, or computed here on host threads (JRAD
  // chained up the batch first when asked)
  if (pre) {
    for (int jb = 0; jb < njobs; jb++) {
      plans[jb] = &pre[jb];
      layers[jb].jrad = pre[jb].jrad;
    }
  } else {
    if (chain_jrad)
      for (int jb = 1; jb < njobs; jb++) layers[jb].jrad = layer_jrad(&layers[jb - 1]);
    std::vector<int> prc(njobs, 0);
    parallel_for(njobs, [&](int jb) { prc[jb] = plan_layer(run, &layers[jb], &own[jb]); });
    for (int jb = 0; jb < njobs; jb++) {
      if (prc[jb]) return prc[jb];
      plans[jb] = &own[jb];
    }
  }
  // 2. offsets of every job's panels, windows, segments and R2/R3 rows
  int64_t r2tot = 0, r3tot = 0;
  size_t npan = 0, nwin_all = 0, nseg = 0;
  for (int jb = 0; jb < njobs; jb++) {
    const LayerPlan &P = *plans[jb];
    DevLayer &Y = ys[jb];
    Y.pan0 = (int)npan;
    Y.win0 = (int)nwin_all;
    Y.seg0 = (int)nseg;
    Y.r2o = r2tot;
    Y.r3o = r3tot;
    npan += P.panels.size();
    nwin_all += P.wins.size();
    nseg += P.segs.size();
    r2tot += P.n2;
    r3tot += P.n3;
  }
  pan.resize(npan);
  win.resize(nwin_all);
  seg.resize(nseg);
  // 3. each job's device record and table rows, in parallel
  std::vector<int64_t> jline(njobs, 0);
  parallel_for(njobs, [&](int jb) {
    const lblx_layer *ly = &layers[jb];
    const LayerPlan &P = *plans[jb];
    DevLayer &Y = ys[jb];
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
      Y.wk[m] = m < ly->nmol ? ly->wk[m] : 0.;
    }
    Y.nmol = ly->nmol;
    Y.voigt = run->voigt ? 1 : 0;
    lc_interval(ly->tave, Y);
    Y.pavp0 = ly->pave / 1013.25;
    Y.pavp2 = Y.pavp0 * Y.pavp0;
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
    cps[jb] = P.cont;
    Y.icntnm = P.cont.on ? 1 : 0;
    Y.v1abs = P.cont.v1abs;
    Y.v2abs = P.cont.v2abs;
    Y.dvabs = P.cont.dvabs;
    Y.nptabs = P.cont.nptabs;
    DevPanel *dp = pan.data() + Y.pan0;
    for (const PanelPlan &q : P.panels) *dp++ = DevPanel{q.vft, q.w, q.nhi, q.istop, q.seg0, q.nseg, 0, q.out0};
    DevWin *dw = win.data() + Y.win0;
    for (const WinPlan &w : P.wins) *dw++ = DevWin{w.v1r4, w.v2r4, w.nptr4, Y.seg0 + w.seg0, w.nseg, jb};
    DevSeg *ds = seg.data() + Y.seg0;
    for (const RadSeg &g : P.segs) *ds++ = DevSeg{g.i1, g.i2, g.radvi, g.rdel};
    JobLines JL;
    jline[jb] = job_lines(c, P, ly, &JL);
    Y.b4 = JL.b4;
    Y.nb4 = JL.nb4;
    const int nlo = JL.nlo, nhi = JL.nlo + JL.nn - 1;
    Y.nlo = JL.nlo;
    Y.nn = JL.nn;
    Y.blk0 = Y.nn > 0 ? c->h_blk[nlo] : 0;
    Y.nbk = Y.nn > 0 ? c->h_blk[nhi] - Y.blk0 + 1 : 0;
    Y.od0 = od_off ? od_off[jb] : 0;
    double shiftm = c->max_pshift * P.rhorat + 2. * P.dv;
    const double dvg[3] = {P.dv, P.dvr2, P.dvr3};
    for (int G = 1; G <= 3; G++)
      margins[3 * jb + G - 1] = (double)(4 * (P.nalfmax + 1) + 8) * dvg[G - 1] + shiftm;
  });
  int maxnn = 0, maxnbk = 0, abs_stride = 0, maxnb4 = 0;
  int64_t maxline = 0;
  for (int jb = 0; jb < njobs; jb++) {
    maxline = std::max(maxline, jline[jb]);
    maxnn = std::max(maxnn, ys[jb].nn);
    maxnbk = std::max(maxnbk, ys[jb].nbk);
    maxnb4 = std::max(maxnb4, ys[jb].nb4);
    if (plans[jb]->cont.on) abs_stride = std::max(abs_stride, plans[jb]->cont.nptabs + 4);
  }
  if (seg.empty()) seg.push_back(DevSeg{0, 0, 0., 0.});
  const int nwin = (int)win.size();
  far_wing_order(c, J);
  abs_stride = (abs_stride + 63) / 64 * 64;
  for (int jb = 0; jb < njobs; jb++) ys[jb].abs0 = (int64_t)jb * abs_stride;
  // tile maps
  const int tiles[3] = {c->tiles[0], c->tiles[1], c->tiles[2]};
  auto &tjob = J.tjob;
  auto &tloc = J.tloc;
  for (int G = 0; G < 3; G++) {
    J.tiles[G] = tiles[G];
    tjob[G].clear();
    tloc[G].clear();
  }
  for (int jb = 0; jb < njobs; jb++) {
    const LayerPlan &P = *plans[jb];
    const int64_t ng[3] = {P.g_last + 1, P.n2, P.n3};
    for (int G = 0; G < 3; G++) {
      int nt = (int)((ng[G] + tiles[G] - 1) / tiles[G]);
      for (int t = 0; t < nt; t++) {
        tjob[G].push_back(jb);
        tloc[G].push_back(t);
      }
    }
  }
  Strides &S = J.S;
  S.line = (maxline + 64 + 63) / 64 * 64;
  S.blk = nblk + 2;
  S.t9blk = n / 1000 + 16;
  J.r2tot = r2tot;
  J.r3tot = r3tot;
  J.maxline = maxline;
  J.maxnn = maxnn;
  J.maxnbk = maxnbk;
  J.abs_stride = abs_stride;
  J.maxnb4 = maxnb4;
  return LBLX_OK;
}

// A tangent-linear request riding on an OD batch: the channels (the
// temperature, then the molecules), per job the tangent constants, per
// (job, channel) the continuum plans stepped up and down, and where each
// (job, channel) optical-depth derivative goes (element offsets from odc,
// in the batch's output type).
struct TanReq {
  int P = 1;
  std::vector<TanLayer> tl;
  std::vector<ContPlan> cps;  // [(job P + q) 2 + {0: +h, 1: -h}] (empty without ICNTNM)
  double rh[TAN_NQ] = {0};    // 1 / (2 h) per channel
  void *odc = nullptr;
  std::vector<int64_t> odc_off;  // [job P + q]
};

template <int P>
static void launch_convf4_tan(dim3 g, hipStream_t s, const Tape9 &T9, const F4TConst *fc, const float4 *fd,
                              const DevLayer *dY, Strides S, const DevWin *dw, const F4TPart *parts,
                              const uint64_t *am, int64_t mwords, const int32_t *ws, float *r4p) {
  hipLaunchKernelGGL(k_convf4_tan<P>, g, dim3(CFT_T), 0, s, T9, fc, fd, dY, S, dw, parts, am, mwords, ws, r4p);
}

// The tangent passes of a batch, after its forward launches (which left
// the scatter records, TAPE9, the compacted LINF4 lines, the accepted
// far-wing lines and the forward R4 rows in the context's work buffers).
static int tangent_launch(lblx_ctx *c, const lblx_run *run, const JobsPrep &J, const TanReq &R, const Recs &recs,
                          const Tape9 &T9, const ShrinkWork &W, int64_t mwords, bool f32out, void *od_base) {
  hipStream_t s = c->stream;
  const int njobs = J.njobs;
  const int64_t NJ = njobs;
  const Strides S = J.S;
  const int PO = R.P;                  // output channels
  const int P = PO > 1 ? PO + 1 : PO;  // internal channels (U when species are asked for)
  const int nwin = (int)J.win.size();
  const int abs_stride = J.abs_stride;
  const bool cont = abs_stride > 0;
  if (PO < 1 || P > TAN_NQ || (int64_t)R.tl.size() != NJ || (int64_t)R.odc_off.size() != NJ * PO || !R.odc ||
      (cont && (int64_t)R.cps.size() != 2 * NJ * PO))
    return LBLX_E_ARG;
  for (int jb = 0; jb < njobs; jb++)
    if (R.tl[jb].nchan != PO || R.tl[jb].nint != P) return LBLX_E_ARG;
  if (!c->t_tabs.p) {  // F1 FG XVER F2 F3 and their slopes per table step, FP32
    std::vector<double> t5(5 * 2601);
    shape_tables_host(t5.data());
    const int ord[5] = {0, 3, 4, 1, 2};  // from F1 F2 F3 FG XVER
    std::vector<float> tt(10 * 2601);
    for (int a = 0; a < 5; a++) {
      const double *f = t5.data() + (size_t)ord[a] * 2601;
      for (int i = 0; i < 2601; i++) {
        // the shapes are even in z: zero slope at the centre
        const double sl = i == 0 ? 0. : (i == 2600 ? f[i] - f[i - 1] : 0.5 * (f[i + 1] - f[i - 1]));
        tt[(size_t)a * 2601 + i] = (float)f[i];
        tt[(size_t)(5 + a) * 2601 + i] = (float)sl;
      }
    }
    if (c->t_tabs.ensure(tt.size() * 4)) return LBLX_E_HIP;
    HIPCHK(hipMemcpy(c->t_tabs.p, tt.data(), tt.size() * 4, hipMemcpyHostToDevice));
  }
  // tile maps of the tangent tiles (k_tan_seg's widths; LDS holds the
  // grid's tables with their slopes)
  const size_t tabb[3] = {6u * 2601 * 4, 2u * 2601 * 4, 2u * 2601 * 4};
  int tw[3];
  std::vector<int32_t> tmv[3];
  for (int G = 0; G < 3; G++) {
    if (G == 0) tw[G] = PO - 1 <= 3 ? TSeg<1, 3>::TILE : TSeg<1, 7>::TILE;
    else tw[G] = G == 1 ? TSeg<2, 3>::TILE : TSeg<3, 3>::TILE;
    std::vector<int32_t> tj, tlc;
    for (int jb = 0; jb < njobs; jb++) {
      const LayerPlan &Pl = *J.plans[jb];
      const int64_t ng[3] = {Pl.g_last + 1, Pl.n2, Pl.n3};
      const int nt = (int)((ng[G] + tw[G] - 1) / tw[G]);
      for (int t = 0; t < nt; t++) {
        tj.push_back(jb);
        tlc.push_back(t);
      }
    }
    tmv[G] = tj;
    tmv[G].insert(tmv[G].end(), tlc.begin(), tlc.end());
  }
  // far-wing parts: a window heavier than the total over the resident
  // workgroups is shared out over several (heaviest parts first)
  std::vector<F4TPart> parts;
  std::vector<int32_t> wfirst(nwin), wnp(nwin);
  {
    std::vector<double> cost;
    window_costs(c, J, cost);
    double tot = 0.;
    for (double x : cost) tot += x;
    const double per = tot / (double)std::max(1, 4 * c->ncu);
    std::vector<std::pair<double, F4TPart>> order;
    int dst = 0;
    for (int w = 0; w < nwin; w++) {
      int np = per > 0. ? (int)std::ceil(cost[w] / per) : 1;
      np = std::min(64, std::max(1, np));
      wfirst[w] = dst;
      wnp[w] = np;
      for (int k = 0; k < np; k++) order.push_back({cost[w] / np, F4TPart{w, k, np, dst + k}});
      dst += np;
    }
    std::stable_sort(order.begin(), order.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    parts.reserve(order.size());
    for (auto &o : order) parts.push_back(o.second);
  }
  const int nparts = (int)parts.size();
  auto need = [&](DevBuf &b, size_t bytes) { return b.ensure(bytes) != 0; };
  if (need(c->t_parts, (size_t)nparts * sizeof(F4TPart) + (size_t)nwin * 8) ||
      need(c->t_r4p, (size_t)nparts * P * R4ROW * 4) ||
      need(c->t_tlay, NJ * sizeof(TanLayer)) || need(c->t_tl, NJ * S.line * sizeof(TanLine)) ||
      need(c->t_t9, NJ * S.line * 4 * P * 4) || need(c->t_r4c, (size_t)nwin * P * R4ROW * 8) ||
      need(c->t_mem, NJ * S.line * sizeof(MemTan)) || need(c->t_memq, NJ * S.line) ||
      need(c->t_f4c, NJ * S.line * sizeof(F4TConst)) || need(c->t_f4d, NJ * S.line * P * sizeof(float4)) ||
      need(c->t_r3c, (size_t)P * J.r3tot * 8) || need(c->t_r2c, (size_t)P * J.r2tot * 8) ||
      need(c->t_odcoff, NJ * PO * 8) || need(c->t_rh, TAN_NQ * 8) ||
      (cont && (need(c->t_cps, 2 * NJ * PO * sizeof(ContPlan)) || need(c->t_absab, 2 * NJ * PO * abs_stride * 8) ||
                need(c->t_absc, NJ * PO * abs_stride * 8)))) {
    set_err("tangent work buffer allocation failed");
    return LBLX_E_HIP;
  }
  for (int G = 0; G < 3; G++)
    if (need(c->t_tmap[G], tmv[G].size() * 8)) return LBLX_E_HIP;
  c->stage_begin();
  HIPCHK(c->h2d(c->t_tlay.p, R.tl.data(), NJ * sizeof(TanLayer), s));
  HIPCHK(c->h2d(c->t_parts.p, parts.data(), (size_t)nparts * sizeof(F4TPart), s));
  {
    int32_t *wf = (int32_t *)(c->t_parts.as<char>() + (size_t)nparts * sizeof(F4TPart));
    HIPCHK(c->h2d(wf, wfirst.data(), (size_t)nwin * 4, s));
    HIPCHK(c->h2d(wf + nwin, wnp.data(), (size_t)nwin * 4, s));
  }
  HIPCHK(c->h2d(c->t_odcoff.p, R.odc_off.data(), NJ * PO * 8, s));
  HIPCHK(c->h2d(c->t_rh.p, R.rh, TAN_NQ * 8, s));
  if (cont) HIPCHK(c->h2d(c->t_cps.p, R.cps.data(), 2 * NJ * PO * sizeof(ContPlan), s));
  TileMap tm[3];
  for (int G = 0; G < 3; G++) {
    const size_t nt = tmv[G].size() / 2;
    int32_t *base = c->t_tmap[G].as<int32_t>();
    HIPCHK(c->h2d(base, tmv[G].data(), nt * 8, s));
    tm[G].job = base;
    tm[G].loc = base + nt;
    tm[G].lo = base + 2 * nt;
    tm[G].hi = base + 3 * nt;
  }
  c->stage_end(s);
  const DevLayer *dY = c->w_dlay.as<DevLayer>();
  const TanLayer *dT = c->t_tlay.as<TanLayer>();
  const DevPanel *dp = c->w_pan.as<DevPanel>();
  const DevWin *dw = c->w_win.as<DevWin>();
  const DevSeg *dg = c->w_seg.as<DevSeg>();
  DevLines L = c->lines();
  TanLine *tl = c->t_tl.as<TanLine>();
  float *t9 = c->t_t9.as<float>();
  double *r4c = c->t_r4c.as<double>();
  int th = c->tbegin(ST_TANGENT);
  if (J.maxnn > 0)
    hipLaunchKernelGGL(k_near_tan, dim3((J.maxnn + 255) / 256, njobs), dim3(256), 0, s, L, dY, dT, S,
                       (const int *)c->w_keep.p, (const int32_t *)c->w_kidx.p, tl);
  const int lgrid = (int)((J.maxline + 1 + 255) / 256);
  MemTan *mt = c->t_mem.as<MemTan>();
  int8_t *mq = c->t_memq.as<int8_t>();
  hipLaunchKernelGGL(k_shr_mem_tan, dim3(lgrid, njobs), dim3(256), 0, s, L, dY, dT, S, W, mt, mq);
  hipLaunchKernelGGL(k_shr_emit_tan, dim3(lgrid, njobs), dim3(256), 0, s, dT, S, W, T9, (const MemTan *)mt,
                     (const int8_t *)mq, t9);
  F4TConst *fc = c->t_f4c.as<F4TConst>();
  float4 *fd = c->t_f4d.as<float4>();
  hipLaunchKernelGGL(k_f4_tan_lines, dim3(lgrid, njobs), dim3(256), 0, s, T9, (const float *)t9, dY, dT, S, fc, fd);
  {
    const dim3 g(nparts);
    const uint64_t *am = c->t_amask.as<uint64_t>();
    const int32_t *ws = c->t_wspan.as<int32_t>();
    const F4TConst *cfc = fc;
    const float4 *cfd = fd;
    const F4TPart *pp = c->t_parts.as<F4TPart>();
    float *r4p = c->t_r4p.as<float>();
    switch (P) {
      case 1: launch_convf4_tan<1>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 2: launch_convf4_tan<2>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 3: launch_convf4_tan<3>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 4: launch_convf4_tan<4>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 5: launch_convf4_tan<5>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 6: launch_convf4_tan<6>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 7: launch_convf4_tan<7>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      case 8: launch_convf4_tan<8>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
      default: launch_convf4_tan<9>(g, s, T9, cfc, cfd, dY, S, dw, pp, am, mwords, ws, r4p); break;
    }
    const int32_t *wf = (const int32_t *)(c->t_parts.as<char>() + (size_t)nparts * sizeof(F4TPart));
    hipLaunchKernelGGL(k_r4_tan_sum, dim3((unsigned)((P * R4ROW + 255) / 256), nwin), dim3(256), 0, s, dY, dT, dw,
                       dg, wf, wf + nwin, (const float *)r4p, P, (const double *)c->w_r4.as<double>(), r4c);
  }
  if (cont) {  // the output channels' continuum (U carries none)
    const int64_t nplan = 2 * NJ * PO;
    const int64_t nth = nplan * abs_stride;
    hipLaunchKernelGGL(k_contnm, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, s, c->t_cps.as<ContPlan>(),
                       c->cont_tabs, abs_stride, (int)nplan, c->t_absab.as<double>());
    const int64_t nd = NJ * PO * abs_stride;
    hipLaunchKernelGGL(k_cont_diff, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s,
                       (const double *)c->t_absab.as<double>(), abs_stride, njobs, PO,
                       (const double *)c->t_rh.as<double>(), c->t_absc.as<double>());
  }
  const float *tabs = c->t_tabs.as<float>();
  const double *absc = cont ? c->t_absc.as<double>() : nullptr;
  double *r3c = c->t_r3c.as<double>(), *r2c = c->t_r2c.as<double>();
  const int64_t *odcoff = c->t_odcoff.as<int64_t>();
  for (int G = 3; G >= 1; G--) {
    const int T = tw[G - 1];
    const int nt = (int)(tmv[G - 1].size() / 2);
    hipLaunchKernelGGL(k_tile_ranges, dim3((nt + 255) / 256), dim3(256), 0, s, c->d_vnu.as<double>(), dY, tm[G - 1],
                       nt, T, G, c->w_margin.as<double>(), (const int32_t *)nullptr, S);
    const int nsp = PO - 1;
    const int NSPsel = nsp == 0 ? 0 : (nsp <= 3 ? 3 : 7);
    const size_t shm = tabb[G - 1];
    const double *vr = c->d_vnu.as<double>();
    const double *mgn = c->w_margin.as<double>();
#define LBLX_TSEG(GG, OUTT, NSPV, RIN, SIN, ROUT, SOUT)                                                             \
  do {                                                                                                             \
    HIPCHK(hipFuncSetAttribute((const void *)k_tan_seg<GG, OUTT, NSPV>, hipFuncAttributeMaxDynamicSharedMemorySize,  \
                               (int)shm));                                                                         \
    hipLaunchKernelGGL((k_tan_seg<GG, OUTT, NSPV>), dim3(nt), dim3(TSEG_T), shm, s, recs, (const TanLine *)tl,       \
                       tm[GG - 1], vr, (const int32_t *)c->w_kidx.p, mgn, tabs, dY, dT, S, dp, dw,                 \
                       (const double *)r4c, absc, abs_stride, dg,                                                  \
                       (const double *)(RIN), (int64_t)(SIN), (double *)(ROUT), (int64_t)(SOUT),                   \
                       (const OUTT *)od_base, (OUTT *)R.odc, odcoff);                                              \
  } while (0)
#define LBLX_TSEG_N(GG, OUTT, RIN, SIN, ROUT, SOUT)                      \
  do {                                                                  \
    if (NSPsel == 0) LBLX_TSEG(GG, OUTT, 0, RIN, SIN, ROUT, SOUT);      \
    else if (NSPsel == 3) LBLX_TSEG(GG, OUTT, 3, RIN, SIN, ROUT, SOUT); \
    else LBLX_TSEG(GG, OUTT, 7, RIN, SIN, ROUT, SOUT);                  \
  } while (0)
    if (G == 3) LBLX_TSEG_N(3, double, nullptr, 0, r3c, J.r3tot);
    else if (G == 2) LBLX_TSEG_N(2, double, r3c, J.r3tot, r2c, J.r2tot);
    else if (f32out) LBLX_TSEG_N(1, float, r2c, J.r2tot, nullptr, 0);
    else LBLX_TSEG_N(1, double, r2c, J.r2tot, nullptr, 0);
#undef LBLX_TSEG_N
#undef LBLX_TSEG
  }
  c->tend(th);
  HIPCHK(hipGetLastError());
  (void)run;
  return LBLX_OK;
}

// Allocate, upload and launch a planned batch.  od_base + od_off[j]
// receives job j's npts values.
static int launch_jobs(lblx_ctx *c, const lblx_run *run, const JobsPrep &J, void *od_base, bool f32out,
                       double *v1p, lblx_stats *st, const TanReq *tan = nullptr) {
  // FP32 split: line shapes evaluated and accumulated in FP32, positions,
  // grid offsets and the panel combine in FP64; OD written as float.
  const bool split = run->precision == LBLX_FP32_SPLIT;
  hipStream_t s = c->stream;
  const int njobs = J.njobs;
  const int nblk = c->nblk;
  const auto &plans = J.plans;
  const auto &ys = J.ys;
  const auto &pan = J.pan;
  const auto &win = J.win;
  const auto &seg = J.seg;
  const auto &margins = J.margins;
  const auto &cps = J.cps;
  const auto &tjob = J.tjob;
  const auto &tloc = J.tloc;
  const int *tiles = J.tiles;
  const Strides S = J.S;
  const int64_t r2tot = J.r2tot, r3tot = J.r3tot, maxline = J.maxline;
  const int maxnn = J.maxnn, maxnbk = J.maxnbk, abs_stride = J.abs_stride, maxnb4 = J.maxnb4;
  const int nwin = (int)win.size();
  c->last_line_stride = S.line;
  const bool voigt = run->voigt != 0;
  // ILBLF4 = 0: no far wing; R4 stays zero, so LNCOR1 rejects against
  // DPTMIN alone (oprop.f90:1156-1160) and XINT(R4) adds nothing
  const bool nof4 = run->ilblf4 == 0;
  if (tan && voigt) return LBLX_E_UNSUPPORTED;  // the tangent follows the tabulated shapes
  if (tan && c->ncpl > 0) return LBLX_E_UNSUPPORTED;  // the tangent has no line-coupling terms
  // below 25 cm-1 + the far-wing bound: the VVW mirror in CONVF4 (and RSYM
  // for a first panel at or below 0 cm-1)
  bool mir = false, fold = false;
  for (int jb = 0; jb < njobs; jb++) {
    mir = mir || (!nof4 && plans[jb]->v1l4 <= 25.);
    fold = fold || plans[jb]->vft0 <= 0.;
  }
  if (tan && (mir || fold)) return LBLX_E_UNSUPPORTED;  // the tangent has neither
  if (tan && nof4) return LBLX_E_UNSUPPORTED;   // the tangent differentiates LBLF4's decisions
  if (voigt) {
    int rv = ensure_voigt(c);
    if (rv) return rv;
  }
  const int64_t NJ = njobs;
  auto need = [&](DevBuf &b, size_t bytes) { return b.ensure(bytes) != 0; };
  if (need(c->w_dlay, NJ * sizeof(DevLayer)) || need(c->w_pan, pan.size() * sizeof(DevPanel)) ||
      need(c->w_win, win.size() * sizeof(DevWin)) || need(c->w_word, win.size() * 4) || need(c->w_seg, seg.size() * sizeof(DevSeg)) ||
      need(c->w_r4, (size_t)nwin * R4ROW * 8) ||
      need(c->w_cx, NJ * S.line * 8) || need(c->w_csp, NJ * S.line * 8) ||
      (c->ncpl > 0 && (need(c->w_cspp, NJ * S.line * 8) || need(c->w_t9spp, NJ * S.line * 8) ||
                       need(c->w_cpl, NJ * S.line * 8))) ||
      need(c->w_cal, NJ * S.line * 8) || need(c->w_cad, NJ * S.line * 8) ||
      need(c->w_cfl, NJ * S.line) || need(c->w_g, NJ * S.line * 4) ||
      need(c->w_jmp, NJ * S.line * 4 * 3) || need(c->w_wsum, NJ * S.line * 2) ||
      need(c->w_bcnt, NJ * S.blk * 4 * 3) || need(c->w_etab, NJ * S.blk * 4 * (EBLK + 1)) || need(c->w_t9s, NJ * S.blk * 4 * 6) ||
      need(c->w_jobi, NJ * 4 * 4) || need(c->w_bat, NJ * S.line * 4 * 3) ||
      need(c->w_t9vnu, NJ * S.line * 8) || need(c->w_t9sp, NJ * S.line * 8) ||
      need(c->w_t9alfa, NJ * S.line * 8) || need(c->w_t9epp, NJ * S.line * 8) ||
      need(c->w_t9mol, NJ * S.line * 4) || need(c->w_t9n, NJ * 8) ||
      need(c->w_t9blk, NJ * S.t9blk * 16) || need(c->w_t9bvhi, NJ * S.t9blk * 8) ||
      need(c->w_near, NJ * S.line * (sizeof(NearHot) + sizeof(NearCold))) || need(c->w_keep, NJ * S.line * 4) || need(c->w_kidx, (NJ * S.line + 2 * NJ * S.blk) * 4) ||
      need(c->w_blkmax, NJ * S.blk * 4) || need(c->w_pread, NJ * S.blk * 4) ||
      need(c->w_dirty, NJ * S.blk * 4) ||
      need(c->w_changed, 64 * 4) || need(c->w_rec, NJ * S.line * (sizeof(SRec1) + 2 * sizeof(SRecA) + (voigt ? sizeof(SRecV) : 0))) ||
      need(c->w_r3, (size_t)r3tot * 8) || need(c->w_r2, (size_t)r2tot * 8) ||
      need(c->w_cnt, NJ * 24) || need(c->w_margin, NJ * 3 * 8) ||
      (abs_stride > 0 && (need(c->w_absrb, NJ * abs_stride * 8) ||
                          need(c->w_cps, NJ * sizeof(ContPlan))))) {
    set_err("work buffer allocation failed");
    return LBLX_E_HIP;
  }
  for (int G = 0; G < 3; G++)
    if (need(c->w_tmap[G], tjob[G].size() * 16)) return LBLX_E_HIP;
  // the tangent passes read back the compacted lines' records and the
  // far-wing lines each window accepted (one word per 64-line chunk)
  const int64_t mwords = S.line / 64 + 1;
  if (tan && (need(c->t_craw, NJ * S.line * 4) || need(c->t_amask, (size_t)nwin * mwords * 8) ||
              need(c->t_wspan, (size_t)nwin * 8))) {
    set_err("tangent work buffer allocation failed");
    return LBLX_E_HIP;
  }
  c->stage_begin();
  HIPCHK(c->h2d(c->w_dlay.p, ys.data(), NJ * sizeof(DevLayer), s));
  HIPCHK(c->h2d(c->w_pan.p, pan.data(), pan.size() * sizeof(DevPanel), s));
  HIPCHK(c->h2d(c->w_win.p, win.data(), win.size() * sizeof(DevWin), s));
  HIPCHK(c->h2d(c->w_word.p, J.worder.data(), win.size() * 4, s));
  HIPCHK(c->h2d(c->w_seg.p, seg.data(), seg.size() * sizeof(DevSeg), s));
  HIPCHK(c->h2d(c->w_margin.p, margins.data(), margins.size() * 8, s));
  if (abs_stride > 0) {  // CONTNM: ABSRB rows of every job
    HIPCHK(c->h2d(c->w_cps.p, cps.data(), NJ * sizeof(ContPlan), s));
    int thc = c->tbegin(ST_CONTNM);
    int64_t nth = NJ * abs_stride;
    hipLaunchKernelGGL(k_contnm, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, s,
                       c->w_cps.as<ContPlan>(), c->cont_tabs, abs_stride, njobs,
                       c->w_absrb.as<double>());
    c->tend(thc);
    HIPCHK(hipGetLastError());
  }
  TileMap tm[3];
  for (int G = 0; G < 3; G++) {
    size_t nt = tjob[G].size();
    int32_t *base = c->w_tmap[G].as<int32_t>();
    HIPCHK(c->h2d(base, tjob[G].data(), nt * 4, s));
    HIPCHK(c->h2d(base + nt, tloc[G].data(), nt * 4, s));
    tm[G].job = base;
    tm[G].loc = base + nt;
    tm[G].lo = base + 2 * nt;
    tm[G].hi = base + 3 * nt;
  }
  c->stage_end(s);
  const DevLayer *dY = c->w_dlay.as<DevLayer>();
  DevLines L = c->lines();
  ShrinkWork W;
  W.cx = c->w_cx.as<double>();
  W.csp = c->w_csp.as<double>();
  W.cspp = c->ncpl > 0 ? c->w_cspp.as<double>() : nullptr;
  W.cal = c->w_cal.as<double>();
  W.cad = c->w_cad.as<double>();
  W.cfl = c->w_cfl.as<int8_t>();
  W.g = c->w_g.as<int32_t>();
  W.jmp = c->w_jmp.as<int32_t>();
  W.grp = W.jmp + NJ * S.line;
  W.gout = W.jmp + 2 * NJ * S.line;
  W.wsum = c->w_wsum.as<int16_t>();
  int32_t *bb = c->w_bcnt.as<int32_t>();
  W.bcnt = bb;
  W.bstop = bb + NJ * S.blk;
  W.bp = bb + 2 * NJ * S.blk;
  int32_t *ji = c->w_jobi.as<int32_t>();
  W.npass = ji;
  W.stop = ji + NJ;
  W.nbat = ji + 2 * NJ;
  W.etab = c->w_etab.as<int32_t>();
  W.estop = W.etab + NJ * S.blk * EBLK;
  int32_t *bat = c->w_bat.as<int32_t>();
  W.bat_s = bat;
  W.bat_t = bat + NJ * S.line;
  W.bat_off = bat + 2 * NJ * S.line;
  int32_t *t9s = c->w_t9s.as<int32_t>();
  W.t9nxt = t9s;
  W.t9nem = t9s + NJ * S.blk;
  W.t9nbt = t9s + 2 * NJ * S.blk;
  W.t9out = t9s + 3 * NJ * S.blk;
  W.t9bat = t9s + 4 * NJ * S.blk;
  W.t9k = t9s + 5 * NJ * S.blk;
  W.craw = tan ? c->t_craw.as<int32_t>() : nullptr;
  Tape9 T9;
  T9.vnu = c->w_t9vnu.as<double>();
  T9.sp = c->w_t9sp.as<double>();
  T9.alfa = c->w_t9alfa.as<double>();
  T9.epp = c->w_t9epp.as<double>();
  T9.spp = c->ncpl > 0 ? c->w_t9spp.as<double>() : nullptr;
  T9.mol = c->w_t9mol.as<int32_t>();
  T9.n = c->w_t9n.as<int32_t>();
  T9.nblk = c->w_t9n.as<int32_t>() + NJ;
  T9.bstart = c->w_t9blk.as<int32_t>();
  T9.bcount = c->w_t9blk.as<int32_t>() + NJ * S.t9blk;
  T9.bvhi = c->w_t9bvhi.as<double>();
  unsigned long long *cnt = c->w_cnt.as<unsigned long long>();  // [3][njobs]: k1, k3, lines
  hipLaunchKernelGGL(k_fill_i32, dim3((6 * NJ + 255) / 256), dim3(256), 0, s, (int *)cnt, 6 * NJ, 0);

  NearHot *nearh = c->w_near.as<NearHot>();
  NearCold *nearc = (NearCold *)(c->w_near.as<char>() + (size_t)NJ * S.line * sizeof(NearHot));
  // 1-2. LINF4 + SHRINK
  int th = c->tbegin(ST_SHRINK);
  const int lgrid = (int)((maxline + 1 + 255) / 256);
  const int bgrid = std::max(maxnb4, 1);
  if (nof4) {  // no TAPE9: its line and block counts are zero
    hipLaunchKernelGGL(k_fill_i32, dim3((int)((2 * NJ + 255) / 256)), dim3(256), 0, s, T9.n, 2 * NJ, 0);
  } else {
  hipLaunchKernelGGL(k_shr_count, dim3(bgrid, njobs), dim3(256), 0, s, L, dY, S, W);
  hipLaunchKernelGGL(k_shr_scan, dim3(njobs), dim3(256), 0, s, dY, S, W);
  hipLaunchKernelGGL(k_shr_compact, dim3(bgrid, njobs), dim3(256), 0, s, L, dY, S, W);
  hipLaunchKernelGGL(k_shr_ewalk, dim3(bgrid, njobs), dim3(64), 0, s, dY, S, W);
  hipLaunchKernelGGL(k_shr_t9sim, dim3((bgrid + 63) / 64, njobs), dim3(64), 0, s, dY, S, W, c->t9_all_wave);
  hipLaunchKernelGGL(k_fill_i32, dim3((int)((NJ * S.blk + 255) / 256)), dim3(256), 0, s, W.t9k, NJ * S.blk, -1);
  {  // a job's start tables in LDS when they fit (three ints per block)
    int lcap = std::min(maxnb4, (int)((150u << 10) / 12u));
    if (lcap < 1) lcap = 1;
    const size_t shm = (size_t)lcap * 12;
    HIPCHK(hipFuncSetAttribute((const void *)k_shr_t9chain, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
    hipLaunchKernelGGL(k_shr_t9chain, dim3(njobs), dim3(64), shm, s, dY, S, W, T9, lcap);
  }
  hipLaunchKernelGGL(k_shr_t9replay, dim3(bgrid, njobs), dim3(64), 0, s, dY, S, W, T9);
  hipLaunchKernelGGL(k_shr_bwalk, dim3(bgrid, njobs), dim3(64), 0, s, S, W);
  hipLaunchKernelGGL(k_shr_emit, dim3(lgrid, njobs), dim3(256), 0, s, dY, S, W, T9);
  hipLaunchKernelGGL(k_t9_vhi, dim3((int)((S.t9blk + 255) / 256), njobs), dim3(256), 0, s, S, T9);
  }
  c->tend(th);
  if (maxnn > 0) {
    // LNCOR1 reads only the line table and the layer/panel tables, so it runs
    // on the side stream beside CONVF4, which is bound by FP64 issue and
    // moves little HBM traffic (beside SHRINK it would slow SHRINK's
    // streaming passes); the rejection fixpoint waits for it
    hipStream_t s2 = c->side;
    HIPCHK(hipEventRecord(c->ev_fork, s));
    HIPCHK(hipStreamWaitEvent(s2, c->ev_fork, 0));
    const int ngrid = (maxnn + 255) / 256;
    int thl = c->tbegin(ST_LNCOR1, s2);
    hipLaunchKernelGGL(k_lncor1, dim3(ngrid, njobs), dim3(256), 0, s2, L, dY, S,
                       c->w_pan.as<DevPanel>(), nearh, nearc);
    hipLaunchKernelGGL(k_fill_i32, dim3((int)((NJ * S.blk + 255) / 256)), dim3(256), 0, s2,
                       c->w_pread.as<int>(), NJ * S.blk, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3((int)((NJ * S.blk + 255) / 256)), dim3(256), 0, s2,
                       c->w_blkmax.as<int>(), NJ * S.blk, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3((int)((NJ * S.blk + 255) / 256)), dim3(256), 0, s2,
                       c->w_dirty.as<int>(), NJ * S.blk, 1);
    c->tend(thl, s2);
    HIPCHK(hipEventRecord(c->ev_join, s2));
  }
  // 3. far-wing windows
  th = c->tbegin(ST_CONVF4);
  {
    uint64_t *am = tan ? c->t_amask.as<uint64_t>() : nullptr;
    int32_t *ws = tan ? c->t_wspan.as<int32_t>() : nullptr;
    if (nof4 && nwin > 0)
      hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)(((int64_t)nwin * R4ROW + 255) / 256)), dim3(256), 0, s,
                         c->w_r4.as<double>(), (int64_t)nwin * R4ROW, 0.0);
    else if (nof4)
      ;
    else {
      // coupled lines (CPL) and the mirror of lines below 25 cm-1 (MIR)
      // each take their own build of the kernel
      using KF = void (*)(Tape9, const DevLayer *, Strides, const DevWin *, const DevSeg *, const int32_t *,
                          double *, unsigned long long *, uint64_t *, int64_t, int32_t *);
      const bool cp = c->ncpl > 0;
      const KF kd[2][2] = {{k_convf4<double, false, false>, k_convf4<double, false, true>},
                           {k_convf4<double, true, false>, k_convf4<double, true, true>}};
      const KF kf[2][2] = {{k_convf4<float, false, false>, k_convf4<float, false, true>},
                           {k_convf4<float, true, false>, k_convf4<float, true, true>}};
      const KF k = split ? kf[cp][mir] : kd[cp][mir];
      hipLaunchKernelGGL(k, dim3(nwin), dim3(CF4_T), 0, s, T9, dY, S, c->w_win.as<DevWin>(),
                         c->w_seg.as<DevSeg>(), c->w_word.as<int32_t>(), c->w_r4.as<double>(), cnt + NJ, am, mwords,
                         ws);
    }
  }
  c->tend(th);
  Recs recs;
  {
    char *rb = c->w_rec.as<char>();
    const size_t nl = (size_t)NJ * S.line;
    recs.r1 = (SRec1 *)rb;
    recs.r2 = (SRecA *)(rb + nl * sizeof(SRec1));
    recs.r3 = recs.r2 + nl;
    recs.v = voigt ? (SRecV *)(rb + nl * (sizeof(SRec1) + 2 * sizeof(SRecA))) : nullptr;
    recs.cpl = c->ncpl > 0 ? c->w_cpl.as<double>() : nullptr;
  }
  c->fix_passes = 0;
  c->fix_seq = 0;
  if (maxnn > 0) {
    const int ngrid = (maxnn + 255) / 256;
    // 4-5. block-read panel fixpoint, after LNCOR1 on the side stream
    HIPCHK(hipStreamWaitEvent(s, c->ev_join, 0));
    th = c->tbegin(ST_REJECT);
    // Pass it raises ch[it] when a read panel moved; the fixpoint holds at
    // the first pass after the first that moves none.  The first passes
    // are queued without a host check: from the third on, a pass whose
    // predecessor moved nothing is a no-op (nothing is dirty) and returns
    // at once, so ch[it] of the last queued pass is 0 exactly when the
    // fixpoint has been reached.
    constexpr int QUEUED = 4;
    int *ch = c->w_changed.as<int>();
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, ch, 64, 0);
    bool converged = false;
    int passes = 0;
    const int maxpass = c->fix_maxpass;
    for (int it = 0; it < maxpass; it++) {
      const int *skip = it >= 2 ? ch + it - 1 : nullptr;
      hipLaunchKernelGGL(k_reject, dim3(ngrid, njobs), dim3(256), 0, s, (const NearHot *)nearh, L, dY,
                         S, c->w_pread.as<int>(), c->w_pan.as<DevPanel>(), c->w_win.as<DevWin>(),
                         c->w_r4.as<double>(), c->w_keep.as<int>(), c->w_blkmax.as<int>(),
                         c->w_dirty.as<int>(), skip);
      hipLaunchKernelGGL(k_pread_scan, dim3(njobs), dim3(256), 0, s, dY, S, c->w_blkmax.as<int>(),
                         c->w_pread.as<int>(), ch + it, c->w_dirty.as<int>(), skip);
      if (it >= QUEUED - 1 || it == maxpass - 1) {
        int flags[64];
        HIPCHK(hipMemcpyAsync(flags, ch, (it + 1) * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (!flags[it]) {
          converged = true;  // this pass already used the final read panels
          passes = it + 1;
          while (passes > 1 && !flags[passes - 2]) passes--;  // queued no-op passes
          break;
        }
      }
    }
    // Not settled within the budget (a read panel that moves back and forth
    // between passes): solve the recurrence block by block instead.
    if (!converged) {
      passes = maxpass;
      hipLaunchKernelGGL(k_reject_seq, dim3(njobs), dim3(64), 0, s, (const NearHot *)nearh, L, dY, S,
                         c->w_pread.as<int>(), c->w_pan.as<DevPanel>(), c->w_win.as<DevWin>(),
                         c->w_r4.as<double>(), c->w_keep.as<int>(), c->w_blkmax.as<int>());
    }
    c->fix_passes = passes;
    c->fix_seq = converged ? 0 : 1;
    c->tend(th);
    // 6. CNVFNV scatter records
    th = c->tbegin(ST_SCATTER);
    int32_t *kidx = c->w_kidx.as<int32_t>(), *bkc = kidx + NJ * S.line, *bkp = bkc + NJ * S.blk;
    hipLaunchKernelGGL(k_keep_blk, dim3(maxnbk, njobs), dim3(256), 0, s, dY, L, S, c->w_keep.as<int>(), bkc);
    hipLaunchKernelGGL(k_keep_scan, dim3(njobs), dim3(256), 0, s, dY, S, (const int32_t *)bkc, bkp, kidx);
    hipLaunchKernelGGL(k_scatter_rec, dim3(maxnbk, njobs), dim3(256), 0, s, (const NearHot *)nearh,
                       (const NearCold *)nearc, c->w_keep.as<int>(), L, dY, S, c->w_pread.as<int>(),
                       c->w_pan.as<DevPanel>(), (const int32_t *)bkp, kidx, recs, st ? cnt : nullptr,
                       cnt + 2 * NJ);  // TAPE6 counters with st
    c->tend(th);
  }
  // 7. tiles: R3, R2, R1 (+ PANEL combine)
  const double *tabs = c->d_tabs.as<double>();
  const double *d_abs = abs_stride > 0 ? c->w_absrb.as<double>() : nullptr;
  for (int G = 3; G >= 1; G--) {
    const int T = tiles[G - 1];
    const int CH = c->lbatch[G - 1] > 0 ? c->lbatch[G - 1] : ((G == 1 && !split) ? 16 : 32);
    const int nt = (int)tjob[G - 1].size();
    int tht = c->tbegin(G == 3 ? ST_TILE3 : (G == 2 ? ST_TILE2 : ST_TILE1));
    hipLaunchKernelGGL(k_tile_ranges, dim3((nt + 255) / 256), dim3(256), 0, s, c->d_vnu.as<double>(), dY,
                       tm[G - 1], nt, T, G, c->w_margin.as<double>(), (const int32_t *)c->w_kidx.p, S);
    const size_t esz = split ? sizeof(float) : sizeof(double);
    // FP32 shapes: accumulator in FP64 LDS for the grids in LBLX_FP32_LDS64
    const bool l64 = !split || ((LBLX_FP32_LDS64 >> G) & 1);
    size_t shm = (size_t)(T + 1) * (l64 ? 8 : esz) + (size_t)(voigt ? 0 : (G == 1 ? 3 : 1) * 2601) * esz;
    const Recs rec = recs;
    const DevPanel *dp = c->w_pan.as<DevPanel>();
    const DevWin *dw = c->w_win.as<DevWin>();
    const DevSeg *dg = c->w_seg.as<DevSeg>();
    const double *r4 = c->w_r4.as<double>();
    double *r3 = c->w_r3.as<double>(), *r2 = c->w_r2.as<double>();
#define LBLX_TILE_L(GG, OUTT, ACCT, VV, LACCT, VTAB, RIN, ROUT, ODP, ABS)                         \
  do {                                                                                             \
    if (c->ncpl > 0) LBLX_TILE_C(GG, OUTT, ACCT, VV, LACCT, true, VTAB, RIN, ROUT, ODP, ABS);       \
    else LBLX_TILE_C(GG, OUTT, ACCT, VV, LACCT, false, VTAB, RIN, ROUT, ODP, ABS);                 \
  } while (0)
#define LBLX_TILE_C(GG, OUTT, ACCT, VV, LACCT, CP, VTAB, RIN, ROUT, ODP, ABS)                     \
  do {                                                                                             \
    HIPCHK(hipFuncSetAttribute((const void *)k_tile<GG, OUTT, ACCT, VV, LACCT, CP>,                \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));             \
    hipLaunchKernelGGL((k_tile<GG, OUTT, ACCT, VV, LACCT, CP>), dim3(nt), dim3(TILE_T), shm, s, rec, \
                       tm[GG - 1], T, CH, tabs, (const ACCT *)(VTAB), dY, S, dp, dw, r4,           \
                       (const double *)(ABS), dg, (const double *)(RIN), (double *)(ROUT),         \
                       (OUTT *)(ODP));                                                            \
  } while (0)
#define LBLX_TILE_V(GG, OUTT, ACCT, VV, VTAB, RIN, ROUT, ODP, ABS)                                \
  do {                                                                                             \
    if (l64) LBLX_TILE_L(GG, OUTT, ACCT, VV, double, VTAB, RIN, ROUT, ODP, ABS);                   \
    else LBLX_TILE_L(GG, OUTT, ACCT, VV, ACCT, VTAB, RIN, ROUT, ODP, ABS);                         \
  } while (0)
#define LBLX_TILE(GG, OUTT, ACCT, RIN, ROUT, ODP, ABS)                                             \
  do {                                                                                             \
    if (voigt)                                                                                     \
      LBLX_TILE_V(GG, OUTT, ACCT, true, (sizeof(ACCT) == 4 ? c->d_vtabf.p : c->d_vtab.p), RIN,     \
                  ROUT, ODP, ABS);                                                                 \
    else                                                                                           \
      LBLX_TILE_V(GG, OUTT, ACCT, false, (sizeof(ACCT) == 4 ? c->d_vtabf.p : c->d_vtab.p), RIN,    \
                  ROUT, ODP, ABS);                                                                 \
  } while (0)
    if (G == 3) {
      if (split) LBLX_TILE(3, double, float, nullptr, r3, nullptr, d_abs);
      else LBLX_TILE(3, double, double, nullptr, r3, nullptr, d_abs);
    } else if (G == 2) {
      if (split) LBLX_TILE(2, double, float, r3, r2, nullptr, nullptr);
      else LBLX_TILE(2, double, double, r3, r2, nullptr, nullptr);
    } else if (!f32out) {
      LBLX_TILE(1, double, double, r2, nullptr, od_base, nullptr);
    } else {
      LBLX_TILE(1, float, float, r2, nullptr, od_base, nullptr);
    }
#undef LBLX_TILE
#undef LBLX_TILE_V
#undef LBLX_TILE_L
#undef LBLX_TILE_C
    c->tend(tht);
  }
  HIPCHK(hipGetLastError());
  if (tan) {
    int rt = tangent_launch(c, run, J, *tan, recs, T9, W, mwords, f32out, od_base);
    if (rt) return rt;
  }
  if (v1p)
    for (int jb = 0; jb < njobs; jb++) v1p[jb] = plans[jb]->panels[0].vft + (double)NSHIFT * plans[jb]->dv;
  if (st) {
    std::vector<unsigned long long> h(3 * NJ);
    std::vector<int32_t> t9n(NJ);
    HIPCHK(hipMemcpyAsync(h.data(), cnt, 3 * NJ * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(t9n.data(), T9.n, NJ * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int jb = 0; jb < njobs; jb++) {
      std::memset(&st[jb], 0, sizeof(lblx_stats));
      st[jb].k1_updates = (int64_t)h[jb];
      st[jb].k3_updates = (int64_t)h[NJ + jb];
      st[jb].lincnt = (int64_t)h[2 * NJ + jb];
      st[jb].nlins4 = t9n[jb];
      st[jb].npanel = (int64_t)plans[jb]->panels.size();
      st[jb].npts = plans[jb]->npts;
      st[jb].nwin4 = nof4 ? 0 : (int64_t)plans[jb]->wins.size();
      st[jb].nfix = c->fix_passes;
      st[jb].fixseq = c->fix_seq;
    }
  }
  return LBLX_OK;
}

// OPDPTH for a batch of jobs (layers): every stage one launch for all jobs.
static int run_jobs(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int njobs, bool chain_jrad,
                    void *od_base, const int64_t *od_off, bool f32out, double *v1p, lblx_stats *st,
                    const LayerPlan *pre = nullptr, const TanReq *tan = nullptr) {
  c->hbegin();
  // the job tables live with the context, so their storage is reused
  if (!c->jprep) c->jprep = std::shared_ptr<void>(new JobsPrep(), [](void *p) { delete (JobsPrep *)p; });
  JobsPrep &J = *(JobsPrep *)c->jprep.get();
  int rc = prep_jobs(c, run, layers, njobs, chain_jrad, od_off, pre, J);
  c->hend();
  if (rc) return rc;
  return launch_jobs(c, run, J, od_base, f32out, v1p, st, tan);
}

extern "C" {

int lblx_od_layer_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layer, void *d_od,
                         int64_t n_out, double *v1p, lblx_stats *st) {
  if (!c || !run || !layer || !d_od) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  int64_t npts = 0;
  int rc = lblx_od_size(run, layer, &npts);
  if (rc) return rc;
  if (n_out < npts) return LBLX_E_SIZE;
  int64_t off = 0;
  rc = run_jobs(c, run, layer, 1, false, d_od, &off, run->precision == LBLX_FP32_SPLIT, v1p, st);
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
  int64_t off = 0;
  rc = run_jobs(c, run, layer, 1, false, c->w_od.p, &off, esz == 4, v1p, st);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(od, c->w_od.p, (size_t)npts * esz, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBLX_OK;
}

int lblx_od_column_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay,
                          void *d_od, const int64_t *od_off, double *v1p, lblx_stats *st) {
  if (!c || !run || !layers || nlay <= 0 || !d_od || !od_off) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  return run_jobs(c, run, layers, nlay, true, d_od, od_off, run->precision == LBLX_FP32_SPLIT, v1p, st);
}

// Copy job `job`'s shrunk far-wing line list (TAPE9) of the last run to
// host memory (diagnostics and tests).
int lblx_debug_tape9(lblx_ctx *c, int job, double *vnu, double *sp, double *alfa, double *epp,
                     int32_t *mol, int cap, int32_t *n_out) {
  if (!c || !n_out || job < 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const int64_t line = c->last_line_stride;
  int32_t n = 0;
  HIPCHK(hipMemcpy(&n, c->w_t9n.as<int32_t>() + job, 4, hipMemcpyDeviceToHost));
  *n_out = n;
  int k = n < cap ? n : cap;
  if (k > 0) {
    int64_t o = job * line;
    HIPCHK(hipMemcpy(vnu, c->w_t9vnu.as<double>() + o, k * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(sp, c->w_t9sp.as<double>() + o, k * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(alfa, c->w_t9alfa.as<double>() + o, k * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(epp, c->w_t9epp.as<double>() + o, k * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(mol, c->w_t9mol.as<int32_t>() + o, k * 4, hipMemcpyDeviceToHost));
  }
  return LBLX_OK;
}

int lblx_debug_tape9_spp(lblx_ctx *c, int job, double *spp, int cap, int32_t *n_out) {
  if (!c || !n_out || job < 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  int32_t n = 0;
  HIPCHK(hipMemcpy(&n, c->w_t9n.as<int32_t>() + job, 4, hipMemcpyDeviceToHost));
  *n_out = n;
  const int k = n < cap ? n : cap;
  if (k > 0 && c->ncpl == 0) std::memset(spp, 0, (size_t)k * 8);  // no coupled lines: SPP is 0 throughout
  else if (k > 0)
    HIPCHK(hipMemcpy(spp, c->w_t9spp.as<double>() + job * c->last_line_stride, k * 8, hipMemcpyDeviceToHost));
  return LBLX_OK;
}

int lblx_contnm(lblx_ctx *c, const lblx_run *run, lblx_layer *layer, double *absrb, int cap,
                int *nptabs, double *v1abs, double *dvabs) {
  if (!c || !run || !layer || !absrb || !nptabs) return LBLX_E_ARG;
  if (run->icntnm == 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  LayerPlan P;
  int rc = plan_layer(run, layer, &P);
  if (rc) return rc;
  const ContPlan &C = P.cont;
  if (cap < C.nptabs) return LBLX_E_SIZE;
  int stride = (C.nptabs + 4 + 63) / 64 * 64;
  if (c->w_cps.ensure(sizeof(ContPlan)) || c->w_absrb.ensure((size_t)stride * 8)) return LBLX_E_HIP;
  HIPCHK(hipMemcpyAsync(c->w_cps.p, &C, sizeof(ContPlan), hipMemcpyHostToDevice, c->stream));
  int th = c->tbegin(ST_CONTNM);
  hipLaunchKernelGGL(k_contnm, dim3((stride + 255) / 256), dim3(256), 0, c->stream, c->w_cps.as<ContPlan>(),
                     c->cont_tabs, stride, 1, c->w_absrb.as<double>());
  c->tend(th);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(absrb, c->w_absrb.as<double>() + 1, (size_t)C.nptabs * 8, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  *nptabs = C.nptabs;
  if (v1abs) *v1abs = C.v1abs;
  if (dvabs) *dvabs = C.dvabs;
  return LBLX_OK;
}

// Replays RADMRG's panel reads (xmerge.f90:2683-2712) and RADNN's slot
// walk (xmerge.f90:3105-3130) over the panel headers of the old (merged)
// and new layers: base[p] = old point held in carry-over slot 1 while new
// panel p is merged.  Grids on which RADNN would read a slot that holds no
// old point yet (or the end padding while old panels remain) are refused.
static int merge_bases(const LayerPlan &O, const LayerPlan &K, int itype, double v2,
                       std::vector<int64_t> &base) {
  const int LL = itype + 1, LLM1 = std::max(itype, 1);
  const int npo = (int)O.panels.size();
  int64_t g0 = 0, npe = 0;
  double v2po = 0., dvpo = 0.;
  int po = 0;
  base.resize(K.panels.size());
  for (size_t pk = 0; pk < K.panels.size(); pk++) {
    const int nlim = K.panels[pk].nhi - NSHIFT;
    const double v2p = K.panels[pk].v2p;
    if (v2po <= v2p + dvpo) {
      while (po < npo) {
        npe += O.panels[po].nhi - NSHIFT;
        v2po = O.panels[po].v2p;
        dvpo = O.dv;
        po++;
        if (!(v2po <= v2p + dvpo)) break;
      }
    }
    const bool padded = v2po + dvpo >= v2;
    const int64_t lim = (padded && po == npo) ? npe + 2 : npe;
    int64_t ipl = 0;
    for (int nl = 1; nl <= LL; nl++) {
      ipl = nl - LLM1 - (nl > 1 ? 1 : 0);
      int64_t cnt = nlim >= nl ? (nlim - nl) / LL + 1 : 0;
      ipl += cnt * LLM1;
      if (cnt > 0 && ipl + (nl > 1 ? 2 : 0) > lim) return LBLX_E_UNSUPPORTED;
    }
    base[pk] = g0;
    const int64_t npl = ipl;   // RADO(NPL) becomes RADO(0)
    g0 += npl;
    npe -= npl;
  }
  return LBLX_OK;
}

static int merge_itype(double dvl, double dvk, int *itype) {
  if (dvl == dvk) *itype = 0;
  else if (dvl > dvk) *itype = fint(dvk / (dvl - dvk) + 0.5);
  else return LBLX_E_COARSER;
  if (*itype > 10) return LBLX_E_UNSUPPORTED;
  return LBLX_OK;
}

// Panel table of one merge launch: VFT and the carry-over base per panel.
static void merge_panels(const LayerPlan &K, const std::vector<int64_t> *base,
                         std::vector<DevPanel> &out) {
  for (size_t i = 0; i < K.panels.size(); i++)
    out.push_back(DevPanel{K.panels[i].vft, 0, 0, 0, 0, 0, 0, base ? (*base)[i] : 0});
}

// MergeArgs of one layer: grid, temperatures and RADMRG's weights.
// Whether the merge of layer K onto layer O is the identity map of RADNN:
// ITYPE 0, the same panels, and new panel p reading old point 2400 p on.
static bool merge_is_identity(const LayerPlan &O, const LayerPlan &K, int itype, const DevPanel *kpan,
                              const DevPanel *opan) {
  if (itype != 0 || K.npts != O.npts || K.dv != O.dv || K.panels.size() != O.panels.size()) return false;
  for (size_t p = 0; p < K.panels.size(); p++)
    if (kpan[p].out0 != (int64_t)2400 * (int64_t)p || kpan[p].vft != opan[p].vft) return false;
  return true;
}

static MergeArgs merge_args(const LayerPlan &K, const LayerPlan *O, int itype, double tave, double tzu,
                            double tbnd, double emis) {
  MergeArgs A{};
  A.dv = K.dv;
  A.xkt = tave / RADCN2;
  A.xkta = tzu / RADCN2;
  A.xktbnd = tbnd / RADCN2;
  A.emis = emis;
  A.n = K.npts;
  A.nold = O ? O->npts : 0;
  A.first = O == nullptr;
  A.itype = itype;
  double ap = 1.0 / ((double)itype + 1.0);
  for (int jpg = 1; jpg <= itype; jpg++) {  // RADMRG (xmerge.f90:2788-2796)
    double pp = 1.0 - (ap * (double)jpg);
    int ipl = jpg + 1;
    A.a1[ipl] = -pp * (pp - 1.0) * (pp - 2.0) / 6.0;
    A.a2[ipl] = (pp * pp - 1.0) * (pp - 2.0) * 0.5;
    A.a3[ipl] = -pp * (pp + 1.0) * (pp - 2.0) * 0.5;
    A.a4[ipl] = pp * (pp * pp - 1.0) / 6.0;
  }
  return A;
}

// XMERGE (IEMIT=1, IPATHL=1) for one layer on device buffers.
// Upwelling (down = false): tbnd is the surface of the first layer
// (EMINIT).  Downwelling: tbnd > 0 is EMBND after a later layer's merge.
static int merge_launch(lblx_ctx *c, const LayerPlan &K, const LayerPlan *O, int itype, double tave,
                        double tzu, double tbnd, double emis, bool down, const void *d_od, bool od32,
                        const double *d_rold, const double *d_told, double *d_rnew, double *d_tnew,
                        const DevPanel *d_pan) {
  MergeArgs A = merge_args(K, O, itype, tave, tzu, (O == nullptr && !down) ? tbnd : 0., emis);
  A.down = down ? 1 : 0;
  if (down && O != nullptr && tbnd > 0.) A.xktend = tbnd / RADCN2;
  int th = c->tbegin(ST_MERGE);
  dim3 grid((unsigned)((A.n + 255) / 256));
  if (od32)
    hipLaunchKernelGGL(k_merge<float>, grid, dim3(256), 0, c->stream, A, d_pan, (const float *)d_od,
                       d_rold, d_told, d_rnew, d_tnew);
  else
    hipLaunchKernelGGL(k_merge<double>, grid, dim3(256), 0, c->stream, A, d_pan, (const double *)d_od,
                       d_rold, d_told, d_rnew, d_tnew);
  c->tend(th);
  HIPCHK(hipGetLastError());
  return LBLX_OK;
}

int lblx_merge_size(const lblx_run *run, const lblx_layer *layer_old, const lblx_layer *layer,
                    int64_t *n) {
  if (!run || !layer || !n) return LBLX_E_ARG;
  lblx_layer tmp = *layer;
  LayerPlan K;
  int rc = plan_layer(run, &tmp, &K);
  if (rc) return rc;
  if (layer_old) {
    lblx_layer t2 = *layer_old;
    LayerPlan O;
    if ((rc = plan_layer(run, &t2, &O))) return rc;
    int itype;
    if ((rc = merge_itype(O.dv, K.dv, &itype))) return rc;
    std::vector<int64_t> base;
    if ((rc = merge_bases(O, K, itype, K.v2, base))) return rc;
  }
  *n = K.npts;
  return LBLX_OK;
}

static int xmerge_path(lblx_ctx *c, const lblx_run *run, const lblx_layer *layer_old,
                       const lblx_layer *layer, int ipathl, double tzu, double tbnd, double emis,
                       const void *d_od, const double *d_rold, const double *d_told, double *d_rnew,
                       double *d_tnew) {
  if (!c || !run || !layer || !d_od || !d_rnew || !d_tnew) return LBLX_E_ARG;
  if (ipathl != 1 && ipathl != 3) return LBLX_E_UNSUPPORTED;
  if (layer_old && (!d_rold || !d_told)) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  lblx_layer tmp = *layer;
  LayerPlan K, O;
  int rc = plan_layer(run, &tmp, &K);
  if (rc) return rc;
  std::vector<int64_t> base;
  int itype = 0;
  if (layer_old) {
    lblx_layer t2 = *layer_old;
    if ((rc = plan_layer(run, &t2, &O))) return rc;
    if ((rc = merge_itype(O.dv, K.dv, &itype))) return rc;
    if ((rc = merge_bases(O, K, itype, K.v2, base))) return rc;
  }
  std::vector<DevPanel> pan;
  merge_panels(K, layer_old ? &base : nullptr, pan);
  if (c->w_mpan.ensure(pan.size() * sizeof(DevPanel))) return LBLX_E_HIP;
  c->stage_begin();
  HIPCHK(c->h2d(c->w_mpan.p, pan.data(), pan.size() * sizeof(DevPanel), c->stream));
  c->stage_end(c->stream);
  return merge_launch(c, K, layer_old ? &O : nullptr, itype, layer->tave, tzu, tbnd, emis, ipathl == 3, d_od,
                      run->precision == LBLX_FP32_SPLIT, d_rold, d_told, d_rnew, d_tnew,
                      c->w_mpan.as<DevPanel>());
}

int lblx_xmerge_device(lblx_ctx *c, const lblx_run *run, const lblx_layer *layer_old,
                       const lblx_layer *layer, double tzu, double tbnd, double emis,
                       const void *d_od, const double *d_rold, const double *d_told, double *d_rnew,
                       double *d_tnew) {
  return xmerge_path(c, run, layer_old, layer, 1, tzu, tbnd, emis, d_od, d_rold, d_told, d_rnew, d_tnew);
}

int lblx_xmerge_path_device(lblx_ctx *c, const lblx_run *run, const lblx_layer *layer_old,
                            const lblx_layer *layer, int ipathl, double tz, double tbnd, double emis,
                            const void *d_od, const double *d_rold, const double *d_told, double *d_rnew,
                            double *d_tnew) {
  return xmerge_path(c, run, layer_old, layer, ipathl, tz, tbnd, emis, d_od, d_rold, d_told, d_rnew, d_tnew);
}

// OPDPTH for every layer of a path in one batch, then XMERGE layer by
// layer in path order, on the last layer's grid: IPATHL = 1, upwelling
// radiance and transmittance at the top from the surface up (tzu: each
// layer's upper boundary temperature, tbnd/emis the surface, EMINIT);
// IPATHL = 3, downwelling radiance at the observer from the layer next to
// it outward (tz: each layer's boundary temperature on the observer's
// side, TZL; tbnd/emis a boundary beyond the last layer, EMBND).
static int spectrum_path(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay, int ipathl,
                         const double *tzu, double tbnd, double emis, double *d_rad, double *d_tra,
                         int64_t n_out, double *v1p_top, lblx_stats *st) {
  if (!c || !run || !layers || nlay <= 0 || !tzu || !d_rad || !d_tra) return LBLX_E_ARG;
  if (ipathl != 1 && ipathl != 3) return LBLX_E_UNSUPPORTED;
  const bool down = ipathl == 3;
  HIPCHK(hipSetDevice(c->device));
  std::vector<LayerPlan> plans(nlay);
  std::vector<int64_t> off(nlay);
  int64_t tot = 0, nmax = 0;
  {  // JRAD chained up the column, then the layer plans on host threads
    std::vector<lblx_layer> tmp(layers, layers + nlay);
    for (int l = 1; l < nlay; l++) tmp[l].jrad = layer_jrad(&tmp[l - 1]);
    std::vector<int> prc(nlay, 0);
    parallel_for(nlay, [&](int l) { prc[l] = plan_layer(run, &tmp[l], &plans[l]); });
    for (int l = 0; l < nlay; l++)
      if (prc[l]) return prc[l];
  }
  for (int l = 0; l < nlay; l++) {
    off[l] = tot;
    tot += plans[l].npts;
    nmax = std::max(nmax, plans[l].npts);
  }
  if (n_out < plans[nlay - 1].npts) return LBLX_E_SIZE;
  // merge tables first: a refused grid fails before any device work
  std::vector<int> itype(nlay, 0);
  std::vector<DevPanel> pan;
  std::vector<int> pan0(nlay);
  for (int l = 0; l < nlay; l++) {
    pan0[l] = (int)pan.size();
    std::vector<int64_t> base;
    if (l > 0) {
      int rc = merge_itype(plans[l - 1].dv, plans[l].dv, &itype[l]);
      if (rc) return rc;
      if ((rc = merge_bases(plans[l - 1], plans[l], itype[l], plans[l].v2, base))) return rc;
    }
    merge_panels(plans[l], l > 0 ? &base : nullptr, pan);
  }
  const bool f32 = run->precision == LBLX_FP32_SPLIT;
  const size_t esz = f32 ? 4 : 8;
  if (c->w_od.ensure((size_t)tot * esz) || c->w_rt.ensure((size_t)nmax * 8 * 4) ||
      c->w_mpan.ensure(pan.size() * sizeof(DevPanel)))
    return LBLX_E_HIP;
  int rc = run_jobs(c, run, layers, nlay, true, c->w_od.p, off.data(), f32, nullptr, st, plans.data());
  if (rc) return rc;
  // runs of layers whose merge after the first is the identity map
  std::vector<int> rstart;
  std::vector<RunLayer> rl(nlay);
  for (int l = 0; l < nlay; l++) {
    rl[l].xkt = layers[l].tave / RADCN2;
    rl[l].xkta = tzu[l] / RADCN2;
    rl[l].od = off[l];
    rl[l].rout = -1;
    if (l == 0 || !merge_is_identity(plans[l - 1], plans[l], itype[l], pan.data() + pan0[l],
                                     pan.data() + pan0[l - 1]))
      rstart.push_back(l);
  }
  rstart.push_back(nlay);
  if (c->w_mrun.ensure((size_t)nlay * sizeof(RunLayer))) return LBLX_E_HIP;
  c->stage_begin();
  HIPCHK(c->h2d(c->w_mpan.p, pan.data(), pan.size() * sizeof(DevPanel), c->stream));
  HIPCHK(c->h2d(c->w_mrun.p, rl.data(), (size_t)nlay * sizeof(RunLayer), c->stream));
  c->stage_end(c->stream);
  double *rt = c->w_rt.as<double>();
  double *R[2] = {rt, rt + nmax}, *T[2] = {rt + 2 * nmax, rt + 3 * nmax};
  for (size_t q = 0; q + 1 < rstart.size(); q++) {
    const int l = rstart[q], nrun = rstart[q + 1] - l, last = l + nrun - 1;
    double *rn = (last == nlay - 1) ? d_rad : R[q & 1];
    double *tn = (last == nlay - 1) ? d_tra : T[q & 1];
    const double *ro = q > 0 ? R[(q - 1) & 1] : nullptr;
    const double *to = q > 0 ? T[(q - 1) & 1] : nullptr;
    MergeArgs A = merge_args(plans[l], l > 0 ? &plans[l - 1] : nullptr, itype[l], layers[l].tave, tzu[l],
                             (l == 0 && !down) ? tbnd : 0., emis);
    A.down = down ? 1 : 0;
    // EMBND after the last merge of a downwelling path (RADMRG with TBND,
    // xmerge.f90:137-141, 2980); a one-layer path is EMINIT's alone
    if (down && last == nlay - 1 && nlay > 1 && tbnd > 0.) A.xktend = tbnd / RADCN2;
    const RunLayer *drl = c->w_mrun.as<RunLayer>() + l;
    const DevPanel *dp = c->w_mpan.as<DevPanel>() + pan0[l];
    int th = c->tbegin(ST_MERGE);
    dim3 grid((unsigned)((A.n + 255) / 256));
    if (f32)
      hipLaunchKernelGGL(k_merge_run<float>, grid, dim3(256), 0, c->stream, A, drl, nrun, dp,
                         c->w_od.as<float>(), ro, to, rn, tn, (double *)nullptr);
    else
      hipLaunchKernelGGL(k_merge_run<double>, grid, dim3(256), 0, c->stream, A, drl, nrun, dp,
                         c->w_od.as<double>(), ro, to, rn, tn, (double *)nullptr);
    c->tend(th);
    HIPCHK(hipGetLastError());
  }
  if (v1p_top) *v1p_top = plans[nlay - 1].panels[0].vft + (double)NSHIFT * plans[nlay - 1].dv;
  return LBLX_OK;
}

int lblx_spectrum_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay,
                         const double *tzu, double tbnd, double emis, double *d_rad, double *d_tra,
                         int64_t n_out, double *v1p_top, lblx_stats *st) {
  return spectrum_path(c, run, layers, nlay, 1, tzu, tbnd, emis, d_rad, d_tra, n_out, v1p_top, st);
}

int lblx_spectrum_path_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay, int ipathl,
                              const double *tz, double tbnd, double emis, double *d_rad, double *d_tra,
                              int64_t n_out, double *v1p_end, lblx_stats *st) {
  return spectrum_path(c, run, layers, nlay, ipathl, tz, tbnd, emis, d_rad, d_tra, n_out, v1p_end, st);
}

// ABSMRG's walk for a merge of layer K's OD onto the accumulated OD on layer
// O's grid (xmerge.f90:437-685): ITYPE, the alignment search (labels
// 100-120), and the new point at which the old file ends.
static int absmrg_plan(const LayerPlan &O, const LayerPlan &K, AbsArgs *A) {
  *A = AbsArgs{};
  A->n = K.npts;
  A->n_old = O.npts;
  A->t_eof = -1;
  int itype = 0;
  int rc = merge_itype(O.dv, K.dv, &itype);
  if (rc) return rc;
  A->itype = itype;
  if (itype == 0) {  // panel by panel (labels 40-70)
    if (O.panels.size() != K.panels.size() || O.npts != K.npts) return LBLX_E_UNSUPPORTED;
    for (size_t p = 0; p < K.panels.size(); p++)
      if (O.panels[p].nhi != K.panels[p].nhi) return LBLX_E_UNSUPPORTED;
    return LBLX_OK;
  }
  const double ap = 1.0 / ((double)itype + 1.0);
  for (int jpg = 1; jpg <= itype; jpg++) {
    const double P = 1.0 - (ap * (double)jpg);
    A->a1[jpg] = -P * (P - 1.0) * (P - 2.0) / 6.0;
    A->a2[jpg] = (P * P - 1.0) * (P - 2.0) * 0.5;
    A->a3[jpg] = -P * (P + 1.0) * (P - 2.0) * 0.5;
    A->a4[jpg] = P * (P * P - 1.0) / 6.0;
  }
  const double v1po = O.panels[0].vft + (double)NSHIFT * O.dv, dvpo = O.dv;
  const double v1p = K.panels[0].vft + (double)NSHIFT * K.dv, dvp = K.dv;
  const int nlim0 = K.panels[0].nhi - NSHIFT, nlimo0 = O.panels[0].nhi - NSHIFT;
  const double dif = dvp * 0.01;
  int64_t nvs = 1, ii = 1;
  if (!(std::fabs(v1po - v1p) < dif)) {
    for (;;) {
      nvs = nvs + 1;
      const double v1pn = v1po + dvpo * (double)(nvs - 1);
      if (std::fabs(v1pn - v1p) < dif) break;
      if (v1pn < v1p) {
        if (nvs > nlimo0) return LBLX_E_UNSUPPORTED;
        continue;
      }
      bool found = false;
      for (;;) {
        ii = ii + 1;
        const double v1pp = v1p + dvp * (double)(ii - 1);
        if (std::fabs(v1pn - v1pp) < dif) {
          found = true;
          break;
        }
        if (!(v1pp < v1pn)) break;
      }
      if (found) break;
      if (nvs > nlimo0 || ii > nlim0) return LBLX_E_UNSUPPORTED;
    }
  }
  A->g0 = ii - 1;
  A->nvs0 = nvs;
  const int LL = itype + 1;
  // NVS after the walk's increment at point t >= 1 is nvs(t) + 1 with
  // nvs(t) = NVS0 + (t - 1) - (t - 1) / LL; the old file ends when it
  // passes N_old - 2
  auto nvs_at = [&](int64_t t) { return A->nvs0 + (t - 1) - (t - 1) / LL; };
  const int64_t tmax = A->n - 1 - A->g0;
  if (tmax >= 1 && nvs_at(tmax) >= A->n_old - 2) {
    int64_t lo = 1, hi = tmax;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (nvs_at(mid) >= A->n_old - 2) hi = mid;
      else lo = mid + 1;
    }
    A->t_eof = lo;
    // the fill covers the rest of that new panel; a later panel would be
    // merged from stale buffer contents
    const int64_t g = A->g0 + lo;
    const int64_t p = g / (NLIM1 - 1);
    if ((p + 1) * (NLIM1 - 1) < A->n) return LBLX_E_UNSUPPORTED;
  }
  if (A->nvs0 + 2 > A->n_old) return LBLX_E_UNSUPPORTED;
  return LBLX_OK;
}

static int absmrg_launch(lblx_ctx *c, const AbsArgs &A, const void *od, bool f32, const double *old,
                         double *out) {
  const unsigned nb = (unsigned)((A.n + 255) / 256);
  int th = c->tbegin(ST_MERGE);
  if (f32) hipLaunchKernelGGL(k_absmrg<float>, dim3(nb), dim3(256), 0, c->stream, A, (const float *)od, old, out);
  else hipLaunchKernelGGL(k_absmrg<double>, dim3(nb), dim3(256), 0, c->stream, A, (const double *)od, old, out);
  c->tend(th);
  HIPCHK(hipGetLastError());
  return LBLX_OK;
}

int lblx_absmrg_device(lblx_ctx *c, const lblx_run *run, const lblx_layer *layer_old,
                       const lblx_layer *layer, const void *d_od, const double *d_acc_old,
                       double *d_acc_new) {
  if (!c || !run || !layer || !d_od || !d_acc_new || (layer_old && !d_acc_old)) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  LayerPlan K, O;
  lblx_layer tk = *layer;
  int rc = plan_layer(run, &tk, &K);
  if (rc) return rc;
  AbsArgs A{};
  if (layer_old) {
    lblx_layer to = *layer_old;
    if ((rc = plan_layer(run, &to, &O))) return rc;
    if ((rc = absmrg_plan(O, K, &A))) return rc;
  } else {
    A.n = K.npts;
    A.first = 1;
  }
  return absmrg_launch(c, A, d_od, run->precision == LBLX_FP32_SPLIT, d_acc_old, d_acc_new);
}

int lblx_column_od_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay, double *d_od,
                          int64_t n_out, double *v1p_top, lblx_stats *st) {
  if (!c || !run || !layers || nlay <= 0 || !d_od) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  std::vector<LayerPlan> plans(nlay);
  std::vector<int64_t> off(nlay);
  int64_t tot = 0, nmax = 0;
  {  // JRAD chained up the column, then the layer plans on host threads
    std::vector<lblx_layer> tmp(layers, layers + nlay);
    for (int l = 1; l < nlay; l++) tmp[l].jrad = layer_jrad(&tmp[l - 1]);
    std::vector<int> prc(nlay, 0);
    parallel_for(nlay, [&](int l) { prc[l] = plan_layer(run, &tmp[l], &plans[l]); });
    for (int l = 0; l < nlay; l++)
      if (prc[l]) return prc[l];
  }
  for (int l = 0; l < nlay; l++) {
    off[l] = tot;
    tot += plans[l].npts;
    nmax = std::max(nmax, plans[l].npts);
  }
  if (n_out < plans[nlay - 1].npts) return LBLX_E_SIZE;
  std::vector<AbsArgs> A(nlay);
  A[0] = AbsArgs{};
  A[0].n = plans[0].npts;
  A[0].first = 1;
  for (int l = 1; l < nlay; l++) {
    int rc = absmrg_plan(plans[l - 1], plans[l], &A[l]);
    if (rc) return rc;
  }
  const bool f32 = run->precision == LBLX_FP32_SPLIT;
  const size_t esz = f32 ? 4 : 8;
  if (c->w_od.ensure((size_t)tot * esz) || c->w_rt.ensure((size_t)nmax * 8 * 2)) return LBLX_E_HIP;
  int rc = run_jobs(c, run, layers, nlay, true, c->w_od.p, off.data(), f32, nullptr, st, plans.data());
  if (rc) return rc;
  double *acc[2] = {c->w_rt.as<double>(), c->w_rt.as<double>() + nmax};
  for (int l = 0; l < nlay; l++) {
    const void *od = (const char *)c->w_od.p + (size_t)off[l] * esz;
    double *dst = (l == nlay - 1) ? d_od : acc[l & 1];
    if ((rc = absmrg_launch(c, A[l], od, f32, l > 0 ? acc[(l - 1) & 1] : nullptr, dst))) return rc;
  }
  if (v1p_top) *v1p_top = plans[nlay - 1].panels[0].vft + (double)NSHIFT * plans[nlay - 1].dv;
  return LBLX_OK;
}

// PNLINT's control flow over the panels of one layer (oprop.f90:1596-1800),
// without the values: the output panels and, for every run of output
// points computed in one call, the input panel and FJ1DIF.
struct DvoPlan {
  DvoArgs A;
  std::vector<DvoSeg> segs;
  double v1out = 0.;
};
static int dvout_plan(const LayerPlan &P, double v2, double dvout, DvoPlan *D) {
  D->A = DvoArgs{};
  D->segs.clear();
  DvoArgs &A = D->A;
  A.n_in = P.npts;
  if (!(dvout > 0.)) return LBLX_E_ARG;
  const double DVP = P.dv;
  double atype = 9.999E09;
  if (DVP == dvout) atype = 0.;
  if (dvout > DVP) atype = DVP / (dvout - DVP) + 0.5;
  if (dvout < DVP) atype = -dvout / (DVP - dvout) - 0.5;
  if (std::fabs(dvout - DVP) < 1.E-8) atype = 0.;
  D->v1out = P.panels[0].vft + (double)NSHIFT * P.dv;
  if (atype == 0.) {
    A.pass = 1;
    A.n_out = P.npts;
    return LBLX_OK;
  }
  for (int jp = 0; jp <= 100; jp++) {
    const double p = 0.01 * (double)jp;
    A.a1[jp] = -p * (p - 1.0) * (p - 2.0) / 6.0;
    A.a2[jp] = (p * p - 1.0) * (p - 2.0) * 0.5;
    A.a3[jp] = -p * (p + 1.0) * (p - 2.0) * 0.5;
    A.a4[jp] = p * (p * p - 1.0) / 6.0;
  }
  const int LIMOUT = 2400;
  A.ratdv = dvout / DVP;
  double v1po = 0., v2po = 0.;
  int nlim1 = 1, nlim2 = 0, ilast = 0;
  int64_t emitted = 0;
  const int np = (int)P.panels.size();
  for (int q = 0; q < np; q++) {
    const PanelPlan &Q = P.panels[q];
    const double V1P = Q.vft + (double)NSHIFT * P.dv;
    double V2P = Q.v2p;
    const int istop = q == np - 1;
    if (istop != Q.istop) return LBLX_E_UNSUPPORTED;
    int npp = 1;
    if (v1po == 0.0) {
      v1po = V1P;
      nlim1 = 1;
    }
    if (istop) V2P = V2P + 2. * DVP;
    const int64_t gq = (int64_t)q * (NLIM1 - 1);
    for (;;) {
      v2po = v1po + (double)LIMOUT * dvout;
      if (V2P <= v2po + DVP && ilast == 0 && npp <= 0) break;
      if ((v1po + (LIMOUT - 1) * dvout) > v2) {
        nlim2 = fint((v2 - v1po) / dvout + 1.);
        v2po = v1po + (double)(nlim2 - 1) * dvout;
        if (v2po < v2) {
          v2po = v2po + dvout;
          nlim2 = nlim2 + 1;
        }
        ilast = 1;
        if (!istop && v2po > V2P - DVP) {
          nlim2 = fint(((V2P - DVP - v1po) / dvout) + 1.);
          v2po = v1po + (double)(nlim2 - 1) * dvout;
          if (v2po + dvout < V2P - DVP) {
            nlim2 = nlim2 + 1;
            v2po = v2po + dvout;
          }
          ilast = 0;
        }
      } else {
        nlim2 = LIMOUT;
        v2po = v1po + (double)(nlim2 - 1) * dvout;
        if (v2po > V2P - DVP) {
          nlim2 = fint(((V2P - DVP - v1po) / dvout) + 1.);
          v2po = v1po + (double)(nlim2 - 1) * dvout;
          if (v2po + dvout < V2P - DVP) {
            nlim2 = nlim2 + 1;
            v2po = v2po + dvout;
          }
        }
        ilast = 0;
      }
      if (nlim2 > LIMOUT || nlim2 < 0) return LBLX_E_UNSUPPORTED;
      if (nlim1 <= nlim2) {
        DvoSeg S;
        S.obase = emitted;
        S.o0 = emitted + nlim1 - 1;
        S.gq = gq;
        S.fj1 = (v1po - V1P) / DVP + 1. + 2.;
        if (!D->segs.empty() && D->segs.back().o0 >= S.o0) return LBLX_E_UNSUPPORTED;
        D->segs.push_back(S);
      }
      if (nlim2 == LIMOUT || ilast == 1) {
        emitted += nlim2;
        nlim1 = 1;
        npp = 0;
        v1po = v2po + dvout;
        if ((v1po + (double)LIMOUT * dvout) > (V2P - DVP)) npp = 1;
      } else {
        nlim1 = nlim2 + 1;
        npp = -1;
      }
      if (ilast != 1) continue;
      v1po = 0.0;
      break;
    }
    if (ilast == 1 && !istop) return LBLX_E_UNSUPPORTED;  // output ended before the layer
  }
  if (D->segs.empty()) return LBLX_E_UNSUPPORTED;
  A.n_out = emitted;
  A.nseg = (int)D->segs.size();
  return LBLX_OK;
}

int lblx_od_dvout_size(const lblx_run *run, const lblx_layer *layer, double dvout, int64_t *n) {
  if (!run || !layer || !n) return LBLX_E_ARG;
  LayerPlan P;
  lblx_layer t = *layer;
  int rc = plan_layer(run, &t, &P);
  if (rc) return rc;
  DvoPlan D;
  if ((rc = dvout_plan(P, layer->v2, dvout, &D))) return rc;
  *n = D.A.n_out;
  return LBLX_OK;
}

int lblx_od_layer_dvout_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layer, double dvout,
                               double *d_out, int64_t cap, int64_t *n_out, double *v1_out) {
  if (!c || !run || !layer || !d_out) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  LayerPlan P;
  lblx_layer t = *layer;
  int rc = plan_layer(run, &t, &P);
  if (rc) return rc;
  DvoPlan D;
  if ((rc = dvout_plan(P, layer->v2, dvout, &D))) return rc;
  if (cap < D.A.n_out) return LBLX_E_SIZE;
  const bool f32 = run->precision == LBLX_FP32_SPLIT;
  const size_t esz = f32 ? 4 : 8;
  if (c->w_od.ensure((size_t)P.npts * esz) || c->w_dvo.ensure(std::max<size_t>(D.segs.size(), 1) * sizeof(DvoSeg)))
    return LBLX_E_HIP;
  int64_t off = 0;
  rc = run_jobs(c, run, layer, 1, false, c->w_od.p, &off, f32, nullptr, nullptr, &P);
  if (rc) return rc;
  if (!D.segs.empty())
    HIPCHK(hipMemcpyAsync(c->w_dvo.p, D.segs.data(), D.segs.size() * sizeof(DvoSeg), hipMemcpyHostToDevice,
                          c->stream));
  const unsigned nb = (unsigned)((D.A.n_out + 255) / 256);
  if (f32)
    hipLaunchKernelGGL(k_dvout<float>, dim3(nb), dim3(256), 0, c->stream, D.A, c->w_dvo.as<DvoSeg>(),
                       c->w_od.as<float>(), d_out);
  else
    hipLaunchKernelGGL(k_dvout<double>, dim3(nb), dim3(256), 0, c->stream, D.A, c->w_dvo.as<DvoSeg>(),
                       c->w_od.as<double>(), d_out);
  HIPCHK(hipGetLastError());
  if (n_out) *n_out = D.A.n_out;
  if (v1_out) *v1_out = D.v1out;
  return LBLX_OK;
}

int lblx_scan_size(const lblx_scan *scan, double v1_in, double dv_in, int64_t n_in, int64_t *n_out,
                   double *v1_out, double *dv_out) {
  if (!scan || !n_out) return LBLX_E_ARG;
  ScanPlan P;
  int rc = plan_scan(scan, v1_in, dv_in, n_in, &P);
  if (rc) return rc;
  *n_out = P.nout;
  if (v1_out) *v1_out = P.v1out;
  if (dv_out) *dv_out = P.dvout;
  return LBLX_OK;
}

int lblx_scanfn_size(const lblx_scanfn *scan, int iemit, double v1_in, double dv_in, int64_t n_in,
                     int64_t *n_out, double *v1_out, double *dv_out) {
  if (!scan || !n_out) return LBLX_E_ARG;
  ScanFnPlan P;
  int rc = plan_scanfn(scan, iemit, v1_in, dv_in, n_in, &P);
  if (rc) return rc;
  *n_out = P.nout;
  if (v1_out) *v1_out = P.v1out;
  if (dv_out) *dv_out = P.dvo;
  return LBLX_OK;
}

int lblx_scanfn_device(lblx_ctx *c, const lblx_scanfn *scan, int iemit, const double *d_in, double v1_in,
                       double dv_in, int64_t n_in, double *d_out, int64_t n_out) {
  if (!c || !scan || !d_in || !d_out) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  ScanFnPlan P;
  int rc = plan_scanfn(scan, iemit, v1_in, dv_in, n_in, &P);
  if (rc) return rc;
  if (n_out < P.nout) return LBLX_E_SIZE;
  hipStream_t s = c->stream;
  const size_t ng = P.zpk.size();
  if (c->f_s.ensure(P.K * sizeof(double)) || c->f_xf.ensure(P.xf.size() * sizeof(double)) ||
      c->f_zp.ensure(ng * sizeof(double)) || c->f_q.ensure(ng * sizeof(int32_t)))
    return LBLX_E_HIP;
  if (P.g_hi >= P.K) return LBLX_E_ARG;  // the walk's groups lie inside the shrunk input
  int th = c->tbegin(ST_SCAN);
  HIPCHK(hipMemcpyAsync(c->f_xf.p, P.xf.data(), P.xf.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->f_zp.p, P.zpk.data(), ng * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->f_q.p, P.q.data(), ng * sizeof(int32_t), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_sfn_shrink, dim3((unsigned)((P.K + 255) / 256)), dim3(256), 0, s, d_in, P.o, P.R, P.K,
                     P.jtrem, P.jabs, c->f_s.as<double>());
  SfnArgs A;
  A.z0 = (P.vg0 - P.vft) / P.dvo;
  A.ratio = P.dvsc / P.dvo;
  A.zbound = P.zbound;
  A.zslope = P.zslope;
  A.xnorm = P.xnorm;
  A.g_lo = P.g_lo;
  A.g_hi = P.g_hi;
  A.j0 = 32 + 1;  // PANLSC's first point NLO = NSHIFT + 1
  A.nfmax = (int)P.xf.size();
  hipLaunchKernelGGL(k_sfn_conv, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, s, A,
                     c->f_s.as<double>(), c->f_zp.as<double>(), c->f_q.as<int32_t>(), c->f_xf.as<double>(),
                     P.nout, d_out);
  HIPCHK(hipGetLastError());
  c->tend(th);
  HIPCHK(hipStreamSynchronize(s));  // the host table is freed on return
  return LBLX_OK;
}

int lblx_fftscn_device(lblx_ctx *c, const lblx_scan *scan, const double *d_in, double v1_in,
                       double dv_in, int64_t n_in, double *d_out, int64_t cap, int64_t *n_out,
                       double *v1_out, double *dv_out) {
  if (!c || !scan || !d_in || !d_out) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  ScanPlan P;
  int rc = plan_scan(scan, v1_in, dv_in, n_in, &P);
  if (rc) return rc;
  if (cap < P.nout) return LBLX_E_SIZE;
  const int64_t N = P.N;
  hipStream_t s = c->stream;
  if (c->s_z[0].ensure(N * sizeof(double2)) || c->s_z[1].ensure(N * sizeof(double2))) return LBLX_E_HIP;
  if (c->s_twN != N) {  // twiddles exp(-2 pi i m / N), m < N/2
    std::vector<double2> tw(N / 2);
    const double two_pi = 4.0 * std::asin(1.0);
    for (int64_t m = 0; m < N / 2; m++) {
      double ang = two_pi * (double)m / (double)N;
      tw[m] = make_double2(std::cos(ang), -std::sin(ang));
    }
    if (c->s_tw.ensure(tw.size() * sizeof(double2))) return LBLX_E_HIP;
    HIPCHK(hipMemcpy(c->s_tw.p, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    c->s_twN = N;
  }
  int th = c->tbegin(ST_SCAN);
  double2 *z0 = c->s_z[0].as<double2>(), *z1 = c->s_z[1].as<double2>();
  const unsigned gN = (unsigned)((N + 255) / 256), gH = (unsigned)((N / 2 + 255) / 256);
  hipLaunchKernelGGL(k_scan_load, dim3(gN), dim3(256), 0, s, d_in, P.n0, P.M, P.ktot, N, P.fill, z0);
  double2 *cur = z0, *oth = z1;
  if (P.jfn != 0) {
    for (int64_t ns = 1; ns < N; ns <<= 1) {
      hipLaunchKernelGGL(k_fft_pass, dim3(gH), dim3(256), 0, s, cur, oth, c->s_tw.as<double2>(), N, ns, -1);
      std::swap(cur, oth);
    }
    ScanFilter F{N, P.k4, P.DX, P.A, P.dvb, P.r4, P.jfn, P.fix};
    hipLaunchKernelGGL(k_scan_filter, dim3((unsigned)((N / 2 + 1 + 255) / 256)), dim3(256), 0, s, F, cur);
    for (int64_t ns = 1; ns < N; ns <<= 1) {
      hipLaunchKernelGGL(k_fft_pass, dim3(gH), dim3(256), 0, s, cur, oth, c->s_tw.as<double2>(), N, ns, +1);
      std::swap(cur, oth);
    }
  }
  const double rn = P.jfn != 0 ? 1.0 / (double)N : 1.0;
  const int64_t off = P.n1 - 1;
  if (scan->dvout > 0) {
    std::vector<double> cof(4 * 203);
    const double xnumcf = 201.;
    for (int ip = 0; ip <= 202; ip++) {  // INTERP coefficients (postsub.f90:2966-2974)
      double p = ((double)ip - 1.0) / (xnumcf - 1.0), pp = p * p;
      cof[ip] = -p / 2.0 * (1 - p) * (1 - p);
      cof[203 + ip] = 1.0 - pp * (3.0 - 2.0 * p) + pp / 2.0 * (1.0 - p);
      cof[406 + ip] = pp * (3.0 - 2.0 * p) + p / 2.0 * (1.0 - p) * (1.0 - p);
      cof[609 + ip] = -pp / 2 * (1.0 - p);
    }
    if (c->s_g.ensure(P.g.size() * 8) || c->s_ip.ensure(P.ip.size() * 4) || c->s_cof.ensure(cof.size() * 8))
      return LBLX_E_HIP;
    HIPCHK(hipMemcpyAsync(c->s_g.p, P.g.data(), P.g.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->s_ip.p, P.ip.data(), P.ip.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->s_cof.p, cof.data(), cof.size() * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scan_out, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, s, cur, off, P.ntot,
                       rn, P.nout, c->s_g.as<int64_t>(), c->s_ip.as<int32_t>(), c->s_cof.as<double>(), d_out);
    HIPCHK(hipStreamSynchronize(s));  // host vectors above are freed on return
  } else {
    hipLaunchKernelGGL(k_scan_out, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, s, cur, off, P.ntot,
                       rn, P.nout, (const int64_t *)nullptr, (const int32_t *)nullptr, (const double *)nullptr,
                       d_out);
  }
  c->tend(th);
  HIPCHK(hipGetLastError());
  if (n_out) *n_out = P.nout;
  if (v1_out) *v1_out = P.v1out;
  if (dv_out) *dv_out = P.dvout;
  return LBLX_OK;
}

// INTPDR tables of the allocation-free scan, planned on the host into
// fixed storage and copied into `scratch`; the lock serialises host threads
// over them (a call returns after its copies have completed).
static std::mutex g_na_mu;
static int64_t g_na_g[NA_OUTMAX];
static int32_t g_na_ip[NA_OUTMAX];

int lblx_fftscn_noalloc_device(lblx_ctx *c, const lblx_scan *scan, const double *d_in, double v1_in,
                               double dv_in, int64_t n_in, double *d_out, int64_t cap, int64_t *n_out,
                               double *v1_out, double *dv_out) {
  if (!c || !scan || !d_in || !d_out) return LBLX_E_ARG;
  if (!c->scratch) return LBLX_E_HIP;
  HIPCHK(hipSetDevice(c->device));
  std::lock_guard<std::mutex> lk(g_na_mu);
  ScanPlan P;  // its table vectors stay empty: the tables go to g_na_g / g_na_ip
  const ScanTab tab{g_na_g, g_na_ip, NA_OUTMAX};
  int rc = plan_scan(scan, v1_in, dv_in, n_in, &P, &tab);
  if (rc == LBLX_E_SCRATCH) set_err("lblx_fftscn_noalloc_device: more than 262144 output points");
  if (rc) return rc;
  if (P.N > NA_NMAX) {
    set_err("lblx_fftscn_noalloc_device: FFT of " + std::to_string(P.N) + " points, above 524288");
    return LBLX_E_SCRATCH;
  }
  if (cap < P.nout) return LBLX_E_SIZE;
  const int64_t N = P.N;
  hipStream_t s = c->stream;
  char *base = (char *)c->scratch;
  double2 *z0 = (double2 *)base, *z1 = (double2 *)(base + NA_OFF_Z1);
  double2 *tw = (double2 *)(base + NA_OFF_TW);
  int64_t *g = (int64_t *)(base + NA_OFF_G);
  int32_t *ip = (int32_t *)(base + NA_OFF_IP);
  double *cof = (double *)(base + NA_OFF_COF);
  // every element the kernels read is written first in this call: the
  // planes over all N points by k_scan_load, the twiddles for m < N/2, the
  // tables for the nout output points
  const unsigned gN = (unsigned)((N + 255) / 256), gH = (unsigned)((N / 2 + 255) / 256);
  hipLaunchKernelGGL(k_scan_load, dim3(gN), dim3(256), 0, s, d_in, P.n0, P.M, P.ktot, N, P.fill, z0);
  double2 *cur = z0, *oth = z1;
  if (P.jfn != 0) {
    hipLaunchKernelGGL(k_scan_twiddle, dim3(gH), dim3(256), 0, s, tw, N);
    for (int64_t ns = 1; ns < N; ns <<= 1) {
      hipLaunchKernelGGL(k_fft_pass, dim3(gH), dim3(256), 0, s, cur, oth, (const double2 *)tw, N, ns, -1);
      std::swap(cur, oth);
    }
    ScanFilter F{N, P.k4, P.DX, P.A, P.dvb, P.r4, P.jfn, P.fix};
    hipLaunchKernelGGL(k_scan_filter, dim3((unsigned)((N / 2 + 1 + 255) / 256)), dim3(256), 0, s, F, cur);
    for (int64_t ns = 1; ns < N; ns <<= 1) {
      hipLaunchKernelGGL(k_fft_pass, dim3(gH), dim3(256), 0, s, cur, oth, (const double2 *)tw, N, ns, +1);
      std::swap(cur, oth);
    }
  }
  const double rn = P.jfn != 0 ? 1.0 / (double)N : 1.0;
  const int64_t off = P.n1 - 1;
  if (scan->dvout > 0) {
    HIPCHK(hipMemcpyAsync(g, g_na_g, (size_t)P.nout * sizeof(int64_t), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(ip, g_na_ip, (size_t)P.nout * sizeof(int32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scan_cof, dim3(1), dim3(256), 0, s, cof);
    hipLaunchKernelGGL(k_scan_out, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, s, cur, off, P.ntot,
                       rn, P.nout, (const int64_t *)g, (const int32_t *)ip, (const double *)cof, d_out);
    HIPCHK(hipStreamSynchronize(s));  // the host tables are reused by the next call
  } else {
    hipLaunchKernelGGL(k_scan_out, dim3((unsigned)((P.nout + 255) / 256)), dim3(256), 0, s, cur, off, P.ntot,
                       rn, P.nout, (const int64_t *)nullptr, (const int32_t *)nullptr, (const double *)nullptr,
                       d_out);
  }
  HIPCHK(hipGetLastError());
  if (n_out) *n_out = P.nout;
  if (v1_out) *v1_out = P.v1out;
  if (dv_out) *dv_out = P.dvout;
  return LBLX_OK;
}

// The tangent request of a batch of layers whose input JRAD is set
// (chained up the column): channel 0 the temperature, channel q the
// molecule species[q - 1]; derivative (l, q) at odc + odc_off[l P + q].
static int tan_request(const lblx_run *run, const lblx_layer *lay, int nlay, const int *species, int nspec,
                       void *odc, const int64_t *odc_off, TanReq &R) {
  const int P = 1 + nspec;
  if (nspec < 0 || nspec > 7 || (nspec > 0 && !species)) return LBLX_E_ARG;
  for (int q = 1; q < P; q++) {
    if (species[q - 1] < 1 || species[q - 1] > 7) return LBLX_E_ARG;
    for (int k = 0; k < q - 1; k++)
      if (species[k] == species[q - 1]) return LBLX_E_ARG;
  }
  R.P = P;
  R.odc = odc;
  R.tl.assign(nlay, TanLayer{});
  R.odc_off.assign(odc_off, odc_off + (size_t)nlay * P);
  const double hT = 0.1, hq = 1e-3;  // steps of the continuum's central differences
  R.rh[0] = 1. / (2. * hT);
  for (int q = 1; q < P; q++) R.rh[q] = 1. / (2. * hq);
  R.cps.clear();
  if (run->icntnm != 0) R.cps.resize((size_t)2 * nlay * P);
  for (int l = 0; l < nlay; l++) {
    TanLayer &T = R.tl[l];
    const lblx_layer &y = lay[l];
    for (int q = 1; q < P; q++)
      if (species[q - 1] > y.nmol) return LBLX_E_ARG;
    T.tave = y.tave;
    tips_dlnscor(T.tave, 0.01, T.dlnscor);
    double wtot = y.wbroad;
    for (int m = 0; m < y.nmol; m++) wtot += y.wk[m];
    for (int m = 0; m < 7; m++) T.chan[m] = 0;
    for (int q = 0; q < TAN_NQ; q++) T.wq[q] = 0.;
    for (int q = 1; q < P; q++) {
      const int m = species[q - 1] - 1;
      T.chan[m] = q;
      T.wq[q] = wtot > 0. ? y.wk[m] / wtot : 0.;
    }
    T.nchan = P;
    T.nint = P > 1 ? P + 1 : P;
    if (run->icntnm == 0) continue;
    const int jrad = layer_jrad(&y);  // the JRAD the layer's own plan takes
    ContPlan base;
    int rc = plan_cont_layer(run, &y, jrad, &base);
    if (rc) return rc;
    for (int q = 0; q < P; q++)
      for (int sg = 0; sg < 2; sg++) {
        lblx_layer t = y;
        const double h = sg == 0 ? 1. : -1.;
        if (q == 0) t.tave = t.tave + h * hT;
        else t.wk[species[q - 1] - 1] *= std::exp(h * hq);
        ContPlan &cp = R.cps[((size_t)l * P + q) * 2 + sg];
        rc = plan_cont_layer(run, &t, jrad, &cp);
        if (rc) return rc;
        if (cp.nptabs != base.nptabs || cp.v1abs != base.v1abs) return LBLX_E_UNSUPPORTED;
      }
  }
  return LBLX_OK;
}

// Jacobians of the TOA radiance with respect to each layer's temperature
// and the log amounts of chosen molecules: each layer's optical-depth
// derivative, then the exact radiative part, dR_l/dx on the layer grid
// carried to the top through every layer above as J <- T_k L4(J).
// The optical-depth derivatives: aj = 0: forward differences of OPDPTH
// (T + dT, q e^dlnq).  aj = 1: the reference's analytic-Jacobian runs
// (lblrtm.f90:2930-2975): T + 1 K, and for molecule m a single-species run
// (every other amount zero, O2's amount moved to the broadener, CONTNM with
// ICFLG = m) whose optical depth is dtau/dln q itself.  aj = 2: the
// tangent-linear optical depths of lblx_tangent.hip, computed beside the
// column's own OPDPTH run.  The radiative part (sources and merge tangents)
// is the same exact tangent in all three.
// d_dLdR given (with aj = 2): reverse mode instead of the Jacobian planes:
// d_grad[q nlay + l] = sum over the top grid of dL/dR dR/dx_(q,l).
static int jacobian_run(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay, const double *tzu,
                        double tbnd, double emis, const int *species, int nspec, double dT, double dlnq,
                        int aj, double *d_rad, double *d_jac, int64_t n_out, double *v1p_top,
                        const double *d_dLdR = nullptr, double *d_grad = nullptr) {
  const bool adj = d_dLdR != nullptr;
  if (!c || !run || !layers || nlay <= 0 || !tzu || (!adj && (!d_rad || !d_jac)) || (adj && (!d_grad || aj != 2)))
    return LBLX_E_ARG;
  if (nspec < 0 || nspec > 7 || (nspec > 0 && !species) || !(dT > 0.) || !(dlnq > 0.)) return LBLX_E_ARG;
  if (aj < 0 || aj > 2) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const int P = 1 + nspec;
  const bool tl = aj == 2;
  std::vector<LayerPlan> plans(nlay);
  std::vector<lblx_layer> jobs;
  std::vector<int> icflg;
  jobs.reserve((size_t)nlay * (1 + P));
  {  // JRAD chained up the column (each job plans again from its input flag)
    std::vector<lblx_layer> tmp(layers, layers + nlay);
    for (int l = 1; l < nlay; l++) tmp[l].jrad = layer_jrad(&tmp[l - 1]);
    jobs.assign(tmp.begin(), tmp.end());
    icflg.assign(nlay, 0);
    std::vector<int> prc(nlay, 0);
    parallel_for(nlay, [&](int l) { prc[l] = plan_layer(run, &tmp[l], &plans[l]); });
    for (int l = 0; l < nlay; l++) {
      if (prc[l]) return prc[l];
      layers[l].jrad = tmp[l].jrad;
    }
  }
  for (int l = 0; l < nlay; l++)
    for (int q = 0; q < P; q++) {
      lblx_layer t = jobs[l];
      int flg = 0;
      if (q > 0 && (species[q - 1] < 1 || species[q - 1] > t.nmol)) return LBLX_E_ARG;
      for (int k = 0; tl && k < q - 1; k++)  // one channel per molecule
        if (species[k] == species[q - 1]) return LBLX_E_ARG;
      if (tl) continue;  // the column's own run carries the tangents
      if (q == 0) {
        t.tave = t.tave + dT;
      } else {
        const int m = species[q - 1];
        if (m < 1 || m > t.nmol) return LBLX_E_ARG;
        if (aj) {
          const double wsave = t.wk[m - 1], w7 = t.nmol >= 7 ? t.wk[6] : 0.;
          for (int k = 0; k < t.nmol; k++) t.wk[k] = 0.;
          t.wk[m - 1] = wsave;
          t.wbroad = t.wbroad + w7;
          flg = m;
        } else {
          t.wk[m - 1] = t.wk[m - 1] * std::exp(dlnq);
        }
      }
      jobs.push_back(t);
      icflg.push_back(flg);
    }
  const int nj = (int)jobs.size();
  // every job's plan (the species runs with their ICFLG)
  if ((int)c->jplans.size() < nj) c->jplans.resize(nj);
  std::vector<LayerPlan> &jplans = c->jplans;
  {
    std::vector<int> prc(nj, 0);
    parallel_for(nj, [&](int j) {
      lblx_layer t = jobs[j];
      prc[j] = plan_layer(run, &t, &jplans[j], icflg[j]);
    });
    for (int j = 0; j < nj; j++)
      if (prc[j]) return prc[j];
  }
  const int nplane = nlay * (1 + P);  // the column's layers, then (layer, parameter)
  std::vector<int64_t> off(nplane);
  int64_t tot = 0, nmax = 0, tot_base = 0;
  for (int j = 0; j < nplane; j++) {
    const LayerPlan &K = plans[j < nlay ? j : (j - nlay) / P];
    off[j] = tot;
    tot += K.npts;
    if (j < nlay) tot_base += K.npts;
    nmax = std::max(nmax, K.npts);
  }
  const int64_t ntop = plans[nlay - 1].npts;
  if (n_out < ntop) return LBLX_E_SIZE;
  // merge tables (as lblx_spectrum_device)
  std::vector<int> itype(nlay, 0), pan0(nlay);
  std::vector<DevPanel> pan;
  for (int l = 0; l < nlay; l++) {
    pan0[l] = (int)pan.size();
    std::vector<int64_t> base;
    if (l > 0) {
      int rc = merge_itype(plans[l - 1].dv, plans[l].dv, &itype[l]);
      if (rc) return rc;
      if ((rc = merge_bases(plans[l - 1], plans[l], itype[l], plans[l].v2, base))) return rc;
    }
    merge_panels(plans[l], l > 0 ? &base : nullptr, pan);
  }
  const bool f32 = run->precision == LBLX_FP32_SPLIT;
  const size_t esz = f32 ? 4 : 8;
  const int64_t plane = nmax;
  // FP32 tangent planes (forward mode); reverse mode: two adjoint rows and
  // the blocks' partial sums
  const int64_t nblk_max = (nmax + ADJ_T - 1) / ADJ_T;
  const size_t jbytes = adj ? (size_t)2 * nmax * 8 : (size_t)P * nlay * plane * 4;
  const size_t jbytes_b = adj ? (size_t)nblk_max * P * 8 : jbytes;
  if (c->w_od.ensure((size_t)tot * esz) || c->w_rt.ensure((size_t)nmax * 8 * 4) ||
      c->w_mpan.ensure(pan.size() * sizeof(DevPanel)) || c->j_rall.ensure((size_t)tot_base * 8) ||
      c->j_a.ensure(jbytes) || c->j_b.ensure(jbytes_b))
    return LBLX_E_HIP;
  // 1. OPDPTH for the column and every perturbed layer, one batch (aj = 2:
  // the column alone, with its tangent-linear optical depths)
  TanReq req;
  if (tl) {
    int rq = tan_request(run, jobs.data(), nlay, species, nspec, c->w_od.p, off.data() + nlay, req);
    if (rq) return rq;
  }
  int rc = run_jobs(c, run, jobs.data(), nj, false, c->w_od.p, off.data(), f32, nullptr, nullptr,
                    jplans.data(), tl ? &req : nullptr);
  if (rc) return rc;
  hipStream_t s = c->stream;
  c->stage_begin();
  HIPCHK(c->h2d(c->w_mpan.p, pan.data(), pan.size() * sizeof(DevPanel), s));
  c->stage_end(s);
  const DevPanel *dpan = c->w_mpan.as<DevPanel>();
  double *rall = c->j_rall.as<double>();
  double *tt[2] = {c->w_rt.as<double>(), c->w_rt.as<double>() + nmax};
  // 2. XMERGE up the column, keeping every layer's radiance: runs of
  // same-grid layers in one pass each, as the spectrum does
  {
    std::vector<int> rstart;
    std::vector<RunLayer> rl(nlay);
    for (int l = 0; l < nlay; l++) {
      rl[l] = RunLayer{layers[l].tave / RADCN2, tzu[l] / RADCN2, off[l], off[l]};
      if (l == 0 || !merge_is_identity(plans[l - 1], plans[l], itype[l], pan.data() + pan0[l],
                                       pan.data() + pan0[l - 1]))
        rstart.push_back(l);
    }
    rstart.push_back(nlay);
    if (c->w_mrun.ensure((size_t)nlay * sizeof(RunLayer))) return LBLX_E_HIP;
    c->stage_begin();
    HIPCHK(c->h2d(c->w_mrun.p, rl.data(), (size_t)nlay * sizeof(RunLayer), s));
    c->stage_end(s);
    int thm = c->tbegin(ST_MERGE);
    for (size_t q = 0; q + 1 < rstart.size(); q++) {
      const int l = rstart[q], nrun = rstart[q + 1] - l, last = l + nrun - 1;
      MergeArgs A = merge_args(plans[l], l > 0 ? &plans[l - 1] : nullptr, itype[l], layers[l].tave, tzu[l],
                               l == 0 ? tbnd : 0., emis);
      const double *ro = l > 0 ? rall + off[l - 1] : nullptr;
      const double *to = l > 0 ? tt[(q - 1) & 1] : nullptr;
      const RunLayer *drl = c->w_mrun.as<RunLayer>() + l;
      dim3 grid((unsigned)((A.n + 255) / 256));
      if (f32)
        hipLaunchKernelGGL(k_merge_run<float>, grid, dim3(256), 0, s, A, drl, nrun, dpan + pan0[l],
                           c->w_od.as<float>(), ro, to, rall + off[last], tt[q & 1], rall);
      else
        hipLaunchKernelGGL(k_merge_run<double>, grid, dim3(256), 0, s, A, drl, nrun, dpan + pan0[l],
                           c->w_od.as<double>(), ro, to, rall + off[last], tt[q & 1], rall);
    }
    c->tend(thm);
    HIPCHK(hipGetLastError());
  }
  if (d_rad) HIPCHK(hipMemcpyAsync(d_rad, rall + off[nlay - 1], (size_t)ntop * 8, hipMemcpyDeviceToDevice, s));
  if (adj) {
    // 3'. reverse mode: lam from the top down; each layer's parameters take
    // the sum of lam times their source terms, and lam moves to the layer
    // below through the transposed merge
    int tha = c->tbegin(ST_MERGE);
    double *lam[2] = {c->j_a.as<double>(), c->j_a.as<double>() + nmax};
    double *part = c->j_b.as<double>();
    HIPCHK(hipMemcpyAsync(lam[0], d_dLdR, (size_t)ntop * 8, hipMemcpyDeviceToDevice, s));
    int cur = 0;
    for (int l = nlay - 1; l >= 0; l--) {
      MergeArgs A = merge_args(plans[l], l > 0 ? &plans[l - 1] : nullptr, itype[l], layers[l].tave, tzu[l],
                               l == 0 ? tbnd : 0., emis);
      JacSrc S{};
      S.A = A;
      S.tave = layers[l].tave;
      S.nparam = P;
      for (int q = 0; q < P; q++) {
        S.direct[q] = 1;
        S.rh[q] = 1.0;
        S.odp[q] = off[nlay + l * P + q] - off[l];
      }
      const void *od = (const char *)c->w_od.p + (size_t)off[l] * esz;
      const double *rold = l > 0 ? rall + off[l - 1] : nullptr;
      const int nb = (int)((A.n + ADJ_T - 1) / ADJ_T);
      if (f32)
        hipLaunchKernelGGL(k_adj_src<float>, dim3(nb), dim3(ADJ_T), 0, s, S, dpan + pan0[l], (const float *)od,
                           (const float *)od, rold, (const double *)lam[cur], part);
      else
        hipLaunchKernelGGL(k_adj_src<double>, dim3(nb), dim3(ADJ_T), 0, s, S, dpan + pan0[l], (const double *)od,
                           (const double *)od, rold, (const double *)lam[cur], part);
      hipLaunchKernelGGL(k_adj_sum, dim3(P), dim3(ADJ_T), 0, s, (const double *)part, nb, P, d_grad + l,
                         (int64_t)nlay);
      if (l == 0) break;
      const int64_t nold = plans[l - 1].npts;
      hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((nold + 255) / 256)), dim3(256), 0, s, lam[cur ^ 1], nold, 0.0);
      const dim3 g((unsigned)((A.n + 255) / 256));
      if (f32)
        hipLaunchKernelGGL(k_adj_prop<float>, g, dim3(256), 0, s, A, dpan + pan0[l], (const float *)od,
                           (const double *)lam[cur], lam[cur ^ 1]);
      else
        hipLaunchKernelGGL(k_adj_prop<double>, g, dim3(256), 0, s, A, dpan + pan0[l], (const double *)od,
                           (const double *)lam[cur], lam[cur ^ 1]);
      cur ^= 1;
    }
    c->tend(tha);
    HIPCHK(hipGetLastError());
    if (v1p_top) *v1p_top = plans[nlay - 1].panels[0].vft + (double)NSHIFT * plans[nlay - 1].dv;
    return LBLX_OK;
  }
  // 3. sources and tangent propagation.  The layers at the top whose merge
  // is the identity map (ITYPE 0, same panels: J <- T_k J pointwise) are
  // handled from the top down with a running transmittance product (K0:
  // the lowest of them; merge K0 - 1 is a real one).
  int K0 = nlay;
  for (int l = nlay - 1; l >= 1; l--) {
    if (!merge_is_identity(plans[l - 1], plans[l], itype[l], pan.data() + pan0[l], pan.data() + pan0[l - 1]))
      break;
    K0 = l;
  }
  int th = c->tbegin(ST_MERGE);
  float *ja = c->j_a.as<float>(), *jb = c->j_b.as<float>();
  auto src_launch = [&](int l, auto *dst, int64_t stride, const double *scale) {
    using JP = std::remove_pointer_t<decltype(dst)>;
    MergeArgs A = merge_args(plans[l], l > 0 ? &plans[l - 1] : nullptr, itype[l], layers[l].tave, tzu[l],
                             l == 0 ? tbnd : 0., emis);
    const dim3 grid((unsigned)((A.n + 255) / 256));
    const void *od = (const char *)c->w_od.p + (size_t)off[l] * esz;
    JacSrc S{};
    S.A = A;
    S.tave = layers[l].tave;
    S.stride = stride;
    S.nparam = P;
    for (int q = 0; q < P; q++) {
      S.direct[q] = (tl || (aj && q > 0)) ? 1 : 0;
      S.rh[q] = tl ? 1.0 : (q == 0 ? 1.0 / dT : (aj ? 1.0 : 1.0 / dlnq));
      S.odp[q] = off[nlay + l * P + q] - off[l];  // relative to the layer's own optical depth
    }
    const double *rold = l > 0 ? rall + off[l - 1] : nullptr;
    if (f32)
      hipLaunchKernelGGL((k_jac_src<float, JP>), grid, dim3(256), 0, s, S, dpan + pan0[l], (const float *)od,
                         (const float *)od, rold, dst, scale);
    else
      hipLaunchKernelGGL((k_jac_src<double, JP>), grid, dim3(256), 0, s, S, dpan + pan0[l], (const double *)od,
                         (const double *)od, rold, dst, scale);
  };
  const dim3 gtop((unsigned)((ntop + 255) / 256));
  const int64_t qstr = (int64_t)nlay * n_out;  // between parameter planes of d_jac
  double *ttop = c->w_rt.as<double>() + 3 * nmax;  // transmittance product of the top region
  double *st = c->w_rt.as<double>() + 2 * nmax;  // pending product of a run of identity merges
  auto trans_launch = [&](int l, double *prod, int64_t n) {
    const void *od = (const char *)c->w_od.p + (size_t)off[l] * esz;
    const dim3 g((unsigned)((n + 255) / 256));
    if (f32) hipLaunchKernelGGL(k_jac_trans<float>, g, dim3(256), 0, s, (const float *)od, n, prod);
    else hipLaunchKernelGGL(k_jac_trans<double>, g, dim3(256), 0, s, (const double *)od, n, prod);
  };
  // the top region first (top down, its sources scaled by the product of
  // the transmittances above them), so that its product is known below
  if (K0 < nlay) {
    hipLaunchKernelGGL(k_fill_f64, gtop, dim3(256), 0, s, ttop, ntop, 1.0);
    for (int l = nlay - 1; l >= K0; l--) {
      src_launch(l, d_jac + (int64_t)l * n_out, qstr, ttop);
      trans_launch(l, ttop, ntop);
    }
  }
  // Below it, layer by layer up to K0 - 1.  A run of layers a..b whose
  // merges a+1..b are identity maps is taken at once: the sources of its
  // layers j > a are scaled by the product of the run's transmittances
  // above j, and the tangents from layers <= a carry the whole run's
  // product into the next real merge (st, applied at the taps).  The last
  // layer (K0 - 1) writes into d_jac scaled by the top region's product.
  int nsc = 0;
  for (int l = 0; l < K0;) {
    const bool to_out = K0 < nlay && l == K0 - 1;
    if (l > 0) {
      MergeArgs A = merge_args(plans[l], &plans[l - 1], itype[l], layers[l].tave, tzu[l], 0., emis);
      const dim3 grid((unsigned)((A.n + 255) / 256));
      const void *od = (const char *)c->w_od.p + (size_t)off[l] * esz;
      const double *sp = nsc > 0 ? st : nullptr;
      if (to_out) {
        JacOut<double> O{d_jac, n_out, qstr, P, ttop};
        if (f32)
          hipLaunchKernelGGL((k_jac_prop<float, double>), grid, dim3(256), 0, s, A, dpan + pan0[l],
                             (const float *)od, (const float *)ja, plane, l * P, sp, nsc, O);
        else
          hipLaunchKernelGGL((k_jac_prop<double, double>), grid, dim3(256), 0, s, A, dpan + pan0[l],
                             (const double *)od, (const float *)ja, plane, l * P, sp, nsc, O);
      } else {
        JacOut<float> O{jb, (int64_t)P * plane, plane, P, nullptr};
        if (f32)
          hipLaunchKernelGGL((k_jac_prop<float, float>), grid, dim3(256), 0, s, A, dpan + pan0[l],
                             (const float *)od, (const float *)ja, plane, l * P, sp, nsc, O);
        else
          hipLaunchKernelGGL((k_jac_prop<double, float>), grid, dim3(256), 0, s, A, dpan + pan0[l],
                             (const double *)od, (const float *)ja, plane, l * P, sp, nsc, O);
      }
      nsc = 0;
      std::swap(ja, jb);
    }
    if (to_out) {
      src_launch(l, d_jac + (int64_t)l * n_out, qstr, ttop);
      break;
    }
    src_launch(l, ja + (int64_t)l * P * plane, plane, nullptr);
    int b = l;
    while (b + 1 < K0 && merge_is_identity(plans[b], plans[b + 1], itype[b + 1], pan.data() + pan0[b + 1],
                                           pan.data() + pan0[b]))
      b++;
    if (b > l) {
      const int64_t ng = plans[l].npts;
      hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, s, st, ng, 1.0);
      for (int j = b; j > l; j--) {
        src_launch(j, ja + (int64_t)j * P * plane, plane, st);
        trans_launch(j, st, ng);
      }
      nsc = (l + 1) * P;
    }
    l = b + 1;
  }
  if (K0 == nlay) {
    // no identity region at the top: a pending product still applies to
    // the tangents from below its run; the planes go to d_jac as they are
    hipLaunchKernelGGL(k_jac_out, gtop, dim3(256), 0, s, (const float *)ja, plane, nlay * P, P,
                       nsc > 0 ? (const double *)st : nullptr, nsc, ntop, d_jac, n_out, qstr);
  }
  c->tend(th);
  HIPCHK(hipGetLastError());
  if (v1p_top) *v1p_top = plans[nlay - 1].panels[0].vft + (double)NSHIFT * plans[nlay - 1].dv;
  return LBLX_OK;
}

int lblx_jacobian_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay,
                         const double *tzu, double tbnd, double emis, const int *species, int nspec,
                         double dT, double dlnq, double *d_rad, double *d_jac, int64_t n_out,
                         double *v1p_top) {
  return jacobian_run(c, run, layers, nlay, tzu, tbnd, emis, species, nspec, dT, dlnq, 0, d_rad, d_jac,
                      n_out, v1p_top);
}

int lblx_jacobian_aj_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay,
                            const double *tzu, double tbnd, double emis, const int *species, int nspec,
                            double *d_rad, double *d_jac, int64_t n_out, double *v1p_top) {
  return jacobian_run(c, run, layers, nlay, tzu, tbnd, emis, species, nspec, 1.0, 1.0, 1, d_rad, d_jac,
                      n_out, v1p_top);
}

int lblx_od_tangent_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay, const int *species,
                           int nspec, void *d_od, void *d_dod, int64_t n_out) {
  if (!c || !run || !layers || nlay <= 0 || !d_od || !d_dod || n_out <= 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const int P = 1 + nspec;
  for (int l = 1; l < nlay; l++) layers[l].jrad = layer_jrad(&layers[l - 1]);
  for (int l = 0; l < nlay; l++) {
    int64_t n = 0;
    int rc = lblx_od_size(run, &layers[l], &n);
    if (rc) return rc;
    if (n > n_out) return LBLX_E_SIZE;
  }
  std::vector<int64_t> off(nlay), doff((size_t)nlay * P);
  for (int l = 0; l < nlay; l++) {
    off[l] = (int64_t)l * n_out;
    for (int q = 0; q < P; q++) doff[(size_t)l * P + q] = ((int64_t)l * P + q) * n_out;
  }
  TanReq req;
  int rc = tan_request(run, layers, nlay, species, nspec, d_dod, doff.data(), req);
  if (rc) return rc;
  const bool f32 = run->precision == LBLX_FP32_SPLIT;
  // JRAD is already chained above: each layer plans from its own input flag
  return run_jobs(c, run, layers, nlay, false, d_od, off.data(), f32, nullptr, nullptr, nullptr, &req);
}

int lblx_jacobian_tl_device(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay,
                            const double *tzu, double tbnd, double emis, const int *species, int nspec,
                            double *d_rad, double *d_jac, int64_t n_out, double *v1p_top) {
  return jacobian_run(c, run, layers, nlay, tzu, tbnd, emis, species, nspec, 1.0, 1.0, 2, d_rad, d_jac,
                      n_out, v1p_top);
}

int lblx_backward_batch(lblx_ctx *c, const lblx_run *run, int nprof, const int *nlay, lblx_layer *layers,
                        const double *tzu, const double *tbnd, const double *emis, const int *species, int nspec,
                        const double *d_dLdR, const int64_t *rad_off, double *d_rad, double *d_grad,
                        const int64_t *grad_off, double *v1p_top) {
  if (!c || !run || nprof <= 0 || !nlay || !layers || !tzu || !tbnd || !emis || !d_dLdR || !rad_off || !d_grad ||
      !grad_off)
    return LBLX_E_ARG;
  int64_t l0 = 0;
  for (int p = 0; p < nprof; p++) {
    if (nlay[p] <= 0) return LBLX_E_ARG;
    int64_t n_out = 0;
    int rc = lblx_od_size(run, &layers[l0 + nlay[p] - 1], &n_out);
    if (rc) return rc;
    rc = jacobian_run(c, run, layers + l0, nlay[p], tzu + l0, tbnd[p], emis[p], species, nspec, 1.0, 1.0, 2,
                      d_rad ? d_rad + rad_off[p] : nullptr, nullptr, n_out, v1p_top ? v1p_top + p : nullptr,
                      d_dLdR + rad_off[p], d_grad + grad_off[p]);
    if (rc) return rc;
    l0 += nlay[p];
  }
  return LBLX_OK;
}

int lblx_jacobians_batch_device(lblx_ctx *c, const lblx_run *run, int nprof, const int *nlay,
                                lblx_layer *layers, const double *tzu, const double *tbnd,
                                const double *emis, const int *species, int nspec, int aj, double dT,
                                double dlnq, double *d_rad, const int64_t *rad_off, double *d_jac,
                                const int64_t *jac_off, double *v1p_top) {
  if (!c || !run || nprof <= 0 || !nlay || !layers || !tzu || !tbnd || !emis || !d_rad || !rad_off || !d_jac ||
      !jac_off)
    return LBLX_E_ARG;
  if (aj < 0 || aj > 2) return LBLX_E_ARG;
  if (aj) dT = dlnq = 1.0;
  int64_t l0 = 0;
  for (int p = 0; p < nprof; p++) {
    if (nlay[p] <= 0) return LBLX_E_ARG;
    int64_t n_out = 0;
    int rc = lblx_od_size(run, &layers[l0 + nlay[p] - 1], &n_out);
    if (rc) return rc;
    rc = jacobian_run(c, run, layers + l0, nlay[p], tzu + l0, tbnd[p], emis[p], species, nspec, dT, dlnq,
                      aj, d_rad + rad_off[p], d_jac + jac_off[p], n_out,
                      v1p_top ? v1p_top + p : nullptr);
    if (rc) return rc;
    l0 += nlay[p];
  }
  return LBLX_OK;
}

// Per-job device bytes of run_jobs (line-table-sized work arrays).
// Device bytes one job takes in run_jobs at a per-job line stride of
// about nline records.
static size_t job_bytes(const lblx_ctx *c, int64_t nline) {
  const size_t line = ((size_t)nline + 127) / 64 * 64;
  return line * (32 + 1 + 4 + 4 * 3 + 2 + 12 + 36 + sizeof(NearHot) + sizeof(NearCold) + 8 + sizeof(SRec1) + 2 * sizeof(SRecA) + sizeof(SRecV) +
                 (c->ncpl > 0 ? 24 : 0)) +
         (size_t)(c->nblk + 2) * 4 * (3 + EBLK + 1 + 3 + 6) + (size_t)(c->nlines / 1000 + 16) * 24;
}

// Many profiles in one call: OPDPTH for the layers of a group of profiles
// in one batch, then XMERGE for all profiles of the group with one launch
// per layer index (ragged columns), the chain of merge launches captured
// once as a HIP graph and replayed on later calls of the same shape.
int lblx_spectra_batch_device(lblx_ctx *c, const lblx_run *run, int nprof, const int *nlay,
                              lblx_layer *layers, const double *tzu, const double *tbnd,
                              const double *emis, double *d_rad, double *d_tra, const int64_t *out_off,
                              double *v1p_top) {
  if (!c || !run || nprof <= 0 || !nlay || !layers || !tzu || !tbnd || !emis || !d_rad || !d_tra || !out_off)
    return LBLX_E_ARG;
  if (c->nlines <= 0) return LBLX_E_NOLINES;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  c->hbegin();
  std::vector<int> first(nprof + 1, 0);
  for (int p = 0; p < nprof; p++) {
    if (nlay[p] <= 0) return LBLX_E_ARG;
    first[p + 1] = first[p] + nlay[p];
  }
  const int ntot = first[nprof];
  std::vector<LayerPlan> plans(ntot);
  std::vector<int64_t> jlines(ntot, 0);
  std::vector<int> itype(ntot, 0), pan0(ntot, 0);
  std::vector<std::vector<DevPanel>> ppan(nprof);
  std::vector<int> prc(nprof, 0);
  // host planning of one profile: layer plans with JRAD chained up the
  // column, line ranges, and the merge panels of each layer
  auto plan_prof = [&](int p) -> int {
    int jr = layers[first[p]].jrad;
    for (int l = first[p]; l < first[p + 1]; l++) {
      layers[l].jrad = jr;
      int rc = plan_layer(run, &layers[l], &plans[l]);
      if (rc) return rc;
      JobLines JL;
      jlines[l] = job_lines(c, plans[l], &layers[l], &JL);
      jr = layers[l].jrad;
      pan0[l] = (int)ppan[p].size();
      std::vector<int64_t> base;
      if (l > first[p]) {
        rc = merge_itype(plans[l - 1].dv, plans[l].dv, &itype[l]);
        if (rc) return rc;
        if ((rc = merge_bases(plans[l - 1], plans[l], itype[l], plans[l].v2, base))) return rc;
      }
      merge_panels(plans[l], l > first[p] ? &base : nullptr, ppan[p]);
    }
    if (v1p_top) {
      const LayerPlan &T = plans[first[p + 1] - 1];
      v1p_top[p] = T.panels[0].vft + (double)NSHIFT * T.dv;
    }
    return 0;
  };
  {  // profiles are independent: plan them on host threads
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = (int)std::min<unsigned>((unsigned)nprof, std::min(16u, hw));
    std::atomic<int> next{0};
    auto worker = [&]() {
      for (int p = next++; p < nprof; p = next++) prc[p] = plan_prof(p);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nth; t++) pool.emplace_back(worker);
    worker();
    for (auto &t : pool) t.join();
  }
  for (int p = 0; p < nprof; p++)
    if (prc[p]) return prc[p];
  std::vector<DevPanel> pan;
  for (int p = 0; p < nprof; p++) {
    const int off = (int)pan.size();
    for (int l = first[p]; l < first[p + 1]; l++) pan0[l] += off;
    pan.insert(pan.end(), ppan[p].begin(), ppan[p].end());
  }
  c->hend();
  if (c->w_mpan.ensure(pan.size() * sizeof(DevPanel))) return LBLX_E_HIP;
  HIPCHK(hipMemcpy(c->w_mpan.p, pan.data(), pan.size() * sizeof(DevPanel), hipMemcpyHostToDevice));
  const DevPanel *dpan = c->w_mpan.as<DevPanel>();
  // groups of whole profiles within the device-memory budget
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  const size_t budget = std::min<size_t>(free_b / 2, (size_t)96 << 30);
  const bool f32 = run->precision == LBLX_FP32_SPLIT;
  const size_t esz = f32 ? 4 : 8;
  struct Group {
    int p0, p1, njobs, maxl;
    std::vector<int64_t> off, nmax, rt_off;
    int64_t tot, rt_tot;
  };
  std::vector<Group> groups;
  int64_t max_tot = 0, max_rt = 0;
  int max_jobs = 0;
  for (int p0 = 0; p0 < nprof;) {
    // whole profiles while njobs x (bytes per job at the group's widest
    // line range) fits the budget
    Group G{};
    int p1 = p0, njobs = 0;
    int64_t wl = 0;
    while (p1 < nprof) {
      int64_t w = wl;
      for (int l = first[p1]; l < first[p1 + 1]; l++) w = std::max(w, jlines[l]);
      if (njobs > 0 && (size_t)(njobs + nlay[p1]) * job_bytes(c, w) > budget) break;
      wl = w;
      njobs += nlay[p1++];
    }
    G.p0 = p0;
    G.p1 = p1;
    G.njobs = njobs;
    G.off.resize(njobs);
    G.nmax.assign(p1 - p0, 0);
    G.rt_off.assign(p1 - p0, 0);
    G.tot = G.rt_tot = 0;
    G.maxl = 0;
    for (int p = p0; p < p1; p++) {
      for (int l = first[p]; l < first[p + 1]; l++) {
        G.off[l - first[p0]] = G.tot;
        G.tot += plans[l].npts;
        G.nmax[p - p0] = std::max(G.nmax[p - p0], plans[l].npts);
      }
      G.rt_off[p - p0] = G.rt_tot;
      G.rt_tot += 4 * G.nmax[p - p0];
      G.maxl = std::max(G.maxl, nlay[p]);
    }
    max_tot = std::max(max_tot, G.tot);
    max_rt = std::max(max_rt, G.rt_tot);
    max_jobs = std::max(max_jobs, njobs);
    groups.push_back(std::move(G));
    p0 = p1;
  }
  // the buffers every group's descriptors point into, sized once
  if (c->w_od.ensure((size_t)max_tot * esz) || c->b_rt.ensure((size_t)max_rt * 8) ||
      c->b_jobs.ensure((size_t)max_jobs * sizeof(MergeJob)) ||
      c->b_rl.ensure((size_t)max_jobs * sizeof(RunLayer)))
    return LBLX_E_HIP;
  // Host plan of one group: its OPDPTH jobs and the merge descriptors of
  // every layer index (kstart/kcount: the launch of layer index k).
  struct GroupPlan {
    JobsPrep J;
    std::vector<MergeJob> mj;
    std::vector<RunLayer> rl;
    std::vector<int> kstart, kcount;
    std::vector<int64_t> key;
    int rc = 0;
  };
  auto plan_group = [&](int g, GroupPlan &P) {
    const Group &G = groups[g];
    P.rc = prep_jobs(c, run, layers + first[G.p0], G.njobs, false, G.off.data(), plans.data() + first[G.p0],
                     P.J);
    if (P.rc) return;
    // runs of same-grid layers of each profile (identity merges after the
    // first); launch k takes the k-th run of every profile
    P.mj.clear();
    P.rl.assign(G.njobs, RunLayer{});
    std::vector<std::vector<int>> rs(G.p1 - G.p0);
    int nstep = 0;
    for (int p = G.p0; p < G.p1; p++) {
      std::vector<int> &r = rs[p - G.p0];
      for (int k = 0; k < nlay[p]; k++) {
        const int l = first[p] + k;
        P.rl[l - first[G.p0]] = RunLayer{layers[l].tave / RADCN2, tzu[l] / RADCN2, G.off[l - first[G.p0]], -1};
        if (k == 0 || !merge_is_identity(plans[l - 1], plans[l], itype[l], pan.data() + pan0[l],
                                         pan.data() + pan0[l - 1]))
          r.push_back(k);
      }
      r.push_back(nlay[p]);
      nstep = std::max(nstep, (int)r.size() - 1);
    }
    P.kstart.assign(nstep + 1, 0);
    P.kcount.assign(nstep, 0);
    P.key = {(int64_t)g, (int64_t)f32, (int64_t)nstep, (int64_t)(intptr_t)c->b_jobs.p};
    const RunLayer *drl = c->b_rl.as<RunLayer>();
    for (int k = 0; k < nstep; k++) {
      P.kstart[k] = (int)P.mj.size();
      int64_t nk = 0;
      for (int p = G.p0; p < G.p1; p++) {
        const std::vector<int> &r = rs[p - G.p0];
        if (k + 1 >= (int)r.size()) continue;
        const int l = first[p] + r[k], nrun = r[k + 1] - r[k];
        double *rt = c->b_rt.as<double>() + G.rt_off[p - G.p0];
        const int64_t m = G.nmax[p - G.p0];
        double *R[2] = {rt, rt + m}, *T[2] = {rt + 2 * m, rt + 3 * m};
        const bool last = r[k + 1] == nlay[p];
        MergeJob M{};
        M.A = merge_args(plans[l], r[k] > 0 ? &plans[l - 1] : nullptr, itype[l], layers[l].tave, tzu[l],
                         r[k] == 0 ? tbnd[p] : 0., emis[p]);
        M.pan = dpan + pan0[l];
        M.od = c->w_od.p;
        M.rl = drl + (l - first[G.p0]);
        M.nrun = nrun;
        M.rold = k > 0 ? R[(k - 1) & 1] : nullptr;
        M.told = k > 0 ? T[(k - 1) & 1] : nullptr;
        M.rnew = last ? d_rad + out_off[p] : R[k & 1];
        M.tnew = last ? d_tra + out_off[p] : T[k & 1];
        P.mj.push_back(M);
        nk = std::max(nk, plans[l].npts);
      }
      P.kcount[k] = (int)P.mj.size() - P.kstart[k];
      P.key.push_back(P.kcount[k]);
      P.key.push_back(nk);
    }
  };
  // Group g + 1 is planned on a host thread while group g runs.  A plan
  // slot is rewritten only after the uploads made from it have completed
  // (event after them).
  GroupPlan gp[2];
  hipEvent_t up_done[2] = {c->ev_get(), c->ev_get()};
  auto put_events = [&]() {
    c->evpool.push_back(up_done[0]);
    c->evpool.push_back(up_done[1]);
  };
  c->hbegin();
  plan_group(0, gp[0]);
  c->hend();
  const int ng = (int)groups.size();
  int rc = LBLX_OK;
  for (int g = 0; g < ng && rc == LBLX_OK; g++) {
    GroupPlan &P = gp[g & 1];
    std::thread helper;
    if (g + 1 < ng) {
      GroupPlan &Q = gp[(g + 1) & 1];
      hipEvent_t ev = up_done[(g + 1) & 1];
      const bool used = g >= 1;
      helper = std::thread([&, g, ev, used]() {
        if (used) (void)hipEventSynchronize(ev);
        plan_group(g + 1, Q);
      });
    }
    auto body = [&]() -> int {
      if (P.rc) return P.rc;
      // 1. optical depths of every layer of the group's profiles
      int r = launch_jobs(c, run, P.J, c->w_od.p, f32, nullptr, nullptr);
      if (r) return r;
      // 2. merge descriptors, then the chain of merge launches as a graph
      // (instantiated once per shape)
      HIPCHK(hipMemcpyAsync(c->b_jobs.p, P.mj.data(), P.mj.size() * sizeof(MergeJob), hipMemcpyHostToDevice,
                            s));
      HIPCHK(hipMemcpyAsync(c->b_rl.p, P.rl.data(), P.rl.size() * sizeof(RunLayer), hipMemcpyHostToDevice, s));
      HIPCHK(hipEventRecord(up_done[g & 1], s));
      hipGraphExec_t exec = nullptr;
      for (auto &e : c->graphs)
        if (e.key == P.key) exec = e.exec;
      if (!exec) {
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        const MergeJob *dj = c->b_jobs.as<MergeJob>();
        const int nstep = (int)P.kcount.size();
        for (int k = 0; k < nstep; k++) {
          const int64_t nk = P.key[4 + 2 * k + 1];
          dim3 grid((unsigned)((nk + 255) / 256), (unsigned)P.kcount[k]);
          if (f32) hipLaunchKernelGGL(k_merge_batch<float>, grid, dim3(256), 0, s, dj + P.kstart[k]);
          else hipLaunchKernelGGL(k_merge_batch<double>, grid, dim3(256), 0, s, dj + P.kstart[k]);
        }
        HIPCHK(hipStreamEndCapture(s, &graph));
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(ie);
        if (c->graphs.size() >= 64) {  // bounded cache
          (void)hipGraphExecDestroy(c->graphs.front().exec);
          c->graphs.erase(c->graphs.begin());
        }
        c->graphs.push_back(lblx_ctx::GraphEntry{P.key, exec});
      }
      int th = c->tbegin(ST_MERGE);
      HIPCHK(hipGraphLaunch(exec, s));
      c->tend(th);
      return LBLX_OK;
    };
    rc = body();
    c->hbegin();  // host time the device waits for the next group's plan
    if (helper.joinable()) helper.join();
    c->hend();
  }
  // uploads and merges done before the host plans are released
  hipError_t se = hipStreamSynchronize(s);
  put_events();
  if (rc) return rc;
  HIPCHK(se);
  HIPCHK(hipGetLastError());
  return LBLX_OK;
}

// Tile widths of the near-wing scatter, chosen by timing the three tile
// stages on the given layers for each candidate (coordinate search).
int lblx_autotune(lblx_ctx *c, const lblx_run *run, lblx_layer *layers, int nlay, int *tiles_out) {
  if (!c || !run || !layers || nlay <= 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  std::vector<int64_t> off(nlay);
  int64_t tot = 0;
  for (int l = 0; l < nlay; l++) {
    lblx_layer t = layers[l];
    LayerPlan P;
    int rc = plan_layer(run, &t, &P);
    if (rc) return rc;
    off[l] = tot;
    tot += P.npts;
  }
  const size_t esz = run->precision == LBLX_FP32_SPLIT ? 4 : 8;
  if (c->w_od.ensure((size_t)tot * esz)) return LBLX_E_HIP;
  const bool was = c->timing;
  auto trial = [&](int stage) -> double {
    std::vector<lblx_layer> tmp(layers, layers + nlay);
    c->tcollect();
    c->timing = true;
    for (int i = 0; i < ST_N; i++) {
      c->stage_ms[i] = 0.;
      c->stage_n[i] = 0;
    }
    double best = 1e30;
    for (int rep = 0; rep < 2; rep++) {  // first pass grows buffers
      std::vector<lblx_layer> t2 = tmp;
      if (run_jobs(c, run, t2.data(), nlay, true, c->w_od.p, off.data(), esz == 4, nullptr, nullptr)) return 1e30;
      c->tcollect();
      best = rep ? std::min(best, c->stage_ms[stage]) : best;
      c->stage_ms[stage] = 0.;
    }
    return best;
  };
  const int cand[3][4] = {{1024, 2048, 4096, 8192}, {1024, 2048, 4096, 8192}, {512, 1024, 2048, 4096}};
  const int stage[3] = {ST_TILE1, ST_TILE2, ST_TILE3};
  for (int G = 0; G < 3; G++) {
    int bestw = c->tiles[G];
    double bt = 1e30;
    for (int w : cand[G]) {
      c->tiles[G] = w;
      double t = trial(stage[G]);
      if (t < bt) {
        bt = t;
        bestw = w;
      }
    }
    c->tiles[G] = bestw;
    // then the line batch of the tile kernel at that width
    const int chs[4] = {8, 16, 32, 64};
    int bestc = c->lbatch[G];
    double bc = 1e30;
    for (int ch : chs) {
      c->lbatch[G] = ch;
      double t = trial(stage[G]);
      if (t < bc) {
        bc = t;
        bestc = ch;
      }
    }
    c->lbatch[G] = bestc;
  }
  c->timing = was;
  for (int i = 0; i < ST_N; i++) {
    c->stage_ms[i] = 0.;
    c->stage_n[i] = 0;
  }
  if (tiles_out)
    for (int G = 0; G < 3; G++) tiles_out[G] = c->tiles[G];
  return LBLX_OK;
}

int lblx_voigt_tables(double *f, double *coef) {
  if (!f || !coef) return LBLX_E_ARG;
  std::vector<double> *cf = nullptr;
  const std::vector<double> &t = voigt_host(&cf);
  std::memcpy(f, t.data(), (size_t)3 * 101 * 2601 * 8);
  std::memcpy(coef, cf->data(), 6 * 101 * 8);
  return LBLX_OK;
}

static bool tuning_ok(const lblx_tuning *t) {
  const int lo[3] = {256, 128, 64};
  for (int G = 0; G < 3; G++) {
    if (t->tiles[G] < lo[G] || t->tiles[G] > 8192 || (t->tiles[G] & 63)) return false;
    if (t->line_batch[G] < 0 || t->line_batch[G] > 64) return false;
  }
  return true;
}

int lblx_get_tuning(lblx_ctx *c, lblx_tuning *t) {
  if (!c || !t) return LBLX_E_ARG;
  for (int G = 0; G < 3; G++) {
    t->tiles[G] = c->tiles[G];
    t->line_batch[G] = c->lbatch[G];
  }
  return LBLX_OK;
}

int lblx_set_tuning(lblx_ctx *c, const lblx_tuning *t) {
  if (!c || !t || !tuning_ok(t)) return LBLX_E_ARG;
  for (int G = 0; G < 3; G++) {
    c->tiles[G] = t->tiles[G];
    c->lbatch[G] = t->line_batch[G];
  }
  return LBLX_OK;
}

// The device a tuning belongs to: its architecture, name and compute units.
static std::string tuning_device_key(lblx_ctx *c) {
  hipDeviceProp_t pr{};
  if (hipGetDeviceProperties(&pr, c->device) != hipSuccess) return "unknown";
  std::string name = pr.gcnArchName;
  name += "|";
  name += pr.name;
  name += "|" + std::to_string(pr.multiProcessorCount);
  for (char &ch : name)
    if (ch == ' ' || ch == '\n') ch = '_';
  return name;
}

int lblx_tuning_save(lblx_ctx *c, const char *path) {
  if (!c || !path) return LBLX_E_ARG;
  FILE *f = std::fopen(path, "w");
  if (!f) return LBLX_E_ARG;
  std::fprintf(f, "lblx-tuning 1 %s\n", tuning_device_key(c).c_str());
  std::fprintf(f, "tiles %d %d %d\nline_batch %d %d %d\n", c->tiles[0], c->tiles[1], c->tiles[2], c->lbatch[0],
               c->lbatch[1], c->lbatch[2]);
  return std::fclose(f) == 0 ? LBLX_OK : LBLX_E_ARG;
}

int lblx_tuning_load(lblx_ctx *c, const char *path) {
  if (!c || !path) return LBLX_E_ARG;
  FILE *f = std::fopen(path, "r");
  if (!f) return LBLX_E_ARG;
  char key[512] = {0};
  int ver = 0;
  lblx_tuning t{};
  const int ok = std::fscanf(f, "lblx-tuning %d %511s tiles %d %d %d line_batch %d %d %d", &ver, key, &t.tiles[0],
                             &t.tiles[1], &t.tiles[2], &t.line_batch[0], &t.line_batch[1], &t.line_batch[2]);
  std::fclose(f);
  if (ok != 8 || ver != 1) return LBLX_E_ARG;
  if (tuning_device_key(c) != key) return LBLX_E_UNSUPPORTED;  // tuned on another device
  return lblx_set_tuning(c, &t);
}

int lblx_set_tiles(lblx_ctx *c, const int *tiles) {
  if (!c || !tiles) return LBLX_E_ARG;
  const int lo[3] = {256, 128, 64}, hi[3] = {8192, 8192, 8192};
  for (int G = 0; G < 3; G++)
    if (tiles[G] < lo[G] || tiles[G] > hi[G] || (tiles[G] & 63)) return LBLX_E_ARG;
  for (int G = 0; G < 3; G++) c->tiles[G] = tiles[G];
  return LBLX_OK;
}

int lblx_timing(lblx_ctx *c, int enable) {
  if (!c) return LBLX_E_ARG;
  c->tcollect();
  c->timing = enable != 0;
  for (int i = 0; i < ST_N; i++) {
    c->stage_ms[i] = 0.;
    c->stage_n[i] = 0;
  }
  return LBLX_OK;
}

int lblx_timing_get(lblx_ctx *c, int stage, const char **name, double *ms, int64_t *count) {
  if (!c || stage < 0 || stage >= ST_N) return LBLX_E_ARG;
  c->tcollect();
  if (name) *name = kStageNames[stage];
  if (ms) *ms = c->stage_ms[stage];
  if (count) *count = c->stage_n[stage];
  return LBLX_OK;
}

int lblx_sync(lblx_ctx *c) {
  if (!c) return LBLX_E_ARG;
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBLX_OK;
}

void *lblx_stream(lblx_ctx *c) { return c ? (void *)c->stream : nullptr; }

void *lblx_dev_alloc(lblx_ctx *c, int64_t bytes) {
  if (!c || bytes <= 0) return nullptr;
  if (hipSetDevice(c->device) != hipSuccess) return nullptr;
  void *p = nullptr;
  if (hipMalloc(&p, (size_t)bytes) != hipSuccess) return nullptr;
  return p;
}

int lblx_dev_free(lblx_ctx *c, void *p) {
  if (!c) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  if (p) HIPCHK(hipFree(p));
  return LBLX_OK;
}

int lblx_copy(lblx_ctx *c, void *dst, const void *src, int64_t bytes, int kind) {
  if (!c || bytes < 0) return LBLX_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  if (kind == 3) {  // device copy through a kernel (traffic calibration)
    int64_t n = bytes / 8;
    hipLaunchKernelGGL(k_copy_f64, dim3(2048), dim3(256), 0, c->stream, (double *)dst,
                       (const double *)src, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return LBLX_OK;
  }
  hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                    : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  HIPCHK(hipMemcpyAsync(dst, src, (size_t)bytes, k, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return LBLX_OK;
}

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
  int64_t nbrd = 0;  // lines with a broadener flag
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
    // BRD_MOL_FLG_IN(7,250) then BRD_MOL_DAT(21,250): per molecule j the
    // width, its temperature exponent and the shift (RDLIN,
    // oprop.f90:801-810).  RDLIN and RDLNFL convert an O2 line's shift to
    // the foreign shift whenever its own O2 flag is set, whatever IBRD
    // (oprop.f90:818-827, 3473-3477); the other broadener data enter only
    // with IBRD > 0 (LNCOR1 :1069-1096, LINF4 :3164-3221).
    const char *fl = b + o;
    const char *dat = fl + (size_t)R * 7 * 4;
    const size_t base = pshift.size() - R;
    for (int i = 0; i < n; i++) {
      int32_t f[7];
      std::memcpy(f, fl + (size_t)i * 28, 28);
      bool any = false;
      for (int j = 0; j < 7; j++) any = any || f[j] != 0;
      if (any) nbrd++;
      const int m = mol[base + i] % 100;
      if (m == 7 && iflg[base + i] >= 0 && f[6] > 0) {
        float shft7;
        std::memcpy(&shft7, dat + ((size_t)i * 21 + 6 * 3 + 2) * 4, 4);
        const double rvmr = 0.21;
        pshift[base + i] = (float)(((double)pshift[base + i] - rvmr * (double)shft7) / (1.0 - rvmr));
      }
    }
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
  int rc = lblx_load_lines(c, &L);
  if (rc == LBLX_OK) c->nbrd = nbrd;
  return rc;
}
