// Device-side data layouts shared by the host orchestration and the
// kernels of the layer optical-depth path.  A "job" is one layer of one
// profile; every stage launches once for all jobs of a batch.
#pragma once
#include <cstdint>

namespace lblx {

constexpr int EBLK = 64;   // batch lengths (in TAPE3 blocks) tabled per block (a batch holds < LIMIN4 + one block of passing lines)
constexpr int R4ROW = 2504;

// Resident line table (TAPE3 records in file order, struct-of-arrays).
struct DevLines {
  int32_t n;            // records
  int32_t nblk;         // TAPE3 blocks
  const int32_t *bstart;  // [nblk+1] first record of each block
  const double *bvmax;    // [nblk] block-header VMAX
  const double *vnu;
  const float *sp, *alfa, *epp, *hwhm, *tmpalf, *pshift;
  const int32_t *mol;
  const int32_t *blk;     // block of each record
  // Line coupling (null when no line has IFLG != 0).  cidx per record:
  // -1 IFLG = 0; -2 another IFLG > 0 (never rejected, no coefficients);
  // 2k + (IFLG == 3) for IFLG 1 or 3 with coefficient set k.  cco[8k..8k+7]
  // = A(1..4), B(1..4) at TEMPLC = 200, 250, 296, 340 K: Y or GAMMA1 in A,
  // G or GAMMA2 in B (oprop.f90:1039-1048).
  const int32_t *cidx;
  const double *cco;
};

// Per-job constants and offsets (a flattened LayerPlan).
struct DevLayer {
  double dv, dvr2, dvr3, dvr4, bound4, alfmax;
  double vft0, vbot, vtop, v1l4, v2l4;
  double rhorat, tratio, xkt, betacr_l, beta4, betacr4;
  double dptmn_hirac, dptfc, dptmn4;
  double scor[7][10], alfd1[7][10], rhoslf[8], wk[8];
  int32_t nmol, jrad, max1, npanel, nwin;
  int32_t voigt;          // exact-Voigt build (oprop_voigt.f90) tables and CONVF4
  // line-coupling coefficients at TAVE: value = C(an) + (C(hi) - C(lo)) rect
  // dif, 0-based indices into the four temperatures; [0] LNCOR1's interval,
  // [1] LINF4's (oprop.f90:981-990, 3077-3082)
  int32_t lc_lo[2], lc_an[2];
  double lc_rect[2], lc_dif[2];
  double pavp0, pavp2;    // P/P0 and its square
  int64_t g_last, g2max, g3max, n1, n2, n3;
  // continuum grid (ICNTNM): absrb[1..nptabs] on v1abs + (k-1)*dvabs
  double v1abs, v2abs, dvabs;
  int32_t nptabs, icntnm;
  int64_t abs0;           // offset of this job's ABSRB row (1-based entries)
  // offsets into concatenated plan tables
  int32_t pan0, win0, seg0;
  // far-wing line range: first TAPE3 block with VMAX >= V1L4, and the
  // number of blocks from there that can hold the stop line (the first
  // passing line above V2L4; LINF4 reads no block after it)
  int32_t b4, nb4;
  // near-line range (RDLIN): records [nlo, nlo+nn), blocks [blk0, blk0+nbk)
  int32_t nlo, nn, blk0, nbk;
  // outputs and grids
  int64_t od0, r2o, r3o;
};

struct DevPanel {   // mirror of PanelPlan
  double vft;
  int32_t w, nhi, istop, seg0, nseg, pad;
  int64_t out0;
};

struct DevWin {     // mirror of WinPlan, plus its job
  double v1r4, v2r4;
  int32_t nptr4, seg0, nseg, job;
};

struct DevSeg {
  int32_t i1, i2;
  double radvi, rdel;
};

// Strides of the per-job slices of the work arrays.
struct Strides {
  int64_t line;   // >= records + 64
  int64_t blk;    // >= blocks + 2
  int64_t t9blk;  // TAPE9 blocks per job
};

// SHRINK work arrays (compacted passing lines of LINF4 and the chain tables).
struct ShrinkWork {
  double *cx, *csp, *cal, *cad;  // centre, strength (0 below 1e-35), widths
  double *cspp;           // coupling strength SPP (0 below 1e-35 or uncoupled)
  // bit 0: CO2, bit 1: positive strength, bit 2: coupled (SPP != 0, emitted
  // alone), bits 3-5: molecule
  int8_t *cfl;
  int32_t *g;             // first-group end from each start
  int32_t *jmp;           // [line] f: next group end after a group end
  int16_t *wsum;          // [line] lines the group after a group end emits
  int32_t *grp, *gout;    // [output slot] first and last line of the group emitted there (grp -1: a group's second slot)
  int32_t *bcnt, *bstop;  // per TAPE3 block: passing lines, stop line
  int32_t *bp;            // block prefix of passing lines
  int32_t *npass, *stop;                  // per job
  int32_t *etab;          // [blk][EBLK] emitted lines of a batch of k+1 blocks
  int32_t *estop;         // [blk] emitted lines of the batch from the block to the stop line (-1: not tabled)
  int32_t *bat_s, *bat_t, *bat_off, *nbat;  // batches
  // [blk] a TAPE9 block started at the block: where the next one starts
  // (-1: none, -2: not tabled), the lines it emits, its batches; and for
  // the starts on the job's chain: output offset, batch offset, index (-1: not a start)
  int32_t *t9nxt, *t9nem, *t9nbt, *t9out, *t9bat, *t9k;
  int32_t *craw;          // [line] TAPE3 record of each compacted line (null: not kept)
};

// Shrunk far-wing line list (TAPE9) and its block table, per job.
struct Tape9 {
  double *vnu, *sp, *alfa, *epp, *spp;
  int32_t *mol;
  int32_t *n;            // [job] lines written
  int32_t *bstart, *bcount, *nblk;  // TAPE9 block table
  double *bvhi;
};

// Near-line record after LNCOR1 (one per TAPE3 record in the near range)
// as two arrays: the fields every rejection pass reads, and the ones only
// the scatter records of the surviving lines need.
struct NearHot {
  double vnu;     // shifted centre
  double recalf;  // 1/ALFV
  double speak;   // SUI * RECALF, rejection peak
  int32_t ff;     // first panel whose buffer holds the line (firstfit)
  int32_t keep;   // LNCOR1 keeps the record (not skipped)
};
struct NearCold {
  double sp, zetai;  // corrected strength, zeta
  double sppsp;      // line coupling SPP/SP (0 uncoupled)
  int32_t izeta, pad;
};

// Scatter records of the near lines, one array per grid so that each tile
// kernel reads only its own fields: the R1 record (with the FG and XVER
// weights), the R2 and R3 records, and the exact-Voigt build's zeta
// column and fraction (all grids).  glo is the line's first point on the
// grid (G2 = G1 - j2shift - 1800 p, G3 = G1 - j3shift - 2250 p); npts = 0
// means no contribution.  Only the lines that survive the rejection tests
// have records, stored contiguously in file order.
struct SRec1 {
  double zfl, zslope, str, strd, strver;
  int32_t glo, npts;
};
struct SRecA {
  double zfl, zslope, str;
  int32_t glo, npts;
};
struct SRecV {
  double zdif;  // ZETDIF
  int32_t izm, pad;
};
struct Recs {
  SRec1 *r1;
  SRecA *r2, *r3;
  SRecV *v;  // written and read only in the exact-Voigt build
  double *cpl;  // SPPSP of the kept lines whose npts carries NPTS_CPL
};
// npts flag of a coupled line's scatter records: the tile adds
// CLC * SPPSP * ZF times the line's value (CNVFNV, oprop.f90:1322-1351)
constexpr int32_t NPTS_CPL = 1 << 30;

}  // namespace lblx
