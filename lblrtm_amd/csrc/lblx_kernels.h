// Device-side data layouts shared by the host orchestration and the
// kernels of the layer optical-depth path.
#pragma once
#include <cstdint>

namespace lblx {

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
};

// Layer constants the kernels need (a flattened subset of LayerPlan).
struct DevLayer {
  double dv, dvr2, dvr3, dvr4, bound4, alfmax;
  double vft0, vbot, vtop, v1l4, v2l4;
  double rhorat, tratio, xkt, betacr_l, beta4, betacr4;
  double dptmn_hirac, dptfc, dptmn4;
  double scor[7][10], alfd1[7][10], rhoslf[8], wk[8];
  int32_t nmol, jrad, max1, npanel, nwin;
  int64_t g_last, g2max, g3max, n1, n2, n3;
  // continuum grid (ICNTNM): absrb[0..nptabs) on v1abs + k*dvabs
  double v1abs, v2abs, dvabs;
  int32_t nptabs, icntnm;
};

struct DevPanel {   // mirror of PanelPlan
  double vft;
  int32_t w, nhi, istop, seg0, nseg, pad;
  int64_t out0;
};

struct DevWin {     // mirror of WinPlan
  double v1r4, v2r4;
  int32_t nptr4, seg0, nseg, p_first;
};

struct DevSeg {
  int32_t i1, i2;
  double radvi, rdel;
};

// Per-line output of the LINF4 correction (far-wing preparation).
struct F4Line {
  double x, sp, al, ad;  // shifted centre, strength, Lorentz and Doppler widths
  int32_t mol;           // molecule (2 = CO2 handled separately)
  int32_t pass;          // enters SHRINK
};

// Shrunk far-wing line list (TAPE9) and its block table.
struct Tape9 {
  double *vnu, *sp, *alfa, *epp;
  int32_t *mol;
  int32_t *n;            // [1] lines written
  int32_t *bstart, *bcount, *nblk;  // TAPE9 block table
  double *bvhi;
};

// Near-line record after LNCOR1 (one per TAPE3 record in the near range).
struct NearLine {
  double vnu;      // shifted centre
  double recalf;   // 1/ALFV
  double sp;       // corrected strength (0 = skipped)
  double speak;    // SUI * RECALF, rejection peak
  double zetai;
  int32_t izeta;
  int32_t ff;      // first panel whose buffer holds the line (firstfit)
  int32_t keep;    // survives every rejection test
  int32_t proc;    // panel in which CNVFNV processes it
};

// Scatter record of one kept line: indices and weights for R1/R2/R3.
struct ScatterRec {
  double zfl1, zfl2, zfl3, zslope;
  double strf1, strf2, strf3, strd, strver;
  int64_t g1lo;    // global fine index of JMIN1
  int64_t s2, s3;  // global shifts: G2 = G1 - s2, G3 = G1 - s3
  int32_t npts;    // JMAX1 - JMIN1 + 1 (0 = no contribution)
  int32_t pad;
};

}  // namespace lblx
