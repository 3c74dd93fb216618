/*
 * lblx.h -- C ABI of the MI355X monochromatic path for LBLRTM.
 *
 * The reference has no plugin API: the hot path sits behind three Fortran
 * subroutines with COMMON-block state, called from the driver XLAYER:
 *   OPDPTH(MPTS)                  src/lblrtm.f90:6135  (CONTNM, LINF4, HIRAC1)
 *   XMERGE(NPTS,LFILE,MFILE,JPATHL) src/xmerge.f90:25   (EMINIT/RADMRG)
 *   FFTSCN(IFILE,JFILE)           src/fftscn_dbl.f90:25
 * The entry points below replace those bodies.  Each takes explicitly what
 * the Fortran routine reads from COMMON (/FILHDR/, /MANE/, /CNTSCL/) and
 * returns in caller-owned memory what the Fortran routine writes to its
 * unit (TAPE10 panels of layer optical depth, TAPE12 radiance and
 * transmittance).  All wavenumbers are double.  Every call returns
 * LBLX_OK (0) or a negative LBLX_E* code; the library never exits, the
 * Fortran shim maps a code to the reference's STOP text.
 *
 * Threading: one context per host thread.  Calls that fill host memory
 * return with the results in place; the *_device calls queue their work on
 * the context's HIP stream and return (lblx_sync waits for it), so the
 * next call's host planning overlaps the queued kernels.  A context owns
 * device memory, the resident line table and its HIP stream, plus a second
 * internal stream for work that overlaps (LNCOR1 beside CONVF4); that work
 * is always joined back into the context stream within the same call, so
 * waiting for the context stream waits for everything.  The library plans
 * on a process-wide pool of up to 15 host worker threads.
 */
#ifndef LBLX_H
#define LBLX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* error codes                                                         */
enum {
  LBLX_OK = 0,
  LBLX_E_ARG = -1,          /* bad argument (null pointer, size)           */
  LBLX_E_NODEV = -2,        /* no HIP device / library built without GPU   */
  LBLX_E_HIP = -3,          /* a HIP runtime call failed                   */
  LBLX_E_NREC = -4,         /* 'RDLIN; NREC GT LIMIN' (oprop.f90:763)      */
  LBLX_E_UNSUPPORTED = -5,  /* input outside the supported subset          */
  LBLX_E_SIZE = -6,         /* output buffer too small                     */
  LBLX_E_NOLINES = -7,      /* no line table loaded                        */
  LBLX_E_DVRATIO = -8,      /* layer DV ratio above 2.5 (lblrtm.f90:5479)  */
  LBLX_E_COARSER = -9,      /* 'RADMRG; ITYPE LT 0' (xmerge.f90:2741)      */
  LBLX_E_SCRATCH = -10      /* request larger than the fixed scratch of an
                               allocation-free entry point                 */
};

/* Precision of the line-shape evaluation and panel accumulation.        */
enum {
  LBLX_FP64 = 0,            /* plain drop-in, everything in double         */
  LBLX_FP32_SPLIT = 1       /* FP32 shape and accumulation, FP64 positions */
};

/* ------------------------------------------------------------------ */
/* Line table in the LNFL TAPE3 block layout (src/struct_types.f90:28-44):
 * nblk blocks of LBLX_LINREC slots; block b holds nrec[b] records in
 * slots [b*LBLX_LINREC, b*LBLX_LINREC + nrec[b]).  vmin/vmax are the
 * block-header wavenumbers RDLIN uses to skip blocks (oprop.f90:768).
 * mol = molecule + 100*isotope, iflg < 0 marks a line-coupling record.
 * Line coupling (LNCOR1/LINF4, oprop.f90:789-795, 1039-1061, 3118-3135):
 * a line with iflg 1 (first-order Y and G) or 3 (reduced widths GAMMA1,
 * GAMMA2) takes its coefficients at 200, 250, 296 and 340 K from the
 * record after it in the same block (iflg < 0): Y or GAMMA1 in vnu, alfa,
 * mol (the bits of a float) and tmpalf; G or GAMMA2 in sp, epp, hwhm and
 * pshift.  A line with iflg != 0 is never rejected by LNCOR1.  Lines with
 * iflg != 0 are refused (LBLX_E_UNSUPPORTED) by the tangent-linear entry
 * points; a coupled line without its coefficient record is refused by
 * lblx_load_lines.                                                       */
#define LBLX_LINREC 250
typedef struct lblx_lines {
  int32_t nblk;
  const int32_t *nrec;              /* [nblk]                              */
  const double *vmin, *vmax;        /* [nblk]                              */
  const double *vnu;                /* [nblk*250] line centre, cm-1        */
  const float *sp;                  /* intensity at 296 K, cm/molecule     */
  const float *alfa;                /* air-broadened HWHM at 1 atm, 296 K  */
  const float *epp;                 /* lower-state energy, cm-1            */
  const int32_t *mol;               /* mol + 100*iso                       */
  const float *hwhm;                /* self-broadened HWHM                 */
  const float *tmpalf;              /* width temperature exponent          */
  const float *pshift;              /* pressure shift, cm-1/atm            */
  const int32_t *iflg;              /* coupling flag                       */
} lblx_lines;

/* Run-level TAPE5 record 1.2/1.3 values (src/lblrtm.f90:674-676,909-951)
 * after the reference's defaulting (ALFAL0<=0 -> 0.04, SAMPLE<=0 -> 4,
 * DPTMIN<0 -> 2e-4, DPTFAC<0 -> 1e-3).                                   */
typedef struct lblx_run {
  double v1, v2;            /* band, cm-1                                  */
  double sample;            /* SAMPLE: average halfwidth / DV              */
  double alfal0;            /* ALFAL0                                      */
  double avmass;            /* AVMASS                                      */
  double dptmin, dptfac;    /* line rejection thresholds                   */
  int32_t ilblf4;           /* ILBLF4 0/1/2                                */
  int32_t icntnm;           /* ICNTNM 0..6 (lblrtm.f90:688-760)            */
  int32_t irad;             /* IRAD != 0 -> JRAD = -1 (lblrtm.f90:832-833) */
  int32_t precision;        /* LBLX_FP64 or LBLX_FP32_SPLIT                */
  double cntscl[7];         /* XSELF XFRGN XCO2C XO3CN XO2CN XN2CN XRAYL,
                               read when icntnm = 6 (record 1.2a)          */
  int32_t ifrgnx;           /* IFRGNX = 1: MT_CKD foreign closure table    */
  int32_t iaersl;           /* IAERSL (Rayleigh only for 0)                */
  int32_t voigt;            /* 0: HIRAC1 of src/oprop.f90 (approximate
                               Voigt: F1/F2/F3 + FG + XVER with VOICON
                               weights); 1: the tabulated exact-Voigt build
                               src/oprop_voigt.f90 (voigt_init tables over
                               101 zeta, CNVFNV:1207-1393, CONVF4:3807-3999) */
  int32_t ibrd;             /* IBRD of record 1.2 (lblrtm.f90:676): > 0 uses
                               the TAPE3 broadener widths and shifts
                               (LNCOR1 oprop.f90:1069-1096, LINF4
                               :3164-3221).  Not modelled: an OD call with
                               ibrd > 0 on a line table that carries
                               broadener flags returns LBLX_E_UNSUPPORTED;
                               without flags IBRD changes nothing           */
} lblx_run;

/* One layer, as /FILHDR/ holds it when OPDPTH is called.                */
#define LBLX_MXMOL 64
typedef struct lblx_layer {
  double pave, tave;        /* mb, K                                       */
  double wk[LBLX_MXMOL];    /* column amounts of molecules 1..nmol         */
  double wbroad;            /* broadening column                           */
  double dv;                /* layer fine-grid spacing (PATH, lblrtm.f90:5407-5500) */
  double v1, v2;            /* layer band (normally the run band)          */
  int32_t nmol;
  int32_t jrad;             /* radiation flag carried across layers:
                               1 initially, set to 0 by OPDPTH when
                               V1/kT < 5 (lblrtm.f90:6211)                  */
} lblx_layer;

/* Counters HIRAC1/LBLF4 print on TAPE6 (oprop.f90:547-631), also the
 * denominators of the roofline figures.                                   */
typedef struct lblx_stats {
  int64_t nlin;             /* lines read with non-zero strength (NLIN)    */
  int64_t lincnt;           /* lines after rejection (LINCNT)              */
  int64_t nlins4;           /* effective lines after SHRINK                */
  int64_t ilin4;            /* far-wing lines after rejection (ILIN4)      */
  int64_t npanel;           /* panels emitted                              */
  int64_t npts;             /* output points                               */
  int64_t k1_updates;       /* near-wing (R1,R2,R3) point updates          */
  int64_t k3_updates;       /* far-wing R4 point updates                   */
  int64_t nwin4;            /* far-wing (LBLF4) windows computed           */
  int64_t nfix;             /* passes of the block-read fixpoint that
                               decides LNCOR1's rejections (the whole
                               batch's count; 0 without near lines)        */
  int64_t fixseq;           /* 1: the fixpoint had not settled within its
                               pass budget (LBLX_FIX_MAXPASS, default 16)
                               and the decisions were taken block by block
                               in RDLIN's order instead; same results      */
} lblx_stats;

/* ------------------------------------------------------------------ */
/* Context                                                              */
typedef struct lblx_ctx lblx_ctx;

int lblx_create(lblx_ctx **ctx, int device);
void lblx_destroy(lblx_ctx *ctx);
const char *lblx_strerror(int code);
int lblx_device_count(void);

/* Parse an LNFL TAPE3 (sequential unformatted, 4-byte record markers) and
 * keep the line table resident on the device.  Replaces the per-layer
 * re-read of TAPE3 in RDLIN/RDLNFL (oprop.f90:734-776, 3371-3507).       */
int lblx_load_tape3(lblx_ctx *ctx, const char *path);
/* Same from host arrays in the TAPE3 block layout.                        */
int lblx_load_lines(lblx_ctx *ctx, const lblx_lines *lines);

/* Layer DV selection of PATH (lblrtm.f90:5407-5500) for nlay layers:
 * fills dv[nlay] and itype[nlay].  w* are per-layer column amounts
 * (wk is [nlay][nmol], row-major).  iemit is IEMIT of record 1.2.        */
int lblx_plan_dv(const lblx_run *run, int nlay, int nmol, const double *pave,
                 const double *tave, const double *wk, const double *wbroad,
                 int iemit, double *dv, int32_t *itype);

/* Number of output points OPDPTH writes for a layer (sum of panel NLIM). */
int lblx_od_size(const lblx_run *run, const lblx_layer *layer, int64_t *npts);

/* Panel headers OPDPTH writes for the layer (PANEL, src/oprop.f90:1418-
 * 1451): V1P, V2P and NLIM of each panel of <= 2400 points; the panels'
 * points are the lblx_od_layer output in order.  npanel receives the
 * count; with cap below it nothing else is written and LBLX_E_SIZE is
 * returned (a count query).  Needs no device.                            */
int lblx_od_panels(const lblx_run *run, const lblx_layer *layer, int cap, double *v1p, double *v2p,
                   int32_t *nlim, int *npanel);

/* The reference's panel files (TAPE10 layer optical depth, TAPE11/TAPE12
 * merged results) written byte for byte as its BUFOUT records would hold
 * them: sequential unformatted records with 4-byte markers; an optional
 * file header record (the /FILHDR/ words, filhdr_bytes long, written as
 * given); per panel a header record V1P, V2P (f64), DVP, NLIM (f64 and
 * i64 when word = 8, the dbl build; f32 and i32 when word = 4), then the
 * panel's values of a (the optical depth, or the radiance) and, when b is
 * not NULL, of b (the transmittance, EMOUT src/xmerge.f90:3737-3790), as
 * REALs of `word` bytes; then, with endfil, ENDFIL's record of six -99
 * (src/lblrtm.f90:1368-1377).  a and b hold sum(nlim) values in panel
 * order.  append != 0 adds to an existing file.                          */
int lblx_write_panels(const char *path, int append, int word, const void *filhdr, int64_t filhdr_bytes,
                      int npanel, const double *v1p, const double *v2p, double dvp, const int32_t *nlim,
                      const double *a, const double *b, int endfil);

/* OPDPTH for one layer (IHIRAC=1, IOD=0): continuum (ICNTNM) + LINF4 +
 * HIRAC1/LBLF4 panels.  od receives npts values on the grid
 * v1p + k*dv (v1p returned); double for LBLX_FP64, float for
 * LBLX_FP32_SPLIT.  layer->jrad is updated as OPDPTH updates FSCDID(9).  */
int lblx_od_layer(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layer,
                  void *od, int64_t n_out, double *v1p, lblx_stats *st);

/* As lblx_od_layer, with od in device memory (HBM) of the context's
 * device; the call returns when the work is queued on the context stream.
 * Used when the next stage (merge, scan) also runs on the device.        */
int lblx_od_layer_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layer,
                         void *d_od, int64_t n_out, double *v1p, lblx_stats *st);

/* OPDPTH for the nlay layers of one column in one batch: every stage is
 * one launch for all layers.  Layer k's radiation flag is the one layer
 * k-1 leaves (FSCDID(9) persists across XLAYER's layer loop); layer 0
 * starts from layers[0].jrad.  Layer k's optical depth goes to
 * d_od + od_off[k] elements (double, or float for LBLX_FP32_SPLIT).       */
int lblx_od_column_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers,
                          int nlay, void *d_od, const int64_t *od_off, double *v1p,
                          lblx_stats *st);

/* CONTNM (src/contnm.f90:20-990) for one layer: ABSRB(1..nptabs) on the
 * 1 cm-1 grid v1abs + (k-1)*dvabs of OPDPTH (src/lblrtm.f90:6214-6226),
 * with the radiation flag OPDPTH sets (layer->jrad is updated).  The same
 * ABSRB enters every lblx_od_* call with run->icntnm != 0.  Continua of
 * -10..4340 cm-1 (MT_CKD H2O, CO2, the N2 rotational band below 350 cm-1,
 * the O2 and N2 fundamentals, Rayleigh); a band reaching above 4340 cm-1
 * with the N2, O2 or O3 continuum on returns LBLX_E_UNSUPPORTED.        */
int lblx_contnm(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layer, double *absrb,
                int cap, int *nptabs, double *v1abs, double *dvabs);

/* XMERGE for one layer, IEMIT=1 upwelling (IPATHL=1), LTE, no aerosols
 * (xmerge.f90:25-167 -> EMINIT/RADMRG, EMIN, RADNN).  layer_old == NULL:
 * the layer at the surface (EMINIT), R = E + T*emis*B(tbnd).  Otherwise
 * RADMRG: R = T*L4(R_old) + E, T = T*L4(T_old), where d_rold/d_told hold
 * the merged result of the layers below on layer_old's grid.  tzu is the
 * layer's upper boundary temperature (TZU).  d_od: the layer's optical
 * depth as lblx_od_layer_device writes it (double, or float for
 * FP32_SPLIT).  Outputs on layer's grid (lblx_merge_size points).        */
int lblx_merge_size(const lblx_run *run, const lblx_layer *layer_old,
                    const lblx_layer *layer, int64_t *n);
int lblx_xmerge_device(lblx_ctx *ctx, const lblx_run *run, const lblx_layer *layer_old,
                       const lblx_layer *layer, double tzu, double tbnd, double emis,
                       const void *d_od, const double *d_rold, const double *d_told,
                       double *d_rnew, double *d_tnew);

/* The spectrum of a column: OPDPTH for all nlay layers (surface first) in
 * one batch, then XMERGE layer by layer; TOA upwelling radiance and
 * transmittance on the top layer's grid in d_rad/d_tra (n_out >= its
 * size).  tzu[nlay]: upper boundary temperatures; tbnd, emis: surface
 * temperature and emissivity (record 1.4).                                */
int lblx_spectrum_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                         const double *tzu, double tbnd, double emis, double *d_rad,
                         double *d_tra, int64_t n_out, double *v1p_top, lblx_stats *st);

/* XMERGE for either looking direction (xmerge.f90:25-167, RADMRG
 * :2587-2990, RADNN :2991-3240).  ipathl = 1 is lblx_xmerge_device.
 * ipathl = 3, downwelling radiance at an observer looking up (layers taken
 * outward from the observer, each new layer on a grid no coarser than the
 * last): EMIN with tz = the layer's boundary temperature on the observer's
 * side (TZL); layer_old == NULL is EMINIT's layer emission, R = E, T = T;
 * otherwise RADNN's IPATH_FLG = 3, R = L4(R_old) + E*L4(T_old),
 * T = T*L4(T_old), and tbnd > 0 adds EMBND (xmerge.f90:3652-3733),
 * R += T*emis*B(tbnd): XMERGE passes TBND to the path's last layer only
 * (LAYER = LH2, :140).  Other ipathl: LBLX_E_UNSUPPORTED.                 */
int lblx_xmerge_path_device(lblx_ctx *ctx, const lblx_run *run, const lblx_layer *layer_old,
                            const lblx_layer *layer, int ipathl, double tz, double tbnd, double emis,
                            const void *d_od, const double *d_rold, const double *d_told,
                            double *d_rnew, double *d_tnew);

/* A path's spectrum in either direction: OPDPTH for all nlay layers in one
 * batch, then XMERGE in path order (layers[0] next to the observer for
 * ipathl = 3, at the surface for ipathl = 1); radiance and transmittance on
 * the last layer's grid.  tz[l]: TZU (ipathl 1) or TZL (ipathl 3); tbnd,
 * emis: the surface (ipathl 1, EMINIT) or a boundary beyond the last layer
 * (ipathl 3, EMBND after the last merge of a path of two or more layers;
 * tbnd = 0: space).                                                      */
int lblx_spectrum_path_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                              int ipathl, const double *tz, double tbnd, double emis, double *d_rad,
                              double *d_tra, int64_t n_out, double *v1p_end, lblx_stats *st);

/* Jacobians of the TOA upwelling radiance (the forward/backward pair of
 * lblx_spectrum_device) with respect to every layer's average temperature
 * and the log column amounts of nspec molecules (species[] 1-based ids).
 * Each layer's optical-depth derivative is a forward difference of OPDPTH
 * (steps dT kelvin, dlnq in ln amount; the perturbed layers run in one
 * batch with the column); the radiative part is exact: dR_l/dx on the
 * layer grid, carried to the top through the merge of every layer above.
 * d_rad: TOA radiance (n_out >= top-grid points); d_jac: (1 + nspec) x
 * nlay planes of n_out, plane (q, l) at (q * nlay + l) * n_out, q = 0
 * temperature (per K), q >= 1 species q (per unit ln amount).  Comparison
 * role of the reference's analytic Jacobians (xmerge.f90:3792-6147).   */
int lblx_jacobian_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                         const double *tzu, double tbnd, double emis, const int *species,
                         int nspec, double dT, double dlnq, double *d_rad, double *d_jac,
                         int64_t n_out, double *v1p_top);

/* The same outputs in the reference's analytic-Jacobian formulation (AJ
 * path, src/lblrtm.f90:2930-2975, src/xmerge.f90:5786-6147): the layer
 * temperature derivative from a +1 K OPDPTH run, and d tau / d ln q of
 * molecule m from a single-species run (every other amount zero, the O2
 * amount added to the broadener, CONTNM's ICFLG = m: no N2 continuum and
 * the H2O continuum's ln q derivative term), whose optical depth is the
 * derivative itself.  Same layout as lblx_jacobian_device.               */
int lblx_jacobian_aj_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                            const double *tzu, double tbnd, double emis, const int *species, int nspec,
                            double *d_rad, double *d_jac, int64_t n_out, double *v1p_top);

/* The same outputs with the layer optical-depth derivatives taken
 * analytically, in one OPDPTH run of the column: its tangent-linear passes
 * differentiate LNCOR1/CNVFNV (strengths, Lorentz, Doppler and Voigt
 * widths, zeta interpolation, pressure-shifted centres; oprop.f90:
 * 995-1320), LINF4/SHRINK/CONVF4 (each TAPE9 line from its group's
 * members, oprop.f90:3093-3640, 3878-4039) and the radiation term on the
 * forward run's own decisions (kept lines, groups, accepted far-wing lines,
 * panels), and take the continuum (CONTNM) by central differences.  The
 * molecules' other-line width change through the total amount is left out
 * (below W_m / W_tot of their self-broadened share).  species[] must be
 * distinct; the exact-Voigt build is not supported (LBLX_E_UNSUPPORTED).
 * Replaces the per-parameter OPDPTH runs of the reference's analytic
 * Jacobians (src/lblrtm.f90:2930-2975) for the same layout as
 * lblx_jacobian_device.                                                  */
int lblx_jacobian_tl_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                            const double *tzu, double tbnd, double emis, const int *species, int nspec,
                            double *d_rad, double *d_jac, int64_t n_out, double *v1p_top);

/* The tangent-linear layer optical depths alone: OPDPTH for nlay layers
 * (JRAD chained up the list, as in a column) with layer l's optical depth
 * at d_od + l * n_out and its derivatives at d_dod + (l * (1 + nspec) + q)
 * * n_out, q = 0 temperature (per K), q >= 1 species q (per unit ln
 * amount); elements double, or float under LBLX_FP32_SPLIT.  The layer-
 * level counterpart of lblx_jacobian_tl_device (oprop.f90:25-668 with its
 * derivatives, as there).                                                */
int lblx_od_tangent_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                           const int *species, int nspec, void *d_od, void *d_dod, int64_t n_out);

/* Reverse mode for a batch of profiles (the layout of
 * lblx_jacobians_batch_device): given dL/dR on each profile's top grid
 * (d_dLdR + rad_off[p]), the gradient of L with respect to every layer's
 * temperature and the log amounts of the nspec molecules, profile p's at
 * d_grad + grad_off[p] as (1 + nspec) x nlay[p] values, [q * nlay[p] + l]
 * (q = 0 temperature, per K; q >= 1 species q, per unit ln amount).  The
 * layer optical depths and their tangent-linear derivatives are those of
 * lblx_jacobian_tl_device; the radiative part runs backwards: lam = dL/dR
 * from the top, each layer's parameters take sum(lam dR_l/dx), and lam
 * moves to the layer below through the transpose of the merge (T_l and
 * the carry-over interpolation of XMERGE, xmerge.f90:2788-2796).  Memory
 * and work grow with layers x points, not with layers^2 x parameters as
 * the Jacobian planes do; equal to dL/dR contracted with those planes.
 * d_rad (may be null): the TOA radiance.  No counterpart in the reference,
 * whose retrievals contract the AJ Jacobian files on the host.           */
int lblx_backward_batch(lblx_ctx *ctx, const lblx_run *run, int nprof, const int *nlay,
                        lblx_layer *layers, const double *tzu, const double *tbnd,
                        const double *emis, const int *species, int nspec,
                        const double *d_dLdR, const int64_t *rad_off, double *d_rad,
                        double *d_grad, const int64_t *grad_off, double *v1p_top);

/* A batch of independent profiles (ragged: nlay[p] layers each, surface
 * first, concatenated in layers/tzu; each profile's own band in its
 * layers' v1/v2).  Layer optical depths run in batches of whole profiles
 * sized to the device memory, and the merges of a batch run one launch per
 * run of same-grid layers over all its profiles, the launch chain captured
 * as a HIP graph and replayed on later calls of the same shape.  Profile p's TOA
 * radiance/transmittance at d_rad/d_tra + out_off[p] (its top-grid size,
 * lblx_od_size of its last layer); v1p_top[p] its first wavenumber.      */
int lblx_spectra_batch_device(lblx_ctx *ctx, const lblx_run *run, int nprof, const int *nlay,
                              lblx_layer *layers, const double *tzu, const double *tbnd,
                              const double *emis, double *d_rad, double *d_tra,
                              const int64_t *out_off, double *v1p_top);

/* Forward and adjoint for a batch of independent profiles (the layout of
 * lblx_spectra_batch_device): profile p's TOA radiance at d_rad +
 * rad_off[p] and its Jacobian planes (lblx_jacobian_device's layout with
 * its own layer count and top-grid size) at d_jac + jac_off[p].  aj = 0
 * takes OD derivatives by forward differences (dT, dlnq), aj = 1 the
 * reference's analytic-Jacobian runs (lblx_jacobian_aj_device), aj = 2 the
 * tangent-linear optical depths (lblx_jacobian_tl_device; dT and dlnq
 * unused for aj >= 1).  Each profile's perturbed layers run as one OPDPTH batch; the
 * profiles are queued one after another on the context's stream, the next
 * one planned on the host while the last one runs.  No counterpart call in
 * the reference, whose AJ runs are one TAPE5 per profile
 * (src/lblrtm.f90:2757-3110).                                            */
int lblx_jacobians_batch_device(lblx_ctx *ctx, const lblx_run *run, int nprof, const int *nlay,
                                lblx_layer *layers, const double *tzu, const double *tbnd,
                                const double *emis, const int *species, int nspec, int aj, double dT,
                                double dlnq, double *d_rad, const int64_t *rad_off, double *d_jac,
                                const int64_t *jac_off, double *v1p_top);

/* Tile widths (points) of the R1, R2, R3 near-wing scatter: chosen by
 * timing candidates on the given layers (lblx_autotune), or set directly
 * (multiples of 64 up to 8192).  They change speed, not results.         */
int lblx_autotune(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay, int *tiles_out);
int lblx_set_tiles(lblx_ctx *ctx, const int *tiles);

/* The tuning of the near-wing scatter: tile widths (as lblx_set_tiles) and
 * each grid's line batch, the lines a wave takes per chunk (1..64, 0: the
 * built-in 16 for the FP64 R1 tile, 32 otherwise; a 1024-thread tile
 * workgroup takes 16 chunks per step, 128..1024 lines for 8..64).
 * lblx_autotune times both on the given layers (widths first, then 8, 16,
 * 32, 64 lines at the chosen width).  lblx_tuning_save writes the current
 * tuning with the device's architecture, name and compute-unit count to a
 * text file; lblx_tuning_load sets it from such a file, refusing one made
 * on another device (LBLX_E_UNSUPPORTED) or malformed (LBLX_E_ARG).  No
 * counterpart in the reference; none of this changes results beyond the
 * order of floating-point additions in the LDS accumulators.              */
typedef struct lblx_tuning {
  int32_t tiles[3];         /* R1, R2, R3 tile widths                       */
  int32_t line_batch[3];    /* lines per wave chunk, R1, R2, R3             */
} lblx_tuning;
int lblx_get_tuning(lblx_ctx *ctx, lblx_tuning *t);
int lblx_set_tuning(lblx_ctx *ctx, const lblx_tuning *t);
int lblx_tuning_save(lblx_ctx *ctx, const char *path);
int lblx_tuning_load(lblx_ctx *ctx, const char *path);

/* OPDPTH with DVOUT (IOD = 1 or 4, src/lblrtm.f90:1062-1076): the layer's
 * optical depth regridded by PNLINT (src/oprop.f90:1564-1806) onto a DVOUT
 * grid starting at the layer's first point, in output panels of 2400
 * points, the last ending at or above V2.  lblx_od_dvout_size gives the
 * number of output points (needs no device).                            */
int lblx_od_dvout_size(const lblx_run *run, const lblx_layer *layer, double dvout, int64_t *n);
int lblx_od_layer_dvout_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layer, double dvout,
                               double *d_out, int64_t cap, int64_t *n_out, double *v1_out);

/* IEMIT = 0: layer optical depths summed up the column (XMERGE ->
 * ABINIT for the first layer, ABSMRG after, src/xmerge.f90:270-372,
 * 437-685, IPATHL = 1).  lblx_absmrg_device merges one layer: d_od is the
 * layer's OD (double, or float under LBLX_FP32_SPLIT) on its grid,
 * d_acc_old the accumulated OD on layer_old's grid (NULL with layer_old
 * NULL: ABINIT), d_acc_new (double) receives the sum on the layer's grid;
 * a coarser layer grid returns LBLX_E_COARSER (ABSMRG's STOP).
 * lblx_column_od_device runs OPDPTH for every layer in one batch and the
 * merge chain, writing the column OD on the top layer's grid.            */
int lblx_absmrg_device(lblx_ctx *ctx, const lblx_run *run, const lblx_layer *layer_old,
                       const lblx_layer *layer, const void *d_od, const double *d_acc_old,
                       double *d_acc_new);
int lblx_column_od_device(lblx_ctx *ctx, const lblx_run *run, lblx_layer *layers, int nlay,
                          double *d_od, int64_t n_out, double *v1p_top, lblx_stats *st);

/* FFTSCN (src/fftscn_dbl.f90:25-789): record 10.1 of TAPE5.             */
typedef struct lblx_scan {
  double hwhm;              /* HWHM field: OPD in cm when jfnin < 0        */
  double v1, v2;            /* band of the scanned output                  */
  double dvout;             /* output spacing (INTPDR); 0: boxcar grid     */
  double param;             /* record 10.1 param (0: none)                 */
  int32_t jemit;            /* 0 transmittance, 1 radiance                 */
  int32_t jfnin;            /* scanning function: 0 boxcar; +-1,2,3,4,5,6  */
  int32_t mratin;           /* boxcar ratio (0: 12, < 0: no boxcar)        */
  int32_t ivx;              /* 0 auto, 1 interferogram domain              */
  int32_t nofix;            /* 1: no boxcar deconvolution                  */
  int32_t pad;
} lblx_scan;

/* Output size of FFTSCN for an input spectrum on v1_in + k*dv_in,
 * k < n_in (the TAPE12 the scan reads).                                  */
int lblx_scan_size(const lblx_scan *scan, double v1_in, double dv_in, int64_t n_in,
                   int64_t *n_out, double *v1_out, double *dv_out);

/* FFTSCN on device buffers: load V1Z..V2Z, boxcar by M, real FFT of the
 * padded record set, apodization x boxcar deconvolution, inverse FFT, then
 * INTPDR's four-point interpolation onto dvout.  Supported: JEMIT 0/1,
 * |JFNIN| in {0..6} with the interferogram-domain functions (IVX = 1),
 * input spacing below 2 HWHM / 12.                                       */
int lblx_fftscn_device(lblx_ctx *ctx, const lblx_scan *scan, const double *d_in, double v1_in,
                       double dv_in, int64_t n_in, double *d_out, int64_t cap, int64_t *n_out,
                       double *v1_out, double *dv_out);

/* lblx_fftscn_device without allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and supported inputs as lblx_fftscn_device; the
 * scan's work buffers live in a fixed device array of 23 MiB that each
 * device holds until the process ends, so the FFT length is limited to
 * 2^19 points and the output (dvout > 0) to 2^18 points: larger requests
 * return LBLX_E_SCRATCH.  The call allocates nothing (no hipMalloc,
 * hipFree or pinned host memory).  Every call on a device shares that
 * array: calls that overlap on one device (from two contexts or threads)
 * race.  With dvout > 0 it returns after the scan has completed.         */
int lblx_fftscn_noalloc_device(lblx_ctx *ctx, const lblx_scan *scan, const double *d_in, double v1_in,
                               double dv_in, int64_t n_in, double *d_out, int64_t cap, int64_t *n_out,
                               double *v1_out, double *dv_out);

/* SCANFN (src/postsub.f90:25-560): record 8.1 of TAPE5, the scanning
 * function convolved in the spectral domain.  jfn 1 triangle, 2 Gauss,
 * 3 sinc^2, 4 sinc (-3, -4: hwhm holds the first zero crossing);
 * jemit -1 absorption (1 - T), 0 transmission, 1 emission; sampl > 0
 * points per half width, < 0 minus the output spacing, 0 the function's
 * default; npanel is the point count of the input file's panels (0: 2400),
 * which fixes where SHRKSC's running average starts.                    */
typedef struct lblx_scanfn {
  double hwhm;              /* HWHM (cm-1)                                 */
  double v1, v2;            /* band of the scanned output                  */
  double sampl;             /* SAMPL field                                 */
  int32_t jemit;            /* -1, 0, 1                                    */
  int32_t jfn;              /* +-3, +-4, 1, 2                              */
  int32_t jvar;             /* 0 (a variable slit, 1, is refused)          */
  int32_t npanel;           /* input panel length (0: 2400)                */
} lblx_scanfn;

/* Output grid of SCANFN for an input spectrum on v1_in + k*dv_in, k < n_in:
 * PANLSC's points from v1 on dv_out up to v2.  iemit is the input file's
 * IEMIT: 0 an optical depth (TAPE10-like; jemit 0 or -1 convolve
 * exp(-OD)), 1 a radiance file whose radiance (jemit 1) or transmittance
 * (jemit 0, -1) record d_in holds (JTREM, postsub.f90:325-335).
 * Refused (LBLX_E_UNSUPPORTED): jfn 0, 5, 6 (CNVRCT, CNVVRC, CNVVRL),
 * jvar 1, and input spacing above hwhm/12 (SCNINT interpolates first).   */
int lblx_scanfn_size(const lblx_scanfn *scan, int iemit, double v1_in, double dv_in, int64_t n_in,
                     int64_t *n_out, double *v1_out, double *dv_out);

/* SCANFN on device buffers: RDSCAN's transform, SHRKSC's running average
 * (postsub.f90:1276-1373), then CONVSC's convolution with the tabulated
 * function (:1516-1598) gathered per output point.  d_out gets n_out
 * values (lblx_scanfn_size).                                             */
int lblx_scanfn_device(lblx_ctx *ctx, const lblx_scanfn *scan, int iemit, const double *d_in,
                       double v1_in, double dv_in, int64_t n_in, double *d_out, int64_t n_out);

/* Wait for the context stream; the stream handle (hipStream_t) for callers
 * that order their own device work after the library's.                  */
int lblx_sync(lblx_ctx *ctx);
void *lblx_stream(lblx_ctx *ctx);

/* Stage timers: HIP events around each stage, on the stream that runs it
 * (linf4, shrink, convf4, lncor1, reject, scatter_rec, tile_r3, tile_r2,
 * tile_r1, merge, contnm, scan), the device analogue of the TAPE6 timing
 * table (oprop.f90:547-620), "host": wall milliseconds the host spends
 * planning layers and staging descriptors, and "tangent": the
 * tangent-linear optical-depth passes of lblx_jacobian_tl_device.  lblx_timing(ctx, 1) clears and enables;
 * lblx_timing_get returns the summed milliseconds and launch count of a
 * stage index (0 .. LBLX_NSTAGE-1).                                       */
#define LBLX_NSTAGE 14

/* Tables of the exact-Voigt build (voigt_init, src/oprop_voigt.f90:
 * 11739-12008) as this library computes them on the host: f = F1, F2, F3,
 * each [101 zeta columns][2601] (entry = Fortran IZ - 1); coef = a_1 a_2
 * a_3 b_1 b_2 b_3, 101 each.  Needs no device.                          */
int lblx_voigt_tables(double *f, double *coef);

/* Diagnostics: the shrunk far-wing line list (TAPE9 content, LINF4 +
 * SHRINK, oprop.f90:2934-3640) of job `job` of the last OPDPTH call.      */
int lblx_debug_tape9(lblx_ctx *ctx, int job, double *vnu, double *sp, double *alfa,
                     double *epp, int32_t *mol, int cap, int32_t *n_out);
/* The same list's line-coupling strengths SPP (0 for the lines SHRINK sums). */
int lblx_debug_tape9_spp(lblx_ctx *ctx, int job, double *spp, int cap, int32_t *n_out);

/* Device memory of the context's device for callers without their own
 * allocator; lblx_copy kind: 0 host->device, 1 device->host, 2 device->
 * device, 3 device->device through a plain copy kernel (the reference
 * kernel for HBM traffic counters); ordered on the context stream,
 * returns when complete.                                                  */
void *lblx_dev_alloc(lblx_ctx *ctx, int64_t bytes);
int lblx_dev_free(lblx_ctx *ctx, void *p);
int lblx_copy(lblx_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind);
int lblx_timing(lblx_ctx *ctx, int enable);
int lblx_timing_get(lblx_ctx *ctx, int stage, const char **name, double *ms,
                    int64_t *count);

#ifdef __cplusplus
}
#endif

#endif /* LBLX_H */
