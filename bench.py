"""Throughput of the MI355X monochromatic path on the 50-layer column.

Workload (one step = one spectrum): the 50-layer atmosphere, 500,000
seeded synthetic lines over 500-2500 cm-1 in 7 molecules, continuum on
(ICNTNM=1), every layer's optical depth (OPDPTH: CONTNM, LINF4, SHRINK,
LBLF4 far wing, HIRAC1 near wing and panel combine) on the layer DV grids
of PATH, then XMERGE up the column to the upwelling radiance and
transmittance at the top (surface 295 K, emissivity 0.98), FP64.  Inputs
(line table, layer parameters) are resident on the device before the
timed region; per-layer optical depths stay in HBM.

Prints one JSON line (the bench contract) with the dominant kernel's
roofline, measured with the library's HIP-event stage timers over the
timed steps, and the CPU oracle timed on this host on the same spectrum.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spectra/s FP32-mixed vs plain FP64 at ≤1e-4 rel; Jacobian/forward time"
HBM_PEAK_GBS = 8000.0
REC1_BYTES = 48      # R1 scatter record (SRec1)
REC_BYTES = 32       # R2 / R3 scatter record (SRecA)
T9_BYTES = 36        # shrunk line: centre, strength, two widths, molecule
LINE_BYTES = 36      # TAPE3 line read per layer (centre f64, six f32, mol)


def stage_bytes(stage, stats, sizes, nwin):
    """Algorithmic HBM bytes of one step for a stage (sum over layers)."""
    npts = sum(sizes)
    if stage == "merge":
        # read the layer OD, read R and T of the layer below, write R and T
        return 8 * npts + 16 * sum(sizes[:-1]) + 16 * npts
    if stage == "contnm":
        return 8 * 2007 * len(sizes)
    if stage == "tile_r1":
        # write OD, read R2 (npts/4) and the scatter record of every kept line
        return 8 * npts + 8 * npts / 4 + REC1_BYTES * sum(s["lincnt"] for s in stats)
    if stage == "tile_r2":
        return 8 * npts / 4 + 8 * npts / 16 + REC_BYTES * sum(s["lincnt"] for s in stats)
    if stage == "tile_r3":
        return 8 * npts / 16 + REC_BYTES * sum(s["lincnt"] for s in stats) + 2504 * 8 * nwin
    if stage == "convf4":
        return T9_BYTES * sum(s["nlins4"] for s in stats) + 2504 * 8 * nwin
    if stage == "shrink":
        return LINE_BYTES * 500000 * len(sizes) + T9_BYTES * sum(s["nlins4"] for s in stats)
    return LINE_BYTES * 500000 * len(sizes)


FP64_PEAK_TFLOPS = 78.6
# FP64 operations of one far-wing point update in the reference's CONVF4
# loop (oprop.f90:3989-4006): XJJ step, XM, XMSQ, ZVSQ, ALFLI2 + XMSQ and
# the division, R4 + F4FN - SUBTR
CF4_FLOPS_PER_UPDATE = 8


def compute_view(stage, stats, per):
    """The far-wing stage against the FP64 peak: it does no HBM-bound work
    (its inputs are the shrunk lines; its output one R4 row per window)."""
    if stage != "convf4":
        return None
    upd = sum(s["k3_updates"] for s in stats)
    tf = CF4_FLOPS_PER_UPDATE * upd / (per[stage] * 1e-3) / 1e12
    return {"bound": "fp64", "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": tf / FP64_PEAK_TFLOPS, "updates": upd, "flops_per_update": CF4_FLOPS_PER_UPDATE}


# stage -> kernel of the FP64 column (names as rocprofv3 lists them)
STAGE_KERNEL = {
    "convf4": "k_convf4<double, false, false>",
    "tile_r1": "k_tile<1, double, double, false, double, false>",
    "tile_r2": "k_tile<2, double, double, false, double, false>",
    "tile_r3": "k_tile<3, double, double, false, double, false>",
    "merge": "k_merge_run<double>",
    "scatter_rec": "k_scatter_rec",
    "lncor1": "k_lncor1",
}
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic_column50.md")


def traffic_of(stage):
    """HBM bytes per dispatch of the stage's kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (calibrated), or None."""
    k = STAGE_KERNEL.get(stage)
    if k is None or not os.path.exists(TRAFFIC_TABLE):
        return None
    for line in open(TRAFFIC_TABLE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0].strip("`") == k:
            try:
                return float(cells[4])
            except ValueError:
                return None
    return None


def bt(rad, v):
    """Brightness temperature of radiance rad (W/(cm2 sr cm-1)) at v."""
    return 1.4387752 * v / np.log1p(1.191042722e-12 * v ** 3 / rad)


def _oracle_spectrum(job):
    """One process of the CPU baseline: the oracle (single-threaded C
    restatement of OPDPTH and XMERGE) on the whole column spectrum.  The
    inputs are rebuilt here from the seeded generator and the layer grids
    the parent planned; the timed part is the spectrum alone.  Process 0
    also returns the radiance and the optical depths of the picked layers
    for the accuracy figures."""
    idx, dvs, picks = job
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from lblrtm_amd import abi, synth
    tab = synth.config1_lines()
    atm = synth.atmosphere()
    run = abi.make_run(500.0, 2500.0, icntnm=1)
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    t0 = time.time()
    jrad = 1
    m = None
    ods = {}
    for l, dv in enumerate(dvs):
        lay = abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv,
                             500.0, 2500.0)
        lay.jrad = jrad
        od, _, _, jrad = O.od_layer(run, lay, tab)
        pan = O.panels()
        if idx == 0 and l in picks:
            ods[l] = od
        if m is None:
            m = O.merge_first(od, pan, dv, lay.tave, tzu[l], 295.0, 0.98)
        else:
            m, _ = O.merge_next(m, od, pan, dv, lay.tave, tzu[l], run.v2)
    t1 = time.time()
    return t0, t1, (m.rad if idx == 0 else None), ods


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(dvs, picks):
    """The oracle on the column spectrum in one process per host core (at
    most 16, the CPU share of one GPU on the box), each computing the whole
    spectrum; aggregate spectra/s = processes / wall time from the first
    start to the last end.  Returns the baseline record and the first
    process's radiance and picked-layer optical depths."""
    import multiprocessing as mp
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    nproc = max(1, min(16, ncpu))
    with mp.get_context("spawn").Pool(nproc) as pool:
        res = pool.map(_oracle_spectrum, [(i, list(dvs), list(picks)) for i in range(nproc)])
    start = min(r[0] for r in res)
    end = max(r[1] for r in res)
    single = float(np.median([r[1] - r[0] for r in res]))
    rec = {"value": nproc / (end - start), "unit": "spectra/s", "cores": nproc, "kind": "port",
           "cpu_model": cpu_model(), "single_core_s_per_spectrum": single,
           "sample": "the whole column spectrum (50 layers, 500k lines, continuum, merge) with the C oracle "
                     "in each of %d processes, one per host core, %.1f s wall" % (nproc, end - start)}
    return rec, res[0][2], res[0][3]


def accuracy(ref_rad, ref_ods, rads, ods, v):
    """Errors of the GPU spectra and layer optical depths against the
    oracle's: max and RMS brightness temperature, and the OD tolerance's
    two measures (relative where OD > 1e-6, absolute below)."""
    out = {}
    for name in rads:
        d = np.abs(bt(rads[name], v) - bt(ref_rad, v))
        rel, absd = 0.0, 0.0
        for l, ref in ref_ods.items():
            got = ods[name][l]
            big = np.abs(ref) > 1e-6
            if big.any():
                rel = max(rel, float(np.max(np.abs(got[big] - ref[big]) / np.abs(ref[big]))))
            if (~big).any():
                absd = max(absd, float(np.max(np.abs(got[~big] - ref[~big]))))
        out[name] = {"max_dbt_K": float(d.max()), "rms_dbt_K": float(np.sqrt(np.mean(d ** 2))),
                     "max_rel_dod": rel, "max_abs_dod_below_1e-6": absd}
    return out


def batch_block(ctx, lib, abi, synth, nprof, steps, nadj):
    profs = synth.batch_profiles(nprof)
    run = abi.make_run(500.0, 2500.0, icntnm=1, precision=abi.FP32_SPLIT)
    items, sizes, npts = [], [], 0
    for atm in profs:
        r = abi.make_run(atm["v1"], atm["v2"], icntnm=1, precision=abi.FP32_SPLIT)
        dv, _ = lib.plan_dv(r, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
        lays = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l],
                               dv[l], atm["v1"], atm["v2"]) for l in range(len(dv))]
        npts += sum(lib.od_size(run, L) for L in lays)
        sizes.append(lib.od_size(run, lays[-1]))
        items.append((lays, atm["tlev"][1:], atm["tlev"][0], 0.98))
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    rad = ctx.alloc(8 * int(sum(sizes)))
    tra = ctx.alloc(8 * int(sum(sizes)))
    t0 = time.perf_counter()
    tiles = ctx.autotune(run, items[0][0][::4])
    t_tune = time.perf_counter() - t0
    ctx.spectra_batch_device(run, items, rad.ptr, tra.ptr, offs)  # warm-up: buffers, graphs
    ctx.sync()
    print("batch: %d profiles, %d layers, warm-up done" % (nprof, sum(len(i[0]) for i in items)),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.spectra_batch_device(run, items, rad.ptr, tra.ptr, offs)
    ctx.sync()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    rad.free()
    tra.free()
    adj = bwd = None
    P = 4
    if nadj > 0:  # forward and Jacobians (tangent-linear OD) of the first nadj profiles
        sub = items[:nadj]
        jsz = [P * len(it[0]) * n for it, n in zip(sub, sizes[:nadj])]
        roff = np.concatenate([[0], np.cumsum(sizes[:nadj])[:-1]]).astype(np.int64)
        joff = np.concatenate([[0], np.cumsum(jsz)[:-1]]).astype(np.int64)
        rad = ctx.alloc(8 * int(sum(sizes[:nadj])))
        jac = ctx.alloc(8 * int(sum(jsz)))
        ctx.jacobians_batch_device(run, sub, [1, 2, 3], rad.ptr, roff, jac.ptr, joff, tl=True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.jacobians_batch_device(run, sub, [1, 2, 3], rad.ptr, roff, jac.ptr, joff, tl=True)
        ctx.sync()
        ms_adj = (time.perf_counter() - t0) * 1000.0
        rad.free()
        jac.free()
        adj = {"value": nadj * 1000.0 / ms_adj, "unit": "spectra/s", "ms_per_step": ms_adj, "profiles": nadj,
               "columns_per_layer": P,
               "workload": "the first %d profiles above: TOA radiance and its Jacobians (T, ln q of H2O, "
                           "CO2, O3 at every layer; tangent-linear OD), FP32 split" % nadj}
        # reverse mode over every profile (dL/dR = 1: the band-integrated
        # radiance's gradient); memory grows with layers x points only
        roff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        gsz = [P * len(it[0]) for it in items]
        goff = np.concatenate([[0], np.cumsum(gsz)[:-1]]).astype(np.int64)
        dl = ctx.alloc(8 * int(sum(sizes)))
        dl.from_host(np.ones(int(sum(sizes))))
        grad = ctx.alloc(8 * int(sum(gsz)))
        ctx.backward_batch_device(run, items[:2], [1, 2, 3], dl.ptr, roff[:2], grad.ptr, goff[:2])
        ctx.sync()
        t0 = time.perf_counter()
        ctx.backward_batch_device(run, items, [1, 2, 3], dl.ptr, roff, grad.ptr, goff)
        ctx.sync()
        ms_bwd = (time.perf_counter() - t0) * 1000.0
        dl.free()
        grad.free()
        bwd = {"value": nprof * 1000.0 / ms_bwd, "unit": "spectra/s", "ms_per_step": ms_bwd, "profiles": nprof,
               "columns_per_layer": P,
               "workload": "all %d profiles: TOA radiance and the gradient of a scalar of it (dL/dR given) "
                           "with respect to T, ln q of H2O, CO2, O3 at every layer, reverse mode, FP32 split"
                           % nprof}
    return {"value": nprof * 1000.0 / ms, "unit": "spectra/s", "ms_per_step": ms, "steps": steps,
            "forward_and_adjoint": adj, "backward": bwd,
            "profiles": nprof, "layers": int(sum(len(i[0]) for i in items)), "points": int(npts),
            "tiles": list(tiles), "line_batch": list(ctx.tuning()[1]), "autotune_s": t_tune,
            "workload": "ragged profiles: 20-80 layers, psfc 600-1050 hPa, 200-800 cm-1 windows "
                        "in 500-2500 cm-1, FP32 split, continuum on, upwelling TOA radiance"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-jacobian", action="store_true")
    ap.add_argument("--batch-profiles", type=int, default=256,
                    help="profiles in the batched FP32-split form (0 skips it)")
    ap.add_argument("--batch-steps", type=int, default=2)
    ap.add_argument("--batch-adjoint", type=int, default=8,
                    help="profiles of the batch also run forward and adjoint (0 skips it)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run the plain copy kernel over this many MB (the reference "
                         "dispatch for scaling HBM traffic counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus %d is not supported "
                                   "(see DESIGN.md for how the path would shard)" % a.gpus}))
        sys.exit(2)
    from lblrtm_amd import abi, lib, synth

    tab = synth.config1_lines()
    atm = synth.atmosphere()
    run = abi.make_run(500.0, 2500.0, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    tbnd, emis = 295.0, 0.98
    nlay = len(dv)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]),
                             atm["wbroad"][l], dv[l], 500.0, 2500.0) for l in range(nlay)]
    sizes = [lib.od_size(run, L) for L in layers]
    ntop = sizes[-1]

    ctx = lib.Context(0)
    if a.calib_copy_mb > 0:
        nb = a.calib_copy_mb << 20
        src, dst = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy_kernel(dst, src, nb)
        ctx.sync()
        src.free()
        dst.free()
    ctx.load_lines(tab)
    rad = ctx.alloc(8 * ntop)
    tra = ctx.alloc(8 * ntop)

    def step(r, stats=False):
        lays = [abi.Layer.from_buffer_copy(x) for x in layers]
        return ctx.spectrum_device(r, lays, tzu, tbnd, emis, rad.ptr, tra.ptr, ntop, stats=stats)

    def timed(r):
        for _ in range(a.warmup):
            step(r)
        ctx.sync()
        ctx.timing(True)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            step(r)
        ctx.sync()
        t1 = time.perf_counter()
        tim = ctx.timings()
        ctx.timing(False)
        return (t1 - t0) * 1000.0 / a.steps, {k: v[0] / a.steps for k, v in tim.items() if v[1] > 0}

    v1top, stats = step(run, stats=True)
    ms, per = timed(run)
    r64 = rad.to_host(np.float64)
    # FP32 split form of the same spectrum, timed the same way
    run32 = abi.make_run(500.0, 2500.0, icntnm=1, precision=abi.FP32_SPLIT)
    ms32, per32 = timed(run32)
    r32 = rad.to_host(np.float64)
    # instrument scan of the precision form: Hamming, 1 cm OPD, onto 0.1 cm-1
    scan = abi.make_scan(1.0, 510.0, 2490.0, jemit=1, jfnin=-6, dvout=0.1)
    nsc, _, _ = lib.scan_size(scan, v1top, layers[-1].dv, ntop)
    sbuf = ctx.alloc(8 * nsc)
    ctx.fftscn_device(scan, rad.ptr, v1top, layers[-1].dv, ntop, sbuf.ptr, nsc)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.fftscn_device(scan, rad.ptr, v1top, layers[-1].dv, ntop, sbuf.ptr, nsc)
    ctx.sync()
    ms_scan = (time.perf_counter() - t0) * 1000.0 / a.steps
    s32 = sbuf.to_host(np.float64)
    rad.from_host(r64)
    n2, v1s, dvs = ctx.fftscn_device(scan, rad.ptr, v1top, layers[-1].dv, ntop, sbuf.ptr, nsc)
    s64 = sbuf.to_host(np.float64)
    vs = v1s + dvs * np.arange(nsc)
    dbt_scan = np.abs(bt(s32, vs) - bt(s64, vs))
    v = v1top + np.arange(ntop) * layers[-1].dv
    dbt = np.abs(bt(r32, v) - bt(r64, v))
    # spectral-domain scan (SCANFN) of the FP64 radiance: sinc, 0.5 cm-1 HWHM
    sfn = abi.make_scanfn(0.5, 520.0, 2480.0, jemit=1, jfn=4)
    nsf, _, _ = lib.scanfn_size(sfn, 1, v1top, layers[-1].dv, ntop)
    fbuf = ctx.alloc(8 * nsf)
    ctx.scanfn_device(sfn, 1, rad.ptr, v1top, layers[-1].dv, ntop, fbuf.ptr, nsf)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        ctx.scanfn_device(sfn, 1, rad.ptr, v1top, layers[-1].dv, ntop, fbuf.ptr, nsf)
    ctx.sync()
    ms_scanfn = (time.perf_counter() - t0) * 1000.0 / a.steps
    fbuf.free()
    # Jacobians of the FP64 spectrum: T and H2O/CO2/O3 at every layer
    jac_ms = jac_aj_ms = jac_tl_ms = bwd_ms = None
    jac32 = None
    if not a.no_jacobian:
        nj = 4 * nlay * ntop
        jbuf = ctx.alloc(8 * nj)
        gbuf = ctx.alloc(8 * 4 * nlay)
        dlbuf = ctx.alloc(8 * ntop)
        dlbuf.from_host(np.random.default_rng(5).normal(size=ntop))
        nrep = max(2, a.steps // 5)

        def time_jac(run_, mode):
            def once():
                L = [abi.Layer.from_buffer_copy(x) for x in layers]
                if mode == "aj":  # the reference's analytic-Jacobian formulation (single-species runs, +1 K)
                    ctx.jacobian_aj_device(run_, L, tzu, tbnd, emis, [1, 2, 3], rad.ptr, jbuf.ptr, ntop)
                elif mode == "tl":  # tangent-linear optical depths
                    ctx.jacobian_tl_device(run_, L, tzu, tbnd, emis, [1, 2, 3], rad.ptr, jbuf.ptr, ntop)
                elif mode == "bwd":  # reverse mode: one gradient
                    ctx.backward_batch_device(run_, [(L, tzu, tbnd, emis)], [1, 2, 3], dlbuf.ptr, [0], gbuf.ptr,
                                              [0], d_rad=rad.ptr)
                else:
                    ctx.jacobian_device(run_, L, tzu, tbnd, emis, [1, 2, 3], 1e-3, 1e-4, rad.ptr, jbuf.ptr, ntop)
            once()
            ctx.sync()
            t0 = time.perf_counter()
            for _ in range(nrep):
                once()
            ctx.sync()
            return (time.perf_counter() - t0) * 1000.0 / nrep

        jac_ms = time_jac(run, "fd")
        jac_aj_ms = time_jac(run, "aj")
        jac_tl_ms = time_jac(run, "tl")
        bwd_ms = time_jac(run, "bwd")
        # the adjoint's own precision form: FP32 line shapes and tangent
        # planes, FP64 positions, against the FP32-split forward
        j32, j32aj, j32tl, b32 = (time_jac(run32, m) for m in ("fd", "aj", "tl", "bwd"))
        jac32 = {"ms": j32, "ratio_vs_forward": j32 / ms32, "aj_ms": j32aj, "aj_ratio_vs_forward": j32aj / ms32,
                 "tl_ms": j32tl, "tl_ratio_vs_forward": j32tl / ms32, "backward_ms": b32,
                 "backward_ratio_vs_forward": b32 / ms32}
        jbuf.free()
        gbuf.free()
        dlbuf.free()
    # batched FP32-split forward: ragged profiles with their own surface
    # pressures and band windows over the same line table
    batch = None
    if a.batch_profiles > 0:
        batch = batch_block(ctx, lib, abi, synth, a.batch_profiles, a.batch_steps,
                            min(a.batch_adjoint, a.batch_profiles))
    # layer optical depths of a few layers across the column, both forms,
    # for the accuracy against the oracle
    picks = (0, 12, 25, 37, 49)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    gods = {}
    for name, r_ in (("fp64", run), ("fp32_split", run32)):
        esz = 4 if name == "fp32_split" else 8
        obuf = ctx.alloc(esz * int(sum(sizes)))
        ctx.opdpth_column_device(r_, [abi.Layer.from_buffer_copy(x) for x in layers], obuf.ptr, offs)
        ctx.sync()
        gods[name] = {l: obuf.to_host(np.float32 if esz == 4 else np.float64, sizes[l], int(offs[l])).astype(np.float64)
                      for l in picks}
        obuf.free()
    dom = max((k for k in per if k != "host"), key=per.get)
    nwin = sum(s["nwin4"] for s in stats)
    nbytes = stage_bytes(dom, stats, sizes, nwin)
    achieved = nbytes / (per[dom] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": 1000.0 / ms,
        "unit": "spectra/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "50-layer column upwelling TOA spectrum: 500k lines 500-2500 cm-1, "
                               "7 molecules, LBLF4 panels, continuum on, XMERGE radiance, FP64",
                   "layers": nlay, "lines": 500000, "points": int(sum(sizes)), "top_points": ntop,
                   "counts": {k: int(sum(s[k] for s in stats))
                              for k in ("nlin", "lincnt", "nlins4", "ilin4", "nwin4", "k1_updates", "k3_updates")},
                   "stage_ms": {k: round(v, 4) for k, v in per.items()}},
        "fp32_split": {"value": 1000.0 / ms32, "ms_per_step": ms32, "ratio_vs_fp64": ms / ms32,
                       "max_dbt_K": float(dbt.max()), "rms_dbt_K": float(np.sqrt(np.mean(dbt ** 2))),
                       "scan": {"ms": ms_scan, "points": nsc, "max_dbt_K": float(dbt_scan.max()),
                                "rms_dbt_K": float(np.sqrt(np.mean(dbt_scan ** 2))),
                                "function": "Hamming, 1 cm OPD, 0.1 cm-1 output"},
                       "scanfn": {"ms": ms_scanfn, "points": nsf,
                                  "function": "SCANFN sinc, 0.5 cm-1 HWHM, 0.125 cm-1 output, FP64 radiance"},
                       "stage_ms": {k: round(v, 4) for k, v in per32.items()}},
        "jacobian": None if jac_ms is None else {
            "ms": jac_ms, "ratio_vs_forward": jac_ms / ms, "columns": 4 * nlay,
            "parameters": "T, ln q of H2O, CO2, O3 at every layer; OD forward differences, exact merge tangents",
            "aj": {"ms": jac_aj_ms, "ratio_vs_forward": jac_aj_ms / ms,
                   "formulation": "reference analytic Jacobian: single-species OPDPTH runs, +1 K"},
            "tl": {"ms": jac_tl_ms, "ratio_vs_forward": jac_tl_ms / ms,
                   "formulation": "tangent-linear optical depths beside one OPDPTH run of the column, "
                                  "exact merge tangents"},
            "backward": {"ms": bwd_ms, "ratio_vs_forward": bwd_ms / ms,
                         "formulation": "reverse mode: gradient of a scalar of the TOA spectrum (dL/dR given) "
                                        "for the same 200 parameters"},
            "fp32_split": jac32},
        "batch": None if batch is None else dict(batch, spectra_ratio_vs_fp64=batch["value"] * ms / 1000.0),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_of(dom),
                     "compute": compute_view(dom, stats, per)},
        "cpu_baseline": None,
    }
    if not a.no_cpu:
        rec, ref_rad, ref_ods = cpu_baseline(dv, picks)
        out["cpu_baseline"] = rec
        out["accuracy"] = dict(accuracy(ref_rad, ref_ods, {"fp64": r64, "fp32_split": r32}, gods, v),
                               against="the oracle's FP64 spectrum from the cpu_baseline leg; optical depths "
                                       "of layers %s" % (list(picks),))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
