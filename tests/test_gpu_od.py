"""GPU parity of the layer optical-depth path (OPDPTH/HIRAC1/LBLF4) with
the CPU oracle on the same seeded inputs, through the C ABI."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth, tape
from parity import assert_od_close

pytestmark = pytest.mark.gpu


def _layer0():
    p, t, wk, wb = synth.config0_layer()
    return p, t, wk, wb


def _run_both(ctx, run, lay, tab):
    ref, v1r, st_r, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
    od, v1p, st = ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
    assert v1p == v1r
    return od, ref, st, st_r


def test_single_layer_workload(gpu_ctx):
    """20,000 lines, 600-1100 cm-1, one layer (the CPU-runnable case)."""
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = _layer0()
    run = abi.make_run(600.0, 1100.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 600.0, 1100.0)
    od, ref, st, st_r = _run_both(gpu_ctx, run, lay, tab)
    print("single layer: npts", len(od), "max rel", np.max(np.abs(od - ref) / np.abs(ref)), flush=True)
    assert_od_close(od, ref, "single layer")
    for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4"):
        assert st[k] == st_r[k], k
    # the shrunk far-wing list (TAPE9) agrees line by line
    cap = 100000
    arrs = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.int32)]
    nb = C.c_int(0)
    blk = np.zeros(1000, np.int32)
    n_o = O.lib().oracle_linf4(C.byref(run), C.byref(abi.Layer.from_buffer_copy(lay)), C.byref(tab.struct()),
                               *[abi.ptr(a, C.c_double) for a in arrs[:4]], abi.ptr(arrs[4], C.c_int32),
                               abi.ptr(blk, C.c_int32), cap, 1000, C.byref(nb))
    gv, gs, ga, ge, gm = gpu_ctx.debug_tape9(0)
    assert len(gv) == n_o
    np.testing.assert_array_equal(gm, arrs[4][:n_o])
    np.testing.assert_allclose(gv, arrs[0][:n_o], rtol=1e-14)
    np.testing.assert_allclose(gs, arrs[1][:n_o], rtol=1e-13)
    np.testing.assert_allclose(ga, arrs[2][:n_o], rtol=1e-13)
    np.testing.assert_allclose(ge, arrs[3][:n_o], rtol=1e-13)


def test_column_workload_layers(gpu_ctx):
    """50-layer column, 500,000 lines over 500-2500 cm-1: every layer in one
    batch on the GPU; five layers across the column checked point by point."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(500.0, 2500.0)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l],
                             dv[l], 500.0, 2500.0) for l in range(len(dv))]
    sizes = [lib.od_size(run, L) for L in layers]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = gpu_ctx.alloc(8 * int(sum(sizes)))
    lays = [abi.Layer.from_buffer_copy(x) for x in layers]
    _, st = gpu_ctx.opdpth_column_device(run, lays, buf.ptr, offs, stats=True)
    allod = buf.to_host(np.float64)
    buf.free()
    jrad = 1
    for l in range(len(dv)):
        inp = abi.Layer.from_buffer_copy(layers[l])
        inp.jrad = jrad
        jrad = lays[l].jrad
        if l not in (0, 12, 25, 37, 49):
            continue
        ref, _, st_r, _ = O.od_layer(run, inp, tab)
        od = allod[offs[l]:offs[l] + sizes[l]]
        print("layer", l, "npts", len(od), "max rel", np.max(np.abs(od - ref) / np.abs(ref)),
              "fixpoint passes", st[l]["nfix"], flush=True)
        assert_od_close(od, ref, "layer %d" % l)
        for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4"):
            assert st[l][k] == st_r[k], (l, k)
    # the block-read fixpoint settles well inside its 16-pass budget
    assert all(x["fixseq"] == 0 for x in st)
    assert max(x["nfix"] for x in st) <= 8, [x["nfix"] for x in st]


def test_radiation_in_line_strengths(gpu_ctx):
    """JRAD = 1 (V1/kT >= 5): radiation term in the strengths and in R4."""
    tab = synth.synth_lines(20000, 1950.0, 2350.0, [1, 2, 3], 31)
    gpu_ctx.load_lines(tab)
    p, t, w, wb = 300.0, 200.0, [1e21, 5e21, 1e19], 1.0e24
    run = abi.make_run(2000.0, 2300.0)
    dv, _ = lib.plan_dv(run, [p], [t], [w], [wb])
    lay = abi.make_layer(p, t, w, wb, dv[0], 2000.0, 2300.0)
    od, ref, st, st_r = _run_both(gpu_ctx, run, lay, tab)
    assert st_r["k3_updates"] > 0
    assert_od_close(od, ref, "jrad=1")


def test_no_rejection_and_ragged_blocks(gpu_ctx):
    tab = synth.synth_lines(1234, 600.0, 700.0, [1, 2, 7], 11)
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = 800.0, 280.0, [1e22, 3e21, 2e24], 1.0e25
    for dmin, dfac in ((0.0, 0.0), (2e-4, 1e-3)):
        run = abi.make_run(610.0, 690.0, dptmin=dmin, dptfac=dfac)
        dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
        lay = abi.make_layer(p, t, wk, wb, dv[0], 610.0, 690.0)
        od, ref, st, st_r = _run_both(gpu_ctx, run, lay, tab)
        assert_od_close(od, ref, "dptmin=%g" % dmin)


def test_band_without_lines(gpu_ctx):
    tab = synth.synth_lines(500, 600.0, 650.0, [1], 5)
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = _layer0()
    run = abi.make_run(900.0, 950.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 900.0, 950.0)
    od, ref, st, st_r = _run_both(gpu_ctx, run, lay, tab)
    assert np.all(ref == 0.0) and np.all(od == 0.0)


def test_tape3_file_and_errors(gpu_ctx, tmp_path):
    tab = synth.synth_lines(3000, 700.0, 760.0, [1], 17)
    path = tmp_path / "TAPE3"
    tape.write_tape3(path, tab)
    gpu_ctx.load_tape3(path)
    p, t, wk, wb = _layer0()
    run = abi.make_run(705.0, 755.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 705.0, 755.0)
    od, ref, st, st_r = _run_both(gpu_ctx, run, lay, tab)
    assert_od_close(od, ref, "tape3 file")
    # a coupled line (IFLG 1) whose next record is a line, not its
    # coefficient record, is refused
    bad = abi.LineTable(tab.nrec, tab.vmin, tab.vmax, vnu=tab.vnu, sp=tab.sp, alfa=tab.alfa,
                        epp=tab.epp, mol=tab.mol, hwhm=tab.hwhm, tmpalf=tab.tmpalf,
                        pshift=tab.pshift, iflg=np.where(np.arange(len(tab.iflg)) == 3, 1, 0))
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.load_lines(bad)
    assert e.value.code == abi.E_UNSUPPORTED
    nrec = tab.nrec.copy()
    nrec[0] = 251
    bad2 = abi.LineTable(nrec, tab.vmin, tab.vmax, vnu=tab.vnu, sp=tab.sp, alfa=tab.alfa,
                         epp=tab.epp, mol=tab.mol, hwhm=tab.hwhm, tmpalf=tab.tmpalf,
                         pshift=tab.pshift, iflg=tab.iflg)
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.load_lines(bad2)
    assert e.value.code == abi.E_NREC


def _tape9_oracle(run, lay, tab):
    cap = 100000
    arrs = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.int32)]
    nb = C.c_int(0)
    blk = np.zeros(1000, np.int32)
    n_o = O.lib().oracle_linf4(C.byref(run), C.byref(abi.Layer.from_buffer_copy(lay)), C.byref(tab.struct()),
                               *[abi.ptr(a, C.c_double) for a in arrs[:4]], abi.ptr(arrs[4], C.c_int32),
                               abi.ptr(blk, C.c_int32), cap, 1000, C.byref(nb))
    return [a[:n_o] for a in arrs]


@pytest.mark.parametrize("case", ["eof", "stop"])
def test_batch_walk_wave_path(monkeypatch, case):
    """LINF4's batch walk with every TAPE9 block simulated by the per-job
    chain wave instead of the per-block tables (the path the tables leave
    to it, forced here by LBLX_SHRINK_WAVE_WALK at context creation): the
    TAPE9 list and the OD equal the oracle's, with the band ending at the
    end of the line file (eof) and before it (the stop line)."""
    monkeypatch.setenv("LBLX_SHRINK_WAVE_WALK", "1")
    ctx = lib.Context(0)
    try:
        if case == "eof":
            tab = synth.config0_lines()
            p, t, wk, wb = _layer0()
            v1, v2 = 600.0, 1100.0
        else:
            tab = synth.synth_lines(1234, 600.0, 700.0, [1, 2, 7], 11)
            p, t, wk, wb = 800.0, 280.0, [1e22, 3e21, 2e24], 1.0e25
            v1, v2 = 610.0, 660.0  # lines continue past V2 + BOUND4
        ctx.load_lines(tab)
        run = abi.make_run(v1, v2)
        dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
        lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
        print("batch walk, wave path:", case, flush=True)
        od, ref, st, st_r = _run_both(ctx, run, lay, tab)
        assert_od_close(od, ref, "wave walk " + case)
        assert st["nlins4"] == st_r["nlins4"]
        ov, osp, oal, oad, om = _tape9_oracle(run, lay, tab)
        gv, gs, ga, ge, gm = ctx.debug_tape9(0)
        assert len(gv) == len(ov) > 0
        np.testing.assert_array_equal(gm, om)
        np.testing.assert_allclose(gv, ov, rtol=1e-14)
        np.testing.assert_allclose(gs, osp, rtol=1e-13)
        np.testing.assert_allclose(ga, oal, rtol=1e-13)
        np.testing.assert_allclose(ge, oad, rtol=1e-13)
    finally:
        ctx.close()


def test_fixpoint_sequential_walk(monkeypatch):
    """LNCOR1's block-read decisions taken block by block (k_reject_seq),
    forced by a one-pass fixpoint budget (LBLX_FIX_MAXPASS=1 at context
    creation): the same lines are kept as by the settled fixpoint (equal
    line and update counts; optical depths equal to the order of the LDS
    accumulation, 1e-12) and the oracle's, on the single-layer workload and
    on five layers of the column workload."""
    monkeypatch.setenv("LBLX_FIX_MAXPASS", "1")
    seq = lib.Context(0)
    monkeypatch.delenv("LBLX_FIX_MAXPASS")
    fix = lib.Context(0)
    try:
        tab = synth.config0_lines()
        p, t, wk, wb = _layer0()
        run = abi.make_run(600.0, 1100.0)
        dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
        lay = abi.make_layer(p, t, wk, wb, dv[0], 600.0, 1100.0)
        for c in (seq, fix):
            c.load_lines(tab)
        print("sequential walk: single layer", flush=True)
        od_s, _, st_s = seq.opdpth(run, abi.Layer.from_buffer_copy(lay))
        od_f, _, st_f = fix.opdpth(run, abi.Layer.from_buffer_copy(lay))
        assert st_s["fixseq"] == 1 and st_f["fixseq"] == 0 and st_f["nfix"] >= 1
        assert (st_s["lincnt"], st_s["k1_updates"]) == (st_f["lincnt"], st_f["k1_updates"])
        np.testing.assert_allclose(od_s, od_f, rtol=1e-12, atol=0)
        ref, _, st_r, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
        assert_od_close(od_s, ref, "sequential walk")
        assert st_s["k1_updates"] == st_r["k1_updates"]

        tab = synth.config1_lines()
        atm = synth.atmosphere()
        run = abi.make_run(500.0, 2500.0)
        dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
        pick = [0, 12, 25, 37, 49]
        layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l],
                                 dv[l], 500.0, 2500.0) for l in pick]
        sizes = [lib.od_size(run, L) for L in layers]
        offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
        outs = []
        for c in (seq, fix):
            c.load_lines(tab)
            print("sequential walk: column layers", pick, flush=True)
            buf = c.alloc(8 * int(sum(sizes)))
            lays = [abi.Layer.from_buffer_copy(x) for x in layers]
            _, st = c.opdpth_column_device(run, lays, buf.ptr, offs, stats=True)
            outs.append((buf.to_host(np.float64), st))
            buf.free()
        assert outs[0][1][0]["fixseq"] == 1 and outs[1][1][0]["fixseq"] == 0
        print("column fixpoint passes", outs[1][1][0]["nfix"], flush=True)
        np.testing.assert_allclose(outs[0][0], outs[1][0], rtol=1e-12, atol=1e-300)
        for l in range(len(pick)):
            for k in ("lincnt", "k1_updates"):
                assert outs[0][1][l][k] == outs[1][1][l][k], (l, k)
    finally:
        seq.close()
        fix.close()


def test_tape3_broadener_words(gpu_ctx, tmp_path):
    """TAPE3 broadener words: an O2 line whose own O2 flag is set has its
    shift converted to the foreign shift on reading, whatever IBRD
    (src/oprop.f90:818-827, 3473-3477); IBRD > 0 on a table carrying
    flags is refused rather than answered with IBRD = 0 widths."""
    tab = synth.synth_lines(3000, 700.0, 760.0, [1, 7], 23)
    n = len(tab.vnu)
    flg = np.zeros((n, 7), np.int32)
    dat = np.zeros((n, 21), np.float32)
    o2 = np.nonzero(tab.mol % 100 == 7)[0][::3]
    flg[o2, 6] = 1
    flg[o2, 0] = 1
    dat[o2, 6 * 3 + 2] = np.linspace(-0.004, 0.004, len(o2)).astype(np.float32)
    path = tmp_path / "TAPE3"
    tape.write_tape3(path, tab, brd_flg=flg, brd_dat=dat)
    back = tape.read_tape3(path)
    assert back.nbrd == len(o2) > 0
    assert np.any(back.pshift != tab.pshift)
    gpu_ctx.load_tape3(path)
    p, t, wb = 700.0, 270.0, 1.0e25
    wk = [1e22, 0.0, 0.0, 0.0, 0.0, 0.0, 2e24]
    run = abi.make_run(705.0, 755.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 705.0, 755.0)
    od, ref, st, st_r = _run_both(gpu_ctx, run, lay, back)
    assert_od_close(od, ref, "broadener shift conversion")
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.opdpth(abi.make_run(705.0, 755.0, ibrd=1), abi.Layer.from_buffer_copy(lay))
    assert e.value.code == abi.E_UNSUPPORTED
    # without broadener flags IBRD changes nothing
    gpu_ctx.load_lines(tab)
    od1, _, _ = gpu_ctx.opdpth(abi.make_run(705.0, 755.0, ibrd=1), abi.Layer.from_buffer_copy(lay))
    od0, _, _ = gpu_ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
    np.testing.assert_allclose(od1, od0, rtol=1e-12, atol=0)  # LDS accumulation order only


@pytest.mark.parametrize("icntnm", [0, 1])
def test_no_far_wing(gpu_ctx, icntnm):
    """ILBLF4 = 0: no LBLF4 far wing (no TAPE9, R4 zero), LNCOR1 rejects
    against DPTMIN alone (oprop.f90:1156-1160); the single-layer workload
    and three layers of the column against the oracle, point by point."""
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = _layer0()
    run = abi.make_run(600.0, 1100.0, ilblf4=0, icntnm=icntnm)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 600.0, 1100.0)
    od, ref, st, st_r = _run_both(gpu_ctx, run, lay, tab)
    print("no far wing (icntnm %d): npts %d, lincnt %d, k3 %d" % (icntnm, len(od), st["lincnt"],
                                                                  st["k3_updates"]), flush=True)
    assert_od_close(od, ref, "ILBLF4 = 0")
    for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4", "lincnt", "nwin4"):
        assert st[k] == st_r[k], k
    assert st["nlins4"] == 0 and st["k3_updates"] == 0
    run1 = abi.make_run(600.0, 1100.0, ilblf4=1, icntnm=icntnm)
    od1, _, _ = gpu_ctx.opdpth(run1, abi.Layer.from_buffer_copy(lay))
    assert np.max(od1 - od) > 0.0  # the far wing's share is gone
    atm = synth.atmosphere()
    runc = abi.make_run(600.0, 1100.0, ilblf4=0, icntnm=icntnm)
    dvc, _ = lib.plan_dv(runc, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    for l in (0, 20, 45):
        L = abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dvc[l],
                           600.0, 1100.0)
        od, ref, st, st_r = _run_both(gpu_ctx, runc, L, tab)
        assert_od_close(od, ref, "ILBLF4 = 0 layer %d" % l)
        assert st["lincnt"] == st_r["lincnt"]
