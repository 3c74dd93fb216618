"""GPU parity of line coupling (IFLG 1 and 3, and IFLG 2 lines that LNCOR1
never rejects) with the CPU oracle through the C ABI: LNCOR1's and
LINF4's coefficient interpolation, CNVFNV's dispersive term in the tiles,
SHRINK's pass-through of coupled lines (the TAPE9 list and its SPP), and
CONVF4's dispersive term with no rejection of coupled lines
(src/oprop.f90:1039-1197, 1322-1351, 3118-3231, 3550-3557, 3909-4006)."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth, tape
from parity import assert_od_close

pytestmark = pytest.mark.gpu

P0, T0 = 1013.25, 296.0
LINE = (700.0, 1e-20, 0.07, 0.35, 0.75, 500.0, 0.001, 2)  # CO2


def _one(line, iflg, y, g):
    f = lambda i, dt: np.array([line[i]], dt)
    return abi.LineTable.from_lines_coupled(f(0, np.float64), f(1, np.float32), f(2, np.float32),
                                            f(5, np.float32), np.array([line[7] + 100], np.int32),
                                            f(3, np.float32), f(4, np.float32), f(6, np.float32),
                                            [iflg], [y], [g])


def _both(ctx, tab, run, p, t, wk, wb, v1, v2, tape9=False):
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    ref, v1r, st_r, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
    od, v1p, st = ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
    assert v1p == v1r
    if tape9:
        t9 = O.linf4(run, lay, tab)
        gv, gs, ga, ge, gm = ctx.debug_tape9(0)
        gspp = ctx.debug_tape9_spp(0)
        assert len(gv) == len(t9["vnu"])
        np.testing.assert_array_equal(gm, t9["mol"])
        np.testing.assert_allclose(gv, t9["vnu"], rtol=1e-14)
        np.testing.assert_allclose(gs, t9["sp"], rtol=1e-13)
        np.testing.assert_allclose(ga, t9["alfa"], rtol=1e-13)
        np.testing.assert_allclose(ge, t9["epp"], rtol=1e-13)
        np.testing.assert_allclose(gspp, t9["spp"], rtol=1e-13, atol=0)
        assert (gspp != 0).sum() == (t9["spp"] != 0).sum()
    return od, ref, st, st_r


@pytest.mark.parametrize("voigt", [0, 1])
@pytest.mark.parametrize("iflg,p,t,y,g", [
    (1, P0, T0, [0.05] * 4, [0.0] * 4),
    (1, 600.0, 231.0, [0.03, -0.02, 0.05, 0.01], [0.3, -0.2, 0.1, 0.4]),
    (1, 300.0, 355.0, [-0.06, 0.02, 0.04, -0.01], [0.1, 0.2, -0.1, 0.0]),
    (3, 700.0, 262.0, [0.06, 0.12, -0.03, 0.09], [-0.06, 0.03, 0.12, 0.0]),
    (2, P0, T0, [0.0] * 4, [0.0] * 4)])
def test_single_coupled_line(gpu_ctx, iflg, p, t, y, g, voigt):
    tab = _one(LINE, iflg, y, g)
    gpu_ctx.load_lines(tab)
    wk, wb = [0.0, 1e20], 1e24
    print("single line iflg", iflg, "p", p, "t", t, "voigt", voigt, flush=True)
    for kw in (dict(dptmin=0.0, dptfac=0.0), dict()):
        run = abi.make_run(660.0, 740.0, voigt=voigt, **kw)
        od, ref, st, st_r = _both(gpu_ctx, tab, run, p, t, wk, wb, 660.0, 740.0, tape9=True)
        assert_od_close(od, ref, "iflg %d" % iflg)
        for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4", "lincnt"):
            assert st[k] == st_r[k], k


def test_weak_lines_not_rejected(gpu_ctx):
    """A line below DPTMIN: IFLG 2 keeps its near wing (LNCOR1 skips the
    test), coupled it keeps both wings (CONVF4 skips it too)."""
    weak = list(LINE)
    weak[1] = 1e-27
    for iflg, y in ((0, 0.0), (2, 0.0), (1, 1e-3)):
        tab = _one(tuple(weak), iflg, [y] * 4, [0.0] * 4)
        gpu_ctx.load_lines(tab)
        run = abi.make_run(660.0, 740.0, dptmin=1e3, dptfac=0.0)
        od, ref, st, st_r = _both(gpu_ctx, tab, run, P0, T0, [0.0, 1e20], 1e24, 660.0, 740.0)
        assert_od_close(od, ref, "weak iflg %d" % iflg)
        assert (np.abs(ref).max() > 0) == (iflg != 0)


@pytest.mark.parametrize("precision,voigt", [(abi.FP64, 0), (abi.FP32_SPLIT, 0), (abi.FP64, 1)])
@pytest.mark.parametrize("p,t", [(500.0, 250.0), (900.0, 290.0), (120.0, 215.0)])
def test_dense_coupled_layer(gpu_ctx, precision, voigt, p, t):
    """20,000 lines over 600-1100 cm-1 (H2O, CO2, O3), 5 % with Y and G, 2 %
    with reduced widths, 1 % IFLG 2: coupled lines sit inside SHRINK's
    groups and LINF4's batches; the TAPE9 list (with SPP), the statistics
    and the optical depth equal the oracle's."""
    tab = synth.coupled_lines(20000, 600.0, 1100.0, [1, 2, 3], 41)
    assert (tab.iflg == 1).sum() > 800 and (tab.iflg == -1).sum() == ((tab.iflg == 1) | (tab.iflg == 3)).sum()
    gpu_ctx.load_lines(tab)
    wk, wb = [1e22, 4e21, 1e18], 1.2e25
    run = abi.make_run(610.0, 1090.0, precision=precision, voigt=voigt)
    print("dense coupled layer p", p, "t", t, "precision", precision, "voigt", voigt, flush=True)
    od, ref, st, st_r = _both(gpu_ctx, tab, run, p, t, wk, wb, 610.0, 1090.0, tape9=precision == abi.FP64)
    rel = np.abs(od - ref)[np.abs(ref) > 1e-6] / np.abs(ref)[np.abs(ref) > 1e-6]
    print("  npts", len(od), "max rel", rel.max(), flush=True)
    assert_od_close(od, ref, "dense coupled")
    if precision == abi.FP64:
        for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4", "lincnt"):
            assert st[k] == st_r[k], k


def test_coupled_column_batch(gpu_ctx):
    """A 50-layer column (200-295 K: every coefficient interval of LNCOR1
    and LINF4) in one batch over 60,000 lines with coupling, 700-1300 cm-1;
    six layers checked against the oracle."""
    tab = synth.coupled_lines(60000, 700.0, 1300.0, [1, 2, 3, 4, 5, 6, 7], 43)
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    v1, v2 = 720.0, 1280.0
    run = abi.make_run(v1, v2)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l], v1, v2)
              for l in range(len(dv))]
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
        if l not in (0, 9, 18, 27, 38, 49):
            continue
        ref, _, st_r, _ = O.od_layer(run, inp, tab)
        od = allod[offs[l]:offs[l] + sizes[l]]
        print("layer", l, "T", round(atm["tave"][l], 1), "npts", len(od), flush=True)
        assert_od_close(od, ref, "layer %d" % l)
        for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4"):
            assert st[l][k] == st_r[k], (l, k)


def test_tape3_round_trip_with_coupling(gpu_ctx, tmp_path):
    """A TAPE3 file holding coupling records, read by the library's own
    TAPE3 reader, gives the oracle's optical depth of the in-memory table."""
    tab = synth.coupled_lines(3000, 700.0, 760.0, [2], 45, frac1=0.2, frac3=0.05)
    path = tmp_path / "TAPE3"
    tape.write_tape3(path, tab)
    gpu_ctx.load_tape3(path)
    run = abi.make_run(705.0, 755.0)
    od, ref, st, st_r = _both(gpu_ctx, tab, run, 500.0, 250.0, [0.0, 1e21], 1e25, 705.0, 755.0)
    assert_od_close(od, ref, "tape3 coupled")


def test_refusals(gpu_ctx):
    """The tangent-linear path refuses coupling."""
    tab = _one(LINE, 1, [0.05] * 4, [0.0] * 4)
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = P0, T0, [0.0, 1e20], 1e24
    run = abi.make_run(660.0, 740.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 660.0, 740.0)
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.od_tangent(run, [abi.Layer.from_buffer_copy(lay)], [2])
    assert e.value.code == abi.E_UNSUPPORTED


@pytest.mark.parametrize("v1", [610.0, 0.2])
def test_coupled_lines_without_far_wing(gpu_ctx, v1):
    """ILBLF4 = 0 with coupled lines: the near wing alone, coupled and IFLG 2
    lines exempt from the DPTMIN test; in the mid IR and near 0 cm-1."""
    if v1 < 1.0:
        tab = synth.coupled_lines(3000, 0.5, 80.0, [1, 3], 61, frac1=0.2, frac3=0.05)
        v2, wk = 60.0, [2e19, 0.0, 1e17]
    else:
        tab = synth.coupled_lines(20000, 600.0, 1100.0, [1, 2, 3], 41)
        v2, wk = 1090.0, [1e22, 4e21, 1e18]
    gpu_ctx.load_lines(tab)
    run = abi.make_run(v1, v2, ilblf4=0, icntnm=1)
    od, ref, st, st_r = _both(gpu_ctx, tab, run, 500.0, 250.0, wk, 1.2e25, v1, v2)
    assert_od_close(od, ref, "coupled, no far wing, v1=%g" % v1)
    for k in ("npanel", "npts", "k1_updates", "lincnt"):
        assert st[k] == st_r[k], k
