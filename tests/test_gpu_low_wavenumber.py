"""GPU parity below 25 cm-1 with the CPU oracle through the C ABI: RSYM's
fold of the near-wing grids when the first panel starts at or below 0
cm-1 (src/oprop.f90:501-505, 2145-2200) and CONVF4's Van Vleck -
Weisskopf mirror of far-wing lines at 0 < nu <= 25 (:4014-4024), in both
builds and precisions, with and without the far wing, with coupled lines
(the mirror negates SPP), and over a far-IR column."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth
from parity import assert_od_close

pytestmark = pytest.mark.gpu


def _both(ctx, tab, run, p, t, wk, wb, v1, v2):
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    ref, v1r, st_r, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
    od, v1p, st = ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
    assert v1p == v1r
    return od, ref, st, st_r


@pytest.mark.parametrize("v1", [0.0, 0.2, 0.215, 3.0, 30.0])
@pytest.mark.parametrize("kw", [dict(), dict(voigt=1), dict(precision=abi.FP32_SPLIT), dict(ilblf4=0),
                                dict(ilblf4=2, icntnm=1)])
def test_far_ir_layer(gpu_ctx, v1, kw):
    """3000 lines of H2O, O3 and O2 over 0-120 cm-1 (some within 25 cm-1
    of 0, some at 0.3 cm-1), one layer from v1 to 90 cm-1."""
    tab = synth.synth_lines(3000, 0.3, 120.0, [1, 3, 7], 51)
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = 700.0, 260.0, [2e19, 0.0, 3e17, 0.0, 0.0, 0.0, 4e24], 1.5e25
    run = abi.make_run(v1, 90.0, **kw)
    print("far-IR layer v1", v1, kw, flush=True)
    od, ref, st, st_r = _both(gpu_ctx, tab, run, p, t, wk, wb, v1, 90.0)
    print("  first points: device", od[:6], "oracle", ref[:6], flush=True)
    assert_od_close(od, ref, "far-IR v1=%g %s" % (v1, kw))
    if kw.get("precision", abi.FP64) == abi.FP64:
        for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4", "lincnt"):
            assert st[k] == st_r[k], k


@pytest.mark.parametrize("v1", [0.1, 0.0])
def test_coupled_lines_near_zero(gpu_ctx, v1):
    """Coupled lines below 25 cm-1: the mirror carries -SPP."""
    tab = synth.coupled_lines(2000, 0.5, 60.0, [1, 3], 53, frac1=0.3, frac3=0.1)
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = 500.0, 240.0, [1e22, 0.0, 1e18], 1.0e25
    for voigt in (0, 1):
        run = abi.make_run(v1, 50.0, voigt=voigt)
        od, ref, st, st_r = _both(gpu_ctx, tab, run, p, t, wk, wb, v1, 50.0)
        assert_od_close(od, ref, "coupled near zero v1=%g voigt=%d" % (v1, voigt))


def test_far_ir_column(gpu_ctx):
    """The 50-layer column over 0.1-80 cm-1 (the first panel at or below 0
    in every layer) in one batch; five layers against the oracle."""
    tab = synth.synth_lines(8000, 0.3, 130.0, [1, 2, 3, 7], 55)
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    v1, v2 = 0.1, 80.0
    run = abi.make_run(v1, v2, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l], v1, v2)
              for l in range(len(dv))]
    sizes = [lib.od_size(run, L) for L in layers]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = gpu_ctx.alloc(8 * int(sum(sizes)))
    lays = [abi.Layer.from_buffer_copy(x) for x in layers]
    gpu_ctx.opdpth_column_device(run, lays, buf.ptr, offs)
    allod = buf.to_host(np.float64)
    buf.free()
    jrad = 1
    for l in range(len(dv)):
        inp = abi.Layer.from_buffer_copy(layers[l])
        inp.jrad = jrad
        jrad = lays[l].jrad
        if l not in (0, 12, 25, 37, 49):
            continue
        ref, _, _, _ = O.od_layer(run, inp, tab)
        print("far-IR column layer", l, "npts", sizes[l], flush=True)
        assert_od_close(allod[offs[l]:offs[l] + sizes[l]], ref, "far-IR layer %d" % l)


def test_tangent_refuses_far_ir(gpu_ctx):
    """The tangent-linear passes have neither the fold nor the mirror: a
    band reaching below 25 cm-1 is refused, not differentiated wrongly."""
    tab = synth.synth_lines(500, 0.3, 60.0, [1], 59)
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = 700.0, 260.0, [2e19], 1.5e25
    for v1, v2 in ((0.1, 40.0), (30.0, 60.0)):
        run = abi.make_run(v1, v2)
        dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
        lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
        with pytest.raises(lib.LblxError) as e:
            gpu_ctx.od_tangent(run, [lay], [1])
        assert e.value.code == abi.E_UNSUPPORTED
