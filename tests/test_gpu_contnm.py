"""GPU parity of the continuum (CONTNM -> ABSRB -> XINT into R3) with the
CPU oracle, through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth
from parity import assert_od_close

pytestmark = pytest.mark.gpu


def _atm_layer(l, v1, v2, dv=0.001):
    atm = synth.atmosphere()
    return abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv,
                          v1, v2)


@pytest.mark.parametrize("icntnm", [1, 3, 5])
def test_absrb_column_layers(gpu_ctx, icntnm):
    """ABSRB of layers across the column (JRAD 0 at the warm surface
    layers' 500 cm-1 edge, 1 above) on 500-2500 and 1200-2000 cm-1."""
    for v1, v2 in [(500.0, 2500.0), (1200.0, 2000.0), (60.0, 700.0)]:
        run = abi.make_run(v1, v2, icntnm=icntnm)
        for l in (0, 7, 20, 35, 49):
            lay = _atm_layer(l, v1, v2)
            lay.jrad = 1
            g, g1, gd = gpu_ctx.contnm(run, lay)
            r, r1, rd = O.contnm(run, _atm_layer(l, v1, v2), lay.jrad)
            print("absrb", icntnm, v1, v2, l, "jrad", lay.jrad, flush=True)
            assert (g1, gd, len(g)) == (r1, rd, len(r))
            np.testing.assert_allclose(g, r, rtol=1e-13, atol=0)


def test_absrb_scale_factors_and_closure_table(gpu_ctx):
    lay = _atm_layer(10, 500.0, 2500.0)
    for run in [abi.make_run(500.0, 2500.0, icntnm=6, cntscl=[0.3, 1.7, 0.5, 1.0, 2.0, 0.25, 1.1]),
                abi.make_run(500.0, 2500.0, icntnm=1, ifrgnx=1),
                abi.make_run(500.0, 2500.0, icntnm=1, iaersl=1)]:
        L = abi.Layer.from_buffer_copy(lay)
        g, _, _ = gpu_ctx.contnm(run, L)
        r, _, _ = O.contnm(run, abi.Layer.from_buffer_copy(lay), L.jrad)
        np.testing.assert_allclose(g, r, rtol=1e-13, atol=0)


def test_unsupported_continuum_band(gpu_ctx):
    run = abi.make_run(4000.0, 4600.0, icntnm=1)
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.contnm(run, _atm_layer(3, 4000.0, 4600.0))
    assert e.value.code == abi.E_UNSUPPORTED


def test_far_infrared_layer_od(gpu_ctx):
    """60-340 cm-1 with the N2 rotational continuum (below 350 cm-1) and
    lines: the layer optical depth against the oracle's, surface and upper
    layers."""
    tab = synth.synth_lines(8000, 30.0, 370.0, [1, 2, 3, 7], 20261211)
    gpu_ctx.load_lines(tab)
    run = abi.make_run(60.0, 340.0, icntnm=1)
    atm = synth.atmosphere()
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    for l in (0, 25, 45):
        lay = abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                             60.0, 340.0)
        od, v1p, st = gpu_ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
        ref, v1r, st_r, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
        print("far-IR layer", l, "points", len(od), flush=True)
        assert v1p == v1r
        assert_od_close(od, ref, "far IR layer %d" % l)


def test_layer_od_with_continuum(gpu_ctx):
    """20,000 lines over 600-1100 cm-1 with the continuum on, and a line-free
    band where the optical depth is the continuum alone."""
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = synth.config0_layer()
    for v1, v2 in [(600.0, 1100.0), (1500.0, 1600.0)]:
        run = abi.make_run(v1, v2, icntnm=1)
        dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
        lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
        ref, v1r, st_r, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
        od, v1p, st = gpu_ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
        print("od+continuum", v1, v2, "npts", len(od), flush=True)
        assert v1p == v1r
        assert_od_close(od, ref, "continuum %g-%g" % (v1, v2))
        assert np.all(od > 0)
