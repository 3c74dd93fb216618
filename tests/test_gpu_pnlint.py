"""GPU PNLINT (OPDPTH onto DVOUT) against the oracle's, through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth
from parity import assert_od_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("frac", [1.0, 2.0, 2.5, 3.7])
def test_single_layer_dvout(gpu_ctx, frac):
    """20,000 lines, 600-1100 cm-1, one layer, DVOUT = DV / frac."""
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = synth.config0_layer()
    run = abi.make_run(600.0, 1100.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 600.0, 1100.0)
    dvout = dv[0] / frac
    ref, pan = O.od_layer_dvout(run, lay, tab, dvout)
    got, v1 = gpu_ctx.od_layer_dvout(run, abi.Layer.from_buffer_copy(lay), dvout)
    print("DVOUT = DV/%.1f: %d points in %d panels" % (frac, len(got), len(pan[0])), flush=True)
    assert len(got) == len(ref) == lib.od_dvout_size(run, lay, dvout)
    assert v1 == pan[0][0]
    big = np.abs(ref) > 1e-6
    assert (np.abs(got - ref)[big] / np.abs(ref[big])).max() < 1e-9
    assert_od_close(got, ref, "DVOUT")


def test_column_top_layer_dvout_fp32(gpu_ctx):
    """Top layer of the 50-layer column, 900-1000 cm-1, FP32 split, DVOUT =
    DV / 2.5, within the OD tolerance of the oracle."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(900.0, 1000.0, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=0)
    l = len(dv) - 1
    lay = abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l], 900.0, 1000.0)
    dvout = dv[l] / 2.5
    ref, pan = O.od_layer_dvout(run, lay, tab, dvout)
    run32 = abi.make_run(900.0, 1000.0, icntnm=1, precision=abi.FP32_SPLIT)
    got, v1 = gpu_ctx.od_layer_dvout(run32, abi.Layer.from_buffer_copy(lay), dvout)
    print("top layer DVOUT FP32: %d points" % len(got), flush=True)
    assert_od_close(got, ref, "DVOUT FP32")
