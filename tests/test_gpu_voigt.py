"""GPU parity of the exact-Voigt build (src/oprop_voigt.f90: zeta-table
CNVFNV and the zeta-dependent CONVF4) with the oracle, FP64 and FP32 split,
through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth
from parity import assert_od_close

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_SPLIT])
def test_single_layer_exact_voigt(gpu_ctx, precision):
    """20,000 lines, 600-1100 cm-1, one layer."""
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = synth.config0_layer()
    run = abi.make_run(600.0, 1100.0, voigt=1, precision=precision)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 600.0, 1100.0)
    ref, v1r, st_r, _ = O.od_layer(abi.make_run(600.0, 1100.0, voigt=1), abi.Layer.from_buffer_copy(lay), tab)
    od, v1p, st = gpu_ctx.opdpth(run, abi.Layer.from_buffer_copy(lay))
    assert v1p == v1r
    rel = np.abs(od - ref) / np.maximum(np.abs(ref), 1e-300)
    print("exact Voigt single layer precision %d: npts %d max rel %.3e" % (precision, len(od), rel.max()),
          flush=True)
    assert_od_close(od, ref, "exact Voigt single layer")
    if precision == abi.FP64:
        big = np.abs(ref) > 1e-6
        assert rel[big].max() < 1e-9
        for k in ("npanel", "npts", "k1_updates", "k3_updates", "nlins4"):
            assert st[k] == st_r[k], k
    # the build differs from the approximate one
    od0, _, _ = gpu_ctx.opdpth(abi.make_run(600.0, 1100.0, precision=precision), abi.Layer.from_buffer_copy(lay))
    assert np.abs(od0 - od).max() > 1e-6 * np.abs(od).max()


def test_column_layers_exact_voigt(gpu_ctx):
    """The 50-layer column over 900-1100 cm-1 with the 500,000-line table:
    all layers in one batch; surface, middle and top layers checked."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(900.0, 1100.0, voigt=1, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l],
                             dv[l], 900.0, 1100.0) for l in range(len(dv))]
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
        if l not in (0, 25, 49):
            continue
        ref, _, _, _ = O.od_layer(run, inp, tab)
        got = allod[offs[l]:offs[l] + sizes[l]]
        print("exact Voigt column layer %d: %d points" % (l, sizes[l]), flush=True)
        assert_od_close(got, ref, "exact Voigt layer %d" % l)
