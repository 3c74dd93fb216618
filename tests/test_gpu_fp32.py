"""FP32 split precision on the GPU: line shapes evaluated and accumulated
in FP32, line positions, grid offsets and panel start wavenumbers in FP64.
Every result is compared with the FP64 oracle (directly for single layers,
and through the FP64 GPU path -- itself pinned to the oracle -- for the
whole column) at the tolerance of the FP64 contract."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth
from parity import OD_ABS, OD_ABS_FLOOR, OD_REL, assert_od_close

pytestmark = pytest.mark.gpu

RADCN1 = 1.191042722E-12
RADCN2 = 1.4387752


def bt(rad, v):
    return RADCN2 * v / np.log1p(RADCN1 * v ** 3 / rad)


def od_errors(got, ref):
    big = np.abs(ref) > OD_ABS_FLOOR
    rel = np.abs(got[big] - ref[big]) / np.abs(ref[big])
    small = np.abs(got[~big] - ref[~big]) if (~big).any() else np.zeros(1)
    return rel.max(), np.sqrt(np.mean(rel ** 2)), small.max()


def test_single_layer_fp32_vs_oracle(gpu_ctx):
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    p, t, wk, wb = synth.config0_layer()
    run64 = abi.make_run(600.0, 1100.0)
    run32 = abi.make_run(600.0, 1100.0, precision=abi.FP32_SPLIT)
    dv, _ = lib.plan_dv(run64, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], 600.0, 1100.0)
    ref, v1r, _, _ = O.od_layer(run64, abi.Layer.from_buffer_copy(lay), tab)
    od, v1p, _ = gpu_ctx.opdpth(run32, abi.Layer.from_buffer_copy(lay))
    assert od.dtype == np.float32 and v1p == v1r
    mx, rms, ab = od_errors(od.astype(np.float64), ref)
    print("fp32 single layer: max rel %.3e rms %.3e max abs below floor %.3e" % (mx, rms, ab), flush=True)
    assert_od_close(od, ref, "fp32 single layer")


def test_column_spectrum_fp32_vs_fp64(gpu_ctx):
    """The 50-layer column with the continuum: per-layer OD within 1e-4 of
    FP64 (five layers against the oracle itself) and TOA brightness
    temperature within 0.01 K."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run64 = abi.make_run(500.0, 2500.0, icntnm=1)
    run32 = abi.make_run(500.0, 2500.0, icntnm=1, precision=abi.FP32_SPLIT)
    dv, _ = lib.plan_dv(run64, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                             500.0, 2500.0) for l in range(len(dv))]
    sizes = [lib.od_size(run64, L) for L in layers]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    b32 = gpu_ctx.alloc(4 * int(sum(sizes)))
    lays = [abi.Layer.from_buffer_copy(x) for x in layers]
    gpu_ctx.opdpth_column_device(run32, lays, b32.ptr, offs)
    od32 = b32.to_host(np.float32).astype(np.float64)
    b32.free()
    jrad = 1
    worst = 0.0
    for l in range(len(dv)):
        inp = abi.Layer.from_buffer_copy(layers[l])
        inp.jrad = jrad
        jrad = lays[l].jrad
        if l not in (0, 12, 25, 37, 49):
            continue
        ref, _, _, _ = O.od_layer(run64, inp, tab)
        got = od32[offs[l]:offs[l] + sizes[l]]
        mx, rms, ab = od_errors(got, ref)
        worst = max(worst, mx)
        print("fp32 layer %d: max rel %.3e rms %.3e max abs below floor %.3e" % (l, mx, rms, ab), flush=True)
        assert_od_close(got, ref, "fp32 layer %d" % l)
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    r64, t64, v1p, dvt = gpu_ctx.spectrum(run64, [abi.Layer.from_buffer_copy(x) for x in layers], tzu,
                                          295.0, 0.98)
    r32, t32, v1p2, _ = gpu_ctx.spectrum(run32, [abi.Layer.from_buffer_copy(x) for x in layers], tzu,
                                         295.0, 0.98)
    assert v1p == v1p2
    v = v1p + np.arange(len(r64)) * dvt
    d = np.abs(bt(r32, v) - bt(r64, v))
    print("fp32 spectrum: max |dBT| %.3e K rms %.3e K" % (d.max(), np.sqrt(np.mean(d ** 2))), flush=True)
    assert d.max() <= 0.01
