"""GPU parity of SCANFN (spectral-domain scanning function, src/postsub.f90:
25-1688) with the CPU oracle through the C ABI: every function, the
first-zero form of sinc/sinc^2, the three JEMIT choices, optical-depth
input, panel layouts that move the average's origin, the unaveraged case
(R = 1), and the 50-layer column's TOA radiance at full size.

The device forms each group's table index from ZF directly and falls back
to CONVSC's running sum only where the two could round to different
entries, so the sums are the oracle's term for term: 1e-13 of the largest
output."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth

pytestmark = pytest.mark.gpu

V1C, DV = 700.0, 1.0e-3
N = 300001


@pytest.mark.parametrize("kw", [dict(hw=0.3, jfn=2), dict(hw=0.5, jfn=1, npanel=777),
                                dict(hw=0.2, jfn=-3), dict(hw=0.015, jfn=1),
                                dict(hw=0.25, jfn=4, sampl=-0.05), dict(hw=0.4, jfn=3, sampl=3.0),
                                dict(hw=0.6, jfn=-4, npanel=1000), dict(hw=0.3, jfn=2, jemit=0),
                                dict(hw=0.3, jfn=1, jemit=-1), dict(hw=0.35, jfn=2, jemit=0, iemit=0)])
def test_scanfn_matches_oracle(gpu_ctx, kw):
    kw = dict(kw)
    rng = np.random.default_rng(11)
    nu = V1C + DV * np.arange(N)
    iemit = kw.pop("iemit", 1)
    jemit = kw.pop("jemit", 1)
    if iemit == 0:
        x = rng.exponential(0.5, N) * (1.0 + np.sin(nu / 0.3)) + np.where(rng.random(N) < 0.1, 1e-5, 0.0)
    elif jemit == 1:
        x = 1e-6 * (1.0 + 0.3 * np.sin(nu / 0.7)) * (1.0 + 0.05 * rng.random(N))
    else:
        x = np.exp(-rng.exponential(0.5, N))
    hw, jfn = kw.pop("hw"), kw.pop("jfn")
    sc = abi.make_scanfn(hw, 820.0, 840.0, jemit=jemit, jfn=jfn, **kw)
    print("scanfn", hw, jfn, jemit, iemit, kw, flush=True)
    got, v1, dvo = gpu_ctx.scanfn(sc, x, V1C, DV, iemit=iemit)
    ref, v1r, dvr = O.scanfn(sc, x, V1C, DV, iemit=iemit)
    assert (len(got), v1, dvo) == (len(ref), v1r, dvr)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())


def test_scanfn_column_radiance(gpu_ctx):
    """The 50-layer column's TOA radiance (5.4e6 points, 500-2500 cm-1)
    scanned with sinc at 0.5 cm-1 HWHM and with the triangle at 0.25 cm-1:
    the device equals the oracle's panel-by-panel scan of the same
    radiance."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(500.0, 2500.0, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                             500.0, 2500.0) for l in range(len(dv))]
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    rad, tra, v1p, dvt = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tzu, 295.0, 0.98)
    for sc in (abi.make_scanfn(0.5, 520.0, 2480.0, jemit=1, jfn=4),
               abi.make_scanfn(0.25, 520.0, 2480.0, jemit=0, jfn=1)):
        x = rad if sc.jemit == 1 else tra
        print("column scanfn jfn %d jemit %d: %d input points" % (sc.jfn, sc.jemit, len(x)), flush=True)
        got, v1, dvo = gpu_ctx.scanfn(sc, x, v1p, dvt)
        ref, v1r, dvr = O.scanfn(sc, x, v1p, dvt)
        assert (len(got), v1, dvo) == (len(ref), v1r, dvr)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
