"""GPU parity of FFTSCN (instrument scan) with the CPU oracle through the C
ABI, and the precision form of the column workload: FP32-split spectrum
scanned to 0.5 cm-1 Hamming resolution against the FP64 one."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth

pytestmark = pytest.mark.gpu

RADCN1 = 1.191042722E-12
RADCN2 = 1.4387752
V1IN, DVIN = 500.0, 3.683e-4


def bt(rad, v):
    return RADCN2 * v / np.log1p(RADCN1 * v ** 3 / rad)


@pytest.mark.parametrize("kw", [dict(jfnin=-6, dvout=0.1), dict(jfnin=-6, dvout=0.0),
                                dict(jfnin=-6, dvout=0.25, nofix=1), dict(jfnin=1, hwhm=0.3, dvout=0.05),
                                dict(jfnin=-3, dvout=0.0), dict(jfnin=-4, dvout=0.1),
                                dict(jfnin=2, hwhm=0.4, mratin=8, dvout=0.1), dict(jfnin=0, hwhm=0.2),
                                dict(jfnin=-6, jemit=0, dvout=0.1)])
def test_scan_matches_oracle(gpu_ctx, kw):
    rng = np.random.default_rng(7)
    n = int(2000.0 / DVIN) + 1
    nu = V1IN + DVIN * np.arange(n)
    x = 1e-6 * (1.0 + 0.3 * np.sin(nu / 7.0)) * (1.0 + 0.05 * rng.random(n))
    if kw.get("jemit", 1) == 0:
        x = np.exp(-rng.exponential(0.5, n))
    args = dict(hwhm=1.0, jemit=1, mratin=0, dvout=0.0, nofix=0)
    args.update(kw)
    sc = abi.make_scan(args.pop("hwhm"), 520.0, 2480.0, **args)
    print("scan", kw, flush=True)
    got, v1, dv = gpu_ctx.fftscn(sc, x, V1IN, DVIN)
    ref, v1r, dvr = O.fftscn(sc, x, V1IN, DVIN)
    assert (len(got), v1, dv) == (len(ref), v1r, dvr)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_scan_without_boxcar(gpu_ctx):
    """MRATIN < 0: the whole loaded range goes through the FFT (16 records)."""
    n = 200001
    nu = V1IN + DVIN * np.arange(n)
    x = np.exp(-(nu - 530.0) ** 2 / 50.0)
    sc = abi.make_scan(1.0, 515.0, 550.0, mratin=-1, dvout=0.05)
    got, v1, dv = gpu_ctx.fftscn(sc, x, V1IN, DVIN)
    ref, v1r, dvr = O.fftscn(sc, x, V1IN, DVIN)
    assert (len(got), v1, dv) == (len(ref), v1r, dvr)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def test_column_scan_fp32_split_vs_fp64(gpu_ctx):
    """50-layer column with continuum, TOA radiance scanned with the Hamming
    function at 1 cm OPD (0.5 cm-1 resolution) onto 0.1 cm-1: FP32 split
    against FP64 within 0.01 K, and the device scan equal to the oracle's
    scan of the same radiance."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run64 = abi.make_run(500.0, 2500.0, icntnm=1)
    run32 = abi.make_run(500.0, 2500.0, icntnm=1, precision=abi.FP32_SPLIT)
    dv, _ = lib.plan_dv(run64, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                             500.0, 2500.0) for l in range(len(dv))]
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    sc = abi.make_scan(1.0, 510.0, 2490.0, jemit=1, jfnin=-6, dvout=0.1)
    out = {}
    for name, run in (("fp64", run64), ("fp32", run32)):
        rad, tra, v1p, dvt = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tzu,
                                              295.0, 0.98)
        s, v1s, dvs = gpu_ctx.fftscn(sc, rad, v1p, dvt)
        out[name] = (s, v1s, dvs, rad, v1p, dvt)
    s64, v1s, dvs, rad64, v1p, dvt = out["fp64"]
    s32 = out["fp32"][0]
    v = v1s + dvs * np.arange(len(s64))
    d = np.abs(bt(s32, v) - bt(s64, v))
    print("scanned fp32 vs fp64: max |dBT| %.3e K, rms %.3e K, %d points" %
          (d.max(), np.sqrt(np.mean(d ** 2)), len(v)), flush=True)
    assert d.max() <= 0.01
    ref, v1r, dvr = O.fftscn(sc, rad64, v1p, dvt)
    assert (v1r, dvr, len(ref)) == (v1s, dvs, len(s64))
    np.testing.assert_allclose(s64, ref, rtol=1e-10)
