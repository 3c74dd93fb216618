"""The Fortran half of the boundary on the GPU: the column driver
(tests/fortran/drive_column.f90, in the role of XLAYER's layer loop) calls
the OPDPTH and XMERGE shims of lblrtm_amd/fortran/lblx_shims.f90, which
call liblblx through lblx_c.  The TAPE10 the shim writes holds the last
layer's optical depth as the oracle computes it, byte for byte as
lblx_write_panels writes the same values; the final merged TAPE11/TAPE12
holds the TOA radiance of lblx_spectrum_device and, within 0.01 K, of the
oracle's OPDPTH + XMERGE chain."""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth, tape
from parity import assert_od_close

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "drive_column")
RADCN1, RADCN2 = 1.191042722E-12, 1.4387752


def bt(rad, v):
    return RADCN2 * v / np.log1p(RADCN1 * v ** 3 / rad)


@pytest.mark.skipif(not os.path.exists(EXE), reason="flang drivers not built")
@pytest.mark.parametrize("ipathl", [1, 3])
def test_fortran_column_through_shims(gpu_ctx, tmp_path, ipathl):
    """ipathl 1: upwelling at the top over a surface; 3: downwelling at the
    surface looking up, with a 240 K black boundary beyond the top layer
    (EMBND)."""
    tab = synth.synth_lines(20000, 575.0, 725.0, [1, 2, 3], 41)
    tape.write_tape3(tmp_path / "TAPE3", tab)
    atm = synth.atmosphere(nlev=11, ptop=100.0, seed=5)
    v1, v2 = 600.0, 700.0
    run = abi.make_run(v1, v2, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    nl, nm = len(dv), atm["wk"].shape[1]
    if ipathl == 1:
        tzu = np.asarray(atm["tlev"][1:], np.float64)
        tbnd, emis = float(atm["tlev"][0]), 0.98
    else:
        tzu = np.asarray(atm["tlev"][:-1], np.float64)  # TZL
        tbnd, emis = 240.0, 1.0
    with open(tmp_path / "layers.txt", "w") as f:
        f.write("%d %d %d %d\n" % (nl, nm, 1, ipathl))
        f.write("%r %r %r %r\n" % (v1, v2, tbnd, emis))
        for l in range(nl):
            vals = [atm["pave"][l], atm["tave"][l], dv[l], tzu[l], atm["wbroad"][l]] + list(atm["wk"][l])
            f.write(" ".join(repr(float(x)) for x in vals) + "\n")
    print("fortran driver: %d layers, %d lines" % (nl, tab.nlines()), flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    final = int(open(tmp_path / "final.txt").read().split()[0])
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l], v1, v2)
              for l in range(nl)]
    # the oracle's chain: every layer's optical depth, merged up the column
    jrad = 1
    m = None
    for l, L in enumerate(layers):
        inp = abi.Layer.from_buffer_copy(L)
        inp.jrad = jrad
        od_ref, _, _, jrad_next = O.od_layer(run, inp, tab)
        pan = O.panels()
        if ipathl == 1:
            m = (O.merge_first(od_ref, pan, L.dv, L.tave, tzu[l], tbnd, emis) if l == 0
                 else O.merge_next(m, od_ref, pan, L.dv, L.tave, tzu[l], v2)[0])
        else:
            last_l = l == nl - 1
            m = (O.merge_first(od_ref, pan, L.dv, L.tave, tzu[l], 0.0, 0.0) if l == 0
                 else O.merge_next(m, od_ref, pan, L.dv, L.tave, tzu[l], v2, ipathl=3,
                                   tbnd=tbnd if last_l else 0.0, emis=emis if last_l else 0.0)[0])
        if l < nl - 1:
            jrad = jrad_next
    # TAPE10: the last layer's optical depth in PANEL's records
    t10 = tape.read_panels(tmp_path / "TAPE10", word=8, nvals=1, filhdr_words=177)
    last = abi.Layer.from_buffer_copy(layers[-1])
    last.jrad = jrad
    pv1, pv2, pnl = lib.od_panels(run, last)
    np.testing.assert_array_equal(t10["v1p"], pv1)
    np.testing.assert_array_equal(t10["v2p"], pv2)
    np.testing.assert_array_equal(t10["nlim"], pnl)
    assert_od_close(t10["values"][0], od_ref, "TAPE10 through the OPDPTH shim")
    lib.write_panels(tmp_path / "W10", pv1, pv2, last.dv, pnl, t10["values"][0], filhdr=t10["filhdr"])
    assert open(tmp_path / "W10", "rb").read() == open(tmp_path / "TAPE10", "rb").read()
    # the final merged file: TOA radiance and transmittance
    mf = tape.read_panels(tmp_path / ("TAPE%d" % final), word=8, nvals=2, filhdr_words=177)
    rad, tra = mf["values"]
    gpu_ctx.load_lines(tab)
    rs, ts, v1p, dvt = gpu_ctx.spectrum_path(run, [abi.Layer.from_buffer_copy(x) for x in layers], ipathl, tzu,
                                             tbnd, emis)
    np.testing.assert_allclose(rad, rs, rtol=1e-10)
    np.testing.assert_allclose(tra, ts, rtol=1e-10, atol=1e-300)
    v = v1p + dvt * np.arange(len(rad))
    d = np.abs(bt(rad, v) - bt(m.rad, v))
    print("fortran column (IPATHL %d): %d points, max |dBT| vs oracle %.2e K" % (ipathl, len(rad), d.max()),
          flush=True)
    assert d.max() <= 0.01
