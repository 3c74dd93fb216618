"""TAPE3 writer/reader round trip and the synthetic inputs."""
import numpy as np

from lblrtm_amd import synth, tape


def test_tape3_roundtrip(tmp_path):
    tab = synth.synth_lines(1234, 600.0, 700.0, [1, 2, 7], 11)
    p = tmp_path / "TAPE3"
    tape.write_tape3(p, tab)
    back = tape.read_tape3(p)
    assert back.nblk == tab.nblk == 5
    assert list(back.nrec) == [250, 250, 250, 250, 234]
    for name in ("vnu", "sp", "alfa", "epp", "mol", "hwhm", "tmpalf", "pshift", "iflg"):
        np.testing.assert_array_equal(getattr(back, name), getattr(tab, name))
    np.testing.assert_array_equal(back.vmax, tab.vmax)


def test_synthetic_distributions():
    tab = synth.config0_lines()
    n = tab.nlines()
    assert n == 20000
    v = tab.vnu[:n]
    assert np.all(np.diff(v) >= 0) and v.min() >= 600 and v.max() <= 1100
    s = tab.sp[:n].astype(np.float64)
    assert 1e-26 <= s.min() and s.max() <= 1e-19
    r = tab.hwhm[:n] / tab.alfa[:n]
    assert r.min() >= 2 - 1e-6 and r.max() <= 5 + 1e-6
    atm = synth.atmosphere()
    assert len(atm["pave"]) == 50
    assert np.all(np.diff(atm["pave"]) < 0)
    assert abs(atm["tave"].min() - 217) < 4
