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


def test_tape3_broadener_words(tmp_path):
    """Broadener words survive the round trip; on reading, an O2 line whose
    own O2 flag is set gets RDLIN's foreign-shift conversion
    (pshift - 0.21 shift_O2) / 0.79 (src/oprop.f90:818-827) and nothing
    else changes."""
    tab = synth.synth_lines(600, 600.0, 700.0, [1, 7], 3)
    n = len(tab.vnu)
    flg = np.zeros((n, 7), np.int32)
    dat = np.zeros((n, 21), np.float32)
    m = tab.mol % 100
    i_o2 = np.nonzero(m == 7)[0][:5]
    i_h2o = np.nonzero(m == 1)[0][:5]
    flg[i_o2, 6] = 1
    flg[i_h2o, 6] = 1      # an H2O line broadened by O2: no conversion
    dat[i_o2, 20] = 0.002
    dat[i_h2o, 20] = 0.002
    p = tmp_path / "TAPE3"
    tape.write_tape3(p, tab, brd_flg=flg, brd_dat=dat)
    back = tape.read_tape3(p)
    assert back.nbrd == 10
    want = tab.pshift.copy()
    want[i_o2] = ((tab.pshift[i_o2].astype(np.float64) - 0.21 * np.float32(0.002).astype(np.float64))
                  / 0.79).astype(np.float32)
    np.testing.assert_array_equal(back.pshift, want)
    for name in ("vnu", "sp", "alfa", "epp", "mol", "hwhm", "tmpalf", "iflg"):
        np.testing.assert_array_equal(getattr(back, name), getattr(tab, name))


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
