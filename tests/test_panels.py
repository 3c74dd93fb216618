"""The panel files of the reference (TAPE10, TAPE11/TAPE12) from the C++
writer, and the Fortran boundary's layout -- CPU only.

* lblx_od_panels gives the oracle's PANEL headers (V1P, V2P bit for bit,
  NLIM) for the single-layer workload and layers of the column.
* lblx_write_panels' records read back with a reader written from the
  /BUFPNL/ layout and BUFOUT's record form (tape.read_panels) for both
  word sizes, with and without the transmittance record, appended, and
  with ENDFIL's sentinel record.
* The bind(c) types of lblrtm_amd/fortran/lblx_c.f90, compiled by flang
  with the reference's double-precision flags, have the offsets and sizes
  the C compiler gives include/lblx.h, and the host-only entry points
  called through the Fortran interfaces give the C results.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth, tape

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FBIN = os.path.join(ROOT, "tests", "fortran", "_build")


def _column_layers(v1, v2):
    atm = synth.atmosphere()
    run = abi.make_run(v1, v2)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    return run, [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                                v1, v2) for l in range(len(dv))]


def test_panel_headers_equal_oracle():
    tab = synth.synth_lines(2000, 600.0, 700.0, [1, 2], 4)
    p, t, wk, wb = synth.config0_layer()
    cases = []
    run = abi.make_run(610.0, 690.0)
    dv, _ = lib.plan_dv(run, [p], [t], [wk], [wb])
    cases.append((run, abi.make_layer(p, t, wk, wb, dv[0], 610.0, 690.0)))
    run1, lays = _column_layers(610.0, 690.0)
    cases += [(run1, lays[l]) for l in (0, 25, 49)]
    for run, lay in cases:
        O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
        ov1, ov2, onl = O.panels()
        v1, v2, nl = lib.od_panels(run, abi.Layer.from_buffer_copy(lay))
        np.testing.assert_array_equal(v1, ov1)
        np.testing.assert_array_equal(v2, ov2)
        np.testing.assert_array_equal(nl, onl)
        assert int(nl.sum()) == lib.od_size(run, lay)


@pytest.mark.parametrize("word", [8, 4])
def test_panel_writer_round_trip(tmp_path, word):
    run, lays = _column_layers(610.0, 690.0)
    lay = lays[40]
    v1, v2, nl = lib.od_panels(run, lay)
    n = int(nl.sum())
    rng = np.random.default_rng(1)
    od = rng.exponential(1.0, n)
    tr = np.exp(-od)
    hdr = bytes(range(256)) * 5 + bytes(136)  # 177 eight-byte words of some header
    assert len(hdr) == 177 * 8
    p10 = tmp_path / "TAPE10"
    lib.write_panels(p10, v1, v2, lay.dv, nl, od, filhdr=hdr[:177 * word], word=word)
    back = tape.read_panels(p10, word=word, nvals=1, filhdr_words=177)
    assert back["filhdr"] == hdr[:177 * word] and not back["endfil"]
    np.testing.assert_array_equal(back["v1p"], v1)
    np.testing.assert_array_equal(back["v2p"], v2)
    np.testing.assert_array_equal(back["nlim"], nl)
    want = od if word == 8 else od.astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(back["values"][0], want)
    assert np.all(back["dvp"] == (lay.dv if word == 8 else np.float32(lay.dv)))
    # TAPE12: radiance and transmittance records, written in two calls (the
    # second appends) and closed by ENDFIL
    p12 = tmp_path / "TAPE12"
    k = len(nl) // 2
    o = int(nl[:k].sum())
    lib.write_panels(p12, v1[:k], v2[:k], lay.dv, nl[:k], od[:o], tr[:o], filhdr=hdr[:177 * word], word=word)
    lib.write_panels(p12, v1[k:], v2[k:], lay.dv, nl[k:], od[o:], tr[o:], word=word, endfil=True, append=True)
    back = tape.read_panels(p12, word=word, nvals=2, filhdr_words=177)
    assert back["endfil"]
    np.testing.assert_array_equal(back["nlim"], nl)
    w2 = tr if word == 8 else tr.astype(np.float32).astype(np.float64)
    np.testing.assert_array_equal(back["values"][0], want)
    np.testing.assert_array_equal(back["values"][1], w2)
    # the record layout itself: 4-byte markers around each record, panel
    # header V1P V2P (f64), DVP, NLIM
    raw = open(p10, "rb").read()
    assert struct.unpack_from("<i", raw, 0)[0] == 177 * word
    h = 8 + 177 * word
    assert struct.unpack_from("<i", raw, h)[0] == 16 + 2 * word
    assert struct.unpack_from("<dd", raw, h + 4) == (v1[0], v2[0])
    assert struct.unpack_from("<q" if word == 8 else "<i", raw, h + 4 + 16 + word)[0] == nl[0]


@pytest.mark.skipif(not os.path.exists(os.path.join(FBIN, "abi_layout_f")), reason="flang drivers not built")
def test_fortran_interface_layout_and_host_calls():
    f = subprocess.run([os.path.join(FBIN, "abi_layout_f")], capture_output=True, text=True, check=True).stdout
    c = subprocess.run([os.path.join(FBIN, "abi_layout_c")], capture_output=True, text=True, check=True).stdout
    fl = [ln.split() for ln in f.splitlines()]
    lay = [" ".join(x) for x in fl if x[0] in ("off", "size")]
    assert lay == c.split("\n")[:-1] and len(lay) > 30
    out = {x[0]: x[1:] for x in fl if x[0] not in ("off", "size")}
    run = abi.make_run(600.0, 1100.0)
    dv, _ = lib.plan_dv(run, [500.0], [250.0], [[1e22]], [1.448766e24])
    assert out["plan_dv"][0] == "0" and float(out["plan_dv"][1]) == dv[0]
    L = abi.make_layer(500.0, 250.0, [1e22], 1.448766e24, dv[0], 600.0, 1100.0)
    assert int(out["od_size"][1]) == lib.od_size(run, L)
    assert int(out["od_panels"][1]) == len(lib.od_panels(run, L)[0])
    assert " ".join(out["strerror"]) == lib.strerror(abi.E_UNSUPPORTED)
