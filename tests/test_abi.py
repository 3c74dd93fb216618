"""The C-ABI library loads, exports what include/lblx.h declares, and its
host-side planner agrees with the oracle (no device work)."""
import ctypes as C
import os
import re

import numpy as np

import oracle_lib as O
from lblrtm_amd import abi, lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lblx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lblx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_error_strings():
    assert lib.strerror(0) == "ok"
    assert "NREC" in lib.strerror(abi.E_NREC)


def test_planner_sizes_match_oracle():
    run = abi.make_run(500.0, 2500.0)
    atm = synth.atmosphere()
    dv, it = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    dvo, ito = O.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    np.testing.assert_array_equal(dv, dvo)
    np.testing.assert_array_equal(it, ito)
    one = lambda v, dt: np.array([v], dt)
    tab = abi.LineTable.from_lines(one(1000.0, np.float64), one(1e-24, np.float32),
                                   one(0.07, np.float32), one(100.0, np.float32),
                                   one(101, np.int32), one(0.2, np.float32),
                                   one(0.7, np.float32), one(0.0, np.float32))
    for l in (0, 20, 49):
        lay = abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l],
                             dv[l], 500.0, 2500.0)
        od, _, st, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
        assert lib.od_size(run, lay) == len(od)


def test_no_device_is_reported_cleanly():
    if lib.load().lblx_device_count() > 0:
        return
    h = C.c_void_p()
    assert lib.load().lblx_create(C.byref(h), 0) == abi.E_NODEV
