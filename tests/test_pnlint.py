"""Pins for the oracle's PNLINT (DVOUT regrid, src/oprop.f90:1564-1806) --
CPU only.

With DVOUT = DV/2 every other output point coincides with an input point
(phase 0: the Lagrange weights are 0, 1, 0, 0), so the output there is the
layer's own optical depth exactly; the points between are the cubic through
their four neighbours, which a smooth (continuum-only) optical depth fixes
to within the interpolation error; DVOUT = DV passes the panels through.
Near the two ends the taps repeat the end points.  The output grid starts at the layer's first point, panels hold 2400 points
and the last ends within one DVOUT above V2.
"""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, synth


def _layer(v1, v2, table, icntnm=0):
    p, t, wk, wb = synth.config0_layer()
    run = abi.make_run(v1, v2, icntnm=icntnm)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    od, v1p, _, _ = O.od_layer(run, abi.Layer.from_buffer_copy(lay), table)
    return run, lay, od, v1p


def _grid_checks(pan, v1p, dvout, v2):
    v1o, v2o, nl = pan
    assert v1o[0] == v1p
    assert np.all(nl[:-1] == 2400)
    np.testing.assert_allclose(np.diff(v1o), 2400 * dvout, rtol=1e-9)
    assert v2 <= v2o[-1] < v2 + dvout * 1.000001


def test_half_dv_hits_input_points():
    tab = synth.config0_lines()
    run, lay, od, v1p = _layer(700.0, 720.0, tab)
    dvout = lay.dv / 2
    out, pan = O.od_layer_dvout(run, lay, tab, dvout)
    _grid_checks(pan, v1p, dvout, 720.0)
    n = min(len(out) // 2, len(od) - 1)
    np.testing.assert_array_equal(out[0:2 * n:2], od[:n])


def test_smooth_od_interpolates():
    # continuum only: lines far outside the band
    one = lambda v, dt: np.array([v], dt)
    tab = abi.LineTable.from_lines(one(3000.0, np.float64), one(1e-22, np.float32),
                                   one(0.07, np.float32), one(100.0, np.float32),
                                   one(101, np.int32), one(0.2, np.float32),
                                   one(0.7, np.float32), one(0.0, np.float32))
    run, lay, od, v1p = _layer(1000.0, 1030.0, tab, icntnm=1)
    assert od.max() > 0
    v = v1p + lay.dv * np.arange(len(od))
    for frac in (2.5, 3.7, 0.999 * 4):
        dvout = lay.dv / frac
        out, pan = O.od_layer_dvout(run, lay, tab, dvout)
        _grid_checks(pan, v1p, dvout, 1030.0)
        vo = v1p + dvout * np.arange(len(out))
        ref = np.interp(vo, v, od)
        # away from the first and last input intervals, where PNLINT repeats
        # the end points as taps
        inside = (vo > v[1]) & (vo < v[-2])
        np.testing.assert_allclose(out[inside], ref[inside], rtol=2e-7)


def test_dvout_equal_dv_passes_panels():
    tab = synth.config0_lines()
    run, lay, od, v1p = _layer(700.0, 710.0, tab)
    out, pan = O.od_layer_dvout(run, lay, tab, lay.dv)
    np.testing.assert_array_equal(out, od)
