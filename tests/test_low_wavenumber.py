"""Pins for the oracle below 25 cm-1 -- CPU only.

At low wavenumber a line's negative-frequency partner is no longer
negligible.  The reference adds it twice over: RSYM folds the near-wing
grids' values at negative wavenumbers onto their mirror images when a
panel starts at or below 0 cm-1 (src/oprop.f90:501-505, 2145-2200), and
CONVF4 convolves each far-wing line at 0 < nu <= 25 a second time at -nu
(:4014-4024).  Together, with the radiation term nu tanh(nu/2kT) on the
grid, a line's optical depth is the Van Vleck - Weisskopf form

    S' [ (V(nu - nu0) - c) + (V(nu + nu0) - c) ] RADFN(nu),

c the value at 25 cm-1 that the far wing subtracts (each term within 25
cm-1 of its centre).  The Voigt profile V comes from scipy's Faddeeva
function, the strength and widths from the line's own fields at 296 K
and 1013.25 mb; grids that start on a point (the fold's weights reduce to
a copy) and between points are both checked, as are the first panel
starting exactly at 0 and the exact-Voigt build.  The fold reaches only
the grids' points from VFT = V1 - 32 DV up, so the lines here keep their
near-wing function (64 halfwidths) above VFT; beyond that the reference
itself drops the part below VFT.
"""
import math

import numpy as np
import pytest
from scipy.special import wofz

import oracle_lib as O
from lblrtm_amd import abi

P0, T0 = 1013.25, 296.0
RADCN2 = 1.4387752
M_H2O = 18.010565


def _line(nu0, s=1e-19, alfa=0.07, hwhm=0.35, n=0.75, epp=100.0, mol=1):
    f = lambda v, dt: np.array([v], dt)
    return abi.LineTable.from_lines(f(nu0, np.float64), f(s, np.float32), f(alfa, np.float32), f(epp, np.float32),
                                    f(mol + 100, np.int32), f(hwhm, np.float32), f(n, np.float32),
                                    f(0.0, np.float32))


def _od(tab, v1, v2, **kw):
    w, wb = 1e20, 1e24
    run = abi.make_run(v1, v2, dptmin=0.0, dptfac=0.0, **kw)
    dv, _ = O.plan_dv(run, [P0], [T0], [[w]], [wb])
    lay = abi.make_layer(P0, T0, [w], wb, dv[0], v1, v2)
    o, v1p, st, jrad = O.od_layer(run, lay, tab)
    return o, v1p + dv[0] * np.arange(len(o)), dv[0], jrad


def _vvw(nu, nu0, mirror=True, alfa=0.07, hwhm=0.35):
    w, wb = 1e20, 1e24
    al = float(np.float32(alfa)) * (1.0 - w / (w + wb)) + float(np.float32(hwhm)) * w / (w + wb)
    kb, c, amu = 1.380649e-23, 2.99792458e8, 1.66053906660e-27
    ad = nu0 / c * math.sqrt(2.0 * math.log(2.0) * kb * T0 / (M_H2O * amu))
    sig = ad / math.sqrt(2.0 * math.log(2.0))
    vp = lambda x: wofz((x + 1j * al) / (sig * math.sqrt(2.0))).real / (sig * math.sqrt(2.0 * math.pi))
    xkt = T0 / RADCN2
    s = float(np.float32(1e-19)) * w * (1.0 + math.exp(-nu0 / xkt))
    cut = al / (math.pi * (al * al + 625.0))
    f = (vp(nu - nu0) - cut) * (np.abs(nu - nu0) < 25.0)
    if mirror:
        f = f + (vp(nu + nu0) - cut) * (np.abs(nu + nu0) < 25.0)
    return s * f * nu * np.tanh(nu / (2.0 * xkt))


@pytest.mark.parametrize("voigt", [0, 1])
@pytest.mark.parametrize("nu0,v1,alfa", [(5.0, 0.2, 0.07), (5.0, 0.215, 0.07), (5.0, 0.0, 0.07), (1.5, 0.06, 0.02),
                                         (1.5, 0.0437, 0.02), (12.0, 0.05, 0.07), (5.0, 0.9, 0.07)])
def test_line_near_zero_is_van_vleck_weisskopf(nu0, v1, alfa, voigt):
    # the grid follows the line's width (ALFAL0), ~7 points per halfwidth
    o, nu, dv, jrad = _od(_line(nu0, alfa=alfa, hwhm=5 * alfa), v1, 40.0, voigt=voigt, alfal0=0.6 * alfa)
    assert jrad == 0
    assert (v1 - 32 * dv <= 0.0) == (v1 < 0.5)  # the fold acts in all but the last case
    if v1 < 0.5:
        assert nu0 - 64.0 * 1.01 * alfa >= v1 - 32 * dv  # the near wing lies above VFT
    want = _vvw(nu, nu0, alfa=alfa, hwhm=5 * alfa)
    big = np.abs(want) > 1e-3 * np.abs(want).max()
    rel = np.abs(o[big] - want[big]) / np.abs(want[big])
    assert rel.max() < 1e-2, rel.max()
    # where the mirror terms carry more than 5 % of the optical depth the
    # oracle follows them to 1 % (without them it would be off by > 5 %)
    bare = _vvw(nu, nu0, mirror=False, alfa=alfa, hwhm=5 * alfa)
    with np.errstate(divide="ignore", invalid="ignore"):
        mm = (np.abs(want) > 1e-6 * np.abs(want).max()) & ((want - bare) / want > 0.05)
    assert mm.sum() > 20
    assert (np.abs(o[mm] - want[mm]) / np.abs(want[mm])).max() < 1e-2


@pytest.mark.parametrize("frac", [0.3, 0.45, 0.7])
def test_fold_reaches_the_point_below_vft(frac):
    """VFT = -frac DVR3: on the coarsest grid the first point at or above 0
    is the second, half a step or more above 0 for frac < 0.5, where the
    fold's first pair reads R(0), the point below VFT (the reference's
    RR arrays extend below R(1)).  The line at 5 cm-1 keeps the VVW form,
    and the result does not depend on what ran before."""
    probe = abi.make_run(1.0, 40.0, alfal0=0.042)
    dv, _ = O.plan_dv(probe, [P0], [T0], [[1e20]], [1e24])
    v1 = (32.0 - 16.0 * frac) * dv[0]
    tab = _line(5.0)
    runs = [_od(tab, v1, 40.0, alfal0=0.042)[0] for _ in range(2)]
    _od(_line(5.0, s=3e-18), 0.0, 40.0, alfal0=0.042)  # something else in between
    runs.append(_od(tab, v1, 40.0, alfal0=0.042)[0])
    o, nu, dvv, _ = _od(tab, v1, 40.0, alfal0=0.042)
    assert dvv == dv[0] and abs(v1 - 32 * dvv + frac * 16 * dvv) < 1e-12
    for r in runs:
        np.testing.assert_array_equal(r, o)
    want = _vvw(nu, 5.0)
    big = np.abs(want) > 1e-3 * np.abs(want).max()
    assert (np.abs(o[big] - want[big]) / np.abs(want[big])).max() < 1e-2
