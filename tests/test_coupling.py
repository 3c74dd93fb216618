"""Pins for the oracle's line coupling (IFLG 1 and 3) -- CPU only.

A coupled line carries first-order coupling coefficients Y and G (IFLG 1)
or reduced-width coefficients GAMMA1 and GAMMA2 (IFLG 3) at 200, 250, 296
and 340 K in the record after it (src/oprop.f90:789-795, 1039-1061,
3118-3135).  LNCOR1 and LINF4 interpolate them linearly in temperature,
scale the strength by 1 + G (P/P0)^2, the width by 1 - GAMMA1 P/P0 -
GAMMA2 (P/P0)^2, and give the line the dispersive term Y (P/P0) (nu - nu')
/ alpha_V times its shape in CNVFNV and CONVF4 (:1322-1351, 3976-4004);
lines with IFLG != 0 are never rejected by LNCOR1 (:1153), coupled ones
never by CONVF4 (:3912), and SHRINK passes coupled lines through on their
own (:3550-3557).  What fixes these:

* piecewise-linear interpolation: the optical depth with tabulated
  coefficients equals the one with the interpolated (or end-interval
  extrapolated) value at every node, between nodes and beyond them;
* the dispersive term against the closed form: in the Lorentz limit the
  coupled line is Rosenkranz's profile S (alpha + Y P x) / (pi (alpha^2 +
  x^2)) with the reference's 25 cm-1 cut (the line minus its value at 25
  cm-1, the CO2 cut 2 - x^2/25^2 times it), times the radiation term;
* G scales the strength exactly; IFLG 3 is the same line with both widths
  scaled by the reduced-width factor;
* SHRINK's known grouping with coupled lines inside groups: the coupled
  line as it stands, then the group's sums over the other lines;
* no rejection: a line too weak for DPTMIN contributes when IFLG != 0 (the
  near wing alone for IFLG 2, the whole line when coupled).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import wofz

import oracle_lib as O
from lblrtm_amd import abi

P0, T0 = 1013.25, 296.0
RADCN2 = 1.4387752
TEMPLC = np.array([200.0, 250.0, 296.0, 340.0])
M_CO2 = 43.98983  # 12C16O2, atomic mass units


def table(lines, iflg, y, g):
    """lines: list of (nu, s, alfa, hwhm, n, epp, shift, mol)."""
    a = list(zip(*lines))
    n = len(lines)
    return abi.LineTable.from_lines_coupled(
        np.array(a[0], np.float64), np.array(a[1], np.float32), np.array(a[2], np.float32),
        np.array(a[5], np.float32), np.array(a[7], np.int32) + 100,
        np.array(a[3], np.float32), np.array(a[4], np.float32), np.array(a[6], np.float32),
        np.asarray(iflg, np.int32), np.broadcast_to(np.asarray(y, float), (n, 4)),
        np.broadcast_to(np.asarray(g, float), (n, 4)))


def od(tab, p, t, mol, v1, v2, w=1e20, wb=1e24, **kw):
    wk = [0.0] * 7
    wk[mol - 1] = w
    run = abi.make_run(v1, v2, **kw)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    o, v1p, st, jrad = O.od_layer(run, lay, tab)
    return o, v1p + dv[0] * np.arange(len(o)), jrad


LINE = (700.0, 1e-20, 0.07, 0.35, 0.75, 500.0, 0.0, 2)  # CO2


def f32(v):
    return np.asarray(v, np.float32).astype(np.float64)


def interp(c, t):
    """Piecewise-linear through (TEMPLC, c), the end intervals extended."""
    k = 0 if t < TEMPLC[1] else (1 if t < TEMPLC[2] else 2)
    return c[k] + (c[k + 1] - c[k]) * (t - TEMPLC[k]) / (TEMPLC[k + 1] - TEMPLC[k])


@pytest.mark.parametrize("t", [185.0, 200.0, 231.0, 250.0, 273.5, 296.0, 318.0, 340.0, 362.0])
@pytest.mark.parametrize("iflg", [1, 3])
def test_coefficients_interpolate_linearly_in_temperature(t, iflg):
    """Tabulated coefficients give the optical depth of the interpolated
    constant ones (LNCOR1's and LINF4's intervals both)."""
    y = [0.03125, -0.015625, 0.046875, 0.0078125]   # float32-exact nodes
    g = [0.25, -0.125, 0.0625, 0.375]
    if iflg == 3:
        y = [0.0625, 0.125, -0.03125, 0.09375]
        g = [-0.0625, 0.03125, 0.125, 0.0]
    p = 700.0
    o_tab, nu, _ = od(table([LINE], [iflg], y, g), p, t, 2, 660.0, 740.0, dptmin=0.0, dptfac=0.0)
    yc, gc = float(f32(interp(y, t))), float(f32(interp(g, t)))
    o_con, _, _ = od(table([LINE], [iflg], [yc] * 4, [gc] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0, dptfac=0.0)
    # float32 rounding of the constant moves the result by ~1e-8
    np.testing.assert_allclose(o_tab, o_con, rtol=1e-6, atol=1e-8 * np.abs(o_con).max())
    # a node off by one interval (the neighbouring interval's line) differs
    k = 0 if t < 250 else (1 if t < 296 else 2)
    k2 = 1 if k != 1 else 2
    yw = y[k2] + (y[k2 + 1] - y[k2]) * (t - TEMPLC[k2]) / (TEMPLC[k2 + 1] - TEMPLC[k2])
    gw = g[k2] + (g[k2 + 1] - g[k2]) * (t - TEMPLC[k2]) / (TEMPLC[k2 + 1] - TEMPLC[k2])
    if abs(yw - yc) + abs(gw - gc) > 1e-3:
        o_w, _, _ = od(table([LINE], [iflg], [yw] * 4, [gw] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0,
                       dptfac=0.0)
        assert np.abs(o_w - o_con).max() > 1e-4 * np.abs(o_con).max()


def voigt_profile(x, al, ad):
    sig = ad / math.sqrt(2.0 * math.log(2.0))
    z = (x + 1j * al) / (sig * math.sqrt(2.0))
    return wofz(z).real / (sig * math.sqrt(2.0 * math.pi))


@pytest.mark.parametrize("voigt", [0, 1])
@pytest.mark.parametrize("p,t,yv,gv", [(P0, T0, 0.05, 0.0), (P0, T0, -0.08, 0.3), (600.0, T0, 0.12, -0.4)])
def test_dispersive_term_closed_form(p, t, yv, gv, voigt):
    """The coupled minus the uncoupled optical depth (same G) is
    Y (P/P0) x / alpha_V times the reference's cut line: S' (V(x) -
    c(x) alpha_L / (pi (alpha_L^2 + 625))), c = 2 - x^2/625 for CO2 in the
    approximate build and 1 in the exact-Voigt build (CO2 only through
    CHI = 1, src/oprop_voigt.f90:3929-3971), times RADFN on the grid (JRAD
    = 0 below V1/kT = 5); at P0 the line is Lorentzian to 1 %, where this
    is Rosenkranz's profile.  G cancels: SPP = S Y P/P0 carries no G."""
    nu0, s, alfa, hwhm, n, epp, shift, mol = LINE
    kw = dict(dptmin=0.0, dptfac=0.0, voigt=voigt)
    o1, nu, jrad = od(table([LINE], [1], [yv] * 4, [gv] * 4), p, t, 2, 660.0, 740.0, **kw)
    o0, _, _ = od(table([LINE], [1], [0.0] * 4, [gv] * 4), p, t, 2, 660.0, 740.0, **kw)
    assert jrad == 0
    d = o1 - o0
    x = nu - nu0
    w, wb = 1e20, 1e24
    rhorat = (p / P0) * (T0 / t)
    rhoslf = rhorat * w / (w + wb)
    tr = (t / T0) ** float(f32(n))
    al = float(f32(alfa)) * tr * (rhorat - rhoslf) + float(f32(hwhm)) * tr * rhoslf
    kb, c, amu = 1.380649e-23, 2.99792458e8, 1.66053906660e-27
    ad = nu0 / c * math.sqrt(2.0 * math.log(2.0) * kb * t / (M_CO2 * amu))
    av = 0.5 * (al + math.sqrt(al * al + 4.0 * ad * ad))
    xkt = t / RADCN2
    sprime = float(f32(s)) * w * (1.0 + math.exp(-nu0 / xkt))  # SCOR = 1, E'' term = 1 at 296 K
    assert t == T0
    cut = (1.0 if voigt else (2.0 - x * x / 625.0)) * al / (math.pi * (al * al + 625.0))
    want = (yv * p / P0) * x / av * sprime * (voigt_profile(x, al, ad) - cut) * nu * np.tanh(nu / (2.0 * xkt))
    m = (np.abs(x) > 0.5 * al) & (np.abs(x) < 10.0)
    err = np.abs(d[m] - want[m]) / np.abs(want[m])
    # the tabulated decomposition follows the Voigt profile to ~1e-3 in the
    # core and ~1e-2 in the near wings, as for an uncoupled line
    assert err.max() < 1e-2, err.max()
    core = m & (np.abs(x) < 3.0 * al)
    assert (np.abs(d[core] - want[core]) / np.abs(want[core])).max() < 3e-3
    # odd about the centre: a sign or centre error fails
    assert np.all(np.sign(d[m]) == np.sign(yv * x[m]))
    # the alternatives a plausible mistake gives: Y without P/P0 at 600 mb,
    # Y over alpha_L^2 (the dispersion without the 1/alpha_V of the table
    # argument), 1 + G in the dispersive strength
    if p != P0:
        assert np.abs(d[m] - want[m] * P0 / p).max() > 0.3 * np.abs(want[m]).max()
    if gv != 0.0:
        f = 1.0 + gv * (p / P0) ** 2
        assert (np.abs(d[m] * f - want[m]) / np.abs(want[m])).max() > 0.05


@pytest.mark.parametrize("p,t", [(P0, T0), (450.0, 230.0)])
def test_g_scales_the_strength(p, t):
    """IFLG 1 with Y = 0 is the line with strength S (1 + G (P/P0)^2);
    IFLG 2 keeps LNCOR1 from rejecting it, like any IFLG != 0."""
    g = 0.5
    o_g, _, _ = od(table([LINE], [1], [0.0] * 4, [g] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0, dptfac=0.0)
    line2 = list(LINE)
    line2[1] = LINE[1] * (1.0 + g * (p / P0) ** 2)
    o_s, _, _ = od(table([tuple(line2)], [2], [0.0] * 4, [0.0] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0,
                   dptfac=0.0)
    np.testing.assert_allclose(o_g, o_s, rtol=2e-7, atol=1e-9 * np.abs(o_s).max())


@pytest.mark.parametrize("p,t", [(P0, T0), (300.0, 260.0)])
def test_iflg3_is_the_reduced_width_line(p, t):
    """IFLG 3 scales the Lorentz width (foreign and self parts alike) by
    1 - GAMMA1 P/P0 - GAMMA2 (P/P0)^2 in LNCOR1 and in LINF4."""
    g1, g2 = 0.15, -0.1
    f = 1.0 - g1 * (p / P0) - g2 * (p / P0) ** 2
    o3, _, _ = od(table([LINE], [3], [g1] * 4, [g2] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0, dptfac=0.0)
    line2 = list(LINE)
    line2[2] = LINE[2] * f
    line2[3] = LINE[3] * f
    o2, _, _ = od(table([tuple(line2)], [2], [0.0] * 4, [0.0] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0,
                  dptfac=0.0)
    # both widths stored as float32 after the scaling: ~1e-7
    np.testing.assert_allclose(o3, o2, rtol=2e-6, atol=1e-8 * np.abs(o2).max())
    o0, _, _ = od(table([LINE], [2], [0.0] * 4, [0.0] * 4), p, t, 2, 660.0, 740.0, dptmin=0.0, dptfac=0.0)
    assert np.abs(o3 - o0).max() > 1e-2 * np.abs(o0).max()


def test_shrink_passes_coupled_lines_through():
    """LINF4 + SHRINK at 296 K / 1013.25 mb on quadruplets 0.001 cm-1
    apart, 0.05 cm-1 between quadruplets, DVR4 = 0.064: after the first
    group (the whole first quadruplet) the groups are [q0] and [q1 q2 q3].
    q2 is coupled: it is emitted as it stands (SPP = SP Y) and then the
    group's sums over q1 (H2O) and q3 (CO2), one effective line each."""
    nq = 40
    lines, iflg = [], []
    rng = np.random.default_rng(5)
    for k in range(nq):
        a = 1040.0 + 0.05 * k
        for j in range(4):
            mol = 2 if j in (2, 3) else 1
            lines.append((a + 0.001 * j, float(10.0 ** rng.uniform(-22, -20)), 0.07, 0.3, 0.7, 300.0, 0.0, mol))
            iflg.append(1 if j == 2 else 0)
    yv = 0.03125
    tab = table(lines, iflg, [yv] * 4, [0.0] * 4)
    w = 1e20
    lay = abi.make_layer(P0, T0, [w, w], 1e24, 0.001, 1030.0, 1070.0)
    run = abi.make_run(1030.0, 1070.0)
    t9 = O.linf4(run, lay, tab)
    beta0 = RADCN2 / T0
    s_line = np.array([float(f32(l[1])) * w * (1.0 + math.exp(-l[0] * beta0)) for l in lines])
    x_line = np.array([l[0] for l in lines])
    # expected sequence: first quadruplet: q2 (coupled), sums {q0,q1}, {q3};
    # then per quadruplet: q0; q2 (coupled), {q1}, {q3}
    want = []
    q = [list(range(4 * k, 4 * k + 4)) for k in range(nq)]
    want += [("c", q[0][2]), ("s", [q[0][0], q[0][1]], 0), ("s", [q[0][3]], 2)]
    for k in range(1, nq):
        want += [("s", [q[k][0]], 0), ("c", q[k][2]), ("s", [q[k][1]], 0), ("s", [q[k][3]], 2)]
    assert len(t9["vnu"]) == len(want)
    for e, wv in enumerate(want):
        if wv[0] == "c":
            i = wv[1]
            assert t9["vnu"][e] == x_line[i]
            assert t9["mol"][e] == 2
            np.testing.assert_allclose(t9["sp"][e], s_line[i], rtol=1e-13)
            assert t9["spp"][e] == t9["sp"][e] * yv  # P/P0 = 1, G = 0
        else:
            ids, mol = wv[1], wv[2]
            ss = s_line[ids].sum()
            assert t9["mol"][e] == mol
            assert t9["spp"][e] == 0.0
            np.testing.assert_allclose(t9["sp"][e], ss, rtol=1e-13)
            np.testing.assert_allclose(t9["vnu"][e], (x_line[ids] * s_line[ids]).sum() / ss, rtol=1e-14)
    # the strength is conserved
    np.testing.assert_allclose(t9["sp"].sum(), s_line.sum(), rtol=1e-13)


def test_lines_with_iflg_are_not_rejected():
    """A line below DPTMIN: rejected everywhere with IFLG 0; with IFLG 2
    LNCOR1 keeps it but CONVF4 (SPP = 0) rejects its far wing, leaving the
    near-wing function alone (the line run without the far wing); coupled
    (IFLG 1, Y tiny) it keeps both."""
    weak = list(LINE)
    weak[1] = 1e-27
    weak = tuple(weak)
    kw = dict(dptmin=1e3, dptfac=0.0)
    o0, _, _ = od(table([weak], [0], [0.0] * 4, [0.0] * 4), P0, T0, 2, 660.0, 740.0, **kw)
    assert np.all(o0 == 0.0)
    o2, _, _ = od(table([weak], [2], [0.0] * 4, [0.0] * 4), P0, T0, 2, 660.0, 740.0, **kw)
    near, _, _ = od(table([weak], [0], [0.0] * 4, [0.0] * 4), P0, T0, 2, 660.0, 740.0, dptmin=0.0,
                    dptfac=0.0, ilblf4=0)
    np.testing.assert_allclose(o2, near, rtol=1e-12, atol=1e-15 * np.abs(near).max())
    o1, _, _ = od(table([weak], [1], [1e-7] * 4, [0.0] * 4), P0, T0, 2, 660.0, 740.0, **kw)
    full, _, _ = od(table([weak], [0], [0.0] * 4, [0.0] * 4), P0, T0, 2, 660.0, 740.0, dptmin=0.0, dptfac=0.0)
    np.testing.assert_allclose(o1, full, rtol=1e-5, atol=1e-6 * np.abs(full).max())
    # the far wing is there (5e-4 of the peak, 500 times the tolerance above)
    assert np.abs(full - near).max() > 2e-4 * np.abs(full).max()


def test_refusals():
    """A coupled line without its coefficient record is refused."""
    lines = [LINE, (701.0,) + LINE[1:]]
    tab = table(lines, [0, 0], [0.0] * 4, [0.0] * 4)
    tab.iflg[0] = 1  # claims coefficients, but the next record is a line
    with pytest.raises(RuntimeError, match="rc=-5"):
        od(tab, P0, T0, 2, 660.0, 740.0)
