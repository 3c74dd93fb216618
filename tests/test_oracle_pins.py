"""Pins of the oracle away from the reference state (296 K, 1013.25 mb).

At the reference state every temperature and pressure correction of
LNCOR1/LINF4 and MOLEC is the identity, so the single-line and table tests
in test_oracle.py cannot see a wrong temperature law.  These tests check:

* the partition sums: Q(T) at every node of the reference's TIPS tables,
  read by this file from the DATA statements of src/oprop.f90 (QT_H2O ..
  QT_O2), and SCOR = Q(296)/Q(T) between nodes against a 4-point Lagrange
  fit made here (numpy), for every isotopologue of molecules 1-7;
* one line at 220 K / 300 mb and 250 K / 500 mb with a pressure shift,
  self broadening, a width temperature exponent and a lower-state energy,
  against the Voigt profile (scipy's Faddeeva function) whose centre,
  Lorentz and Doppler widths and strength this file computes from the
  line-correction formulas of src/oprop.f90:965-993, 1068-1145 and
  MOLEC :2476-2557 with CODATA constants and the isotope masses read from
  the reference's source text.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest
from scipy.special import wofz

import oracle_lib as O
from lblrtm_amd import abi

REF = "/root/reference/src/oprop.f90"
ref_src = pytest.mark.skipif(not os.path.exists(REF), reason="reference sources absent")

ISO_MAX = [6, 10, 9, 5, 6, 4, 3]
QT_NAMES = ["QT_H2O", "QT_CO2", "QT_O3", "QT_N2O", "QT_CO", "QT_CH4", "QT_O2"]
RADCN2 = 1.4387752
P0, T0 = 1013.25, 296.0


def _num(tok):
    return float(tok.replace("D", "E").replace("d", "e"))


def _data_values(text):
    """Numbers of a Fortran DATA value list (continuation ampersands,
    comments and line breaks removed)."""
    text = re.sub(r"!.*", "", text)
    text = text.replace("&", " ")
    return [_num(t) for t in re.findall(r"[-+]?\d*\.\d*(?:[EeDd][-+]?\d+)?|[-+]?\d+", text)]


def reference_tips():
    """{mol: array [niso, 119]} and the 119 node temperatures, parsed from
    the QT_* subroutines and BD_TDAT of src/oprop.f90."""
    src = open(REF).read()
    m = re.search(r"data\s+Tdat\s*/(.*?)/", src, re.S | re.I)
    tdat = np.array(_data_values(m.group(1)))
    assert len(tdat) == 119
    out = {}
    for mol, name in enumerate(QT_NAMES, start=1):
        start = re.search(r"Subroutine\s+%s\b" % name, src, re.I).start()
        end = re.search(r"end\s+subroutine\s+%s\b" % name, src[start:], re.I).end() + start
        body = src[start:end]
        rows = {}
        for mm in re.finditer(r"data\s*\(QofT\(\s*(\d+)\s*,J\),J=1,119\)\s*/(.*?)/", body, re.S | re.I):
            vals = _data_values(mm.group(2))
            assert len(vals) == 119, (name, mm.group(1), len(vals))
            rows[int(mm.group(1))] = vals
        out[mol] = np.array([rows[k] for k in sorted(rows)])
    return tdat, out


def reference_masses():
    """Isotope masses SMASS(m, iso), m = 1..7, from BLOCK DATA Isotop."""
    src = open(REF).read()
    out = {}
    for mol in range(1, 8):
        m = re.search(r"data\s*\(smass\(%d,i\),i=1,(\d+)\)\s*(.*?)/(.*?)/" % mol, src, re.S | re.I)
        vals = _data_values(m.group(3))
        assert len(vals) == int(m.group(1))
        out[mol] = vals
    return out


def lagrange4(tdat, q, t):
    """Q(t) from the 4 nodes around t (the first node >= t is the third of
    them; three nodes at the table ends), fitted here with numpy."""
    i = int(np.searchsorted(tdat, t, side="left"))  # first node >= t (0-based)
    if i < 2:
        idx = [0, 1, 2]
    elif i == len(tdat) - 1:
        idx = [i - 2, i - 1, i]
    else:
        idx = [i - 2, i - 1, i, i + 1]
    x = (tdat[idx] - t) / 25.0
    coef = np.polyfit(x, q[idx], len(idx) - 1)
    return float(np.polyval(coef, 0.0))


def oracle_scor(t):
    sc = np.zeros(70)
    O.lib().tips_scor(t, abi.ptr(sc, C.c_double))
    return sc.reshape(7, 10)


@ref_src
def test_tips_tables_equal_reference_text():
    """Q(T_node) implied by the oracle's SCOR at every node in 70-3000 K
    equals the reference's QofT DATA value there, for every isotopologue
    of molecules 1-7 (a mis-extracted table row fails here)."""
    tdat, q = reference_tips()
    for mol in range(1, 8):
        assert q[mol].shape[0] >= ISO_MAX[mol - 1]
    q296 = {(m, i): lagrange4(tdat, q[m][i - 1], 296.0)
            for m in range(1, 8) for i in range(1, ISO_MAX[m - 1] + 1)}
    nodes = [j for j, t in enumerate(tdat) if 70.0 <= t <= 3000.0]
    assert len(nodes) == 117
    for j in nodes:
        sc = oracle_scor(float(tdat[j]))
        for m in range(1, 8):
            for i in range(1, ISO_MAX[m - 1] + 1):
                q_or = q296[(m, i)] / sc[m - 1, i - 1]
                assert abs(q_or / q[m][i - 1][j] - 1.0) < 1e-11, (m, i, tdat[j])


@ref_src
@pytest.mark.parametrize("t", [75.0, 97.5, 186.3, 220.0, 250.0, 272.5, 296.0, 333.3, 1500.7, 2990.0])
def test_tips_between_nodes_is_four_point_lagrange(t):
    """SCOR between nodes: Q(296)/Q(T), both by the 4-point Lagrange fit
    (3-point in the first and last intervals) of the reference's tables."""
    tdat, q = reference_tips()
    sc = oracle_scor(t)
    for m in range(1, 8):
        for i in range(1, ISO_MAX[m - 1] + 1):
            want = lagrange4(tdat, q[m][i - 1], 296.0) / lagrange4(tdat, q[m][i - 1], t)
            assert abs(sc[m - 1, i - 1] / want - 1.0) < 1e-11, (m, i, t)


# one line per case: (mol, T, P, nu0, S, alfa, hwhm, n, E'', shift, W, WBROAD)
LINE_CASES = [
    (1, 220.0, 300.0, 850.0, 1e-21, 0.07, 0.35, 0.65, 800.0, 0.004, 2.0e22, 4.0e23),
    (1, 250.0, 500.0, 850.0, 1e-21, 0.06, 0.24, 0.75, 1500.0, -0.003, 3.0e22, 6.0e23),
    (2, 260.0, 700.0, 700.0, 3e-21, 0.08, 0.16, 0.72, 300.0, -0.002, 4.0e21, 1.2e24),
]


def _one_line(mol, nu0, s, alfa, hwhm, n, epp, shift):
    one = lambda v, dt: np.array([v], dt)
    return abi.LineTable.from_lines(one(nu0, np.float64), one(s, np.float32), one(alfa, np.float32),
                                    one(epp, np.float32), one(mol + 100, np.int32), one(hwhm, np.float32),
                                    one(n, np.float32), one(shift, np.float32))


@ref_src
@pytest.mark.parametrize("case", LINE_CASES, ids=["H2O-220K-300mb", "H2O-250K-500mb", "CO2-260K-700mb"])
def test_single_line_away_from_reference_state(case):
    """Optical depth of one line against S(T) x Voigt(nu - nu'; aL, aD),
    every factor computed here.  The line's own stored fields are float32,
    so the test uses the float32 values."""
    mol, t, p, nu0, s, alfa, hwhm, n, epp, shift, w, wb = case
    tab = _one_line(mol, nu0, s, alfa, hwhm, n, epp, shift)
    s, alfa, hwhm, n, epp, shift = (float(np.float32(x)) for x in (s, alfa, hwhm, n, epp, shift))
    v1, v2 = nu0 - 40.0, nu0 + 40.0
    wk = [0.0] * 7
    wk[mol - 1] = w
    run = abi.make_run(v1, v2, dptmin=0.0, dptfac=0.0)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    od, v1p, st, jrad = O.od_layer(run, lay, tab)
    nu = v1p + dv[0] * np.arange(len(od))
    # LNCOR1 (src/oprop.f90:965-993, 1068-1145) and MOLEC (:2525-2552)
    rhorat = (p / P0) * (T0 / t)
    tratio = t / T0
    wtot = sum(wk) + wb
    rhoslf = rhorat * w / wtot
    tmpcor = tratio ** n
    al = alfa * tmpcor * (rhorat - rhoslf) + hwhm * tmpcor * rhoslf
    nup = nu0 + rhorat * shift
    mass = reference_masses()[mol][0]
    kb, c, amu = 1.380649e-23, 2.99792458e8, 1.66053906660e-27
    ad = nup / c * math.sqrt(2.0 * math.log(2.0) * kb * t / (mass * amu))
    tdat, q = reference_tips()
    scor = lagrange4(tdat, q[mol][0], 296.0) / lagrange4(tdat, q[mol][0], t)
    xkt, xkt0 = t / RADCN2, T0 / RADCN2
    strength = s * w * scor * math.exp(-epp * (1.0 / xkt - 1.0 / xkt0)) * (1.0 + math.exp(-nup / xkt))
    radfn = lambda v: v * np.tanh(v / (2.0 * xkt))
    # OPDPTH sets JRAD = 0 when V1/kT < 5: the radiation term then
    # multiplies the optical depth on the grid, otherwise the strength
    # (at the unshifted centre, :1021-1030)
    assert jrad == (0 if v1 / xkt < 5.0 else 1)
    if jrad == 1:
        strength *= float(radfn(nu0))
        grid = np.ones_like(nu)
    else:
        grid = radfn(nu)
    sig = ad / math.sqrt(2.0 * math.log(2.0))
    z = ((nu - nup) + 1j * al) / (sig * math.sqrt(2.0))
    expect = strength * wofz(z).real / (sig * math.sqrt(2.0 * math.pi)) * grid
    av = 0.5 * (al + math.sqrt(al * al + 4.0 * ad * ad))
    # the decomposition's own error is ~1e-3 in the core; the alternatives
    # a plausible mistake gives (width law inverted, shift sign, self
    # broadening dropped, E'' or partition sum off) move the core by more
    # than 5 %
    core = np.abs(nu - nup) < 2.0 * av
    rel = np.abs(od[core] - expect[core]) / expect[core]
    assert rel.max() < 3e-3, rel.max()
    c0 = int(np.argmin(np.abs(nu - nup)))
    assert abs(od[c0] / expect[c0] - 1.0) < 1e-3
    # the peak sits at the shifted centre, to a grid step
    assert abs(nu[int(np.argmax(od))] - nup) <= dv[0]
    wing = (np.abs(nu - nup) > 2.0 * av) & (np.abs(nu - nup) < 10.0 * av)
    assert (np.abs(od[wing] - expect[wing]) / expect[wing]).max() < 1.5e-2


@ref_src
def test_single_line_pins_detect_plausible_mistakes():
    """The expectations above separate the oracle from plausible mistakes:
    each variant of the expected core (width law inverted, shift sign
    flipped, self broadening dropped, E'' dropped) differs from the
    oracle's optical depth by far more than the tolerance."""
    mol, t, p, nu0, s, alfa, hwhm, n, epp, shift, w, wb = LINE_CASES[0]
    tab = _one_line(mol, nu0, s, alfa, hwhm, n, epp, shift)
    s, alfa, hwhm, n, epp, shift = (float(np.float32(x)) for x in (s, alfa, hwhm, n, epp, shift))
    v1, v2 = nu0 - 40.0, nu0 + 40.0
    wk = [w, 0, 0, 0, 0, 0, 0]
    run = abi.make_run(v1, v2, dptmin=0.0, dptfac=0.0)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    od, v1p, st, jrad = O.od_layer(run, abi.make_layer(p, t, wk, wb, dv[0], v1, v2), tab)
    nu = v1p + dv[0] * np.arange(len(od))
    tdat, q = reference_tips()
    mass = reference_masses()[mol][0]
    xkt, xkt0 = t / RADCN2, T0 / RADCN2

    def model(width_exp=1.0, shift_sign=1.0, self_on=1.0, epp_on=1.0):
        rhorat = (p / P0) * (T0 / t)
        rhoslf = self_on * rhorat * w / (sum(wk) + wb)
        tmpcor = (t / T0) ** (width_exp * n)
        al = alfa * tmpcor * (rhorat - rhoslf) + hwhm * tmpcor * rhoslf
        nup = nu0 + shift_sign * rhorat * shift
        ad = nup / 2.99792458e8 * math.sqrt(2 * math.log(2) * 1.380649e-23 * t / (mass * 1.66053906660e-27))
        scor = lagrange4(tdat, q[mol][0], 296.0) / lagrange4(tdat, q[mol][0], t)
        sk = s * w * scor * math.exp(-epp_on * epp * (1 / xkt - 1 / xkt0)) * (1 + math.exp(-nup / xkt))
        sk *= nu0 * math.tanh(nu0 / (2 * xkt))
        sig = ad / math.sqrt(2 * math.log(2))
        zz = ((nu - nup) + 1j * al) / (sig * math.sqrt(2))
        return sk * wofz(zz).real / (sig * math.sqrt(2 * math.pi)), nup, al, ad

    good, nup, al, ad = model()
    core = np.abs(nu - nup) < 2.0 * 0.5 * (al + math.sqrt(al * al + 4 * ad * ad))
    err = lambda e: float(np.max(np.abs(od[core] - e[core]) / e[core]))
    assert err(good) < 3e-3
    for bad in (dict(width_exp=-1.0), dict(shift_sign=-1.0), dict(self_on=0.0), dict(epp_on=0.0)):
        assert err(model(**bad)[0]) > 5e-2, bad


@ref_src
def test_linf4_line_correction_away_from_reference_state():
    """LINF4's correction of the far-wing lines (src/oprop.f90:3050-3230) at
    230 K / 400 mb: lines 1.3 cm-1 apart, so every SHRINK group holds one
    line and TAPE9 carries each corrected line itself: centre
    nu + (P/P0)(T0/T) shift, Lorentz width with the (T/T0)^n law and self
    broadening, Doppler width nu' ALFD1 (no radiation term in the
    strength, jrad4 = 0), strength S W SCOR exp(-E'' (c2/T - c2/T0))
    (1 + exp(-c2 nu'/T))."""
    t, p = 230.0, 400.0
    n = 60
    rng = np.random.Generator(np.random.PCG64(3))
    vnu = 700.0 + 1.3 * np.arange(n)
    mol = np.where(np.arange(n) % 3 == 0, 2, 1)
    f32 = lambda a: np.asarray(a, np.float32)
    sp = f32(10.0 ** rng.uniform(-23, -20, n))
    alfa = f32(rng.uniform(0.04, 0.1, n))
    hwhm = f32(alfa * rng.uniform(2, 5, n))
    tmp = f32(rng.uniform(0.5, 0.8, n))
    epp = f32(rng.uniform(0, 2500, n))
    sh = f32(rng.uniform(-0.005, 0.005, n))
    tab = abi.LineTable.from_lines(vnu, sp, alfa, epp, mol + 100, hwhm, tmp, sh)
    wk = [2e22, 6e21, 0, 0, 0, 0, 0]
    wb = 1e24
    v1, v2 = 730.0, 740.0
    run = abi.make_run(v1, v2)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    cap = 1000
    arrs = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.int32)]
    nb = C.c_int(0)
    bc = np.zeros(100, np.int32)
    nt9 = O.lib().oracle_linf4(C.byref(run), C.byref(abi.Layer.from_buffer_copy(lay)), C.byref(tab.struct()),
                               *[abi.ptr(a, C.c_double) for a in arrs[:4]], abi.ptr(arrs[4], C.c_int32),
                               abi.ptr(bc, C.c_int32), cap, 100, C.byref(nb))
    g_vnu, g_sp, g_al, g_ad = (a[:nt9] for a in arrs[:4])
    assert nt9 > 30
    tdat, q = reference_tips()
    masses = reference_masses()
    rhorat = (p / P0) * (T0 / t)
    wtot = sum(wk) + wb
    c2 = RADCN2
    kb, c, amu = 1.380649e-23, 2.99792458e8, 1.66053906660e-27
    # TAPE9 lines in order: match each to its source line by centre
    for k in range(nt9):
        i = int(np.argmin(np.abs(vnu - g_vnu[k])))
        m = int(mol[i])
        nup = vnu[i] + rhorat * float(sh[i])
        rhoslf = rhorat * wk[m - 1] / wtot
        tc = (t / T0) ** float(tmp[i])
        al = float(alfa[i]) * tc * (rhorat - rhoslf) + float(hwhm[i]) * tc * rhoslf
        ad = nup / c * math.sqrt(2 * math.log(2) * kb * t / (masses[m][0] * amu))
        scor = lagrange4(tdat, q[m][0], 296.0) / lagrange4(tdat, q[m][0], t)
        s = float(sp[i]) * wk[m - 1] * scor * math.exp(-float(epp[i]) * (c2 / t - c2 / T0)) * \
            (1 + math.exp(-nup * c2 / t))
        assert abs(g_vnu[k] - nup) < 1e-9, k
        assert abs(g_al[k] / al - 1) < 1e-10, k
        assert abs(g_ad[k] / ad - 1) < 1e-5, k
        assert abs(g_sp[k] / s - 1) < 1e-10, k
