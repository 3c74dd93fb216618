"""Pins for the oracle's continuum (CONTNM, MT_CKD H2O, CO2, O2 and N2
fundamentals, Rayleigh) -- CPU only.

XINT, which carries every continuum from its native grid onto the 1 cm-1
ABSRB grid, is Keys' cubic convolution with a = -1/2; the tests evaluate
that kernel in its textbook form, so the interpolation weights and the
coarse-to-fine index mapping are both pinned.  At the temperatures where
each continuum's temperature law is the identity (296 K for MT_CKD and O2,
246 K for the CO2 band heads, 272 K for N2) the coarse values are the
tabulated ones times closed-form amounts and the radiation term
v tanh(v / 2kT).  Scale-factor linearity, the ICNTNM codes, Rayleigh's
closed form and the continuum-only optical depth close the set.
"""
import ctypes as C
import os

import numpy as np
import pytest

import contdata
import oracle_lib as O
from lblrtm_amd import abi, synth

RADCN2 = 1.4387752
XLOSMT = 2.68675e19
T = contdata.tables()


def keys(x):
    """Keys (1981) cubic-convolution kernel, a = -1/2."""
    x = np.abs(x)
    return np.where(x <= 1, 1.5 * x ** 3 - 2.5 * x ** 2 + 1,
                    np.where(x < 2, -0.5 * x ** 3 + 2.5 * x ** 2 - 4 * x + 2, 0.0))


def interp(vk, ck, v):
    dv = vk[1] - vk[0]
    return np.array([np.sum(ck * keys((x - vk) / dv)) for x in v])


def radterm(v, t):
    return v * np.tanh(v * RADCN2 / (2.0 * t))


def radfn(v, t):
    """RADFN with its branches (src/oprop.f90:2235-2296): x/2 v below
    x = 0.01, v tanh(x/2) up to x = 10, v beyond (x = v/kT)."""
    v = np.asarray(v, np.float64)
    x = v * RADCN2 / t
    e = np.exp(-np.minimum(x, 10.0))
    return np.where(x <= 0.01, 0.5 * x * v, np.where(x <= 10.0, v * (1.0 - e) / (1.0 + e), v))


def layer(v1, v2, t, p=1013.0, wk=(1e22, 4e21, 0.0, 0.0, 0.0, 0.0, 5e23), wb=2e24):
    return abi.make_layer(p, t, list(wk), wb, 0.001, v1, v2)


def run_scaled(v1, v2, scl):
    return abi.make_run(v1, v2, icntnm=6, cntscl=scl)


def one(k):
    s = [0.0] * 7
    s[k] = 1.0
    return s


def absrb(run, lay, jrad=1):
    a, v1abs, dvabs = O.contnm(run, lay, jrad)
    return a, v1abs + dvabs * np.arange(len(a))


def test_absrb_grid_of_opdpth():
    for (v1, v2), (a, b, n) in {(500.0, 2500.0): (497.0, 2503.0, 2007),
                                (600.3, 1100.7): (597.0, 1104.0, 508)}.items():
        x, y, d = C.c_double(), C.c_double(), C.c_double()
        m = C.c_int()
        assert O.lib().oracle_cont_grid(C.byref(layer(v1, v2, 250.0)), C.byref(x), C.byref(y),
                                        C.byref(d), C.byref(m)) == 0
        assert (x.value, y.value, d.value, m.value) == (a, b, 1.0, n)


@pytest.mark.skipif(not os.path.isdir("/root/reference/data"), reason="reference sources absent")
def test_tables_match_reference_files():
    from scipy.io import netcdf_file
    f = netcdf_file("/root/reference/data/absco-ref_wv-mt-ckd.nc", "r", mmap=False)
    for k, name in [("wavenumbers", "LBLX_MTCKD_WVN"), ("self_absco_ref", "LBLX_MTCKD_SELF"),
                    ("for_absco_ref", "LBLX_MTCKD_FOR"), ("self_texp", "LBLX_MTCKD_TEXP"),
                    ("for_closure_absco_ref", "LBLX_MTCKD_FORCL")]:
        np.testing.assert_array_equal(T[name], f.variables[k].data.astype(float))
    f.close()
    # values read off src/contnm.f90 by eye
    assert T["LBLX_FCO2"][-1] == 1.887e-19 and len(T["LBLX_FCO2"]) == 5003
    assert list(T["LBLX_XFACCO2"][:3]) == [1.0, 0.9998, 0.9997] and T["LBLX_XFACCO2"][-1] == 1.0
    assert T["LBLX_N2F_272"][1] == 4.691e-11 and list(T["LBLX_N2F_HDR"]) == [1997.784896, 2901.576661,
                                                                              3.981461525, 228.0]
    assert T["LBLX_O2F_XO2"][1] == 9.744e-09 and T["LBLX_O2F_XO2T"][5] == 467.0
    assert list(T["LBLX_CO2_TDEP"][:2]) == [0.144, 0.361]
    assert list(T["LBLX_N2R_HDR"]) == [-10.0, 350.0, 5.0, 73.0]
    assert T["LBLX_N2R_C296"][2] == 0.4979e-06 and T["LBLX_N2R_SF296"][-1] == 1.6157
    assert T["LBLX_N2R_SF220"][1] == 1.3515


def _check(got, v, want, lo, hi, rtol):
    m = (v >= lo) & (v <= hi)
    np.testing.assert_allclose(got[m], want[m], rtol=rtol, atol=0)


def test_h2o_self_and_foreign_at_reference_state():
    lay = layer(500.0, 2500.0, 296.0)
    wk1, wtot = 1e22, 1e22 + 4e21 + 5e23 + 2e24
    vmr = wk1 / wtot
    vk = T["LBLX_MTCKD_WVN"]
    for k, tab, fac in [(0, T["LBLX_MTCKD_SELF"], vmr), (1, T["LBLX_MTCKD_FOR"], 1.0 - vmr)]:
        a, v = absrb(run_scaled(500.0, 2500.0, one(k)), lay, jrad=1)
        want = wk1 * interp(vk, tab * fac * radterm(vk, 296.0), v)
        _check(a, v, want, 510, 2490, 2e-4)
        a0, _ = absrb(run_scaled(500.0, 2500.0, one(k)), lay, jrad=0)
        _check(a0, v, wk1 * interp(vk, tab * fac, v), 510, 2490, 1e-12)


def test_co2_at_band_head_temperature():
    t = 246.0
    lay = layer(500.0, 2500.0, t, p=700.0)
    rho = (700.0 / 1013.0) * (296.0 / t)
    hdr = T["LBLX_FCO2_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    cfac = np.ones_like(vk)
    m = (vk >= 2000) & (vk <= 2998)
    cfac[m] = T["LBLX_XFACCO2"][((vk[m] - 1998.0) / 2.0 + 1e-5).astype(int) - 1]
    ck = T["LBLX_FCO2"] * cfac * 4e21 * rho * 1e-20
    a, v = absrb(run_scaled(500.0, 2500.0, one(2)), lay, jrad=1)
    _check(a, v, interp(vk, ck * radterm(vk, t), v), 505, 2495, 2e-4)
    a, v = absrb(run_scaled(500.0, 2500.0, one(2)), lay, jrad=0)
    _check(a, v, interp(vk, ck, v), 505, 2495, 1e-12)


def test_o2_fundamental_at_296K():
    lay = layer(1200.0, 2000.0, 296.0, p=500.0)
    amagat = (500.0 / 1013.0) * (273.0 / 296.0)
    hdr = T["LBLX_O2F_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    ck = 5e23 * 1e-20 * amagat * (1e20 / XLOSMT) * T["LBLX_O2F_XO2"] / vk
    a, v = absrb(run_scaled(1200.0, 2000.0, one(4)), lay, jrad=0)
    _check(a, v, interp(vk, ck, v), 1345, 1845, 1e-12)
    # pre_xint clips XINT to the table's own span: 1341..1850 cm-1
    assert np.all(a[(v < 1341) | (v > 1850)] == 0.0)
    assert np.all(a[(v >= 1345) & (v <= 1845)] > 0.0)


def test_n2_fundamental_at_272K():
    t = 272.0
    wk = (1e22, 4e21, 0.0, 0.0, 0.0, 0.0, 5e23)
    wtot = sum(wk) + 2e24
    lay = layer(1900.0, 3000.0, t, wk=wk)
    amagat = (1013.0 / 1013.0) * (273.0 / t)
    xh, xo = wk[0] / wtot, wk[6] / wtot
    xn = 1.0 - xh - xo
    hdr = T["LBLX_N2F_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    c0 = T["LBLX_N2F_272"] / vk
    a_o2 = 1.294 - 0.4545 * t / 296.0
    ck = (xn * wtot / XLOSMT) * amagat * (xn * c0 + xo * a_o2 * c0 + xh * (9.0 / 7.0) * T["LBLX_N2F_AH2O"] * c0)
    a, v = absrb(run_scaled(1900.0, 3000.0, one(5)), lay, jrad=1)
    _check(a, v, interp(vk, ck * radterm(vk, t), v), 2006, 2893, 2e-4)
    assert np.all(a[(v < 1998) | (v > 2901)] == 0.0)


@pytest.mark.parametrize("t,p", [(220.0, 300.0), (250.0, 700.0), (310.0, 1013.0)])
def test_h2o_self_temperature_law(t, p):
    """MT_CKD self continuum away from 296 K: ref x (296/T)^texp x vmr x
    (P/1013)(296/T), foreign ref x (1 - vmr) x the same density ratio
    (src/mt_ckd_h2o_module.f90:141-171)."""
    lay = layer(500.0, 2500.0, t, p=p)
    wk1, wtot = 1e22, 1e22 + 4e21 + 5e23 + 2e24
    vmr = wk1 / wtot
    rho = (p / 1013.0) * (296.0 / t)
    vk = T["LBLX_MTCKD_WVN"]
    self_t = T["LBLX_MTCKD_SELF"] * (296.0 / t) ** T["LBLX_MTCKD_TEXP"]
    for k, ck in [(0, self_t * vmr * rho), (1, T["LBLX_MTCKD_FOR"] * (1.0 - vmr) * rho)]:
        a0, v = absrb(run_scaled(500.0, 2500.0, one(k)), lay, jrad=0)
        _check(a0, v, wk1 * interp(vk, ck, v), 510, 2490, 1e-12)
        a1, _ = absrb(run_scaled(500.0, 2500.0, one(k)), lay, jrad=1)
        _check(a1, v, wk1 * interp(vk, ck * radfn(vk, t), v), 510, 2490, 1e-12)
    # the law is not the identity here: the texp factor moves the result
    if t != 296.0:
        a0, v = absrb(run_scaled(500.0, 2500.0, one(0)), lay, jrad=0)
        m = (v >= 510) & (v <= 2490)
        flat = wk1 * interp(vk, T["LBLX_MTCKD_SELF"] * vmr * rho, v)
        assert np.max(np.abs(a0[m] / flat[m] - 1.0)) > 0.05


@pytest.mark.parametrize("t", [200.0, 280.0])
def test_co2_band_head_temperature_law(t):
    """FRNCO2 away from T_eff = 246 K: the 25 band-head coefficients
    (table entries 1196-1220) scale as (T/246)^tdep, the rest is unchanged
    (src/contnm.f90:1295-1351); the v3 scale factor XFACCO2 applies on
    2000-2998 cm-1 (:338-375)."""
    p = 700.0
    lay = layer(500.0, 2500.0, t, p=p)
    rho = (p / 1013.0) * (296.0 / t)
    hdr = T["LBLX_FCO2_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    tcor = np.ones_like(vk)
    tcor[1195:1220] = (t / 246.0) ** T["LBLX_CO2_TDEP"]
    cfac = np.ones_like(vk)
    m = (vk >= 2000) & (vk <= 2998)
    cfac[m] = T["LBLX_XFACCO2"][((vk[m] - 1998.0) / 2.0 + 1e-5).astype(int) - 1]
    ck = T["LBLX_FCO2"] * tcor * cfac * 4e21 * rho * 1e-20
    a, v = absrb(run_scaled(500.0, 2500.0, one(2)), lay, jrad=0)
    _check(a, v, interp(vk, ck, v), 505, 2495, 1e-12)
    a, v = absrb(run_scaled(500.0, 2500.0, one(2)), lay, jrad=1)
    _check(a, v, interp(vk, ck * radfn(vk, t), v), 505, 2495, 1e-12)
    head = (v > 2390) & (v < 2430)
    flat = interp(vk, T["LBLX_FCO2"] * cfac * 4e21 * rho * 1e-20 * radfn(vk, t), v)
    assert np.max(np.abs(a[head] / flat[head] - 1.0)) > 0.02


@pytest.mark.parametrize("t", [220.0, 260.0])
def test_o2_fundamental_temperature_law(t):
    """O2 fundamental away from 296 K: xo2 x exp(xo2t (1/296 - 1/T)) / v
    (o2_ver_1, src/contnm.f90:7254-7318)."""
    p = 500.0
    lay = layer(1200.0, 2000.0, t, p=p)
    amagat = (p / 1013.0) * (273.0 / t)
    hdr = T["LBLX_O2F_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    law = np.exp(T["LBLX_O2F_XO2T"] * (1.0 / 296.0 - 1.0 / t))
    ck = 5e23 * 1e-20 * amagat * (1e20 / XLOSMT) * T["LBLX_O2F_XO2"] * law / vk
    a, v = absrb(run_scaled(1200.0, 2000.0, one(4)), lay, jrad=0)
    _check(a, v, interp(vk, ck, v), 1345, 1845, 1e-12)
    a1, _ = absrb(run_scaled(1200.0, 2000.0, one(4)), lay, jrad=1)
    _check(a1, v, interp(vk, ck * radfn(vk, t), v), 1345, 1845, 1e-12)
    assert np.max(np.abs(law - 1.0)) > 0.1


@pytest.mark.parametrize("t", [228.0, 250.0, 300.0])
def test_n2_fundamental_temperature_law(t):
    """N2 fundamental between and beyond its 272 K and 228 K tables:
    logarithmic interpolation in 1/T where both are positive, linear in T
    elsewhere; the O2 efficiency 1.294 - 0.4545 T/296 and the H2O
    efficiency (9/7) a_h2o (n2_ver_1, src/contnm.f90:2668-2754)."""
    wk = (1e22, 4e21, 0.0, 0.0, 0.0, 0.0, 5e23)
    wtot = sum(wk) + 2e24
    p = 800.0
    lay = layer(1900.0, 3000.0, t, p=p, wk=wk)
    amagat = (p / 1013.0) * (273.0 / t)
    xh, xo = wk[0] / wtot, wk[6] / wtot
    xn = 1.0 - xh - xo
    hdr = T["LBLX_N2F_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    a272, a228 = T["LBLX_N2F_272"], T["LBLX_N2F_228"]
    xtfac = (1.0 / t - 1.0 / 272.0) / (1.0 / 228.0 - 1.0 / 272.0)
    xtlin = (t - 272.0) / (228.0 - 272.0)
    pos = (a272 > 0) & (a228 > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        c0 = np.where(pos, a272 * (a228 / np.where(pos, a272, 1.0)) ** xtfac, a272 + (a228 - a272) * xtlin) / vk
    a_o2 = 1.294 - 0.4545 * t / 296.0
    ck = (xn * wtot / XLOSMT) * amagat * (xn * c0 + xo * a_o2 * c0 + xh * (9.0 / 7.0) * T["LBLX_N2F_AH2O"] * c0)
    a, v = absrb(run_scaled(1900.0, 3000.0, one(5)), lay, jrad=0)
    want = interp(vk, ck, v)
    m = (v >= 2006) & (v <= 2893)
    # pow of the table ratio: 1e-12 of the band's peak absolute
    np.testing.assert_allclose(a[m], want[m], rtol=1e-12, atol=1e-12 * np.abs(want[m]).max())
    if t == 228.0:  # the 228 K table itself
        np.testing.assert_allclose(c0[pos], (a228 / vk)[pos], rtol=1e-12)


@pytest.mark.parametrize("t", [296.0, 220.0, 250.0, 310.0])
def test_n2_rotational_temperature_law(t):
    """N2 rotational (collision-induced) continuum below 350 cm-1 (xn2_r,
    src/contnm.f90:2497-2567): the 296 K and 220 K tables interpolated
    geometrically in T with tfac = (T - 296)/(220 - 296), the O2 efficiency
    from the interpolated scale factor, (sf_T - 1) 0.79/0.21, and H2O at
    efficiency 1, in units of amagat (contnm.f90:754-790)."""
    wk = (1e22, 4e21, 0.0, 0.0, 0.0, 0.0, 5e23)
    wtot = sum(wk) + 2e24
    p = 900.0
    lay = layer(60.0, 340.0, t, p=p, wk=wk)
    amagat = (p / 1013.0) * (273.0 / t)
    xh, xo = wk[0] / wtot, wk[6] / wtot
    xn = 1.0 - xh - xo
    hdr = T["LBLX_N2R_HDR"]
    vk = hdr[0] + hdr[2] * np.arange(int(hdr[3]))
    tfac = (t - 296.0) / (220.0 - 296.0)
    c0 = T["LBLX_N2R_C296"] * (T["LBLX_N2R_C220"] / T["LBLX_N2R_C296"]) ** tfac
    sf = T["LBLX_N2R_SF296"] * (T["LBLX_N2R_SF220"] / T["LBLX_N2R_SF296"]) ** tfac
    ck = (xn * wtot / XLOSMT) * amagat * c0 * (xn + (sf - 1.0) * 0.79 / 0.21 * xo + xh)
    for jrad in (0, 1):
        a, v = absrb(run_scaled(60.0, 340.0, one(5)), lay, jrad=jrad)
        want = interp(vk, ck * (radfn(vk, t) if jrad else 1.0), v)
        m = (v >= 60) & (v <= 340)
        np.testing.assert_allclose(a[m], want[m], rtol=1e-12)
    if t == 220.0:
        np.testing.assert_allclose(c0, T["LBLX_N2R_C220"], rtol=1e-14)
    if t == 296.0:
        np.testing.assert_array_equal(c0, T["LBLX_N2R_C296"])


def test_rayleigh_closed_form():
    lay = layer(800.0, 2500.0, 250.0)
    wtot = 1e22 + 4e21 + 5e23 + 2e24
    a, v = absrb(run_scaled(800.0, 2500.0, one(6)), lay, jrad=1)
    x = v / 1e4
    want = x ** 4 / (938.076 - 10.8426 * x ** 2) * wtot * 1e-20 / 26867.5
    np.testing.assert_allclose(a, want, rtol=1e-13)
    a0, _ = absrb(run_scaled(800.0, 2500.0, one(6)), lay, jrad=0)
    np.testing.assert_allclose(a0, want / radterm(v, 250.0), rtol=2e-4)
    # no Rayleigh below 820 cm-1 or with aerosols
    lo = run_scaled(500.0, 810.0, one(6))
    assert np.all(absrb(lo, layer(500.0, 810.0, 250.0))[0] == 0.0)
    aer = run_scaled(800.0, 2500.0, one(6))
    aer.iaersl = 1
    assert np.all(absrb(aer, lay)[0] == 0.0)


def test_scale_factor_linearity_and_codes():
    atm = synth.atmosphere()
    lay = abi.make_layer(atm["pave"][3], atm["tave"][3], list(atm["wk"][3]), atm["wbroad"][3], 0.001,
                         500.0, 2500.0)
    full, _ = absrb(abi.make_run(500.0, 2500.0, icntnm=1), lay)
    parts = sum(absrb(run_scaled(500.0, 2500.0, one(k)), lay)[0] for k in range(7))
    np.testing.assert_allclose(full, parts, rtol=1e-12)
    half, _ = absrb(run_scaled(500.0, 2500.0, [0.5] * 7), lay)
    np.testing.assert_allclose(half, 0.5 * full, rtol=1e-14)
    codes = {2: [0, 1, 1, 1, 1, 1, 1], 3: [1, 0, 1, 1, 1, 1, 1], 4: [0, 0, 1, 1, 1, 1, 1],
             5: [1, 1, 1, 1, 1, 1, 0]}
    for c, s in codes.items():
        a, _ = absrb(abi.make_run(500.0, 2500.0, icntnm=c), lay)
        b, _ = absrb(run_scaled(500.0, 2500.0, [float(x) for x in s]), lay)
        np.testing.assert_array_equal(a, b)


def test_unsupported_bands():
    for v1, v2 in [(4000.0, 4500.0), (2000.0, 5000.0)]:
        with pytest.raises(RuntimeError):
            O.contnm(abi.make_run(v1, v2, icntnm=1), layer(v1, v2, 250.0), 1)


def test_continuum_only_optical_depth():
    """No lines in the band: every 16th point of the layer OD (the DVR3
    nodes, where both Lagrange steps of PANEL are the identity) is XINT of
    ABSRB, i.e. the Keys interpolation of the 1 cm-1 continuum."""
    tab = abi.LineTable.from_lines(np.array([100.0]), np.array([1e-22]), np.array([0.07]),
                                   np.array([100.0]), np.array([1]), np.array([0.3]),
                                   np.array([0.7]), np.array([0.0]))
    run = abi.make_run(1500.0, 1700.0, icntnm=1)
    lay = layer(1500.0, 1700.0, 296.0)
    lay.dv = 0.002
    od, v1p, st, jrad = O.od_layer(run, abi.Layer.from_buffer_copy(lay), tab)
    assert jrad == 1
    a, va = absrb(run, lay, jrad=1)
    i = np.arange(0, len(od), 16)
    v = v1p + i * lay.dv
    m = (v > va[0] + 2) & (v < va[-1] - 2)
    np.testing.assert_allclose(od[i][m], interp(va, a, v[m]), rtol=1e-12)
    assert np.all(od > 0)
