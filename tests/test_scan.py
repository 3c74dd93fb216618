"""Pins for the oracle's FFTSCN (instrument scan) and the host planner --
CPU only.

The scan is a convolution with the instrument function, so closed forms
fix it: a constant transmittance scans to H(0) times the constant
(H = the apodization at zero path difference); a line far narrower than
the instrument, scanned without the boxcar deconvolution, gives the
Hamming instrument function 2 a0 L [sinc U + a1/2 (sinc(U+pi) +
sinc(U-pi))], U = 2 pi dv L; a Gaussian band scans to a0 times the band
plus two half-weight copies shifted by 1/2L, which also pins INTPDR's
index walk (an off-by-one shifts it by a whole input step).
"""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib

A0, A1 = 0.53835685, 0.857504
H0 = A0 * (1.0 + A1)
V1IN, DVIN = 500.0, 3.683e-4
NIN = int(2000.0 / DVIN) + 1


def ils(dnu, L=1.0):
    u = 2 * np.pi * dnu * L
    s = lambda x: np.sinc(x / np.pi)
    return 2 * A0 * L * (s(u) + A1 / 2 * (s(u + np.pi) + s(u - np.pi)))


def test_constant_transmittance_scans_to_h0():
    x = np.ones(NIN)
    for dvout in (0.0, 0.1):
        sc = abi.make_scan(1.0, 520.0, 2480.0, jemit=0, dvout=dvout)
        out, v1, dv = O.fftscn(sc, x, V1IN, DVIN)
        assert (len(out), v1, dv) == lib.scan_size(sc, V1IN, DVIN, NIN)
        np.testing.assert_allclose(out, H0, rtol=1e-12)


def test_narrow_line_gives_hamming_instrument_function():
    x = np.zeros(NIN)
    i0 = int((1500.0 - V1IN) / DVIN)
    x[i0] = 1.0 / DVIN
    sc = abi.make_scan(1.0, 1480.0, 1520.0, jemit=1, nofix=1)
    out, v1, dvb = O.fftscn(sc, x, V1IN, DVIN)
    # the boxcar moves the line to the centre of its box
    hw = 1.0 / 2.195676
    v1z = V1IN + int((1480.0 - hw * 20 - V1IN) / DVIN + 1.01) * DVIN
    M = int(2.0 * hw / (12 * DVIN))
    n0 = round((v1z - V1IN) / DVIN)
    k0 = (i0 - n0) // M
    vc = v1z + DVIN * (M - 1) / 2.0 + k0 * dvb
    nu = v1 + dvb * np.arange(len(out))
    want = ils(nu - vc)
    assert abs(dvb - M * DVIN) < 1e-15
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-3 * want.max())
    assert abs(out.max() - 2 * A0) < 2e-3


def test_gaussian_band_scans_to_closed_form():
    """A Gaussian band (sigma 20 cm-1) scanned with the Hamming function is
    a0 [G(v) + a1/2 (G(v - 1/2L) + G(v + 1/2L))]: the cosine term of the
    apodization shifts copies of the band by half the inverse OPD.  On the
    INTPDR grids the tolerance allows INTPDR's 1/200 position steps but not
    a shift by one input point."""
    nu = V1IN + DVIN * np.arange(NIN)
    g = lambda v: np.exp(-(v - 1500.0) ** 2 / (2 * 20.0 ** 2))
    x = g(nu)
    for dvout in (0.0, 0.1, 0.25):
        sc = abi.make_scan(1.0, 510.0, 2490.0, jemit=1, dvout=dvout)
        out, v1, dv = O.fftscn(sc, x, V1IN, DVIN)
        v = v1 + dv * np.arange(len(out))
        want = A0 * (g(v) + A1 / 2 * (g(v - 0.5) + g(v + 0.5)))
        tol = 1e-6 if dvout == 0.0 else 1e-4
        np.testing.assert_allclose(out, want, rtol=0, atol=tol)


@pytest.mark.parametrize("jfnin,mratin,dvout", [(-6, 0, 0.2), (-6, -1, 0.0), (1, 0, 0.05), (2, 8, 0.1),
                                                (-3, 0, 0.0), (-5, 0, 0.1), (-4, 0, 0.0), (0, 0, 0.0)])
def test_planner_matches_oracle_sizes(jfnin, mratin, dvout):
    hw = 0.3 if jfnin > 0 else 1.0
    sc = abi.make_scan(hw, 600.0, 700.0, jemit=1, jfnin=jfnin, mratin=mratin, dvout=dvout)
    x = np.ones(NIN)
    out, v1, dv = O.fftscn(sc, x, V1IN, DVIN)
    assert (len(out), v1, dv) == lib.scan_size(sc, V1IN, DVIN, NIN)


def test_unsupported_scans():
    # a band only a few instrument widths wide: spectral-domain function
    with pytest.raises(lib.LblxError) as e:
        lib.scan_size(abi.make_scan(1.0, 600.0, 601.0, ivx=-1), V1IN, DVIN, NIN)
    assert e.value.code == abi.E_UNSUPPORTED
    # input coarser than 2 HWHM / 12: INTPDR pre-interpolation
    with pytest.raises(lib.LblxError) as e:
        lib.scan_size(abi.make_scan(1.0, 600.0, 700.0), V1IN, 0.1, 20001)
    assert e.value.code == abi.E_UNSUPPORTED
