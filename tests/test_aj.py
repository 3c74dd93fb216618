"""Pins for the analytic-Jacobian continuum of a single-species run (CONTNM
with /CDERIV/ ICFLG, src/contnm.f90:221-239, 314-322) in the oracle -- CPU
only.

For an H2O run the ABSRB carries the ln q derivative of the H2O continuum:
self + foreign + (self - foreign h/(1-h)), h = the H2O share of the column,
so with the scale factors isolating each term it equals 2 self + foreign -
foreign h/(1-h) exactly; any other species run drops the N2 continuum."""
import numpy as np

import oracle_lib as O
from lblrtm_amd import abi, synth


def _absrb(layer, cntscl, icflg):
    run = abi.make_run(layer.v1, layer.v2, icntnm=6, cntscl=cntscl)
    O.set_icflg(icflg)
    try:
        a, _, _ = O.contnm(run, layer, 1)
    finally:
        O.set_icflg(0)
    return a


def test_h2o_run_carries_the_ln_q_derivative():
    p, t, wk, wb = synth.config0_layer()
    w1 = 5e21
    lay = abi.make_layer(1013.0, 290.0, [w1, 0, 0, 0, 0, 0, 0], 2.0e25, 0.005, 800.0, 1200.0)
    h = w1 / (w1 + 2.0e25)
    self_ = _absrb(lay, [1, 0, 0, 0, 0, 0, 0], 0)
    forn = _absrb(lay, [0, 1, 0, 0, 0, 0, 0], 0)
    both = _absrb(lay, [1, 1, 0, 0, 0, 0, 0], 0)
    aj = _absrb(lay, [1, 1, 0, 0, 0, 0, 0], 1)
    np.testing.assert_allclose(both, self_ + forn, rtol=1e-13)
    np.testing.assert_allclose(aj, 2 * self_ + forn - forn * h / (1 - h), rtol=1e-12)
    assert np.abs(aj - both).max() > 1e-3 * np.abs(both).max()


def test_species_run_has_no_n2_continuum():
    lay = abi.make_layer(1013.0, 280.0, [0, 3e21, 0, 0, 0, 0, 0], 2.0e25, 0.005, 2100.0, 2600.0)
    n2 = _absrb(lay, [0, 0, 0, 0, 0, 1, 0], 0)
    assert np.abs(n2).max() > 0
    n2aj = _absrb(lay, [0, 0, 0, 0, 0, 1, 0], 2)
    assert np.abs(n2aj).max() == 0.0
