"""Jacobians of the TOA radiance on the GPU (lblx_jacobian_device) against
finite differences of the forward spectrum through the same C ABI.

A one-sided difference of the spectrum with the Jacobian's own steps
perturbs exactly the optical depths the Jacobian differences, so the two
agree up to the step times the cross derivative d2R/dtau dx (line cores
with steep temperature dependence make that the largest term): this
checks the source terms and the tangent propagation through every layer
above.  Central differences with the same steps check the whole
derivative."""
import numpy as np
import pytest

from lblrtm_amd import abi, lib, synth

pytestmark = pytest.mark.gpu

V1, V2 = 880.0, 960.0
SPECIES = [1, 2, 3]


@pytest.fixture(scope="module", params=[1.0, 1e-4], ids=["column", "thin-column"])
def column(gpu_ctx, request):
    """The 50-layer column, and a thin column -- molecules 1-6 only, amounts
    scaled by 1e-4 -- through which the surface shows, so that every layer's
    tangent crosses the whole merge chain.  (O2 stays out of the thin column:
    its air-to-foreign width conversion assumes its own 21 % share.)"""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(V1, V2, icntnm=1)
    wk = atm["wk"] if request.param == 1.0 else atm["wk"][:, :6] * request.param
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], wk, atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(wk[l]), atm["wbroad"][l], dv[l],
                             V1, V2) for l in range(len(dv))]
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    return run, layers, tzu


def _spec(ctx, run, layers, tzu):
    return ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tzu, 295.0, 0.98)[0]


def test_jacobian_matches_finite_differences(gpu_ctx, column):
    run, layers, tzu = column
    dT, dq = 1e-3, 1e-4
    rad, jac, v1p, dvt = gpu_ctx.jacobian(run, [abi.Layer.from_buffer_copy(x) for x in layers], tzu, 295.0,
                                          0.98, SPECIES, dT=dT, dlnq=dq)
    r0 = _spec(gpu_ctx, run, layers, tzu)
    np.testing.assert_allclose(rad, r0, rtol=1e-12)
    print("jacobian: %d points, planes %s" % (len(rad), jac.shape), flush=True)
    res = {}
    for l in (0, 9, 24, 38, 49):
        for q in range(1 + len(SPECIES)):
            up = [abi.Layer.from_buffer_copy(x) for x in layers]
            dn = [abi.Layer.from_buffer_copy(x) for x in layers]
            if q == 0:
                up[l].tave += dT
                dn[l].tave -= dT
                h = dT
            else:
                m = SPECIES[q - 1] - 1
                up[l].wk[m] *= np.exp(dq)
                dn[l].wk[m] *= np.exp(-dq)
                h = dq
            rp = _spec(gpu_ctx, run, up, tzu)
            rm = _spec(gpu_ctx, run, dn, tzu)
            res[l, q] = ((rp - r0) / h, (rp - rm) / (2 * h))
    for q in range(1 + len(SPECIES)):
        # layers deep in an opaque column have Jacobians at rounding level:
        # errors are measured against the largest column of the parameter
        scale = max(np.abs(res[l, q][0]).max() for l in (0, 9, 24, 38, 49))
        for l in (0, 9, 24, 38, 49):
            fwd, cen = res[l, q]
            j = jac[q, l]
            # line rejection (S / alpha against DPTMIN + DPTFAC R4) flips for a
            # few lines between the base and the perturbed state; those jumps
            # enter both sides but leave second-order residue at their points,
            # hence a loose maximum and tight norms
            e1 = np.abs(j - fwd).max() / scale
            floor = 1e-3 * scale * np.sqrt(len(cen))
            n1 = np.linalg.norm(j - fwd) / max(np.linalg.norm(fwd), floor)
            e2 = np.linalg.norm(j - cen) / max(np.linalg.norm(cen), floor)
            # a line whose rejection (near wing, or far wing over +-25 cm-1)
            # flips between T - dT and T puts a jump divided by dT into the
            # central difference: its median over the significant points is
            # the robust comparison
            big = np.abs(cen) > 1e-2 * scale
            med = np.median(np.abs(j - cen)[big] / np.abs(cen)[big]) if big.any() else 0.0
            # OPDPTH switches the radiation term to panel interpolation below
            # V1/kT = 5 (JRAD = 0): a temperature step across that switch is a
            # jump of the model itself, not a derivative
            t = layers[l].tave
            jump = q == 0 and ((V1 * 1.4387752 / (t - dT) < 5) != (V1 * 1.4387752 / (t + dT) < 5))
            print("layer %2d param %d: max|J| %.3e (param max %.3e)  one-sided max %.2e L2 %.2e  central L2 %.2e "
                  "median %.2e%s" % (l, q, np.abs(fwd).max(), scale, e1, n1, e2, med,
                                     "  (JRAD switch inside +-dT)" if jump else ""), flush=True)
            assert e1 < 2e-2
            assert n1 < 1e-3
            if not jump:
                assert med < 1e-2


def test_analytic_jacobian_formulation(gpu_ctx, column):
    """The reference's analytic-Jacobian formulation (single-species runs,
    +1 K) against the finite-difference Jacobian: the species columns agree
    up to self-broadening, the H2O continuum's second-order terms and lines
    that only the full-amount run rejects (observed relative L2 below 6e-4);
    the temperature columns differ by the 1 K step's curvature (below
    5e-3)."""
    run, layers, tzu = column
    rad, jfd, _, _ = gpu_ctx.jacobian(run, [abi.Layer.from_buffer_copy(x) for x in layers], tzu, 295.0, 0.98,
                                      SPECIES, dT=1e-3, dlnq=1e-4)
    rad2, jaj, _, _ = gpu_ctx.jacobian(run, [abi.Layer.from_buffer_copy(x) for x in layers], tzu, 295.0, 0.98,
                                       SPECIES, aj=True)
    np.testing.assert_allclose(rad, rad2, rtol=1e-12)  # LDS atomic order
    worst = {}
    for q in range(1 + len(SPECIES)):
        a, b = jaj[q], jfd[q]
        scale = np.abs(b).max()
        err = np.sqrt(np.mean((a - b) ** 2)) / scale
        worst[q] = err
        print("AJ vs FD: parameter %d  rel L2 %.3e  (max |J| %.3e)" % (q, err, scale), flush=True)
    for q in range(1, 1 + len(SPECIES)):
        assert worst[q] < 3e-3, (q, worst[q])
    assert worst[0] < 2e-2
