"""Tangent-linear optical depths (lblx_od_tangent_device) and the Jacobians
built on them (lblx_jacobian_tl_device) against central differences of the
forward model through the same C ABI.

Two properties of the forward model shape the comparisons:

* Line shapes are read from their tables at the nearest entry (CNVFNV,
  0.002 half-widths per entry), so the optical depth of one point is a
  step function of the widths at fine scale: a temperature difference
  with a small step measures table steps, not the derivative.  Layer
  optical-depth temperature derivatives are therefore compared with a
  2 K central difference (the steps average out; the error falls as 1/h,
  the fit coefficient tends to 1).
* A line whose rejection (DPTMIN + DPTFAC R4 in LNCOR1 and CONVF4) flips
  inside +-h adds a jump / 2h to the difference at its points.  Such points
  are found by comparing the differences with steps h and 2h, which agree
  to second order where the model is smooth, and are left out; the
  remaining points carry the comparison.
"""
import numpy as np
import pytest

from lblrtm_amd import abi, lib, synth

pytestmark = pytest.mark.gpu

SPECIES = [1, 2, 3]


def _cp(layers):
    return [abi.Layer.from_buffer_copy(x) for x in layers]


def _column(v1, v2, precision=abi.FP64):
    atm = synth.atmosphere()
    run = abi.make_run(v1, v2, icntnm=1, precision=precision)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l], v1, v2)
              for l in range(len(dv))]
    return run, layers, np.asarray(atm["tlev"][1:], np.float64)


@pytest.fixture(scope="module")
def lines(gpu_ctx):
    gpu_ctx.load_lines(synth.config1_lines())
    return True


def _perturb(layers, l, q, h):
    L = _cp(layers)
    if q == 0:
        L[l].tave += h
    else:
        L[l].wk[SPECIES[q - 1] - 1] *= np.exp(h)
    return L


def test_od_tangent_temperature(gpu_ctx, lines):
    """dOD/dT of four layers (880-960 cm-1) against 2 K central differences:
    least-squares fit within 1 %, and within 3 % after a 64-point running
    mean (observed 0.9994-1.0009 and 0.3-0.6 %)."""
    run, layers, _ = _column(880.0, 960.0)
    od, dod, sizes = gpu_ctx.od_tangent(run, _cp(layers), SPECIES)
    assert np.all(np.isfinite(dod))
    ker = np.ones(64) / 64
    for l in (30, 38, 45, 49):
        n = sizes[l]
        outs = [gpu_ctx.od_tangent(run, _perturb(layers, l, 0, h), SPECIES)[0][l, :n] for h in (2.0, -2.0)]
        cen = (outs[0] - outs[1]) / 4.0
        t = dod[l, 0, :n]
        fit = np.sum(t * cen) / np.sum(cen * cen)
        ts, cs = np.convolve(t, ker, "valid"), np.convolve(cen, ker, "valid")
        sm = np.linalg.norm(ts - cs) / np.linalg.norm(cs)
        print("layer %2d dOD/dT: fit %.4f  smoothed rel L2 %.2e" % (l, fit, sm), flush=True)
        assert abs(fit - 1.0) < 1e-2
        assert sm < 3e-2


def test_od_tangent_species(gpu_ctx, lines):
    """d OD/d ln q of H2O, CO2, O3 (500-2500 cm-1, layer 45) at the points of
    moderate optical depth (0.01 to 100), where the radiance sees them:
    median relative error below 1e-3 on points whose two differences
    agree (observed 3e-7, 7e-5, 1.3e-6).  This includes every line's
    width change through the total amount (O2 carries most of it)."""
    run, layers, _ = _column(500.0, 2500.0)
    od, dod, sizes = gpu_ctx.od_tangent(run, _cp(layers), SPECIES)
    l = 45
    n = sizes[l]
    o = od[l, :n]
    for q in (1, 2, 3):
        h = 1e-4
        d = {}
        for k in (1, 2):
            a = gpu_ctx.od_tangent(run, _perturb(layers, l, q, k * h), SPECIES)[0][l, :n]
            b = gpu_ctx.od_tangent(run, _perturb(layers, l, q, -k * h), SPECIES)[0][l, :n]
            d[k] = (a - b) / (2 * k * h)
        t = dod[l, q, :n]
        ok = np.abs(d[1] - d[2]) <= 1e-3 * np.abs(d[1])
        m = ok & (o > 1e-2) & (o <= 1e2) & (np.abs(d[1]) > 1e-3 * o)
        r = np.abs(t[m] - d[1][m]) / np.abs(d[1][m])
        print("layer %d species %d: %d points, median %.2e, share off by >1%% %.2e (left out: %d)" %
              (l, SPECIES[q - 1], int(m.sum()), np.median(r), np.mean(r > 1e-2), int((~ok).sum())), flush=True)
        assert m.sum() > 1000
        assert np.median(r) < 1e-3
        assert np.mean(r > 1e-2) < 2e-2


def test_tl_jacobian_full_band(gpu_ctx, lines):
    """The 200-column Jacobian of the 500-2500 cm-1 column from the
    tangent-linear optical depths against central differences of the
    spectrum (T: 1e-3 relative, ln q: 1e-4), on points whose h and 2h
    differences agree: median below 1e-3, share off by more than 1 % below
    1e-2 (observed medians 2e-5..1.5e-4 (T), 1e-7..7e-5 (species))."""
    run, layers, tzu = _column(500.0, 2500.0)
    print("full-band TL Jacobian: 50 layers, 200 columns", flush=True)
    rad, jac, _, _ = gpu_ctx.jacobian(run, _cp(layers), tzu, 295.0, 0.98, SPECIES, tl=True)
    assert jac.shape == (4, 50, len(rad)) and np.all(np.isfinite(jac))
    spec = lambda L: gpu_ctx.spectrum(run, L, tzu, 295.0, 0.98)[0]
    r0 = spec(_cp(layers))
    np.testing.assert_allclose(rad, r0, rtol=1e-12)
    for l in (0, 25, 42, 45, 47, 49):
        for q in range(4):
            h = 1e-3 * layers[l].tave if q == 0 else 1e-4
            d = {}
            for k in (1, 2):
                d[k] = (spec(_perturb(layers, l, q, k * h)) - spec(_perturb(layers, l, q, -k * h))) / (2 * k * h)
            j = jac[q, l]
            scale = np.abs(d[1]).max()
            # opaque above: the differences are rounding (R 1e-16 / h), the Jacobian ~ 0
            if scale < 1e-9 * r0.max():
                print("layer %2d param %d: max|dR/dx| %.3e (rounding), max|J| %.3e" % (l, q, scale, np.abs(j).max()),
                      flush=True)
                assert np.abs(j).max() < 1e-9 * r0.max()
                continue
            ok = np.abs(d[1] - d[2]) <= 1e-3 * np.abs(d[1])
            big = ok & (np.abs(d[1]) > 1e-2 * scale)
            r = np.abs(j[big] - d[1][big]) / np.abs(d[1][big])
            med = np.median(r) if big.any() else 0.0
            off = np.mean(r > 1e-2) if big.any() else 0.0
            print("layer %2d param %d: max|dR/dx| %.3e  %d points (left out %d)  median %.2e  off>1%% %.2e" %
                  (l, q, scale, int(big.sum()), int((~ok & (np.abs(d[1]) > 1e-2 * scale)).sum()), med, off),
                  flush=True)
            assert med < 1e-3
            assert off < 1e-2


def test_tl_jacobian_fp32_split(gpu_ctx, lines):
    """FP32 split: the tangent-linear Jacobian equals the FP64 one within the
    split's own accuracy (median relative difference below 1e-3 on the
    significant points of every plane)."""
    run64, layers, tzu = _column(880.0, 960.0)
    run32, _, _ = _column(880.0, 960.0, precision=abi.FP32_SPLIT)
    _, j64, _, _ = gpu_ctx.jacobian(run64, _cp(layers), tzu, 295.0, 0.98, SPECIES, tl=True)
    _, j32, _, _ = gpu_ctx.jacobian(run32, _cp(layers), tzu, 295.0, 0.98, SPECIES, tl=True)
    for q in range(4):
        scale = np.abs(j64[q]).max()
        big = np.abs(j64[q]) > 1e-2 * scale
        r = np.abs(j32[q][big] - j64[q][big]) / np.abs(j64[q][big])
        print("param %d: FP32 split vs FP64 median %.2e p99 %.2e" % (q, np.median(r), np.percentile(r, 99)),
              flush=True)
        assert np.median(r) < 1e-3


def test_tl_batch_equals_single(gpu_ctx, lines):
    """lblx_jacobians_batch_device with the tangent-linear optical depths
    equals lblx_jacobian_tl_device profile by profile."""
    profs = synth.batch_profiles(256)[:3]
    items = []
    for atm in profs:
        r = abi.make_run(atm["v1"], atm["v2"], icntnm=1)
        dv, _ = lib.plan_dv(r, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
        lays = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                               atm["v1"], atm["v2"]) for l in range(len(dv))]
        items.append((lays, atm["tlev"][1:], atm["tlev"][0], 0.98))
    run = abi.make_run(500.0, 2500.0, icntnm=1)
    P = 1 + len(SPECIES)
    sizes = [lib.od_size(run, it[0][-1]) for it in items]
    jsz = [P * len(it[0]) * n for it, n in zip(items, sizes)]
    roff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    joff = np.concatenate([[0], np.cumsum(jsz)[:-1]]).astype(np.int64)
    rad = gpu_ctx.alloc(8 * int(sum(sizes)))
    jac = gpu_ctx.alloc(8 * int(sum(jsz)))
    try:
        gpu_ctx.jacobians_batch_device(run, items, SPECIES, rad.ptr, roff, jac.ptr, joff, tl=True)
        rb = rad.to_host(np.float64, int(sum(sizes)))
        jb = jac.to_host(np.float64, int(sum(jsz)))
    finally:
        rad.free()
        jac.free()
    for k, it in enumerate(items):
        rs, js, _, _ = gpu_ctx.jacobian(run, _cp(it[0]), it[1], it[2], it[3], SPECIES, tl=True)
        np.testing.assert_allclose(rb[roff[k]:roff[k] + sizes[k]], rs, rtol=1e-12)
        jk = jb[joff[k]:joff[k] + jsz[k]].reshape(js.shape)
        np.testing.assert_allclose(jk, js, rtol=1e-9, atol=1e-12 * np.abs(js).max())


def test_tl_rejects_bad_requests(gpu_ctx, lines):
    """Repeated species are refused (one channel per molecule), as is the
    exact-Voigt build, whose shapes the tangent does not follow."""
    run, layers, tzu = _column(880.0, 960.0)
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.jacobian(run, _cp(layers), tzu, 295.0, 0.98, [1, 1], tl=True)
    assert e.value.code == abi.E_ARG
    vrun = abi.make_run(880.0, 960.0, icntnm=1)
    vrun.voigt = 1
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.od_tangent(vrun, _cp(layers[:2]), SPECIES)
    assert e.value.code == abi.E_UNSUPPORTED


@pytest.mark.parametrize("scale", [1.0, 1e-4], ids=["column", "thin-column"])
def test_backward_equals_jacobian_contraction(gpu_ctx, lines, scale):
    """lblx_backward_batch (reverse mode) against the tangent-linear Jacobian
    planes contracted with the same seeded dL/dR: the same derivative taken
    the other way round, equal to the FP32 rounding of the forward-mode
    planes (1e-5 of the largest gradient).  The thin column (amounts x 1e-4,
    molecules 1-6) is transparent much deeper, so the adjoint crosses more
    of the merge chain before it meets the layers it differentiates."""
    atm = synth.atmosphere()
    run = abi.make_run(880.0, 960.0, icntnm=1)
    wk = atm["wk"] if scale == 1.0 else atm["wk"][:, :6] * scale
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], wk, atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(wk[l]), atm["wbroad"][l], dv[l], 880.0, 960.0)
              for l in range(len(dv))]
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    rad, jac, _, _ = gpu_ctx.jacobian(run, _cp(layers), tzu, 295.0, 0.98, SPECIES, tl=True)
    g = np.random.default_rng(7).normal(size=len(rad))
    r2, grad = gpu_ctx.backward(run, _cp(layers), tzu, 295.0, 0.98, SPECIES, g)
    np.testing.assert_allclose(r2, rad, rtol=1e-12)
    ref = np.einsum("qln,n->ql", jac, g)
    err = np.abs(grad - ref).max() / np.abs(ref).max()
    print("%s: max|grad| %.3e, max deviation from J^T g %.2e (relative)" % (
        "column" if scale == 1.0 else "thin column", np.abs(ref).max(), err), flush=True)
    assert err < 1e-5
    seen = int(np.sum(np.abs(ref).max(axis=0) > 1e-6 * np.abs(ref).max()))
    print("layers with a gradient above 1e-6 of the largest: %d" % seen, flush=True)
    assert seen >= (20 if scale != 1.0 else 5)


def test_backward_batch_equals_single(gpu_ctx, lines):
    """Two profiles in one lblx_backward_batch call equal their single runs
    (the adjoint's scatter uses atomic adds: equal to rounding)."""
    profs = synth.batch_profiles(256)[3:5]
    items = []
    for atm in profs:
        r = abi.make_run(atm["v1"], atm["v2"], icntnm=1)
        dv, _ = lib.plan_dv(r, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
        lays = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                               atm["v1"], atm["v2"]) for l in range(len(dv))]
        items.append((lays, atm["tlev"][1:], atm["tlev"][0], 0.98))
    run = abi.make_run(500.0, 2500.0, icntnm=1)
    P = 1 + len(SPECIES)
    sizes = [lib.od_size(run, it[0][-1]) for it in items]
    roff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    gsz = [P * len(it[0]) for it in items]
    goff = np.concatenate([[0], np.cumsum(gsz)[:-1]]).astype(np.int64)
    rng = np.random.default_rng(11)
    dl = rng.normal(size=int(sum(sizes)))
    d_dl = gpu_ctx.alloc(8 * len(dl))
    d_g = gpu_ctx.alloc(8 * int(sum(gsz)))
    try:
        d_dl.from_host(dl)
        gpu_ctx.backward_batch_device(run, items, SPECIES, d_dl.ptr, roff, d_g.ptr, goff)
        gb = d_g.to_host(np.float64, int(sum(gsz)))
    finally:
        d_dl.free()
        d_g.free()
    for k, it in enumerate(items):
        _, gs = gpu_ctx.backward(run, _cp(it[0]), it[1], it[2], it[3], SPECIES, dl[roff[k]:roff[k] + sizes[k]])
        gk = gb[goff[k]:goff[k] + gsz[k]].reshape(gs.shape)
        print("profile %d: %d layers, max|grad| %.3e" % (k, len(it[0]), np.abs(gs).max()), flush=True)
        np.testing.assert_allclose(gk, gs, rtol=1e-10, atol=1e-12 * np.abs(gs).max())


def test_tl_jacobian_against_oracle_differences(gpu_ctx, lines):
    """The tangent-linear Jacobian of the 880-960 cm-1 column against central
    differences of the oracle's own OPDPTH + XMERGE chain (CPU restatement of
    the reference): perturbing one layer reruns that layer's OPDPTH in the
    oracle and the merge chain above it.  Points whose h and 2h differences
    disagree (rejection flips) are left out; median below 1e-3 and share off
    by more than 1 % below 2e-2 on the rest."""
    import oracle_lib as O
    tab = synth.config1_lines()
    run, layers, tzu = _column(880.0, 960.0)
    rad, jac, _, _ = gpu_ctx.jacobian(run, _cp(layers), tzu, 295.0, 0.98, SPECIES, tl=True)
    tbnd, emis, v2 = 295.0, 0.98, 960.0

    def layer_od(L, jrad_in):
        inp = abi.Layer.from_buffer_copy(L)
        inp.jrad = jrad_in
        od, _, _, jout = O.od_layer(run, inp, tab)
        return od, O.panels(), jout

    base, jin = [], 1
    for L in layers:
        od, pan, jout = layer_od(L, jin)
        base.append((od, pan, jin))
        jin = jout

    def chain(items):
        m = None
        for k, (od, pan, _) in enumerate(items):
            if k == 0:
                m = O.merge_first(od, pan, layers[k].dv, layers[k].tave, tzu[k], tbnd, emis)
            else:
                m, _ = O.merge_next(m, od, pan, layers[k].dv, layers[k].tave, tzu[k], v2)
        return m.rad

    r0 = chain(base)
    np.testing.assert_allclose(rad, r0, rtol=1e-9)  # GPU spectrum = oracle spectrum

    def spec(l, q, h):
        L = _perturb(layers, l, q, h)[l]
        od, pan, jout = layer_od(L, base[l][2])
        items = list(base)
        items[l] = (od, pan, base[l][2])
        if l + 1 < len(layers) and jout != base[l + 1][2]:  # the next layer's JRAD moved
            od1, pan1, _ = layer_od(layers[l + 1], jout)
            items[l + 1] = (od1, pan1, jout)
        # the merge uses the perturbed layer's temperature too
        saved = layers[l].tave
        layers[l].tave = L.tave
        try:
            return chain(items)
        finally:
            layers[l].tave = saved

    for l, q in ((49, 0), (47, 0), (45, 0), (49, 1), (45, 2), (47, 2), (49, 3), (45, 3)):
        h = 1e-3 * layers[l].tave if q == 0 else 1e-4
        d = {k: (spec(l, q, k * h) - spec(l, q, -k * h)) / (2 * k * h) for k in (1, 2)}
        j = jac[q, l]
        scale = np.abs(d[1]).max()
        ok = np.abs(d[1] - d[2]) <= 1e-3 * np.abs(d[1])
        big = ok & (np.abs(d[1]) > 1e-2 * scale)
        r = np.abs(j[big] - d[1][big]) / np.abs(d[1][big])
        print("oracle differences, layer %2d param %d: %d points, median %.2e, off>1%% %.2e" %
              (l, q, int(big.sum()), np.median(r), np.mean(r > 1e-2)), flush=True)
        assert big.sum() > 100
        assert np.median(r) < 1e-3
        assert np.mean(r > 1e-2) < 2e-2
