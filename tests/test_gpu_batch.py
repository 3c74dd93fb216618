"""Batched profiles on the GPU (lblx_spectra_batch_device): ragged columns
with their own surface pressures and band windows, each equal to the same
profile run alone through lblx_spectrum_device (itself pinned to the
oracle); the merge graph replayed; autotuned tile widths leave results
unchanged."""
import numpy as np
import pytest

from lblrtm_amd import abi, lib, synth

pytestmark = pytest.mark.gpu


def _profile_layers(run, atm):
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    return [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                           atm["v1"], atm["v2"]) for l in range(len(dv))]


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_SPLIT])
def test_batch_equals_single_profiles(gpu_ctx, precision):
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    profs = synth.batch_profiles(6, seed=77)
    run = abi.make_run(500.0, 2500.0, icntnm=1, precision=precision)
    items = []
    for atm in profs:
        r = abi.make_run(atm["v1"], atm["v2"], icntnm=1, precision=precision)
        items.append((_profile_layers(r, atm), atm["tlev"][1:], atm["tlev"][0], 0.98))
    sizes = [lib.od_size(run, it[0][-1]) for it in items]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    rad = gpu_ctx.alloc(8 * int(sum(sizes)))
    tra = gpu_ctx.alloc(8 * int(sum(sizes)))
    for rep in range(3):  # first call captures the merge graph, later ones replay it
        v1p = gpu_ctx.spectra_batch_device(run, items, rad.ptr, tra.ptr, offs)
        rb = rad.to_host(np.float64)
        tb = tra.to_host(np.float64)
        for k, it in enumerate(items):
            r = abi.make_run(profs[k]["v1"], profs[k]["v2"], icntnm=1, precision=precision)
            rs, ts, v1s, _ = gpu_ctx.spectrum(r, [abi.Layer.from_buffer_copy(x) for x in it[0]], it[1], it[2], it[3])
            print("rep %d profile %d: %d layers %.1f-%.1f cm-1, %d points" %
                  (rep, k, len(it[0]), profs[k]["v1"], profs[k]["v2"], len(rs)), flush=True)
            assert v1p[k] == v1s
            # FP64: same kernels, only the order of LDS atomic adds differs;
            # FP32 split: float atomics in a different order, ~1e-9 relative
            tol = 1e-12 if precision == abi.FP64 else 1e-6
            np.testing.assert_allclose(rb[offs[k]:offs[k] + sizes[k]], rs, rtol=tol)
            np.testing.assert_allclose(tb[offs[k]:offs[k] + sizes[k]], ts, rtol=tol, atol=1e-300)
    rad.free()
    tra.free()


def test_autotune_keeps_results(gpu_ctx):
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(900.0, 1000.0, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                             900.0, 1000.0) for l in range(len(dv))]
    tz = atm["tlev"][1:]
    r0 = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tz, 295.0, 0.98)[0]
    tiles = gpu_ctx.autotune(run, layers[::5])
    print("autotuned tiles", tiles, flush=True)
    assert all(t % 64 == 0 for t in tiles)
    r1 = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tz, 295.0, 0.98)[0]
    np.testing.assert_allclose(r1, r0, rtol=1e-12)
    gpu_ctx.set_tiles((8192, 4096, 2048))
    with pytest.raises(lib.LblxError):
        gpu_ctx.set_tiles((100, 1024, 512))


def test_tuning_line_batch_and_persistence(gpu_ctx, tmp_path):
    """The tiles' line batch (lines per wave chunk, 1..64) changes only the
    order of the LDS additions; the autotuner picks widths and batches, the
    tuning file round-trips into a new context of the same device, and a
    file from another device or a bad value is refused."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    atm = synth.atmosphere()
    run = abi.make_run(900.0, 1000.0, icntnm=1)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    layers = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l],
                             900.0, 1000.0) for l in range(len(dv))]
    tz = atm["tlev"][1:]
    base = gpu_ctx.tuning()
    r0 = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tz, 295.0, 0.98)[0]
    for lb in ((8, 8, 8), (64, 64, 64), (1, 13, 47)):
        gpu_ctx.set_tuning(base[0], lb)
        r = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in layers], tz, 295.0, 0.98)[0]
        np.testing.assert_allclose(r, r0, rtol=1e-12)
    gpu_ctx.set_tuning(*base)
    gpu_ctx.autotune(run, layers[::5])
    tuned = gpu_ctx.tuning()
    print("autotuned tiles %s, line batch %s" % tuned, flush=True)
    assert all(b in (8, 16, 32, 64) for b in tuned[1])
    path = tmp_path / "tuning.txt"
    gpu_ctx.tuning_save(path)
    other = lib.Context(0)
    try:
        other.tuning_load(path)
        assert other.tuning() == tuned
        txt = open(path).read().split("\n")
        bad = tmp_path / "other.txt"
        open(bad, "w").write("\n".join(["lblx-tuning 1 gfx000|elsewhere|1"] + txt[1:]))
        with pytest.raises(lib.LblxError) as e:
            other.tuning_load(bad)
        assert e.value.code == abi.E_UNSUPPORTED
        with pytest.raises(lib.LblxError):
            other.set_tuning((8192, 4096, 2048), (16, 65, 32))
    finally:
        other.close()
    gpu_ctx.set_tuning(*base)


@pytest.mark.parametrize("aj", [False, True])
def test_jacobians_batch_equal_single(gpu_ctx, aj):
    """Forward and adjoint of a ragged batch (lblx_jacobians_batch_device):
    each profile's radiance and Jacobian planes equal the same profile's
    lblx_jacobian_device / lblx_jacobian_aj_device alone."""
    gpu_ctx.load_lines(synth.config1_lines())
    run = abi.make_run(500.0, 2500.0, icntnm=1)
    specs = [(22, 1000.0, 1020.0, 1013.0), (26, 700.0, 715.0, 850.0), (20, 1500.0, 1530.0, 600.0)]
    items = []
    for k, (nl, v1, v2, ps) in enumerate(specs):
        atm = synth.atmosphere(nlev=nl + 1, psfc=ps, seed=900 + k)
        atm["v1"], atm["v2"] = v1, v2
        items.append((_profile_layers(run, atm), atm["tlev"][1:], atm["tlev"][0], 0.98))
    species = [1, 2, 3]
    P = 1 + len(species)
    sizes = [lib.od_size(run, it[0][-1]) for it in items]
    jsz = [P * len(it[0]) * n for it, n in zip(items, sizes)]
    roff = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    joff = np.concatenate([[0], np.cumsum(jsz)[:-1]]).astype(np.int64)
    rad = gpu_ctx.alloc(8 * int(sum(sizes)))
    jac = gpu_ctx.alloc(8 * int(sum(jsz)))
    v1p = gpu_ctx.jacobians_batch_device(run, items, species, rad.ptr, roff, jac.ptr, joff, aj=aj)
    rb = rad.to_host(np.float64)
    jb = jac.to_host(np.float64)
    rad.free()
    jac.free()
    for k, it in enumerate(items):
        rs, js, v1s, _ = gpu_ctx.jacobian(run, [abi.Layer.from_buffer_copy(x) for x in it[0]], it[1], it[2], it[3],
                                          species, aj=aj)
        print("aj %d profile %d: %d layers, %d points" % (aj, k, len(it[0]), sizes[k]), flush=True)
        assert v1p[k] == v1s
        np.testing.assert_allclose(rb[roff[k]:roff[k] + sizes[k]], rs, rtol=1e-12)
        # same kernels; the tangent planes are FP32, so an OD difference in the
        # last bits (order of LDS atomic adds) can move a plane value by one
        # FP32 ulp
        jk = jb[joff[k]:joff[k] + jsz[k]].reshape(js.shape)
        np.testing.assert_allclose(jk, js, rtol=1e-5, atol=1e-6 * np.abs(js).max())


def test_batch_with_coupling_and_far_ir(gpu_ctx):
    """A batch mixing profiles in the far IR (first panels at or below 0
    cm-1, the VVW mirror) and mid-IR windows over a line file with coupled
    lines: each profile equals the same profile run alone."""
    tab = synth.coupled_lines(60000, 0.3, 1300.0, [1, 2, 3, 7], 57)
    gpu_ctx.load_lines(tab)
    profs = synth.batch_profiles(4, seed=79)
    bands = [(0.1, 60.0), (10.0, 120.0), (600.0, 900.0), (0.0, 40.0)]
    run = abi.make_run(0.0, 1300.0, icntnm=1)
    items, runs = [], []
    for atm, (v1, v2) in zip(profs, bands):
        atm = dict(atm, v1=v1, v2=v2)
        r = abi.make_run(v1, v2, icntnm=1)
        runs.append(r)
        items.append((_profile_layers(r, atm), atm["tlev"][1:], atm["tlev"][0], 0.98))
    sizes = [lib.od_size(r, it[0][-1]) for r, it in zip(runs, items)]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    rad = gpu_ctx.alloc(8 * int(sum(sizes)))
    tra = gpu_ctx.alloc(8 * int(sum(sizes)))
    v1p = gpu_ctx.spectra_batch_device(run, items, rad.ptr, tra.ptr, offs)
    rb = rad.to_host(np.float64)
    tb = tra.to_host(np.float64)
    for k, it in enumerate(items):
        rs, ts, v1s, _ = gpu_ctx.spectrum(runs[k], [abi.Layer.from_buffer_copy(x) for x in it[0]], it[1], it[2],
                                          it[3])
        print("profile %d: %d layers %.1f-%.1f cm-1, %d points" % (k, len(it[0]), bands[k][0], bands[k][1], len(rs)),
              flush=True)
        assert v1p[k] == v1s
        np.testing.assert_allclose(rb[offs[k]:offs[k] + sizes[k]], rs, rtol=1e-12)
        # far-IR columns are opaque (T down to 1e-185): compare the total
        # optical depth -ln T, whose rounding (LDS atomics in another order)
        # stays ~1e-13 relative
        tk = tb[offs[k]:offs[k] + sizes[k]]
        pos = (tk > 0) & (ts > 0)
        assert np.array_equal(tk > 0, ts > 0)
        np.testing.assert_allclose(-np.log(tk[pos]), -np.log(ts[pos]), rtol=1e-11, atol=1e-12)
    rad.free()
    tra.free()
