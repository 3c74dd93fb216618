"""GPU parity of the layer merge (XMERGE, IEMIT=1 upwelling) and of the
whole spectrum (OPDPTH for every layer, then XMERGE up the column) with
the CPU oracle, through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth

pytestmark = pytest.mark.gpu

RADCN1 = 1.191042722E-12
RADCN2 = 1.4387752
BT_TOL = 0.01   # K, brightness temperature


def bt(rad, v):
    return RADCN2 * v / np.log1p(RADCN1 * v ** 3 / rad)


def _layers(v1, v2, icntnm=0):
    atm = synth.atmosphere()
    run = abi.make_run(v1, v2, icntnm=icntnm)
    dv, it = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=1)
    lays = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l],
                           dv[l], v1, v2) for l in range(len(dv))]
    return run, lays, atm, it


@pytest.mark.parametrize("pair", [(0, 1), (1, 2), (5, 6), (10, 11)])
def test_xmerge_layer_pairs(gpu_ctx, pair):
    """One merge step on the grids of two adjacent layers (ITYPE 0, 2, 4, 5)
    with the same seeded optical depth and carry-over radiance on both."""
    a, b = pair
    tab = synth.config0_lines()
    run, lays, atm, it = _layers(600.0, 1100.0)
    La, Lb = lays[a], lays[b]
    print("xmerge pair", pair, "itype", int(it[b]), flush=True)
    O.od_layer(run, abi.Layer.from_buffer_copy(La), tab)
    pan_a = O.panels()
    O.od_layer(run, abi.Layer.from_buffer_copy(Lb), tab)
    pan_b = O.panels()
    na, nb = int(pan_a[2].sum()), int(pan_b[2].sum())
    assert nb == lib.merge_size(run, La, Lb)
    rng = np.random.default_rng(100 + b)
    od = rng.exponential(0.4, nb)
    ro = 1e-6 * (1.0 + rng.random(na))
    to = rng.random(na)
    bufs = [gpu_ctx.alloc(8 * n) for n in (nb, na, na, nb, nb)]
    for buf, x in zip(bufs[:3], (od, ro, to)):
        buf.from_host(x)
    tzu = float(atm["tlev"][b + 1])
    gpu_ctx.xmerge_device(run, La, Lb, tzu, 0.0, 0.0, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                          bufs[3].ptr, bufs[4].ptr)
    rg, tg = bufs[3].to_host(np.float64), bufs[4].to_host(np.float64)
    ref, q = O.merge_next(O.Merged(pan_a, La.dv, ro, to), od, pan_b, Lb.dv, Lb.tave, tzu, 1100.0)
    assert q == 0
    np.testing.assert_allclose(rg, ref.rad, rtol=1e-12)
    np.testing.assert_allclose(tg, ref.tra, rtol=1e-12, atol=1e-300)
    # the layer at the surface (EMINIT)
    gpu_ctx.xmerge_device(run, None, Lb, tzu, 288.0, 0.98, bufs[0].ptr, None, None,
                          bufs[3].ptr, bufs[4].ptr)
    rg, tg = bufs[3].to_host(np.float64), bufs[4].to_host(np.float64)
    ref = O.merge_first(od, pan_b, Lb.dv, Lb.tave, tzu, 288.0, 0.98)
    np.testing.assert_allclose(rg, ref.rad, rtol=1e-12)
    np.testing.assert_allclose(tg, ref.tra, rtol=1e-14)
    for buf in bufs:
        buf.free()


def test_spectrum_column_workload(gpu_ctx):
    """The 50-layer column, 500,000 lines over 500-2500 cm-1, continuum on:
    TOA upwelling radiance within 0.01 K brightness temperature of the
    oracle's OPDPTH + XMERGE chain, and within 1e-10 of the oracle's merge
    of the GPU's own layer optical depths."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    run, lays, atm, _ = _layers(500.0, 2500.0, icntnm=1)
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    tbnd, emis = float(atm["tlev"][0]), 0.98
    work = [abi.Layer.from_buffer_copy(x) for x in lays]
    rad, tra, v1p, dv = gpu_ctx.spectrum(run, work, tzu, tbnd, emis)
    print("spectrum: top grid points", len(rad), "v1p", v1p, flush=True)
    # the GPU's per-layer optical depths
    sizes = [lib.od_size(run, L) for L in lays]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = gpu_ctx.alloc(8 * int(sum(sizes)))
    gpu_ctx.opdpth_column_device(run, [abi.Layer.from_buffer_copy(x) for x in lays], buf.ptr, offs)
    allod = buf.to_host(np.float64)
    buf.free()
    jrad = 1
    m_ref = m_mix = None
    for l, L in enumerate(lays):
        inp = abi.Layer.from_buffer_copy(L)
        inp.jrad = jrad
        od_ref, _, _, jrad = O.od_layer(run, inp, tab)
        pan = O.panels()
        od_gpu = allod[offs[l]:offs[l] + sizes[l]]
        assert len(od_ref) == len(od_gpu)
        if l == 0:
            m_ref = O.merge_first(od_ref, pan, L.dv, L.tave, tzu[l], tbnd, emis)
            m_mix = O.merge_first(od_gpu, pan, L.dv, L.tave, tzu[l], tbnd, emis)
        else:
            m_ref, q1 = O.merge_next(m_ref, od_ref, pan, L.dv, L.tave, tzu[l], 2500.0)
            m_mix, q2 = O.merge_next(m_mix, od_gpu, pan, L.dv, L.tave, tzu[l], 2500.0)
            assert q1 == 0 and q2 == 0
    assert len(rad) == len(m_ref.rad)
    np.testing.assert_allclose(rad, m_mix.rad, rtol=1e-10)
    np.testing.assert_allclose(tra, m_mix.tra, rtol=1e-10, atol=1e-300)
    v = v1p + np.arange(len(rad)) * dv
    d = np.abs(bt(rad, v) - bt(m_ref.rad, v))
    print("spectrum: max |dBT| %.3e K, max rel dR %.3e" % (d.max(), np.max(np.abs(rad - m_ref.rad) / m_ref.rad)),
          flush=True)
    assert d.max() <= BT_TOL
    np.testing.assert_allclose(tra, m_ref.tra, rtol=1e-4, atol=1e-12)


def test_spectrum_merge_runs_match_layer_chain(gpu_ctx):
    """The spectrum merges runs of same-grid layers in one pass; its TOA
    radiance and transmittance equal the layer-by-layer XMERGE chain
    (lblx_xmerge_device) over the same layer optical depths, up to the
    summation order of the optical depths themselves."""
    tab = synth.config0_lines()
    gpu_ctx.load_lines(tab)
    run, lays, atm, it = _layers(600.0, 1100.0)
    tzu = np.asarray(atm["tlev"][1:], np.float64)
    tbnd, emis = float(atm["tlev"][0]), 0.98
    same = sum(1 for l in range(1, len(lays)) if lays[l].dv == lays[l - 1].dv)
    print("merge runs: layers on the grid of the layer below:", same, flush=True)
    assert same >= 10
    rad, tra, v1p, dv = gpu_ctx.spectrum(run, [abi.Layer.from_buffer_copy(x) for x in lays], tzu, tbnd, emis)
    sizes = [lib.od_size(run, L) for L in lays]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = gpu_ctx.alloc(8 * int(sum(sizes)))
    gpu_ctx.opdpth_column_device(run, [abi.Layer.from_buffer_copy(x) for x in lays], buf.ptr, offs)
    nmax = max(sizes)
    rt = [gpu_ctx.alloc(8 * nmax) for _ in range(4)]
    for l, L in enumerate(lays):
        dst_r, dst_t = rt[2 * (l & 1)], rt[2 * (l & 1) + 1]
        src_r, src_t = rt[2 * ((l - 1) & 1)], rt[2 * ((l - 1) & 1) + 1]
        gpu_ctx.xmerge_device(run, lays[l - 1] if l else None, L, float(tzu[l]), tbnd if l == 0 else 0.0, emis,
                              buf.ptr + 8 * int(offs[l]), src_r.ptr if l else 0, src_t.ptr if l else 0,
                              dst_r.ptr, dst_t.ptr)
    top = 2 * ((len(lays) - 1) & 1)
    r_chain = rt[top].to_host(np.float64, sizes[-1])
    t_chain = rt[top + 1].to_host(np.float64, sizes[-1])
    for b in rt + [buf]:
        b.free()
    np.testing.assert_allclose(rad, r_chain, rtol=1e-12)
    np.testing.assert_allclose(tra, t_chain, rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize("pair", [(0, 1), (1, 2), (5, 6), (10, 11)])
def test_xmerge_downwelling_layer_pairs(gpu_ctx, pair):
    """IPATHL = 3 (downwelling at an observer looking up): one merge step
    on two adjacent layers' grids (ITYPE 0, 2, 4, 5) with EMBND behind it,
    and the layer next to the observer (EMINIT, its own emission)."""
    a, b = pair
    tab = synth.config0_lines()
    run, lays, atm, it = _layers(600.0, 1100.0)
    La, Lb = lays[a], lays[b]
    print("downwelling xmerge pair", pair, "itype", int(it[b]), flush=True)
    O.od_layer(run, abi.Layer.from_buffer_copy(La), tab)
    pan_a = O.panels()
    O.od_layer(run, abi.Layer.from_buffer_copy(Lb), tab)
    pan_b = O.panels()
    na, nb = int(pan_a[2].sum()), int(pan_b[2].sum())
    rng = np.random.default_rng(200 + b)
    od = rng.exponential(0.4, nb)
    ro = 1e-6 * (1.0 + rng.random(na))
    to = rng.random(na)
    bufs = [gpu_ctx.alloc(8 * n) for n in (nb, na, na, nb, nb)]
    for buf, x in zip(bufs[:3], (od, ro, to)):
        buf.from_host(x)
    tzl = float(atm["tlev"][b])
    gpu_ctx.xmerge_path_device(run, La, Lb, 3, tzl, 250.0, 0.9, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr,
                               bufs[3].ptr, bufs[4].ptr)
    rg, tg = bufs[3].to_host(np.float64), bufs[4].to_host(np.float64)
    ref, q = O.merge_next(O.Merged(pan_a, La.dv, ro, to), od, pan_b, Lb.dv, Lb.tave, tzl, 1100.0,
                          ipathl=3, tbnd=250.0, emis=0.9)
    assert q == 0
    np.testing.assert_allclose(rg, ref.rad, rtol=1e-12)
    np.testing.assert_allclose(tg, ref.tra, rtol=1e-12, atol=1e-300)
    gpu_ctx.xmerge_path_device(run, None, Lb, 3, tzl, 288.0, 0.98, bufs[0].ptr, None, None,
                               bufs[3].ptr, bufs[4].ptr)
    rg, tg = bufs[3].to_host(np.float64), bufs[4].to_host(np.float64)
    ref = O.merge_first(od, pan_b, Lb.dv, Lb.tave, tzl, 0.0, 0.0)
    np.testing.assert_allclose(rg, ref.rad, rtol=1e-12)
    np.testing.assert_allclose(tg, ref.tra, rtol=1e-14)
    for buf in bufs:
        buf.free()


@pytest.mark.parametrize("band,cfg", [((600.0, 1100.0), 0), ((500.0, 2500.0), 1)])
def test_spectrum_downwelling_column(gpu_ctx, band, cfg):
    """Downwelling radiance at the surface through the 50-layer column
    (layers from the surface up, TZL = each layer's lower level): equal to
    the oracle's IPATHL = 3 merge chain over the GPU's own layer optical
    depths (1e-10), and for the 600-1100 cm-1 column within 0.01 K of the
    oracle's whole OPDPTH + XMERGE chain."""
    v1, v2 = band
    tab = synth.config0_lines() if cfg == 0 else synth.config1_lines()
    gpu_ctx.load_lines(tab)
    run, lays, atm, _ = _layers(v1, v2, icntnm=cfg)
    tzl = np.asarray(atm["tlev"][:-1], np.float64)
    rad, tra, v1p, dv = gpu_ctx.spectrum_path(run, [abi.Layer.from_buffer_copy(x) for x in lays], 3, tzl,
                                              0.0, 0.0)
    print("downwelling spectrum: %d points" % len(rad), flush=True)
    sizes = [lib.od_size(run, L) for L in lays]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = gpu_ctx.alloc(8 * int(sum(sizes)))
    gpu_ctx.opdpth_column_device(run, [abi.Layer.from_buffer_copy(x) for x in lays], buf.ptr, offs)
    allod = buf.to_host(np.float64)
    buf.free()
    jrad = 1
    m_ref = m_mix = None
    for l, L in enumerate(lays):
        inp = abi.Layer.from_buffer_copy(L)
        pan = lib.od_panels(run, inp)  # the panel headers do not depend on JRAD
        od_gpu = allod[offs[l]:offs[l] + sizes[l]]
        od_ref = None
        if cfg == 0:
            inp.jrad = jrad
            od_ref, _, _, jrad = O.od_layer(run, inp, tab)
            pan = O.panels()
        if l == 0:
            m_mix = O.merge_first(od_gpu, pan, L.dv, L.tave, tzl[l], 0.0, 0.0)
            if od_ref is not None:
                m_ref = O.merge_first(od_ref, pan, L.dv, L.tave, tzl[l], 0.0, 0.0)
        else:
            m_mix, q = O.merge_next(m_mix, od_gpu, pan, L.dv, L.tave, tzl[l], v2, ipathl=3)
            assert q == 0
            if od_ref is not None:
                m_ref, _ = O.merge_next(m_ref, od_ref, pan, L.dv, L.tave, tzl[l], v2, ipathl=3)
    np.testing.assert_allclose(rad, m_mix.rad, rtol=1e-10)
    np.testing.assert_allclose(tra, m_mix.tra, rtol=1e-10, atol=1e-300)
    if m_ref is not None:
        v = v1p + np.arange(len(rad)) * dv
        d = np.abs(bt(rad, v) - bt(m_ref.rad, v))
        print("downwelling: max |dBT| %.3e K" % d.max(), flush=True)
        assert d.max() <= BT_TOL
