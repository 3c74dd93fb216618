"""GPU IEMIT = 0 column optical depth (OPDPTH + ABINIT/ABSMRG) against the
oracle's merge, through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi, lib, synth
from parity import assert_od_close
from test_merge import grid

pytestmark = pytest.mark.gpu


def _column(v1, v2, precision=abi.FP64):
    atm = synth.atmosphere()
    run = abi.make_run(v1, v2, icntnm=1, precision=precision)
    dv, _ = lib.plan_dv(run, atm["pave"], atm["tave"], atm["wk"], atm["wbroad"], iemit=0)
    lays = [abi.make_layer(atm["pave"][l], atm["tave"][l], list(atm["wk"][l]), atm["wbroad"][l], dv[l], v1, v2)
            for l in range(len(dv))]
    return run, lays


def _gpu_column_od(ctx, run, lays):
    n = lib.od_size(run, lays[-1])
    buf = ctx.alloc(8 * n)
    v1p = ctx.column_od_device(run, [abi.Layer.from_buffer_copy(x) for x in lays], buf.ptr, n)
    out = buf.to_host(np.float64)
    buf.free()
    return out, v1p


def test_column_od_equals_oracle_merge_of_gpu_layers(gpu_ctx):
    """500,000 lines over 500-2500 cm-1, 50 layers: the GPU's merge chain of
    its own layer optical depths equals the oracle's ABINIT/ABSMRG walk."""
    gpu_ctx.load_lines(synth.config1_lines())
    run, lays = _column(500.0, 2500.0)
    tot, v1top = _gpu_column_od(gpu_ctx, run, lays)
    sizes = [lib.od_size(run, L) for L in lays]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    buf = gpu_ctx.alloc(8 * int(sum(sizes)))
    v1ps, _ = gpu_ctx.opdpth_column_device(run, [abi.Layer.from_buffer_copy(x) for x in lays], buf.ptr, offs)
    allod = buf.to_host(np.float64)
    buf.free()
    acc, pan_o, dv_o = None, None, None
    for l, L in enumerate(lays):
        od = allod[offs[l]:offs[l] + sizes[l]]
        pan = grid(v1ps[l], L.dv, sizes[l])
        acc = od.copy() if acc is None else O.absmrg(pan_o, acc, dv_o, pan, od, L.dv)
        pan_o, dv_o = pan, L.dv
    print("column OD: %d points on the top grid, v1p %.6f" % (len(tot), v1top), flush=True)
    assert v1top == v1ps[-1]
    np.testing.assert_allclose(tot, acc, rtol=1e-13, atol=0)


def test_column_od_narrow_band_matches_oracle(gpu_ctx):
    """900-1000 cm-1, 50 layers: the GPU column OD within the OD tolerance
    of the oracle's OPDPTH + ABINIT/ABSMRG chain."""
    tab = synth.config1_lines()
    gpu_ctx.load_lines(tab)
    run, lays = _column(900.0, 1000.0)
    tot, v1top = _gpu_column_od(gpu_ctx, run, lays)
    jrad, acc, pan_o, dv_o = 1, None, None, None
    for L in lays:
        inp = abi.Layer.from_buffer_copy(L)
        inp.jrad = jrad
        od, v1p, _, jrad = O.od_layer(run, inp, tab)
        pan = O.panels()
        acc = od.copy() if acc is None else O.absmrg(pan_o, acc, dv_o, pan, od, L.dv)
        pan_o, dv_o = pan, L.dv
    print("narrow column OD: %d points, max %.3e" % (len(tot), tot.max()), flush=True)
    assert_od_close(tot, acc, "column OD 900-1000")


def test_column_od_fp32_split(gpu_ctx):
    gpu_ctx.load_lines(synth.config1_lines())
    run64, lays = _column(500.0, 2500.0)
    run32, _ = _column(500.0, 2500.0, abi.FP32_SPLIT)
    a, _ = _gpu_column_od(gpu_ctx, run64, lays)
    b, _ = _gpu_column_od(gpu_ctx, run32, lays)
    rel = np.abs(b - a) / np.abs(a)
    print("column OD FP32 split vs FP64: max rel %.3e" % rel.max(), flush=True)
    assert_od_close(b, a, "column OD FP32 split")


def test_single_merge_step_and_abinit(gpu_ctx):
    """lblx_absmrg_device on two adjacent layers with random optical depths
    equals the oracle's ABINIT then ABSMRG."""
    gpu_ctx.load_lines(synth.config1_lines())
    run, lays = _column(1000.0, 1100.0)
    la, lb = lays[10], lays[11]
    na, nb = lib.od_size(run, la), lib.od_size(run, lb)
    tmp = gpu_ctx.alloc(8 * (na + nb))
    v1ps, _ = gpu_ctx.opdpth_column_device(run, [abi.Layer.from_buffer_copy(x) for x in (la, lb)], tmp.ptr,
                                           np.array([0, na], np.int64))
    tmp.free()
    rng = np.random.default_rng(5)
    oda, odb = rng.random(na), rng.random(nb)
    bufs = [gpu_ctx.alloc(8 * n) for n in (na, nb, na, nb)]
    bufs[0].from_host(oda)
    bufs[1].from_host(odb)
    gpu_ctx.absmrg_device(run, None, la, bufs[0].ptr, None, bufs[2].ptr)
    np.testing.assert_array_equal(bufs[2].to_host(np.float64), oda)
    gpu_ctx.absmrg_device(run, la, lb, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr)
    got = bufs[3].to_host(np.float64)
    ref = O.absmrg(grid(v1ps[0], la.dv, na), oda, la.dv, grid(v1ps[1], lb.dv, nb), odb, lb.dv)
    print("single ABSMRG step: DV %.4e -> %.4e, %d points" % (la.dv, lb.dv, nb), flush=True)
    np.testing.assert_allclose(got, ref, rtol=1e-15, atol=0)
    # a coarser layer grid is refused before any device work (ABSMRG's STOP)
    assert lays[0].dv > lays[-1].dv
    with pytest.raises(lib.LblxError) as e:
        gpu_ctx.absmrg_device(run, lays[-1], lays[0], bufs[0].ptr, bufs[1].ptr, bufs[2].ptr)
    assert e.value.code == -9
    for b in bufs:
        b.free()
