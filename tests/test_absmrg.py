"""Pins for the oracle's IEMIT = 0 layer merge (ABINIT/ABSMRG,
src/xmerge.f90:270-685) -- CPU only.

On equal grids the merged optical depth is the sum.  On a finer new grid
the old optical depth is interpolated with 4-point Lagrange weights, so a
cubic old OD is reproduced exactly at every interpolated point (a phase or
tap off by one breaks it), and where the old file ends the remaining points
take the mean of its last two points.
"""
import numpy as np
import pytest

import oracle_lib as O
from test_merge import grid


def test_equal_grids_add():
    rng = np.random.default_rng(3)
    pan = grid(600.0, 0.01, 6000)
    a, b = rng.random(6000), rng.random(6000)
    out = O.absmrg(pan, a, 0.01, pan, b, 0.01)
    np.testing.assert_array_equal(out, a + b)


@pytest.mark.parametrize("itype", [1, 2, 3, 4, 5])
def test_finer_grid_interpolates_cubics(itype):
    dvk = 0.003
    dvo = dvk * (itype + 1) / itype
    v1 = 700.0
    no = 5000
    nk = int((no - 1) * dvo / dvk) + 1
    opan = grid(v1, dvo, no)
    kpan = grid(v1, dvk, nk)
    cubic = lambda v: 1.0 + 0.3 * (v - v1) - 0.02 * (v - v1) ** 2 + 1e-3 * (v - v1) ** 3
    vo = v1 + dvo * np.arange(no)
    vk = v1 + dvk * np.arange(nk)
    lay = np.sin(vk)  # the layer's own OD is added point by point
    out = O.absmrg(opan, cubic(vo), dvo, kpan, lay, dvk)
    # away from the first old interval (tap replicated) and the old file's end
    core = (vk > v1 + 2 * dvo) & (vk < vo[-1] - 3 * dvo)
    np.testing.assert_allclose(out[core], cubic(vk[core]) + lay[core], rtol=1e-11, atol=1e-11)
    # coincident points take the old value itself
    co = np.arange(0, nk, itype + 1)
    co = co[co < nk - 3 * (itype + 1)]
    np.testing.assert_allclose(out[co], cubic(vo[co // (itype + 1) * itype]) + lay[co], rtol=1e-13)
    # where the old file runs out: mean of its last two points
    last = (cubic(vo[-1]) + cubic(vo[-2])) * 0.5
    tail = out[-1] - lay[-1]
    assert tail == pytest.approx(last, rel=1e-13)


def test_coarser_grid_is_refused():
    pan = grid(600.0, 0.01, 100)
    pan2 = grid(600.0, 0.02, 50)
    with pytest.raises(RuntimeError):
        O.absmrg(pan, np.ones(100), 0.01, pan2, np.ones(50), 0.02)
