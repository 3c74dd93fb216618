"""The allocation-free FFTSCN entry (lblx_fftscn_noalloc_device) against
lblx_fftscn_device through the C ABI: two calls in a row on the same
device scratch, the second with a smaller input, each equal to the
allocating entry's output; a request beyond the scratch is refused."""
import numpy as np
import pytest

from lblrtm_amd import abi, lib

pytestmark = pytest.mark.gpu

V1IN, DVIN = 500.0, 3.683e-4


def _spectrum(v1, v2, seed):
    rng = np.random.default_rng(seed)
    n = int((v2 - v1) / DVIN) + 1
    nu = v1 + DVIN * np.arange(n)
    return 1e-6 * (1.0 + 0.3 * np.sin(nu / 7.0)) * (1.0 + 0.05 * rng.random(n))


def _both(ctx, sc, x, v1):
    n, _, _ = lib.scan_size(sc, v1, DVIN, len(x))
    bi = ctx.alloc(x.nbytes)
    ba, bb = ctx.alloc(8 * n), ctx.alloc(8 * n)
    try:
        bi.from_host(x)
        na, v1a, dva = ctx.fftscn_device(sc, bi.ptr, v1, DVIN, len(x), ba.ptr, n)
        nb, v1b, dvb = ctx.fftscn_noalloc_device(sc, bi.ptr, v1, DVIN, len(x), bb.ptr, n)
        ctx.sync()
        assert (na, v1a, dva) == (nb, v1b, dvb) == (n, v1a, dva)
        return ba.to_host(np.float64, n), bb.to_host(np.float64, n)
    finally:
        bi.free()
        ba.free()
        bb.free()


@pytest.mark.parametrize("dvout", [0.1, 0.0])
def test_noalloc_scan_equals_allocating_entry(gpu_ctx, dvout):
    big = _spectrum(V1IN, 2500.0, 7)        # the column workload's band
    small = _spectrum(V1IN, 900.0, 8)       # a smaller second input
    for x, v2o in ((big, 2480.0), (small, 880.0)):
        sc = abi.make_scan(1.0, 520.0, v2o, jemit=1, jfnin=-6, dvout=dvout)
        print("noalloc scan: %d input points, dvout %g" % (len(x), dvout), flush=True)
        a, b = _both(gpu_ctx, sc, x, V1IN)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_noalloc_scan_refuses_beyond_scratch(gpu_ctx):
    """No boxcar over 200,001 points needs a 2^20-point FFT, twice the
    scratch: refused with LBLX_E_SCRATCH, nothing written."""
    n = 200001
    x = np.ones(n)
    sc = abi.make_scan(1.0, 515.0, 550.0, mratin=-1, dvout=0.05)
    nout, _, _ = lib.scan_size(sc, V1IN, DVIN, n)
    bi, bo = gpu_ctx.alloc(x.nbytes), gpu_ctx.alloc(8 * nout)
    try:
        bi.from_host(x)
        with pytest.raises(lib.LblxError) as e:
            gpu_ctx.fftscn_noalloc_device(sc, bi.ptr, V1IN, DVIN, n, bo.ptr, nout)
        assert e.value.code == abi.E_SCRATCH
    finally:
        bi.free()
        bo.free()
