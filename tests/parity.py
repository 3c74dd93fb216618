"""Shared parity helpers: the optical-depth tolerance of the north-star
contract (1e-4 relative where OD > 1e-6, 1e-10 absolute below)."""
import numpy as np

OD_REL = 1e-4
OD_ABS_FLOOR = 1e-6
OD_ABS = 1e-10


def assert_od_close(got, ref, what=""):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    big = np.abs(ref) > OD_ABS_FLOOR
    if big.any():
        rel = np.abs(got[big] - ref[big]) / np.abs(ref[big])
        assert rel.max() <= OD_REL, "%s: max rel %.3e at %d" % (what, rel.max(), int(np.argmax(rel)))
    small = ~big
    if small.any():
        d = np.abs(got[small] - ref[small]).max()
        assert d <= OD_ABS, "%s: max abs %.3e below the floor" % (what, d)
