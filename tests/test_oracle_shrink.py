"""Pins for the oracle's far-wing line preparation (LINF4 + SHRINK) --
CPU only.

At the reference state (296 K, 1013.25 mb) every LINF4 correction is the
identity except the (1 + exp(-v/kT)) factor, so the shrunk list (TAPE9) is
fixed in closed form: SHRINK conserves the strength and the strength-
weighted centre of every group, keeps CO2 apart from the other molecules,
closes a group at the first line farther than DVR4/2 from the line that
closed the previous one, and LINF4 hands SHRINK batches of at least 1000 lines read in whole TAPE3
blocks.
"""
import ctypes as C

import numpy as np

import oracle_lib as O
from lblrtm_amd import abi

BETA0 = 1.4387752 / 296.0
W = 1e20


def f32(x):
    """Line strengths are stored as REAL*4 in TAPE3."""
    return np.asarray(x, np.float32).astype(np.float64)


def tape9(lines, v1=1000.0, v2=1100.0, dv=0.001):
    lay = abi.make_layer(1013.25, 296.0, [W, W], 1e24, dv, v1, v2)
    run = abi.make_run(v1, v2)
    cap = 400000
    out = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.int32)]
    blk = np.zeros(10000, np.int32)
    nb = C.c_int(0)
    n = O.lib().oracle_linf4(C.byref(run), C.byref(lay), C.byref(lines.struct()),
                             *[abi.ptr(a, C.c_double) for a in out[:4]], abi.ptr(out[4], C.c_int32),
                             abi.ptr(blk, C.c_int32), cap, 10000, C.byref(nb))
    return [a[:n] for a in out], blk[:nb.value], 64.0 * dv


def table(v, s, mol, shift=None):
    n = len(v)
    return abi.LineTable.from_lines(v, s, np.full(n, 0.07), np.full(n, 500.0), np.asarray(mol) + 100,
                                    np.full(n, 0.3), np.full(n, 0.7),
                                    np.zeros(n) if shift is None else shift)


def test_strength_and_centre_conserved():
    rng = np.random.default_rng(3)
    v = np.sort(rng.uniform(990.0, 1110.0, 6000))
    s = 10 ** rng.uniform(-24, -20, len(v))
    mol = rng.choice([1, 2], len(v))
    (vt, st, _, _, mt), _, _ = tape9(table(v, s, mol))
    sp = f32(s) * W * (1.0 + np.exp(-v * BETA0))
    keep = (v >= 1000.0 - 3 * 0.064 - 25.0)
    for m in (1, 2):
        sel = keep & (mol == m)
        tm = (mt == 2) if m == 2 else (mt != 2)
        np.testing.assert_allclose(st[tm].sum(), sp[sel].sum(), rtol=1e-12)
        np.testing.assert_allclose((st[tm] * vt[tm]).sum(), (sp[sel] * v[sel]).sum(), rtol=1e-12)
    assert len(vt) < keep.sum()


def test_grouping_known_answer():
    """Clusters a, b, c (H2O) and d (CO2) at DVR4/8 spacing, 1 cm-1 apart.
    A group closes at the first line farther than DVR4/2 from the line that
    closed the previous group (oprop.f90:3580-3600), so after a gap the first
    line stands alone: every cluster but the first gives a, (b+c), d; the
    first gives (a+b+c), d."""
    dvr4 = 0.064
    centres = np.arange(976.0, 1124.0, 1.0)
    v, s, mol = [], [], []
    for c in centres:
        for k, m in enumerate([1, 1, 1, 2]):
            v.append(c + k * dvr4 / 8)
            s.append(1e-21 * (k + 1))
            mol.append(m)
    v, s, mol = np.array(v), np.array(s), np.array(mol)
    (vt, st, _, _, mt), _, _ = tape9(table(v, s, mol))
    assert len(vt) == 3 * len(centres) - 1
    sp = f32(s) * W * (1.0 + np.exp(-v * BETA0))
    h2o = vt[mt != 2]
    first = (sp[0] * v[0] + sp[1] * v[1] + sp[2] * v[2]) / (sp[0] + sp[1] + sp[2])
    np.testing.assert_allclose(h2o[0], first, rtol=1e-15)
    np.testing.assert_allclose(h2o[1::2], v[4::4], rtol=1e-15)
    bc = (sp[5::4] * v[5::4] + sp[6::4] * v[6::4]) / (sp[5::4] + sp[6::4])
    np.testing.assert_allclose(h2o[2::2], bc, rtol=1e-15)
    np.testing.assert_allclose(vt[mt == 2], v[3::4], rtol=1e-15)


def test_isolated_lines_and_batches():
    """Lines 0.2 cm-1 apart (> DVR4/2): SHRINK is the identity, and LINF4
    writes TAPE9 in batches of 1000 lines (four 250-line TAPE3 blocks)."""
    v = 974.9 + 0.2 * np.arange(760)
    s = np.full(len(v), 1e-21)
    mol = np.ones(len(v), int)
    (vt, st, _, _, _), blk, _ = tape9(table(v, s, mol))
    # LINF4 reads from V1R4ST - BOUND4 - DVR4 through the first line past
    # V2R4ST + BOUND4 + 2 DVR4 (lblrtm.f90:6251-6255, oprop.f90:3130-3135)
    vhi = 1100.0 + 2 * 0.064 + 25.0 + 2 * 0.064
    last = np.argmax(v > vhi)
    use = (v >= 1000.0 - 3 * 0.064 - 25.0) & (np.arange(len(v)) <= last)
    # singleton groups: (v S) / S, equal to v within an ulp
    np.testing.assert_allclose(vt, v[use], rtol=2e-16, atol=0)
    np.testing.assert_allclose(st, f32(s[use]) * W * (1 + np.exp(-v[use] * BETA0)), rtol=1e-15)
    v2 = 974.9 + 0.05 * np.arange(3000)
    (vt2, _, _, _, _), blk2, _ = tape9(table(v2, np.full(3000, 1e-21), np.ones(3000, int)),
                                       v2=1070.0)
    assert blk2.sum() == len(vt2)
    assert np.all(blk2[:-1] >= 1000)
