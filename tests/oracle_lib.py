"""ctypes loader for the CPU oracle (oracle/liblblx_oracle.so) -- tests only."""
import ctypes as C
import os

import numpy as np

from lblrtm_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liblblx_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        L = C.CDLL(path)
        L.oracle_od_layer.restype = C.c_int
        L.oracle_od_layer.argtypes = [
            C.POINTER(abi.Run), C.POINTER(abi.Layer), C.POINTER(abi.Lines),
            C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double, C.c_int,
            C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64),
            C.POINTER(C.c_double), C.POINTER(abi.Stats)]
        L.oracle_plan_dv.restype = C.c_int
        L.oracle_plan_dv.argtypes = [
            C.POINTER(abi.Run), C.c_int, C.c_int, C.POINTER(C.c_double),
            C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
            C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int32)]
        L.oracle_armstrong.restype = C.c_double
        L.oracle_armstrong.argtypes = [C.c_double, C.c_double]
        L.oracle_voigt_tables.restype = None
        L.oracle_voigt_tables.argtypes = [C.POINTER(C.c_double)] * 2
        L.oracle_shape_tables.restype = None
        L.oracle_shape_tables.argtypes = [C.POINTER(C.c_double)] * 5
        L.tips_scor.restype = None
        L.tips_scor.argtypes = [C.c_double, C.POINTER(C.c_double)]
        L.radfn.restype = C.c_double
        L.radfn.argtypes = [C.c_double, C.c_double]
        L.oracle_linf4.restype = C.c_int
        L.oracle_linf4.argtypes = [
            C.POINTER(abi.Run), C.POINTER(abi.Layer), C.POINTER(abi.Lines),
            C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
            C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
            C.c_int, C.c_int, C.POINTER(C.c_int)]
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.oracle_linf4_spp.restype = C.c_int
        L.oracle_linf4_spp.argtypes = [dp, C.c_int]
        L.oracle_panels.restype = C.c_int
        L.oracle_panels.argtypes = [dp, dp, ip, C.c_int]
        L.oracle_planck.restype = C.c_double
        L.oracle_planck.argtypes = [C.c_double, C.c_double]
        L.oracle_merge_first.restype = C.c_int
        L.oracle_merge_first.argtypes = [C.c_int, dp, ip, C.c_double, dp, C.c_double, C.c_double,
                                         C.c_double, C.c_double, dp, dp]
        L.oracle_merge_next.restype = C.c_int
        L.oracle_merge_next.argtypes = [C.c_int, dp, dp, ip, C.c_double, dp, dp, C.c_int, dp, dp,
                                        ip, C.c_double, dp, C.c_double, C.c_double, C.c_double,
                                        dp, dp]
        L.oracle_merge_next_path.restype = C.c_int
        L.oracle_merge_next_path.argtypes = [C.c_int, dp, dp, ip, C.c_double, dp, dp, C.c_int, dp, dp,
                                             ip, C.c_double, dp, C.c_double, C.c_double, C.c_double,
                                             C.c_int, C.c_double, C.c_double, dp, dp]
        L.oracle_cont_grid.restype = C.c_int
        L.oracle_cont_grid.argtypes = [C.POINTER(abi.Layer), dp, dp, dp, ip]
        L.oracle_contnm.restype = C.c_int
        L.oracle_contnm.argtypes = [C.POINTER(abi.Run), C.POINTER(abi.Layer), C.c_int, dp, C.c_int, ip]
        L.oracle_fftscn.restype = C.c_int
        L.oracle_fftscn.argtypes = [C.POINTER(abi.Scan), dp, C.c_double, C.c_double, C.c_int64, dp,
                                    C.c_int64, C.POINTER(C.c_int64), dp, dp]
        L.oracle_scanfn.restype = C.c_int
        L.oracle_scanfn.argtypes = [C.POINTER(abi.ScanFn), C.c_int, dp, C.c_double, C.c_double, C.c_int64, dp,
                                    C.c_int64, C.POINTER(C.c_int64), dp, dp]
        L.oracle_merge_quirks.restype = C.c_int64
        L.oracle_merge_quirks.argtypes = []
        _LIB = L
    return _LIB


def plan_dv(run, pave, tave, wk, wbroad, iemit=0):
    pave = np.ascontiguousarray(pave, np.float64)
    tave = np.ascontiguousarray(tave, np.float64)
    wk = np.ascontiguousarray(np.atleast_2d(wk), np.float64)
    wbroad = np.ascontiguousarray(wbroad, np.float64)
    nlay, nmol = wk.shape
    dv = np.zeros(nlay)
    it = np.zeros(nlay, np.int32)
    rc = lib().oracle_plan_dv(C.byref(run), nlay, nmol, abi.ptr(pave, C.c_double),
                              abi.ptr(tave, C.c_double), abi.ptr(wk, C.c_double),
                              abi.ptr(wbroad, C.c_double), iemit,
                              abi.ptr(dv, C.c_double), abi.ptr(it, C.c_int32))
    if rc:
        raise RuntimeError("oracle_plan_dv rc=%d" % rc)
    return dv, it


def od_layer(run, layer, table, absrb=None, v1abs=0.0, v2abs=0.0, dvabs=0.0):
    """Run the oracle for one layer; returns (od, v1p, stats, layer.jrad)."""
    lines = table.struct()
    n = C.c_int64(0)
    v1p = C.c_double(0.0)
    st = abi.Stats()
    ab = None
    nab = 0
    if absrb is not None:
        absrb = np.ascontiguousarray(absrb, np.float64)
        ab = abi.ptr(absrb, C.c_double)
        nab = len(absrb)
    # size query, then the real run
    rc = lib().oracle_od_layer(C.byref(run), C.byref(layer), C.byref(lines), ab, v1abs, v2abs,
                               dvabs, nab, None, 0, C.byref(n), C.byref(v1p), C.byref(st))
    if rc not in (0, abi.E_SIZE):
        raise RuntimeError("oracle_od_layer rc=%d" % rc)
    jrad_after = layer.jrad
    out = np.zeros(n.value)
    lay2 = abi.Layer.from_buffer_copy(layer)
    rc = lib().oracle_od_layer(C.byref(run), C.byref(lay2), C.byref(lines), ab, v1abs, v2abs,
                               dvabs, nab, abi.ptr(out, C.c_double), n.value, C.byref(n),
                               C.byref(v1p), C.byref(st))
    if rc:
        raise RuntimeError("oracle_od_layer rc=%d" % rc)
    return out, v1p.value, st.as_dict(), jrad_after


def panels():
    """Panel headers (v1p, v2p, nlim) of the last od_layer run."""
    cap = 100000
    v1 = np.zeros(cap)
    v2 = np.zeros(cap)
    nl = np.zeros(cap, np.int32)
    n = lib().oracle_panels(abi.ptr(v1, C.c_double), abi.ptr(v2, C.c_double), abi.ptr(nl, C.c_int), cap)
    return v1[:n].copy(), v2[:n].copy(), nl[:n].copy()


class Merged:
    """Accumulated XMERGE result: panels of the grid and the R/T arrays."""

    def __init__(self, pan, dv, rad, tra):
        self.pan, self.dv, self.rad, self.tra = pan, dv, rad, tra


def _d(a):
    return abi.ptr(np.ascontiguousarray(a, np.float64), C.c_double)


def merge_first(layer_od, pan, dv, tave, tzu, tbnd, emis):
    v1, v2, nl = pan
    n = int(nl.sum())
    rad = np.zeros(n)
    tra = np.zeros(n)
    od = np.ascontiguousarray(layer_od, np.float64)
    nlc = np.ascontiguousarray(nl, np.int32)
    rc = lib().oracle_merge_first(len(v1), _d(v1), abi.ptr(nlc, C.c_int), dv, abi.ptr(od, C.c_double),
                                  tave, tzu, tbnd, emis, abi.ptr(rad, C.c_double), abi.ptr(tra, C.c_double))
    assert rc == 0
    return Merged(pan, dv, rad, tra)


def merge_next(old, layer_od, pan, dv, tave, tzu, v2run, ipathl=1, tbnd=0.0, emis=0.0):
    """Returns (Merged, quirks) -- quirks counts RADNN reads past the old
    points read so far.  ipathl = 3: downwelling (tzu carries TZL; tbnd > 0
    adds EMBND)."""
    v1o, v2o, nlo = old.pan
    v1, v2, nl = pan
    n = int(nl.sum())
    rad = np.zeros(n)
    tra = np.zeros(n)
    od = np.ascontiguousarray(layer_od, np.float64)
    nloc = np.ascontiguousarray(nlo, np.int32)
    nlc = np.ascontiguousarray(nl, np.int32)
    ro = np.ascontiguousarray(old.rad)
    to = np.ascontiguousarray(old.tra)
    rc = lib().oracle_merge_next_path(len(v1o), _d(v1o), _d(v2o), abi.ptr(nloc, C.c_int), old.dv,
                                      abi.ptr(ro, C.c_double), abi.ptr(to, C.c_double), len(v1), _d(v1),
                                      _d(v2), abi.ptr(nlc, C.c_int), dv, abi.ptr(od, C.c_double), tave, tzu,
                                      v2run, ipathl, tbnd, emis, abi.ptr(rad, C.c_double),
                                      abi.ptr(tra, C.c_double))
    if rc:
        raise RuntimeError("oracle_merge_next rc=%d" % rc)
    return Merged(pan, dv, rad, tra), int(lib().oracle_merge_quirks())


def contnm(run, layer, jrad):
    """ABSRB of CONTNM for one layer: (absrb, v1abs, dvabs)."""
    a, b, d = C.c_double(), C.c_double(), C.c_double()
    n = C.c_int()
    rc = lib().oracle_cont_grid(C.byref(layer), C.byref(a), C.byref(b), C.byref(d), C.byref(n))
    if rc:
        raise RuntimeError("oracle_cont_grid rc=%d" % rc)
    out = np.zeros(n.value)
    rc = lib().oracle_contnm(C.byref(run), C.byref(layer), jrad, abi.ptr(out, C.c_double), n.value,
                             C.byref(n))
    if rc:
        raise RuntimeError("oracle_contnm rc=%d" % rc)
    return out, a.value, d.value


def fftscn(scan, spectrum, v1_in, dv_in, cap=4000000):
    """The oracle's FFTSCN: (out, v1, dv)."""
    x = np.ascontiguousarray(spectrum, np.float64)
    out = np.zeros(cap)
    n = C.c_int64(0)
    v1, dv = C.c_double(0.0), C.c_double(0.0)
    rc = lib().oracle_fftscn(C.byref(scan), abi.ptr(x, C.c_double), v1_in, dv_in, len(x),
                             abi.ptr(out, C.c_double), cap, C.byref(n), C.byref(v1), C.byref(dv))
    if rc:
        raise RuntimeError("oracle_fftscn rc=%d" % rc)
    return out[:n.value].copy(), v1.value, dv.value


def scanfn(scan, spectrum, v1_in, dv_in, iemit=1, cap=4000000):
    """The oracle's SCANFN, panel by panel: (out, v1, dv)."""
    x = np.ascontiguousarray(spectrum, np.float64)
    out = np.zeros(cap)
    n = C.c_int64(0)
    v1, dv = C.c_double(0.0), C.c_double(0.0)
    rc = lib().oracle_scanfn(C.byref(scan), iemit, abi.ptr(x, C.c_double), v1_in, dv_in, len(x),
                             abi.ptr(out, C.c_double), cap, C.byref(n), C.byref(v1), C.byref(dv))
    if rc:
        raise RuntimeError("oracle_scanfn rc=%d" % rc)
    return out[:n.value].copy(), v1.value, dv.value


def voigt_tables():
    """Exact-Voigt build tables: f [3, 101, 2601] (table, zeta index, IZ-1)
    and coefficients a_1 a_2 a_3 b_1 b_2 b_3 [6, 101]."""
    f = np.zeros(3 * 101 * 2601)
    c = np.zeros(6 * 101)
    lib().oracle_voigt_tables(f.ctypes.data_as(C.POINTER(C.c_double)), c.ctypes.data_as(C.POINTER(C.c_double)))
    return f.reshape(3, 101, 2601), c.reshape(6, 101)


def absmrg(old_pan, old_od, dvo, pan, layer_od, dvk):
    """ABINIT/ABSMRG (IEMIT = 0) of the oracle: accumulated OD on the new
    layer's panels.  old_pan / pan: (v1p, v2p, nlim) arrays."""
    v1o, _, nlo = old_pan
    v1, _, nl = pan
    n = int(np.sum(nl))
    out = np.zeros(n)
    nloc = np.ascontiguousarray(nlo, np.int32)
    nlc = np.ascontiguousarray(nl, np.int32)
    oo = np.ascontiguousarray(old_od, np.float64)
    od = np.ascontiguousarray(layer_od, np.float64)
    L = lib()
    L.oracle_absmrg.restype = C.c_int
    rc = L.oracle_absmrg(len(v1o), _d(np.ascontiguousarray(v1o, np.float64)), abi.ptr(nloc, C.c_int),
                         C.c_double(dvo), abi.ptr(oo, C.c_double), len(v1),
                         _d(np.ascontiguousarray(v1, np.float64)), abi.ptr(nlc, C.c_int), C.c_double(dvk),
                         abi.ptr(od, C.c_double), abi.ptr(out, C.c_double))
    if rc < 0:
        raise RuntimeError("oracle_absmrg rc=%d" % rc)
    return out


def od_layer_dvout(run, layer, table, dvout, cap=20000000):
    """OPDPTH with PNLINT onto DVOUT: (values, (v1o, v2o, nlimo))."""
    out = np.zeros(cap)
    pc = 20000
    v1o, v2o = np.zeros(pc), np.zeros(pc)
    nl = np.zeros(pc, np.int32)
    nw = C.c_int64()
    L = lib()
    L.oracle_od_layer_dvout.restype = C.c_int
    lay = abi.Layer.from_buffer_copy(layer)
    lines = table.struct()
    rc = L.oracle_od_layer_dvout(C.byref(run), C.byref(lay), C.byref(lines), C.c_double(dvout),
                                 abi.ptr(out, C.c_double), C.c_int64(cap), C.byref(nw), abi.ptr(v1o, C.c_double),
                                 abi.ptr(v2o, C.c_double), abi.ptr(nl, C.c_int), pc)
    if rc < 0:
        raise RuntimeError("oracle_od_layer_dvout rc=%d" % rc)
    return out[:nw.value], (v1o[:rc], v2o[:rc], nl[:rc])


def set_icflg(icflg):
    """CONTNM's /CDERIV/ ICFLG for the following oracle calls (0: off)."""
    lib().oracle_set_icflg(int(icflg))


def linf4(run, layer, table, cap=2000000):
    """The oracle's shrunk far-wing list (LINF4 + SHRINK) for one layer:
    dict of vnu, sp, alfa, epp, mol, spp and the TAPE9 block counts."""
    arrs = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.int32)]
    blk = np.zeros(100000, np.int32)
    nb = C.c_int(0)
    n = lib().oracle_linf4(C.byref(run), C.byref(abi.Layer.from_buffer_copy(layer)), C.byref(table.struct()),
                           *[abi.ptr(a, C.c_double) for a in arrs[:4]], abi.ptr(arrs[4], C.c_int32),
                           abi.ptr(blk, C.c_int32), cap, len(blk), C.byref(nb))
    if n < 0:
        raise RuntimeError("oracle_linf4 rc=%d" % n)
    k = min(n, cap)
    spp = np.zeros(max(k, 1))
    lib().oracle_linf4_spp(abi.ptr(spp, C.c_double), k)
    return dict(vnu=arrs[0][:k], sp=arrs[1][:k], alfa=arrs[2][:k], epp=arrs[3][:k], mol=arrs[4][:k],
                spp=spp[:k], bcount=blk[:nb.value])
