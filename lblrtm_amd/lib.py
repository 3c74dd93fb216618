"""Python host mirror of the C ABI in include/lblx.h (ctypes over
lblrtm_amd/liblblx.so).

The reference drives the path through OPDPTH (src/lblrtm.f90:6135),
XMERGE (src/xmerge.f90:25) and FFTSCN (src/fftscn_dbl.f90:25); the
methods of Context carry the same names and argument meaning.  There is no
CPU fallback: without the built library or a HIP device every call raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "liblblx.so")
_LIB = None
NSTAGE = 14


class LblxError(RuntimeError):
    def __init__(self, code, where):
        msg = "%s failed: %s (%d)" % (where, strerror(code), code)
        super().__init__(msg)
        self.code = code


def load():
    """Load liblblx.so (raises when it is not built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError("liblblx.so is not built; run python -m lblrtm_amd.build")
    L = C.CDLL(LIBPATH)
    P = C.POINTER
    sig = {
        "lblx_strerror": (C.c_char_p, [C.c_int]),
        "lblx_device_count": (C.c_int, []),
        "lblx_create": (C.c_int, [P(C.c_void_p), C.c_int]),
        "lblx_destroy": (None, [C.c_void_p]),
        "lblx_load_tape3": (C.c_int, [C.c_void_p, C.c_char_p]),
        "lblx_load_lines": (C.c_int, [C.c_void_p, P(abi.Lines)]),
        "lblx_plan_dv": (C.c_int, [P(abi.Run), C.c_int, C.c_int, P(C.c_double), P(C.c_double),
                                   P(C.c_double), P(C.c_double), C.c_int, P(C.c_double),
                                   P(C.c_int32)]),
        "lblx_od_size": (C.c_int, [P(abi.Run), P(abi.Layer), P(C.c_int64)]),
        "lblx_od_layer": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_void_p, C.c_int64,
                                    P(C.c_double), P(abi.Stats)]),
        "lblx_od_layer_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_void_p,
                                           C.c_int64, P(C.c_double), P(abi.Stats)]),
        "lblx_sync": (C.c_int, [C.c_void_p]),
        "lblx_od_column_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int,
                                            C.c_void_p, P(C.c_int64), P(C.c_double),
                                            P(abi.Stats)]),
        "lblx_debug_tape9": (C.c_int, [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double),
                                       P(C.c_double), P(C.c_double), P(C.c_int32), C.c_int,
                                       P(C.c_int32)]),
        "lblx_debug_tape9_spp": (C.c_int, [C.c_void_p, C.c_int, P(C.c_double), C.c_int, P(C.c_int32)]),
        "lblx_timing": (C.c_int, [C.c_void_p, C.c_int]),
        "lblx_timing_get": (C.c_int, [C.c_void_p, C.c_int, P(C.c_char_p), P(C.c_double),
                                      P(C.c_int64)]),
        "lblx_stream": (C.c_void_p, [C.c_void_p]),
        "lblx_dev_alloc": (C.c_void_p, [C.c_void_p, C.c_int64]),
        "lblx_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
        "lblx_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int]),
        "lblx_contnm": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), P(C.c_double), C.c_int,
                                  P(C.c_int), P(C.c_double), P(C.c_double)]),
        "lblx_scan_size": (C.c_int, [P(abi.Scan), C.c_double, C.c_double, C.c_int64, P(C.c_int64),
                                     P(C.c_double), P(C.c_double)]),
        "lblx_fftscn_device": (C.c_int, [C.c_void_p, P(abi.Scan), C.c_void_p, C.c_double, C.c_double,
                                         C.c_int64, C.c_void_p, C.c_int64, P(C.c_int64), P(C.c_double),
                                         P(C.c_double)]),
        "lblx_scanfn_size": (C.c_int, [P(abi.ScanFn), C.c_int, C.c_double, C.c_double, C.c_int64,
                                       P(C.c_int64), P(C.c_double), P(C.c_double)]),
        "lblx_scanfn_device": (C.c_int, [C.c_void_p, P(abi.ScanFn), C.c_int, C.c_void_p, C.c_double,
                                         C.c_double, C.c_int64, C.c_void_p, C.c_int64]),
        "lblx_od_panels": (C.c_int, [P(abi.Run), P(abi.Layer), C.c_int, P(C.c_double), P(C.c_double),
                                     P(C.c_int32), P(C.c_int)]),
        "lblx_write_panels": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_void_p, C.c_int64, C.c_int,
                                        P(C.c_double), P(C.c_double), C.c_double, P(C.c_int32), P(C.c_double),
                                        P(C.c_double), C.c_int]),
        "lblx_fftscn_noalloc_device": (C.c_int, [C.c_void_p, P(abi.Scan), C.c_void_p, C.c_double, C.c_double,
                                                 C.c_int64, C.c_void_p, C.c_int64, P(C.c_int64), P(C.c_double),
                                                 P(C.c_double)]),
        "lblx_jacobian_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, P(C.c_double),
                                           C.c_double, C.c_double, P(C.c_int), C.c_int, C.c_double,
                                           C.c_double, C.c_void_p, C.c_void_p, C.c_int64,
                                           P(C.c_double)]),
        "lblx_spectra_batch_device": (C.c_int, [C.c_void_p, P(abi.Run), C.c_int, P(C.c_int), P(abi.Layer),
                                                P(C.c_double), P(C.c_double), P(C.c_double), C.c_void_p,
                                                C.c_void_p, P(C.c_int64), P(C.c_double)]),
        "lblx_jacobians_batch_device": (C.c_int, [C.c_void_p, P(abi.Run), C.c_int, P(C.c_int), P(abi.Layer),
                                                  P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_int), C.c_int,
                                                  C.c_int, C.c_double, C.c_double, C.c_void_p, P(C.c_int64),
                                                  C.c_void_p, P(C.c_int64), P(C.c_double)]),
        "lblx_autotune": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, P(C.c_int)]),
        "lblx_set_tiles": (C.c_int, [C.c_void_p, P(C.c_int)]),
        "lblx_get_tuning": (C.c_int, [C.c_void_p, P(abi.Tuning)]),
        "lblx_set_tuning": (C.c_int, [C.c_void_p, P(abi.Tuning)]),
        "lblx_tuning_save": (C.c_int, [C.c_void_p, C.c_char_p]),
        "lblx_tuning_load": (C.c_int, [C.c_void_p, C.c_char_p]),
        "lblx_voigt_tables": (C.c_int, [P(C.c_double), P(C.c_double)]),
        "lblx_jacobian_aj_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, P(C.c_double),
                                              C.c_double, C.c_double, P(C.c_int), C.c_int, C.c_void_p,
                                              C.c_void_p, C.c_int64, P(C.c_double)]),
        "lblx_jacobian_tl_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, P(C.c_double),
                                              C.c_double, C.c_double, P(C.c_int), C.c_int, C.c_void_p,
                                              C.c_void_p, C.c_int64, P(C.c_double)]),
        "lblx_od_tangent_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, P(C.c_int), C.c_int,
                                             C.c_void_p, C.c_void_p, C.c_int64]),
        "lblx_backward_batch": (C.c_int, [C.c_void_p, P(abi.Run), C.c_int, P(C.c_int), P(abi.Layer),
                                          P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_int), C.c_int,
                                          C.c_void_p, P(C.c_int64), C.c_void_p, C.c_void_p, P(C.c_int64),
                                          P(C.c_double)]),
        "lblx_od_dvout_size": (C.c_int, [P(abi.Run), P(abi.Layer), C.c_double, P(C.c_int64)]),
        "lblx_od_layer_dvout_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_double, C.c_void_p,
                                                 C.c_int64, P(C.c_int64), P(C.c_double)]),
        "lblx_absmrg_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), P(abi.Layer), C.c_void_p,
                                         C.c_void_p, C.c_void_p]),
        "lblx_column_od_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, C.c_void_p,
                                            C.c_int64, P(C.c_double), C.c_void_p]),
        "lblx_merge_size": (C.c_int, [P(abi.Run), P(abi.Layer), P(abi.Layer), P(C.c_int64)]),
        "lblx_xmerge_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), P(abi.Layer),
                                         C.c_double, C.c_double, C.c_double, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lblx_spectrum_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int,
                                           P(C.c_double), C.c_double, C.c_double, C.c_void_p,
                                           C.c_void_p, C.c_int64, P(C.c_double), P(abi.Stats)]),
        "lblx_spectrum_path_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), C.c_int, C.c_int,
                                                P(C.c_double), C.c_double, C.c_double, C.c_void_p,
                                                C.c_void_p, C.c_int64, P(C.c_double), P(abi.Stats)]),
        "lblx_xmerge_path_device": (C.c_int, [C.c_void_p, P(abi.Run), P(abi.Layer), P(abi.Layer), C.c_int,
                                              C.c_double, C.c_double, C.c_double, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def strerror(code):
    try:
        return load().lblx_strerror(code).decode()
    except Exception:
        return "error %d" % code


def _chk(rc, where):
    if rc != 0:
        raise LblxError(rc, where)


def plan_dv(run, pave, tave, wk, wbroad, iemit=0):
    """Layer DV chain of PATH (lblrtm.f90:5407-5500)."""
    L = load()
    pave = np.ascontiguousarray(pave, np.float64)
    tave = np.ascontiguousarray(tave, np.float64)
    wk = np.ascontiguousarray(np.atleast_2d(wk), np.float64)
    wbroad = np.ascontiguousarray(wbroad, np.float64)
    nlay, nmol = wk.shape
    dv = np.zeros(nlay)
    it = np.zeros(nlay, np.int32)
    _chk(L.lblx_plan_dv(C.byref(run), nlay, nmol, abi.ptr(pave, C.c_double),
                        abi.ptr(tave, C.c_double), abi.ptr(wk, C.c_double),
                        abi.ptr(wbroad, C.c_double), iemit, abi.ptr(dv, C.c_double),
                        abi.ptr(it, C.c_int32)), "lblx_plan_dv")
    return dv, it


def od_panels(run, layer):
    """Panel headers OPDPTH writes for the layer: (v1p, v2p, nlim) arrays."""
    L = load()
    n = C.c_int(0)
    L.lblx_od_panels(C.byref(run), C.byref(layer), 0, None, None, None, C.byref(n))
    v1 = np.zeros(n.value)
    v2 = np.zeros(n.value)
    nl = np.zeros(n.value, np.int32)
    _chk(L.lblx_od_panels(C.byref(run), C.byref(layer), n.value, abi.ptr(v1, C.c_double), abi.ptr(v2, C.c_double),
                          abi.ptr(nl, C.c_int32), C.byref(n)), "lblx_od_panels")
    return v1, v2, nl


def write_panels(path, v1p, v2p, dvp, nlim, a, b=None, filhdr=None, word=8, endfil=False, append=False):
    """The reference's panel file (TAPE10 with a only, TAPE11/12 with a and
    b) through lblx_write_panels; filhdr is the header record's bytes."""
    v1p = np.ascontiguousarray(v1p, np.float64)
    v2p = np.ascontiguousarray(v2p, np.float64)
    nlim = np.ascontiguousarray(nlim, np.int32)
    a = np.ascontiguousarray(a, np.float64)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    hb = None if filhdr is None else bytes(filhdr)
    _chk(load().lblx_write_panels(str(path).encode(), int(append), word, hb, len(hb) if hb else 0, len(nlim),
                                  abi.ptr(v1p, C.c_double), abi.ptr(v2p, C.c_double), dvp,
                                  abi.ptr(nlim, C.c_int32), abi.ptr(a, C.c_double),
                                  abi.ptr(bb, C.c_double) if bb is not None else None, int(endfil)),
         "lblx_write_panels")


def od_size(run, layer):
    n = C.c_int64(0)
    _chk(load().lblx_od_size(C.byref(run), C.byref(layer), C.byref(n)), "lblx_od_size")
    return n.value


def merge_size(run, layer_old, layer):
    """Points of the merged result on layer's grid; raises when RADMRG's
    panel walk over these two grids is not supported."""
    n = C.c_int64(0)
    _chk(load().lblx_merge_size(C.byref(run), C.byref(layer_old) if layer_old is not None else None,
                                C.byref(layer), C.byref(n)), "lblx_merge_size")
    return n.value


def voigt_tables():
    """The library's exact-Voigt tables (host build): f [3, 101, 2601] and
    coefficients a_1 a_2 a_3 b_1 b_2 b_3 [6, 101]."""
    f = np.zeros(3 * 101 * 2601)
    c = np.zeros(6 * 101)
    _chk(load().lblx_voigt_tables(abi.ptr(f, C.c_double), abi.ptr(c, C.c_double)), "lblx_voigt_tables")
    return f.reshape(3, 101, 2601), c.reshape(6, 101)


def od_dvout_size(run, layer, dvout):
    n = C.c_int64()
    _chk(load().lblx_od_dvout_size(C.byref(run), C.byref(layer), dvout, C.byref(n)), "lblx_od_dvout_size")
    return n.value


def scan_size(scan, v1_in, dv_in, n_in):
    """(points, v1, dv) of FFTSCN's output for an input on v1_in + k dv_in."""
    n = C.c_int64(0)
    v1 = C.c_double(0.0)
    dv = C.c_double(0.0)
    _chk(load().lblx_scan_size(C.byref(scan), v1_in, dv_in, n_in, C.byref(n), C.byref(v1), C.byref(dv)),
         "lblx_scan_size")
    return n.value, v1.value, dv.value


def scanfn_size(scan, iemit, v1_in, dv_in, n_in):
    """(points, v1, dv) of SCANFN's output for an input on v1_in + k dv_in."""
    n = C.c_int64(0)
    v1 = C.c_double(0.0)
    dv = C.c_double(0.0)
    _chk(load().lblx_scanfn_size(C.byref(scan), iemit, v1_in, dv_in, n_in, C.byref(n), C.byref(v1),
                                 C.byref(dv)), "lblx_scanfn_size")
    return n.value, v1.value, dv.value


class Context:
    """One device context: resident line table, stream and work buffers."""

    def __init__(self, device=0):
        L = load()
        h = C.c_void_p()
        _chk(L.lblx_create(C.byref(h), device), "lblx_create")
        self._h = h
        self._lib = L

    def close(self):
        if self._h:
            self._lib.lblx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_lines(self, table):
        s = table.struct()
        _chk(self._lib.lblx_load_lines(self._h, C.byref(s)), "lblx_load_lines")

    def load_tape3(self, path):
        _chk(self._lib.lblx_load_tape3(self._h, str(path).encode()), "lblx_load_tape3")

    def opdpth(self, run, layer):
        """Layer optical depth (OPDPTH/HIRAC1) into host memory.
        Returns (od, v1p, stats); layer.jrad is updated in place."""
        n = od_size(run, layer)
        dt = np.float32 if run.precision == abi.FP32_SPLIT else np.float64
        out = np.zeros(n, dt)
        v1p = C.c_double(0.0)
        st = abi.Stats()
        _chk(self._lib.lblx_od_layer(self._h, C.byref(run), C.byref(layer),
                                     out.ctypes.data_as(C.c_void_p), n, C.byref(v1p),
                                     C.byref(st)), "lblx_od_layer")
        return out, v1p.value, st.as_dict()

    def opdpth_device(self, run, layer, dptr, n, stats=False):
        """Layer optical depth into device memory at dptr (n points)."""
        v1p = C.c_double(0.0)
        st = abi.Stats() if stats else None
        _chk(self._lib.lblx_od_layer_device(self._h, C.byref(run), C.byref(layer),
                                            C.c_void_p(dptr), n, C.byref(v1p),
                                            C.byref(st) if stats else None),
             "lblx_od_layer_device")
        return v1p.value, (st.as_dict() if stats else None)

    def opdpth_column_device(self, run, layers, dptr, offsets, stats=False):
        """OPDPTH for all layers of a column in one batch; layer k's optical
        depth at element offsets[k] of the device buffer dptr.  The radiation
        flag is chained from layer to layer as XLAYER does.  Returns
        (v1p per layer, stats per layer or None)."""
        n = len(layers)
        arr = (abi.Layer * n)(*layers)
        off = np.ascontiguousarray(offsets, np.int64)
        v1p = (C.c_double * n)()
        st = (abi.Stats * n)() if stats else None
        _chk(self._lib.lblx_od_column_device(self._h, C.byref(run), arr, n, C.c_void_p(dptr),
                                             abi.ptr(off, C.c_int64), v1p, st),
             "lblx_od_column_device")
        for k in range(n):
            layers[k].jrad = arr[k].jrad
        return list(v1p), ([x.as_dict() for x in st] if stats else None)

    def contnm(self, run, layer):
        """CONTNM for one layer: (absrb, v1abs, dvabs); layer.jrad is
        updated as OPDPTH does."""
        cap = 5050
        out = np.zeros(cap)
        n = C.c_int(0)
        v1 = C.c_double(0.0)
        dv = C.c_double(0.0)
        _chk(self._lib.lblx_contnm(self._h, C.byref(run), C.byref(layer), abi.ptr(out, C.c_double), cap,
                                   C.byref(n), C.byref(v1), C.byref(dv)), "lblx_contnm")
        return out[:n.value].copy(), v1.value, dv.value

    def xmerge_device(self, run, layer_old, layer, tzu, tbnd, emis, d_od, d_rold, d_told,
                      d_rnew, d_tnew):
        """XMERGE of one layer (IEMIT=1, IPATHL=1) on device pointers;
        layer_old None for the layer at the surface (EMINIT)."""
        _chk(self._lib.lblx_xmerge_device(
            self._h, C.byref(run), C.byref(layer_old) if layer_old is not None else None,
            C.byref(layer), tzu, tbnd, emis, C.c_void_p(d_od), C.c_void_p(d_rold or 0),
            C.c_void_p(d_told or 0), C.c_void_p(d_rnew), C.c_void_p(d_tnew)), "lblx_xmerge_device")

    def spectrum_device(self, run, layers, tzu, tbnd, emis, d_rad, d_tra, n_out, stats=False):
        """OPDPTH for every layer (surface first) then XMERGE up the column:
        TOA upwelling radiance/transmittance at device pointers d_rad/d_tra.
        Returns (v1p of the top grid, stats per layer or None)."""
        n = len(layers)
        arr = (abi.Layer * n)(*layers)
        tz = np.ascontiguousarray(tzu, np.float64)
        v1p = C.c_double(0.0)
        st = (abi.Stats * n)() if stats else None
        _chk(self._lib.lblx_spectrum_device(self._h, C.byref(run), arr, n, abi.ptr(tz, C.c_double),
                                            tbnd, emis, C.c_void_p(d_rad), C.c_void_p(d_tra),
                                            n_out, C.byref(v1p), st), "lblx_spectrum_device")
        for k in range(n):
            layers[k].jrad = arr[k].jrad
        return v1p.value, ([x.as_dict() for x in st] if stats else None)

    def xmerge_path_device(self, run, layer_old, layer, ipathl, tz, tbnd, emis, d_od, d_rold, d_told,
                           d_rnew, d_tnew):
        """XMERGE of one layer for IPATHL 1 or 3 (downwelling: tz = TZL,
        tbnd > 0 adds EMBND after a later layer's merge)."""
        _chk(self._lib.lblx_xmerge_path_device(
            self._h, C.byref(run), C.byref(layer_old) if layer_old is not None else None,
            C.byref(layer), ipathl, tz, tbnd, emis, C.c_void_p(d_od), C.c_void_p(d_rold or 0),
            C.c_void_p(d_told or 0), C.c_void_p(d_rnew), C.c_void_p(d_tnew)), "lblx_xmerge_path_device")

    def spectrum_path(self, run, layers, ipathl, tz, tbnd, emis):
        """A path's radiance and transmittance in either direction
        (lblx_spectrum_path_device): (rad, tra, v1p, dv) on the last
        layer's grid."""
        n = od_size(run, layers[-1])
        nl = len(layers)
        arr = (abi.Layer * nl)(*layers)
        t = np.ascontiguousarray(tz, np.float64)
        rad, tra = self.alloc(8 * n), self.alloc(8 * n)
        v1p = C.c_double(0.0)
        try:
            _chk(self._lib.lblx_spectrum_path_device(self._h, C.byref(run), arr, nl, ipathl,
                                                     abi.ptr(t, C.c_double), tbnd, emis, C.c_void_p(rad.ptr),
                                                     C.c_void_p(tra.ptr), n, C.byref(v1p), None),
                 "lblx_spectrum_path_device")
            return rad.to_host(np.float64), tra.to_host(np.float64), v1p.value, layers[-1].dv
        finally:
            rad.free()
            tra.free()

    def spectrum(self, run, layers, tzu, tbnd, emis):
        """Host-side convenience over spectrum_device: returns (radiance,
        transmittance, v1p, dv) of the top of the atmosphere."""
        n = od_size(run, layers[-1])
        rad = self.alloc(8 * n)
        tra = self.alloc(8 * n)
        try:
            v1p, _ = self.spectrum_device(run, layers, tzu, tbnd, emis, rad.ptr, tra.ptr, n)
            return rad.to_host(np.float64), tra.to_host(np.float64), v1p, layers[-1].dv
        finally:
            rad.free()
            tra.free()

    def jacobian_device(self, run, layers, tzu, tbnd, emis, species, dT, dlnq, d_rad, d_jac, n_out):
        """TOA radiance and its Jacobians (lblx_jacobian_device); returns v1p
        of the top grid."""
        n = len(layers)
        arr = (abi.Layer * n)(*layers)
        tz = np.ascontiguousarray(tzu, np.float64)
        sp = (C.c_int * max(1, len(species)))(*species)
        v1p = C.c_double(0.0)
        _chk(self._lib.lblx_jacobian_device(self._h, C.byref(run), arr, n, abi.ptr(tz, C.c_double), tbnd, emis,
                                            sp, len(species), dT, dlnq, C.c_void_p(d_rad), C.c_void_p(d_jac),
                                            n_out, C.byref(v1p)), "lblx_jacobian_device")
        for k in range(n):
            layers[k].jrad = arr[k].jrad
        return v1p.value

    def jacobian_aj_device(self, run, layers, tzu, tbnd, emis, species, d_rad, d_jac, n_out):
        """The reference's analytic-Jacobian formulation (lblx_jacobian_aj_device)."""
        n = len(layers)
        arr = (abi.Layer * n)(*layers)
        tz = np.ascontiguousarray(tzu, np.float64)
        sp = (C.c_int * max(1, len(species)))(*species)
        v1p = C.c_double(0.0)
        _chk(self._lib.lblx_jacobian_aj_device(self._h, C.byref(run), arr, n, abi.ptr(tz, C.c_double), tbnd,
                                               emis, sp, len(species), C.c_void_p(d_rad), C.c_void_p(d_jac),
                                               n_out, C.byref(v1p)), "lblx_jacobian_aj_device")
        for k in range(n):
            layers[k].jrad = arr[k].jrad
        return v1p.value

    def jacobian_tl_device(self, run, layers, tzu, tbnd, emis, species, d_rad, d_jac, n_out):
        """Jacobians from the tangent-linear optical depths (lblx_jacobian_tl_device)."""
        n = len(layers)
        arr = (abi.Layer * n)(*layers)
        tz = np.ascontiguousarray(tzu, np.float64)
        sp = (C.c_int * max(1, len(species)))(*species)
        v1p = C.c_double(0.0)
        _chk(self._lib.lblx_jacobian_tl_device(self._h, C.byref(run), arr, n, abi.ptr(tz, C.c_double), tbnd,
                                               emis, sp, len(species), C.c_void_p(d_rad), C.c_void_p(d_jac),
                                               n_out, C.byref(v1p)), "lblx_jacobian_tl_device")
        for k in range(n):
            layers[k].jrad = arr[k].jrad
        return v1p.value

    def od_tangent(self, run, layers, species):
        """Layer optical depths and their tangent-linear derivatives
        (lblx_od_tangent_device): (od[layer, point], dod[layer, q, point],
        sizes); each layer's values fill its first sizes[l] points."""
        n = max(od_size(run, x) for x in layers)
        P = 1 + len(species)
        arr = (abi.Layer * len(layers))(*layers)
        sp = (C.c_int * max(1, len(species)))(*species)
        esz = 4 if run.precision == abi.FP32_SPLIT else 8
        dt = np.float32 if esz == 4 else np.float64
        od = self.alloc(esz * n * len(layers))
        dod = self.alloc(esz * n * len(layers) * P)
        try:
            _chk(self._lib.lblx_od_tangent_device(self._h, C.byref(run), arr, len(layers), sp, len(species),
                                                  C.c_void_p(od.ptr), C.c_void_p(dod.ptr), n),
                 "lblx_od_tangent_device")
            for k in range(len(layers)):
                layers[k].jrad = arr[k].jrad
            sizes = [od_size(run, x) for x in layers]
            return (od.to_host(dt, n * len(layers)).reshape(len(layers), n),
                    dod.to_host(dt, n * len(layers) * P).reshape(len(layers), P, n), sizes)
        finally:
            od.free()
            dod.free()

    def jacobian(self, run, layers, tzu, tbnd, emis, species, dT=1e-3, dlnq=1e-4, aj=False, tl=False):
        """Host-side convenience: (radiance, jac[q, layer, point], v1p, dv);
        aj selects the reference's analytic-Jacobian formulation, tl the
        tangent-linear optical depths."""
        n = od_size(run, layers[-1])
        P = 1 + len(species)
        rad = self.alloc(8 * n)
        jac = self.alloc(8 * n * P * len(layers))
        try:
            if tl:
                v1p = self.jacobian_tl_device(run, layers, tzu, tbnd, emis, species, rad.ptr, jac.ptr, n)
            elif aj:
                v1p = self.jacobian_aj_device(run, layers, tzu, tbnd, emis, species, rad.ptr, jac.ptr, n)
            else:
                v1p = self.jacobian_device(run, layers, tzu, tbnd, emis, species, dT, dlnq, rad.ptr, jac.ptr, n)
            return (rad.to_host(np.float64), jac.to_host(np.float64).reshape(P, len(layers), n), v1p,
                    layers[-1].dv)
        finally:
            rad.free()
            jac.free()

    def od_layer_dvout(self, run, layer, dvout):
        """OPDPTH + PNLINT onto DVOUT: (od, v1) on the host."""
        n = od_dvout_size(run, layer, dvout)
        buf = self.alloc(8 * n)
        no, v1 = C.c_int64(), C.c_double()
        try:
            _chk(self._lib.lblx_od_layer_dvout_device(self._h, C.byref(run), C.byref(layer), dvout,
                                                      C.c_void_p(buf.ptr), n, C.byref(no), C.byref(v1)),
                 "lblx_od_layer_dvout_device")
            out = buf.to_host(np.float64, no.value)
        finally:
            buf.free()
        return out, v1.value

    def absmrg_device(self, run, layer_old, layer, d_od, d_acc_old, d_acc_new):
        """One IEMIT = 0 merge step (ABINIT when layer_old is None)."""
        _chk(self._lib.lblx_absmrg_device(self._h, C.byref(run), None if layer_old is None else C.byref(layer_old),
                                          C.byref(layer), C.c_void_p(d_od),
                                          C.c_void_p(d_acc_old) if d_acc_old else None, C.c_void_p(d_acc_new)),
             "lblx_absmrg_device")

    def column_od_device(self, run, layers, d_od, n_out):
        """Column optical depth (OPDPTH + ABINIT/ABSMRG) on the top layer's
        grid; returns its first wavenumber."""
        n = len(layers)
        arr = (abi.Layer * n)(*layers)
        v1p = C.c_double()
        _chk(self._lib.lblx_column_od_device(self._h, C.byref(run), arr, n, C.c_void_p(d_od), n_out,
                                             C.byref(v1p), None), "lblx_column_od_device")
        for i in range(n):
            layers[i].jrad = arr[i].jrad
        return v1p.value

    def spectra_batch_device(self, run, profiles, d_rad, d_tra, out_off):
        """Batch of profiles: profiles is a list of (layers, tzu, tbnd, emis).
        Returns v1p of each profile's top grid."""
        nl = [len(p[0]) for p in profiles]
        lays = [abi.Layer.from_buffer_copy(x) for p in profiles for x in p[0]]
        arr = (abi.Layer * len(lays))(*lays)
        tz = np.ascontiguousarray(np.concatenate([np.asarray(p[1], np.float64) for p in profiles]))
        tb = np.ascontiguousarray([p[2] for p in profiles], np.float64)
        em = np.ascontiguousarray([p[3] for p in profiles], np.float64)
        oo = np.ascontiguousarray(out_off, np.int64)
        v1p = np.zeros(len(profiles))
        nla = (C.c_int * len(nl))(*nl)
        _chk(self._lib.lblx_spectra_batch_device(self._h, C.byref(run), len(profiles), nla, arr,
                                                 abi.ptr(tz, C.c_double), abi.ptr(tb, C.c_double),
                                                 abi.ptr(em, C.c_double), C.c_void_p(d_rad), C.c_void_p(d_tra),
                                                 abi.ptr(oo, C.c_int64), abi.ptr(v1p, C.c_double)),
             "lblx_spectra_batch_device")
        return v1p

    def jacobians_batch_device(self, run, profiles, species, d_rad, rad_off, d_jac, jac_off, aj=False,
                               dT=1e-3, dlnq=1e-4, tl=False):
        """Forward and adjoint of a batch of profiles (lblx_jacobians_batch_device):
        profiles is a list of (layers, tzu, tbnd, emis); returns v1p of each
        profile's top grid.  aj: the reference's analytic-Jacobian runs, tl:
        the tangent-linear optical depths."""
        nl = [len(p[0]) for p in profiles]
        lays = [abi.Layer.from_buffer_copy(x) for p in profiles for x in p[0]]
        arr = (abi.Layer * len(lays))(*lays)
        tz = np.ascontiguousarray(np.concatenate([np.asarray(p[1], np.float64) for p in profiles]))
        tb = np.ascontiguousarray([p[2] for p in profiles], np.float64)
        em = np.ascontiguousarray([p[3] for p in profiles], np.float64)
        ro = np.ascontiguousarray(rad_off, np.int64)
        jo = np.ascontiguousarray(jac_off, np.int64)
        sp = (C.c_int * max(1, len(species)))(*species)
        v1p = np.zeros(len(profiles))
        nla = (C.c_int * len(nl))(*nl)
        _chk(self._lib.lblx_jacobians_batch_device(self._h, C.byref(run), len(profiles), nla, arr,
                                                   abi.ptr(tz, C.c_double), abi.ptr(tb, C.c_double),
                                                   abi.ptr(em, C.c_double), sp, len(species),
                                                   2 if tl else (1 if aj else 0),
                                                   dT, dlnq, C.c_void_p(d_rad), abi.ptr(ro, C.c_int64),
                                                   C.c_void_p(d_jac), abi.ptr(jo, C.c_int64),
                                                   abi.ptr(v1p, C.c_double)),
             "lblx_jacobians_batch_device")
        return v1p

    def backward_batch_device(self, run, profiles, species, d_dLdR, rad_off, d_grad, grad_off, d_rad=None):
        """Reverse mode for a batch of profiles (lblx_backward_batch):
        profiles is a list of (layers, tzu, tbnd, emis); dL/dR of profile p
        at d_dLdR + rad_off[p], its gradient [(1 + nspec) x nlay] at d_grad +
        grad_off[p].  Returns v1p of each profile's top grid."""
        nl = [len(p[0]) for p in profiles]
        lays = [abi.Layer.from_buffer_copy(x) for p in profiles for x in p[0]]
        arr = (abi.Layer * len(lays))(*lays)
        tz = np.ascontiguousarray(np.concatenate([np.asarray(p[1], np.float64) for p in profiles]))
        tb = np.ascontiguousarray([p[2] for p in profiles], np.float64)
        em = np.ascontiguousarray([p[3] for p in profiles], np.float64)
        ro = np.ascontiguousarray(rad_off, np.int64)
        go = np.ascontiguousarray(grad_off, np.int64)
        sp = (C.c_int * max(1, len(species)))(*species)
        v1p = np.zeros(len(profiles))
        nla = (C.c_int * len(nl))(*nl)
        _chk(self._lib.lblx_backward_batch(self._h, C.byref(run), len(profiles), nla, arr, abi.ptr(tz, C.c_double),
                                           abi.ptr(tb, C.c_double), abi.ptr(em, C.c_double), sp, len(species),
                                           C.c_void_p(d_dLdR), abi.ptr(ro, C.c_int64),
                                           C.c_void_p(d_rad) if d_rad else None, C.c_void_p(d_grad),
                                           abi.ptr(go, C.c_int64), abi.ptr(v1p, C.c_double)),
             "lblx_backward_batch")
        return v1p

    def backward(self, run, layers, tzu, tbnd, emis, species, dLdR):
        """Host-side convenience for one profile: (radiance, grad[q, layer])."""
        n = od_size(run, layers[-1])
        P = 1 + len(species)
        g = np.ascontiguousarray(dLdR, np.float64)
        assert g.shape == (n,)
        dl = self.alloc(8 * n)
        rad = self.alloc(8 * n)
        grad = self.alloc(8 * P * len(layers))
        try:
            dl.from_host(g)
            self.backward_batch_device(run, [(layers, tzu, tbnd, emis)], species, dl.ptr, [0], grad.ptr, [0],
                                       d_rad=rad.ptr)
            return rad.to_host(np.float64, n), grad.to_host(np.float64, P * len(layers)).reshape(P, len(layers))
        finally:
            dl.free()
            rad.free()
            grad.free()

    def autotune(self, run, layers):
        """Time the near-wing tile widths on these layers; returns (R1, R2, R3)."""
        n = len(layers)
        arr = (abi.Layer * n)(*[abi.Layer.from_buffer_copy(x) for x in layers])
        t = (C.c_int * 3)()
        _chk(self._lib.lblx_autotune(self._h, C.byref(run), arr, n, t), "lblx_autotune")
        return tuple(t)

    def set_tiles(self, tiles):
        t = (C.c_int * 3)(*tiles)
        _chk(self._lib.lblx_set_tiles(self._h, t), "lblx_set_tiles")

    def tuning(self):
        """(tiles, line_batch) of the near-wing scatter."""
        t = abi.Tuning()
        _chk(self._lib.lblx_get_tuning(self._h, C.byref(t)), "lblx_get_tuning")
        return tuple(t.tiles), tuple(t.line_batch)

    def set_tuning(self, tiles, line_batch):
        t = abi.Tuning()
        t.tiles[:] = list(tiles)
        t.line_batch[:] = list(line_batch)
        _chk(self._lib.lblx_set_tuning(self._h, C.byref(t)), "lblx_set_tuning")

    def tuning_save(self, path):
        _chk(self._lib.lblx_tuning_save(self._h, str(path).encode()), "lblx_tuning_save")

    def tuning_load(self, path):
        _chk(self._lib.lblx_tuning_load(self._h, str(path).encode()), "lblx_tuning_load")

    def fftscn_device(self, scan, d_in, v1_in, dv_in, n_in, d_out, cap):
        """FFTSCN on device pointers; returns (points, v1, dv) of the output."""
        n = C.c_int64(0)
        v1 = C.c_double(0.0)
        dv = C.c_double(0.0)
        _chk(self._lib.lblx_fftscn_device(self._h, C.byref(scan), C.c_void_p(d_in), v1_in, dv_in, n_in,
                                          C.c_void_p(d_out), cap, C.byref(n), C.byref(v1), C.byref(dv)),
             "lblx_fftscn_device")
        return n.value, v1.value, dv.value

    def fftscn_noalloc_device(self, scan, d_in, v1_in, dv_in, n_in, d_out, cap):
        """FFTSCN through the allocation-free entry (fixed device scratch);
        returns (points, v1, dv) of the output."""
        n = C.c_int64(0)
        v1 = C.c_double(0.0)
        dv = C.c_double(0.0)
        _chk(self._lib.lblx_fftscn_noalloc_device(self._h, C.byref(scan), C.c_void_p(d_in), v1_in, dv_in, n_in,
                                                  C.c_void_p(d_out), cap, C.byref(n), C.byref(v1), C.byref(dv)),
             "lblx_fftscn_noalloc_device")
        return n.value, v1.value, dv.value

    def scanfn_device(self, scan, iemit, d_in, v1_in, dv_in, n_in, d_out, n_out):
        """SCANFN on device pointers (spectral-domain scanning function)."""
        _chk(self._lib.lblx_scanfn_device(self._h, C.byref(scan), iemit, C.c_void_p(d_in), v1_in, dv_in, n_in,
                                          C.c_void_p(d_out), n_out), "lblx_scanfn_device")

    def scanfn(self, scan, spectrum, v1_in, dv_in, iemit=1):
        """Host-side convenience: SCANFN of a spectrum held in host memory;
        returns (output, v1, dv)."""
        x = np.ascontiguousarray(spectrum, np.float64)
        n, v1, dv = scanfn_size(scan, iemit, v1_in, dv_in, len(x))
        bi, bo = self.alloc(x.nbytes), self.alloc(8 * max(n, 1))
        try:
            bi.from_host(x)
            self.scanfn_device(scan, iemit, bi.ptr, v1_in, dv_in, len(x), bo.ptr, n)
            return bo.to_host(np.float64, n), v1, dv
        finally:
            bi.free()
            bo.free()

    def fftscn(self, scan, spectrum, v1_in, dv_in):
        """Host-side convenience: scan a spectrum held in host memory."""
        x = np.ascontiguousarray(spectrum, np.float64)
        n, _, _ = scan_size(scan, v1_in, dv_in, len(x))
        bi, bo = self.alloc(x.nbytes), self.alloc(8 * max(n, 1))
        try:
            bi.from_host(x)
            n, v1, dv = self.fftscn_device(scan, bi.ptr, v1_in, dv_in, len(x), bo.ptr, n)
            return bo.to_host(np.float64, n), v1, dv
        finally:
            bi.free()
            bo.free()

    def debug_tape9(self, job=0, cap=2000000):
        vnu = np.zeros(cap)
        sp = np.zeros(cap)
        al = np.zeros(cap)
        ad = np.zeros(cap)
        mol = np.zeros(cap, np.int32)
        n = C.c_int32(0)
        _chk(self._lib.lblx_debug_tape9(self._h, job, abi.ptr(vnu, C.c_double), abi.ptr(sp, C.c_double),
                                        abi.ptr(al, C.c_double), abi.ptr(ad, C.c_double),
                                        abi.ptr(mol, C.c_int32), cap, C.byref(n)), "lblx_debug_tape9")
        k = min(n.value, cap)
        return vnu[:k], sp[:k], al[:k], ad[:k], mol[:k]

    def debug_tape9_spp(self, job=0, cap=2000000):
        spp = np.zeros(cap)
        n = C.c_int32(0)
        _chk(self._lib.lblx_debug_tape9_spp(self._h, job, abi.ptr(spp, C.c_double), cap, C.byref(n)),
             "lblx_debug_tape9_spp")
        return spp[:min(n.value, cap)]

    def timing(self, enable=True):
        _chk(self._lib.lblx_timing(self._h, 1 if enable else 0), "lblx_timing")

    def timings(self):
        """{stage: (ms_total, launches)} since timing was enabled."""
        out = {}
        for i in range(NSTAGE):
            nm = C.c_char_p()
            ms = C.c_double()
            cnt = C.c_int64()
            _chk(self._lib.lblx_timing_get(self._h, i, C.byref(nm), C.byref(ms), C.byref(cnt)),
                 "lblx_timing_get")
            out[nm.value.decode()] = (ms.value, cnt.value)
        return out

    def alloc(self, nbytes):
        """Device buffer (DeviceBuffer) of nbytes on the context's device."""
        return DeviceBuffer(self, nbytes)

    def copy_kernel(self, dst, src, nbytes):
        """Plain device-to-device copy by a kernel (lblx_copy kind 3)."""
        _chk(self._lib.lblx_copy(self._h, C.c_void_p(dst.ptr), C.c_void_p(src.ptr), int(nbytes), 3),
             "lblx_copy")

    def sync(self):
        _chk(self._lib.lblx_sync(self._h), "lblx_sync")

    def stream(self):
        return self._lib.lblx_stream(self._h)


class DeviceBuffer:
    """Device memory owned through the library (no second HIP runtime)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctx._lib.lblx_dev_alloc(ctx._h, self.nbytes)
        if not p:
            raise RuntimeError("lblx_dev_alloc(%d) failed" % self.nbytes)
        self.ptr = p

    def to_host(self, dtype, count=None, offset=0):
        """count elements of dtype from element offset (all by default)."""
        dtype = np.dtype(dtype)
        n = self.nbytes // dtype.itemsize - offset if count is None else int(count)
        assert 0 <= offset and (offset + n) * dtype.itemsize <= self.nbytes
        out = np.empty(n, dtype)
        _chk(self.ctx._lib.lblx_copy(self.ctx._h, out.ctypes.data_as(C.c_void_p),
                                     C.c_void_p(self.ptr + offset * dtype.itemsize), n * dtype.itemsize, 1),
             "lblx_copy")
        return out

    def from_host(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _chk(self.ctx._lib.lblx_copy(self.ctx._h, C.c_void_p(self.ptr),
                                     arr.ctypes.data_as(C.c_void_p), arr.nbytes, 0), "lblx_copy")

    def free(self):
        if self.ptr:
            self.ctx._lib.lblx_dev_free(self.ctx._h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
