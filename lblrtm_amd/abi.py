"""ctypes mirror of include/lblx.h (structs and error codes).

The same structures are used to call the HIP library (liblblx.so) and, in
tests only, the CPU oracle library.
"""
import ctypes as C

import numpy as np

LINREC = 250
MXMOL = 64

OK = 0
E_ARG = -1
E_NODEV = -2
E_HIP = -3
E_NREC = -4
E_UNSUPPORTED = -5
E_SIZE = -6
E_NOLINES = -7
E_DVRATIO = -8
E_COARSER = -9
E_SCRATCH = -10

FP64 = 0
FP32_SPLIT = 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class Lines(C.Structure):
    _fields_ = [
        ("nblk", C.c_int32),
        ("nrec", _ip),
        ("vmin", _dp),
        ("vmax", _dp),
        ("vnu", _dp),
        ("sp", _fp),
        ("alfa", _fp),
        ("epp", _fp),
        ("mol", _ip),
        ("hwhm", _fp),
        ("tmpalf", _fp),
        ("pshift", _fp),
        ("iflg", _ip),
    ]


class Run(C.Structure):
    _fields_ = [
        ("v1", C.c_double),
        ("v2", C.c_double),
        ("sample", C.c_double),
        ("alfal0", C.c_double),
        ("avmass", C.c_double),
        ("dptmin", C.c_double),
        ("dptfac", C.c_double),
        ("ilblf4", C.c_int32),
        ("icntnm", C.c_int32),
        ("irad", C.c_int32),
        ("precision", C.c_int32),
        ("cntscl", C.c_double * 7),
        ("ifrgnx", C.c_int32),
        ("iaersl", C.c_int32),
        ("voigt", C.c_int32),
        ("ibrd", C.c_int32),
    ]


class Layer(C.Structure):
    _fields_ = [
        ("pave", C.c_double),
        ("tave", C.c_double),
        ("wk", C.c_double * MXMOL),
        ("wbroad", C.c_double),
        ("dv", C.c_double),
        ("v1", C.c_double),
        ("v2", C.c_double),
        ("nmol", C.c_int32),
        ("jrad", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("nlin", C.c_int64),
        ("lincnt", C.c_int64),
        ("nlins4", C.c_int64),
        ("ilin4", C.c_int64),
        ("npanel", C.c_int64),
        ("npts", C.c_int64),
        ("k1_updates", C.c_int64),
        ("k3_updates", C.c_int64),
        ("nwin4", C.c_int64),
        ("nfix", C.c_int64),
        ("fixseq", C.c_int64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Scan(C.Structure):
    _fields_ = [
        ("hwhm", C.c_double),
        ("v1", C.c_double),
        ("v2", C.c_double),
        ("dvout", C.c_double),
        ("param", C.c_double),
        ("jemit", C.c_int32),
        ("jfnin", C.c_int32),
        ("mratin", C.c_int32),
        ("ivx", C.c_int32),
        ("nofix", C.c_int32),
        ("pad", C.c_int32),
    ]


class Tuning(C.Structure):
    _fields_ = [("tiles", C.c_int32 * 3), ("line_batch", C.c_int32 * 3)]


class ScanFn(C.Structure):
    _fields_ = [
        ("hwhm", C.c_double),
        ("v1", C.c_double),
        ("v2", C.c_double),
        ("sampl", C.c_double),
        ("jemit", C.c_int32),
        ("jfn", C.c_int32),
        ("jvar", C.c_int32),
        ("npanel", C.c_int32),
    ]


def make_scanfn(hwhm, v1, v2, jemit=1, jfn=1, sampl=0.0, jvar=0, npanel=0):
    """TAPE5 record 8.1 of SCANFN (src/postsub.f90:189-190)."""
    s = ScanFn()
    s.hwhm, s.v1, s.v2, s.sampl = hwhm, v1, v2, sampl
    s.jemit, s.jfn, s.jvar, s.npanel = jemit, jfn, jvar, npanel
    return s


def make_scan(hwhm, v1, v2, jemit=1, jfnin=-6, mratin=0, dvout=0.0, ivx=0, nofix=0, param=0.0):
    """TAPE5 record 10.1 of FFTSCN (src/fftscn_dbl.f90:313-315)."""
    s = Scan()
    s.hwhm, s.v1, s.v2, s.dvout, s.param = hwhm, v1, v2, dvout, param
    s.jemit, s.jfnin, s.mratin, s.ivx, s.nofix = jemit, jfnin, mratin, ivx, nofix
    return s


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class LineTable:
    """Host line table in the TAPE3 block layout; owns the numpy arrays the
    Lines struct points into."""

    FIELDS = [("vnu", np.float64), ("sp", np.float32), ("alfa", np.float32),
              ("epp", np.float32), ("mol", np.int32), ("hwhm", np.float32),
              ("tmpalf", np.float32), ("pshift", np.float32), ("iflg", np.int32)]

    def __init__(self, nrec, vmin, vmax, **arrays):
        self.nrec = np.ascontiguousarray(nrec, dtype=np.int32)
        self.vmin = np.ascontiguousarray(vmin, dtype=np.float64)
        self.vmax = np.ascontiguousarray(vmax, dtype=np.float64)
        nblk = len(self.nrec)
        for name, dt in self.FIELDS:
            a = np.ascontiguousarray(arrays[name], dtype=dt)
            assert a.shape == (nblk * LINREC,), (name, a.shape)
            setattr(self, name, a)
        self.nblk = nblk

    @classmethod
    def from_lines(cls, vnu, sp, alfa, epp, mol, hwhm, tmpalf, pshift, iflg=None):
        """Pack sorted per-line arrays into full 250-slot blocks."""
        n = len(vnu)
        nblk = max(1, (n + LINREC - 1) // LINREC)
        tot = nblk * LINREC
        out = {}
        src = dict(vnu=vnu, sp=sp, alfa=alfa, epp=epp, mol=mol, hwhm=hwhm,
                   tmpalf=tmpalf, pshift=pshift,
                   iflg=np.zeros(n, np.int32) if iflg is None else iflg)
        for name, dt in cls.FIELDS:
            a = np.zeros(tot, dtype=dt)
            a[:n] = np.asarray(src[name], dtype=dt)
            out[name] = a
        nrec = np.full(nblk, LINREC, np.int32)
        nrec[-1] = n - (nblk - 1) * LINREC if n else 0
        vmin = np.zeros(nblk)
        vmax = np.zeros(nblk)
        for b in range(nblk):
            seg = out["vnu"][b * LINREC:b * LINREC + nrec[b]]
            if len(seg):
                vmin[b] = seg.min()
                vmax[b] = seg.max()
        return cls(nrec, vmin, vmax, **out)

    @classmethod
    def from_lines_coupled(cls, vnu, sp, alfa, epp, mol, hwhm, tmpalf, pshift, iflg, ycoef, gcoef):
        """Pack sorted lines into 250-slot blocks with line coupling: a line
        with iflg 1 or 3 is followed by its coefficient record (iflg -1),
        ycoef[i] (Y or GAMMA1) and gcoef[i] (G or GAMMA2) at 200, 250, 296
        and 340 K stored in vnu, alfa, mol (float bits), tmpalf and sp, epp,
        hwhm, pshift (src/oprop.f90:789-795, 1039-1048).  A line and its
        record share a block; block VMIN/VMAX are over the lines."""
        n = len(vnu)
        iflg = np.asarray(iflg, np.int32)
        ycoef = np.asarray(ycoef, np.float64).reshape(n, 4)
        gcoef = np.asarray(gcoef, np.float64).reshape(n, 4)
        blocks = [[]]
        for i in range(n):
            cp = iflg[i] in (1, 3)
            need = 2 if cp else 1
            if len(blocks[-1]) + need > LINREC:
                blocks.append([])
            blocks[-1].append(("line", i))
            if cp:
                blocks[-1].append(("coef", i))
        nblk = len(blocks)
        tot = nblk * LINREC
        out = {name: np.zeros(tot, dtype=dt) for name, dt in cls.FIELDS}
        nrec = np.zeros(nblk, np.int32)
        vmin = np.zeros(nblk)
        vmax = np.zeros(nblk)
        src = dict(vnu=vnu, sp=sp, alfa=alfa, epp=epp, mol=mol, hwhm=hwhm, tmpalf=tmpalf, pshift=pshift)
        for b, blk in enumerate(blocks):
            nrec[b] = len(blk)
            vs = []
            for k, (kind, i) in enumerate(blk):
                j = b * LINREC + k
                if kind == "line":
                    for name in src:
                        out[name][j] = src[name][i]
                    out["iflg"][j] = iflg[i]
                    vs.append(float(vnu[i]))
                else:
                    y, g = ycoef[i], gcoef[i]
                    out["vnu"][j] = y[0]
                    out["alfa"][j] = y[1]
                    out["mol"][j] = np.float32(y[2]).view(np.int32)
                    out["tmpalf"][j] = y[3]
                    out["sp"][j] = g[0]
                    out["epp"][j] = g[1]
                    out["hwhm"][j] = g[2]
                    out["pshift"][j] = g[3]
                    out["iflg"][j] = -1
            if vs:
                vmin[b], vmax[b] = min(vs), max(vs)
        return cls(nrec, vmin, vmax, **out)

    def nlines(self):
        return int(self.nrec.sum())

    def struct(self):
        s = Lines()
        s.nblk = self.nblk
        s.nrec = ptr(self.nrec, C.c_int32)
        s.vmin = ptr(self.vmin, C.c_double)
        s.vmax = ptr(self.vmax, C.c_double)
        s.vnu = ptr(self.vnu, C.c_double)
        s.sp = ptr(self.sp, C.c_float)
        s.alfa = ptr(self.alfa, C.c_float)
        s.epp = ptr(self.epp, C.c_float)
        s.mol = ptr(self.mol, C.c_int32)
        s.hwhm = ptr(self.hwhm, C.c_float)
        s.tmpalf = ptr(self.tmpalf, C.c_float)
        s.pshift = ptr(self.pshift, C.c_float)
        s.iflg = ptr(self.iflg, C.c_int32)
        self._s = s
        return s


def make_run(v1, v2, sample=4.0, alfal0=0.04, avmass=36.0, dptmin=2e-4,
             dptfac=1e-3, ilblf4=1, icntnm=0, irad=0, precision=FP64, cntscl=None,
             ifrgnx=0, iaersl=0, voigt=0, ibrd=0):
    r = Run()
    r.v1, r.v2 = v1, v2
    r.sample, r.alfal0, r.avmass = sample, alfal0, avmass
    r.dptmin, r.dptfac = dptmin, dptfac
    r.ilblf4, r.icntnm, r.irad, r.precision = ilblf4, icntnm, irad, precision
    for i, x in enumerate(cntscl if cntscl is not None else [0.0] * 7):
        r.cntscl[i] = x
    r.ifrgnx, r.iaersl = ifrgnx, iaersl
    r.voigt = voigt
    r.ibrd = ibrd
    return r


def make_layer(pave, tave, wk, wbroad, dv, v1, v2, jrad=1):
    L = Layer()
    L.pave, L.tave = pave, tave
    for i, w in enumerate(wk):
        L.wk[i] = w
    L.nmol = len(wk)
    L.wbroad = wbroad
    L.dv, L.v1, L.v2 = dv, v1, v2
    L.jrad = jrad
    return L
