"""TAPE3 (LNFL binary line file) reader and writer.

Layout (sequential unformatted, 4-byte record markers, as the reference
reads it in RDLIN, src/oprop.f90:734-776, with the record types of
src/struct_types.f90:28-44):

  record 1      line-file header: HLINID(10) CHARACTER*8, then the rest of
                /LINHDR/ (BMOLID(64) CHARACTER*8, MOLCNT, MCNTLC, MCNTNL
                INTEGER*4(64), SUMSTR REAL*4(64), LINMOL INTEGER*4, FLINLO,
                FLINHI REAL*8, LINCNT, ILINLC, ILINNL, IREC, IRECTL
                INTEGER*4, HID1(2) CHARACTER*8)
  then per block of up to 250 lines:
  INPUT_HEADER  VMIN, VMAX REAL*8; NREC, NWDS INTEGER*4
  INPUT_BLOCK   NWDS = 9750 four-byte words: VNU REAL*8(250); SP, ALFA,
                EPP REAL*4(250); MOL INTEGER*4(250); HWHM, TMPALF, PSHIFT
                REAL*4(250); IFLG INTEGER*4(250); BRD_MOL_FLG_IN
                INTEGER*4(7,250); BRD_MOL_DAT REAL*4(21,250); SPEED_DEP
                REAL*4(250)
The writer always writes full 250-slot blocks (NWDS = 9750).
"""
import struct

import numpy as np

from .abi import LINREC, LineTable

NWDS = 9750
_BLOCK = np.dtype([
    ("vnu", "<f8", (LINREC,)),
    ("sp", "<f4", (LINREC,)),
    ("alfa", "<f4", (LINREC,)),
    ("epp", "<f4", (LINREC,)),
    ("mol", "<i4", (LINREC,)),
    ("hwhm", "<f4", (LINREC,)),
    ("tmpalf", "<f4", (LINREC,)),
    ("pshift", "<f4", (LINREC,)),
    ("iflg", "<i4", (LINREC,)),
    ("brd_flg", "<i4", (LINREC, 7)),
    ("brd_dat", "<f4", (LINREC, 21)),
    ("speed_dep", "<f4", (LINREC,)),
])
assert _BLOCK.itemsize == NWDS * 4


def _rec(payload):
    n = len(payload)
    return struct.pack("<i", n) + payload + struct.pack("<i", n)


def _header(table):
    hlinid = [b"LBLX    ", b"SYNTH   ", b"TAPE3   ", b"        ", b"        ",
              b"        ", b"        ", b"        ", b"        ", b"       I"]
    bmolid = b"        " * 64
    mol = table.mol[:]
    molcnt = np.zeros(64, np.int32)
    sumstr = np.zeros(64, np.float32)
    for b in range(table.nblk):
        n = table.nrec[b]
        m = mol[b * LINREC:b * LINREC + n] % 100
        s = table.sp[b * LINREC:b * LINREC + n]
        for k in np.unique(m):
            if 1 <= k <= 64:
                molcnt[k - 1] += int((m == k).sum())
                sumstr[k - 1] += float(s[m == k].sum())
    nl = table.nlines()
    vnu = table.vnu
    flinlo = float(table.vmin[0]) if table.nblk else 0.0
    flinhi = float(table.vmax[-1]) if table.nblk else 0.0
    del vnu
    p = b"".join(hlinid) + bmolid
    p += molcnt.tobytes() + np.zeros(64, np.int32).tobytes() * 2 + sumstr.tobytes()
    p += struct.pack("<i", int((molcnt > 0).sum()))
    p += struct.pack("<dd", flinlo, flinhi)
    p += struct.pack("<iiiii", nl, 0, 0, table.nblk, table.nblk)
    p += b"        " * 2
    return p


def write_tape3(path, table, brd_flg=None, brd_dat=None):
    """Write `table` as a TAPE3.  brd_flg [nblk*250, 7] int and brd_dat
    [nblk*250, 21] float (per molecule: width, its temperature exponent,
    shift) fill the broadener words; zero when omitted."""
    with open(path, "wb") as f:
        f.write(_rec(_header(table)))
        for b in range(table.nblk):
            n = int(table.nrec[b])
            f.write(_rec(struct.pack("<ddii", table.vmin[b], table.vmax[b], n, NWDS)))
            blk = np.zeros(1, _BLOCK)
            sl = slice(b * LINREC, (b + 1) * LINREC)
            for name in ("vnu", "sp", "alfa", "epp", "mol", "hwhm", "tmpalf",
                         "pshift", "iflg"):
                blk[name][0] = getattr(table, name)[sl]
            if brd_flg is not None:
                blk["brd_flg"][0] = np.asarray(brd_flg, np.int32)[sl]
            if brd_dat is not None:
                blk["brd_dat"][0] = np.asarray(brd_dat, np.float32)[sl]
            f.write(_rec(blk.tobytes()))


def read_tape3(path):
    """Read a TAPE3 written in the layout above into a LineTable."""
    data = open(path, "rb").read()
    pos = 0

    def rec():
        nonlocal pos
        if pos + 4 > len(data):
            return None
        (n,) = struct.unpack_from("<i", data, pos)
        body = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n:
            raise ValueError("TAPE3: record marker mismatch at %d" % pos)
        pos += 8 + n
        return body

    hdr = rec()
    if hdr is None:
        raise ValueError("TAPE3: empty file")
    if len(hdr) >= 56 and hdr[55:56] == b"^":
        rec()  # optional negative-EPP record (RDLIN, oprop.f90:746-750)
    nrec, vmin, vmax, blocks = [], [], [], []
    while True:
        h = rec()
        if h is None:
            break
        v0, v1, n, nw = struct.unpack_from("<ddii", h, 0)
        if n > LINREC:
            raise ValueError("RDLIN; NREC GT LIMIN")
        body = rec()
        if nw != NWDS or len(body) != NWDS * 4:
            raise ValueError("TAPE3: unsupported block size %d" % nw)
        blocks.append(np.frombuffer(body, _BLOCK)[0])
        nrec.append(n)
        vmin.append(v0)
        vmax.append(v1)
    arrays = {}
    for name in ("vnu", "sp", "alfa", "epp", "mol", "hwhm", "tmpalf", "pshift", "iflg",
                 "brd_flg", "brd_dat"):
        arrays[name] = (np.concatenate([blk[name] for blk in blocks])
                        if blocks else np.zeros(0))
    # RDLIN/RDLNFL: an O2 line whose own O2 broadener flag is set carries
    # its shift converted to the foreign shift, whatever IBRD
    # (src/oprop.f90:818-827, 3473-3477); stored back in single precision
    flg = arrays.pop("brd_flg")
    dat = arrays.pop("brd_dat")
    nbrd = 0
    if len(flg):
        nbrd = int(np.any(flg != 0, axis=1).sum())
        o2 = (arrays["mol"] % 100 == 7) & (arrays["iflg"] >= 0) & (flg[:, 6] > 0)
        if o2.any():
            ps = arrays["pshift"].astype(np.float64)
            sh = dat[:, 6 * 3 + 2].astype(np.float64)
            rvmr = 0.21
            arrays["pshift"] = np.where(o2, ((ps - rvmr * sh) / (1.0 - rvmr)).astype(np.float32),
                                        arrays["pshift"])
    t = LineTable(np.array(nrec), np.array(vmin), np.array(vmax), **arrays)
    t.nbrd = nbrd  # lines carrying broadener flags (IBRD data)
    return t


def read_panels(path, word=8, nvals=1, filhdr_words=None):
    """Read a panel file of the reference (TAPE10: nvals = 1, the optical
    depth; TAPE11/TAPE12: nvals = 2, radiance and transmittance), laid out
    as BUFOUT writes it: an optional file header record of filhdr_words
    words, then per panel the /BUFPNL/ header V1P, V2P (REAL*8), DVP
    (REAL), NLIM (INTEGER) and nvals records of NLIM REALs, default REAL and
    INTEGER being `word` bytes; an ENDFIL record of six -99 may end it
    (src/oprop.f90:1524-1526, src/xmerge.f90:3737-3790,
    src/lblrtm.f90:1368-1377).  Returns a dict: filhdr (bytes or None),
    v1p, v2p, dvp, nlim (arrays per panel), values (list of nvals arrays,
    panels concatenated) and endfil."""
    data = open(path, "rb").read()
    pos = 0
    recs = []
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        body = data[pos + 4:pos + 4 + n]
        (n2,) = struct.unpack_from("<i", data, pos + 4 + n)
        if n2 != n:
            raise ValueError("panel file: record marker mismatch at %d" % pos)
        recs.append(body)
        pos += 8 + n
    fdt = "<f8" if word == 8 else "<f4"
    idt = "<i8" if word == 8 else "<i4"
    out = {"filhdr": None, "v1p": [], "v2p": [], "dvp": [], "nlim": [], "endfil": False}
    k = 0
    if filhdr_words is not None:
        if len(recs[0]) != filhdr_words * word:
            raise ValueError("panel file: header record of %d bytes" % len(recs[0]))
        out["filhdr"] = recs[0]
        k = 1
    vals = [[] for _ in range(nvals)]
    hsize = 16 + 2 * word
    while k < len(recs):
        r = recs[k]
        if len(r) == 6 * word and np.all(np.frombuffer(r, idt) == -99):
            out["endfil"] = True
            k += 1
            break
        if len(r) != hsize:
            raise ValueError("panel file: panel header of %d bytes" % len(r))
        v1, v2 = struct.unpack_from("<dd", r, 0)
        dvp = float(np.frombuffer(r[16:16 + word], fdt)[0])
        nlim = int(np.frombuffer(r[16 + word:16 + 2 * word], idt)[0])
        for j in range(nvals):
            body = recs[k + 1 + j]
            if len(body) != nlim * word:
                raise ValueError("panel file: %d values for NLIM %d" % (len(body) // word, nlim))
            vals[j].append(np.frombuffer(body, fdt).astype(np.float64))
        out["v1p"].append(v1)
        out["v2p"].append(v2)
        out["dvp"].append(dvp)
        out["nlim"].append(nlim)
        k += 1 + nvals
    if k != len(recs):
        raise ValueError("panel file: records after ENDFIL")
    for key in ("v1p", "v2p", "dvp"):
        out[key] = np.array(out[key])
    out["nlim"] = np.array(out["nlim"], np.int64)
    out["values"] = [np.concatenate(v) if v else np.zeros(0) for v in vals]
    return out
