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
