"""The oracle's order-dependent line rejections against a sequential
restatement written here, line by line, from the Fortran.

Two decisions of the reference depend on the order in which lines are
processed, and no closed form fixes them:

* CONVF4 (src/oprop.f90:3878-4039) rejects a shrunk far-wing line when its
  peak SPEAK = A3 |SIV| is at most DPTMN + DPTFC R4(JJ), where R4 is the
  far-wing array *as accumulated from the lines accepted before it*;
* LNCOR1 (src/oprop.f90:995-1193) applies the same test to near lines
  against the current R4 window, and drops a line whose reach lies below
  the current panel (HWF3 ALFV + VNU < VFT); both depend on the panel at
  which RDLIN reads the line's block, which depends on how far CNVFNV got
  with the blocks before it (oprop.f90:416-540, 734-885, 1287-1292).

The restatement below (LBLF4 + CONVF4, the HIRAC1 panel loop with RDLIN,
LNCOR1 and CNVFNV's fit test, the R4 window refills) runs on a small table
of ragged TAPE3 blocks with thresholds raised so that the running-R4 term
decides many lines; its decisions are compared, one by one and in order,
with the decision log the oracle writes (oracle_log).  SHRINK's output
(TAPE9) is taken from the oracle: SHRINK is pinned in
test_oracle_shrink.py.
"""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from lblrtm_amd import abi

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "lblrtm_amd", "csrc", "lblx_refdata.h")
RADCN2 = 1.4387752
BOLTZ, CLIGHT, AVOGAD = 1.3806503E-16, 2.99792458E+10, 6.02214199E+23  # phys_consts.f90
ONEPL = 1.001
P0, TEMP0 = 1013.25, 296.0
LINREC = 250


def _table(name):
    text = open(HDR).read()
    m = re.search(r"static const double %s\[[^=]*=\s*\{(.*?)\};" % name, text, re.S)
    return [float(x) for x in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?", m.group(1))]


AVRAT = _table("LBLX_AVRAT")
SMASS = np.array(_table("LBLX_SMASS")).reshape(7, 10)
ISO_MAX = [6, 10, 9, 5, 6, 4, 3]


def fint(x):
    """Fortran INT(): truncation toward zero."""
    return int(x)


def radfn(v, xkt):
    x = v / xkt
    if x <= 0.01:
        return 0.5 * x * v
    if x <= 10.0:
        e = math.exp(-x)
        return v * (1.0 - e) / (1.0 + e)
    return v


# ---------------------------------------------------------------------
# LBLF4 + RDLIN4 + CONVF4 for one far-wing window (oprop.f90:3641-4039)
class Window:
    def __init__(self, t9, v1r4, v2r4, dvr4, bound4, dptmn, dptfc, win, log):
        vnu, sp, alfa, epp, mol, bstart, bcount = t9
        limout = 2500
        self.v1r4, self.dvr4 = v1r4, dvr4
        self.nptr4 = min(fint((v2r4 - v1r4) / dvr4 + ONEPL), limout)
        self.v2r4 = v1r4 + dvr4 * (self.nptr4 - 1)
        r4 = [0.0] * (limout + 3)
        recpi = 1.0 / (2.0 * math.asin(1.0))
        zsqbnd = 64.0 * 64.0
        a3 = (1.0 + 2.0 * zsqbnd) / ((1.0 + zsqbnd) * (1.0 + zsqbnd))
        b3 = -1.0 / ((1.0 + zsqbnd) * (1.0 + zsqbnd))
        bndsq = bound4 * bound4
        r_bndsq = 1.0 / bndsq
        vnulst = self.v2r4 + bound4
        vlo = self.v1r4 - bound4
        done = False
        for b in range(len(bcount)):
            if done:
                break
            first, last = bstart[b], bstart[b] + bcount[b]
            if vnu[last - 1] < vlo:  # RDLIN4 skips the block
                continue
            for i in range(first, last):
                if sp[i] == 0.0:
                    continue
                alfadi, alfali = epp[i], alfa[i]
                zetai = alfali / (alfali + alfadi)
                iz = fint(100.0 * zetai + ONEPL)
                zetdif = 100.0 * zetai - float(iz - 1)
                alfavi = (AVRAT[iz - 1] + zetdif * (AVRAT[iz] - AVRAT[iz - 1])) * (alfali + alfadi)
                ralfvi = 1.0 / alfavi
                sil = sp[i] * recpi * alfali
                siv = (alfali * ralfvi) * sp[i] * recpi * ralfvi
                speak = a3 * abs(siv)
                jj = min(max(fint((vnu[i] - self.v1r4) / dvr4 + 1.0), 1), self.nptr4)
                if speak <= dptmn + dptfc * r4[jj]:
                    log.append((0, win, i, 0))
                    continue
                log.append((0, win, i, 1))
                xnui = vnu[i] - self.v1r4
                jmin = fint((xnui - bound4) / dvr4 + 2.0)
                if vnu[i] >= vnulst:
                    done = True
                    break
                if jmin > self.nptr4:
                    continue
                jmin = max(jmin, 1)
                jmax = fint((xnui + bound4) / dvr4 + 1.0)
                if jmax < jmin:
                    continue
                jmax = min(jmax, self.nptr4)
                alfli2 = alfali * alfali
                alfvi2 = alfavi * alfavi
                xjj = float(jmin - 1) * dvr4
                f4bnd = sil / (alfli2 + bndsq)
                rec_alfvi2 = 1.0 / alfvi2
                siv_a3, siv_b3 = siv * a3, siv * b3
                for j in range(jmin, jmax + 1):
                    xm = xjj - xnui
                    xmsq = xm * xm
                    zvsq = xmsq * rec_alfvi2
                    sub = (2.0 - xmsq * r_bndsq) * f4bnd if mol[i] == 2 else f4bnd
                    if zvsq <= zsqbnd:
                        f4fn = (siv_a3 + zvsq * siv_b3) - sub
                    else:
                        f4fn = sil / (alfli2 + xmsq) - sub
                    r4[j] = r4[j] + f4fn  # CO2: times CHI = 1 (oprop.f90:4075)
                    xjj = xjj + dvr4
        self.r4 = r4


# ---------------------------------------------------------------------
# HIRAC1's panel loop with RDLIN, LNCOR1 and CNVFNV's fit test
def hirac1_decisions(tab, lay, run, t9, scor):
    V1, V2, DV, T, P = lay.v1, lay.v2, lay.dv, lay.tave, lay.pave
    sample, alfal0 = run.sample, run.alfal0
    xkt = T / RADCN2
    jrad = 1
    if V1 / xkt < 5.0 and lay.jrad != -1:
        jrad = 0
    assert jrad == 0  # the window's R4 is not scaled by the radiation term
    alfav = sample * DV
    DVR4 = 64.0 * alfav / sample
    BOUND4 = 25.0
    V1R4ST, V2R4ST = V1 - 2.0 * DVR4, V2 + 2.0 * DVR4
    DPTMN = run.dptmin / radfn(V2, xkt)
    DPTFC = run.dptfac
    NSHIFT, NLIM1 = 32, 2401
    NLIM2 = NLIM1 // 4 + 1
    NLIM3 = NLIM2 // 4 + 1
    ALFMAX = 4 * sample * DV * 0.04 / alfal0
    nalfmax = fint(ALFMAX / DV)
    NBOUND = fint((2.0 * 64.0) * nalfmax + 0.01)
    MAX1 = NSHIFT + NLIM1 + NSHIFT + NBOUND // 2 + 4 * nalfmax
    DVR3 = (0.032 / 0.002) * DV
    BOUND = float(NBOUND) * DV / 2.0
    VFT = V1 - float(NSHIFT) * DV
    VBOT, VTOP = V1 - BOUND, V2 + BOUND
    N2R3 = NLIM3 + (NSHIFT // 16) - 1 + 3
    # MOLEC (oprop.f90:2525-2552) and LNCOR1's first-call constants (:968-993)
    fad = math.log(2.0) * 2.0 * AVOGAD * BOLTZ / (CLIGHT * CLIGHT)
    wk = [lay.wk[m] for m in range(lay.nmol)]
    wtot = sum(wk) + lay.wbroad
    rhorat = (P / P0) * (TEMP0 / T)
    tratio = T / TEMP0
    betacr = (1.0 / xkt) - (1.0 / (TEMP0 / RADCN2))
    rhoslf = [rhorat * w / wtot for w in wk]
    wavdxf, hwdxf = DV / 0.002, 4.0 / 0.002

    log = []
    win_no = [0]

    def window(v1r4):
        w = Window(t9, v1r4, V2R4ST, DVR4, BOUND4, DPTMN, DPTFC, win_no[0], log)
        win_no[0] += 1
        return w

    W = window(V1R4ST)
    blocks = [b for b in range(tab.nblk)]
    bi = 0
    panels = 0
    kept = {}  # slot -> (vnu', recalf) of lines LNCOR1 keeps
    ilo, ihi, idata = 0, -1, 0
    cur = []   # slots of the current block in IOUT order
    ieof = False
    state = 60
    while True:
        if state == 60:
            if ieof:
                state = 80
                continue
            while bi < len(blocks) and tab.vmax[blocks[bi]] < VBOT:  # RDLIN skips the block
                bi += 1
            if bi >= len(blocks):
                ieof = True
                state = 80
                continue
            b = blocks[bi]
            bi += 1
            n = int(tab.nrec[b])
            cur = [b * LINREC + i for i in range(n)]  # IFLG = 0 for every record
            ilo = 1
            for j in range(1, n + 1):
                ilo = j
                if tab.vnu[cur[j - 1]] >= VBOT:
                    break
            idata = 0
            for j in range(ilo, n + 1):
                if tab.mol[cur[j - 1]] > 0:
                    ihi = j
                    if tab.vnu[cur[j - 1]] > VTOP:
                        break
            if ihi < n:
                idata = 1
            for j in range(ilo, ihi + 1):  # LNCOR1
                s = cur[j - 1]
                keep = 0
                m = int(tab.mol[s]) % 100
                iso = (int(tab.mol[s]) % 1000) // 100
                if 1 <= m <= 7 and 1 <= iso <= ISO_MAX[m - 1]:
                    sui = float(tab.sp[s]) * wk[m - 1] if m <= lay.nmol else 0.0
                    if sui != 0.0:
                        vnu = float(tab.vnu[s]) + rhorat * float(tab.pshift[s])
                        tmpcor = tratio ** float(tab.tmpalf[s])
                        alfl = float(tab.alfa[s]) * tmpcor * (rhorat - rhoslf[m - 1]) + \
                            float(tab.hwhm[s]) * tmpcor * rhoslf[m - 1]
                        alfad = vnu * math.sqrt(fad * T / SMASS[m - 1][iso - 1])
                        zeta = alfl / (alfl + alfad)
                        fz = 100.0 * zeta
                        iz = fint(fz + ONEPL)
                        zd = fz - float(iz - 1)
                        alfv = (AVRAT[iz - 1] + zd * (AVRAT[iz] - AVRAT[iz - 1])) * (alfl + alfad)
                        alfv = min(max(alfv, DV), ALFMAX)
                        if not (64.0 * alfv + vnu < VFT):
                            recalf = 1.0 / alfv
                            epp = float(tab.epp[s])
                            if epp <= -1.001:
                                epp = abs(epp)
                            if epp > -0.999:
                                sui = sui * scor[m - 1][iso - 1] * math.exp(-epp * betacr) * \
                                    (1.0 + math.exp(-vnu / xkt))
                            speak = sui * recalf
                            jj = min(max(fint((vnu - W.v1r4) / DVR4 + 1.0), 1), W.nptr4)
                            if not (speak <= DPTMN + DPTFC * W.r4[jj]):
                                keep = 1
                                kept[s] = (vnu, recalf)
                log.append((1, s, panels, keep))
            state = 70
        elif state == 70:  # CNVFNV: stop at the first kept line that does not fit
            ipanel = None
            if ilo <= ihi:
                ilast = ihi
                for j in range(ilo, ihi + 1):
                    s = cur[j - 1]
                    if s not in kept:
                        continue
                    vnu, recalf = kept[s]
                    zint = (vnu - VFT) / DV
                    bhwdxf = hwdxf / (recalf * wavdxf)
                    if fint(zint + bhwdxf + 1.5) > MAX1:
                        ilast = j - 1
                        ipanel = 1
                        break
                if ipanel is None:
                    ipanel = idata
            else:
                ilast = ilo - 1
                ipanel = 1
            ilo = ilast + 1
            state = 60 if ipanel == 0 else 80
        else:  # 80: PANEL emits R1(NLO..NHI), then the R4 window may move on
            nhi = NLIM1 + NSHIFT - 1
            istop = (VFT + (nhi - 1) * DV) > V2
            panels += 1
            VFT = VFT + float(NLIM1 - 1) * DV
            if istop:
                break
            vf1 = VFT - 2.0 * DVR4
            vf1 = V1 + math.floor((vf1 - V1) / DVR4) * DVR4
            vf2 = VFT + 2.0 * DVR4 + float(N2R3 + 4) * DVR3
            if vf2 > W.v2r4 and W.v2r4 != V2R4ST:
                W = window(vf1)
            state = 70
    return log, panels


# ---------------------------------------------------------------------
def ragged_table(seed, counts, v1, v2, mols):
    """Sorted lines in TAPE3 blocks of the given record counts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = sum(counts)
    vnu = np.sort(rng.uniform(v1, v2, n))
    fields = dict(
        vnu=vnu, sp=10.0 ** rng.uniform(-27.0, -19.0, n), alfa=rng.uniform(0.04, 0.10, n),
        epp=rng.uniform(0.0, 3000.0, n), mol=np.asarray(mols)[rng.integers(0, len(mols), n)] + 100,
        tmpalf=rng.uniform(0.5, 0.8, n), pshift=rng.uniform(-0.005, 0.005, n))
    fields["hwhm"] = fields["alfa"] * rng.uniform(2.0, 5.0, n)
    nblk = len(counts)
    arrays = {}
    for name, dt in abi.LineTable.FIELDS:
        a = np.zeros(nblk * LINREC, dt)
        if name in fields:
            o = 0
            for b, c in enumerate(counts):
                a[b * LINREC:b * LINREC + c] = np.asarray(fields[name][o:o + c], dt)
                o += c
        arrays[name] = a
    vmin = np.array([arrays["vnu"][b * LINREC] for b in range(nblk)])
    vmax = np.array([arrays["vnu"][b * LINREC + c - 1] for b, c in enumerate(counts)])
    return abi.LineTable(np.array(counts), vmin, vmax, **arrays)


def oracle_log_run(run, lay, tab):
    L = O.lib()
    L.oracle_log.argtypes = [C.c_int]
    L.oracle_log.restype = None
    L.oracle_log_get.argtypes = [C.POINTER(C.c_int32), C.c_int64]
    L.oracle_log_get.restype = C.c_int64
    L.oracle_log(1)
    try:
        n = C.c_int64(0)
        v1p = C.c_double(0.0)
        st = abi.Stats()
        rc = L.oracle_od_layer(C.byref(run), C.byref(abi.Layer.from_buffer_copy(lay)), C.byref(tab.struct()),
                               None, 0.0, 0.0, 0.0, 0, None, 0, C.byref(n), C.byref(v1p), C.byref(st))
        assert rc == 0
        cnt = L.oracle_log_get(None, 0)
        buf = np.zeros(4 * cnt, np.int32)
        L.oracle_log_get(abi.ptr(buf, C.c_int32), cnt)
    finally:
        L.oracle_log(0)
    return [tuple(int(x) for x in r) for r in buf.reshape(-1, 4)], st.as_dict()


def tape9(run, lay, tab):
    cap = 200000
    arrs = [np.zeros(cap) for _ in range(4)] + [np.zeros(cap, np.int32)]
    nb = C.c_int(0)
    bc = np.zeros(10000, np.int32)
    n = O.lib().oracle_linf4(C.byref(run), C.byref(abi.Layer.from_buffer_copy(lay)), C.byref(tab.struct()),
                             *[abi.ptr(a, C.c_double) for a in arrs[:4]], abi.ptr(arrs[4], C.c_int32),
                             abi.ptr(bc, C.c_int32), cap, 10000, C.byref(nb))
    bcount = [int(x) for x in bc[:nb.value]]
    bstart = [int(x) for x in np.concatenate([[0], np.cumsum(bcount)[:-1]])]
    vnu, sp, alfa, epp = (a[:n].tolist() for a in arrs[:4])
    return vnu, sp, alfa, epp, arrs[4][:n].tolist(), bstart, bcount


CASES = {
    # (seed, block record counts, band, lines band, molecules, T, P, DPTFAC)
    "one-window": (5, [30, 17, 25, 40, 12, 28, 23, 25], (600.0, 660.0), (570.0, 690.0), [1, 2], 250.0, 500.0, 0.05),
    "window-refills": (9, [9, 31, 14, 22, 18, 27, 11, 25, 20, 23], (600.0, 1500.0), (575.0, 1525.0), [1, 2, 3],
                       230.0, 200.0, 0.2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_inorder_rejections_match_sequential_restatement(name):
    seed, counts, (v1, v2), (lv1, lv2), mols, t, p, dptfac = CASES[name]
    tab = ragged_table(seed, counts, lv1, lv2, mols)
    wk = [0.0] * 7
    for m in mols:
        wk[m - 1] = {1: 3e22, 2: 8e21, 3: 5e18}[m]
    wb = 2.5e24 * (p / 1013.0)
    run = abi.make_run(v1, v2, dptmin=2e-4, dptfac=dptfac)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    got, st = oracle_log_run(run, lay, tab)
    sc = np.zeros(70)
    O.lib().tips_scor(t, abi.ptr(sc, C.c_double))
    want, npanel = hirac1_decisions(tab, lay, run, tape9(run, lay, tab), sc.reshape(7, 10).tolist())
    assert npanel == st["npanel"]
    far_g = [e for e in got if e[0] == 0]
    far_w = [e for e in want if e[0] == 0]
    near_g = [e for e in got if e[0] == 1]
    near_w = [e for e in want if e[0] == 1]
    assert len(far_g) == len(far_w) and len(near_g) == len(near_w)
    for a, b in zip(far_g, far_w):
        assert a == b, ("CONVF4", a, b)
    for a, b in zip(near_g, near_w):
        assert a == b, ("LNCOR1", a, b)
    # the inputs exercise what is being pinned: rejections on both paths,
    # several far-wing windows or read panels, and far-wing rejections that
    # only the running R4 makes (the same line passes against DPTMN alone)
    assert 0 < sum(e[3] for e in far_g) < len(far_g)
    assert 0 < sum(e[3] for e in near_g) < len(near_g)
    assert len({e[2] for e in near_g}) >= 2
    if name == "window-refills":
        assert len({e[1] for e in far_g}) >= 2
    assert st["ilin4"] == sum(e[3] for e in far_g)
    assert st["lincnt"] == sum(e[3] for e in near_g)


def test_restatement_detects_a_dropped_running_r4():
    """A plausible mistake -- testing against DPTMN alone, without the
    accumulated R4 -- changes decisions on the same inputs, so the
    comparison above can see it."""
    seed, counts, (v1, v2), (lv1, lv2), mols, t, p, dptfac = CASES["one-window"]
    tab = ragged_table(seed, counts, lv1, lv2, mols)
    wk = [3e22, 8e21, 0, 0, 0, 0, 0]
    wb = 2.5e24 * (p / 1013.0)
    run = abi.make_run(v1, v2, dptmin=2e-4, dptfac=dptfac)
    dv, _ = O.plan_dv(run, [p], [t], [wk], [wb])
    lay = abi.make_layer(p, t, wk, wb, dv[0], v1, v2)
    got, _ = oracle_log_run(run, lay, tab)
    run0 = abi.make_run(v1, v2, dptmin=2e-4, dptfac=0.0)
    got0, _ = oracle_log_run(run0, lay, tab)
    for kind in (0, 1):  # CONVF4 and LNCOR1
        acc = sum(e[3] for e in got if e[0] == kind)
        acc0 = sum(e[3] for e in got0 if e[0] == kind)
        assert acc0 > acc, kind
