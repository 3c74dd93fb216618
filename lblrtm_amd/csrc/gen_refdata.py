"""Extract the numeric tables the monochromatic path needs from LBLRTM's
Fortran sources (read as text) and write them as a C header.

The tables are physical data, not algorithm: the exact-Voigt build's
halfwidth ratios AVC and the Armstrong/Hummer coefficients
(src/oprop_voigt.f90), the Voigt-width/shape
coefficient table (BLOCK DATA VOICON, src/oprop.f90:1993-2144), the TIPS
partition sums QofT(iso, T) for molecules 1-7 (QT_H2O..QT_O2,
src/oprop.f90:6588-8231) on the 119-point temperature grid Tdat
(BD_TDAT, src/oprop.f90:6296-6316), the isotopologue counts ISO_MAX and
masses SMASS (BLOCK DATA Isotop, src/oprop.f90:2697-2895).

Usage:  python gen_refdata.py /path/to/lblrtm/src  > lblx_refdata.h
The generated header is committed; nothing at run time reads the sources.
"""
import re
import sys

NMOL = 7  # H2O CO2 O3 N2O CO CH4 O2: the molecules the synthetic configs use
QT_NAMES = ["QT_H2O", "QT_CO2", "QT_O3", "QT_N2O", "QT_CO", "QT_CH4", "QT_O2"]

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?"


def _join_continuations(text):
    # Fortran free form: '&' at end of line continues; leading '&' on next
    lines = text.split("\n")
    out = []
    buf = ""
    for ln in lines:
        s = ln.split("!")[0] if not ln.lstrip().startswith("!") else ""
        s = s.rstrip()
        if buf:
            s2 = s.lstrip()
            if s2.startswith("&"):
                s2 = s2[1:]
            buf += " " + s2
        else:
            buf = s
        if buf.endswith("&"):
            buf = buf[:-1]
            continue
        out.append(buf)
        buf = ""
    if buf:
        out.append(buf)
    return out


def _nums(s):
    return [float(x.replace("D", "E").replace("d", "e")) for x in re.findall(NUM, s)]


def _data_values(stmt):
    # 'data NAME / v1, v2, ... /' or 'DATA (X(i,J),J=1,119)/ ... /'
    m = re.search(r"/(.*)/\s*$", stmt)
    if not m:
        raise ValueError("no data list in: " + stmt[:80])
    body = m.group(1)
    vals = []
    for tok in body.split(","):
        tok = tok.strip()
        if not tok:
            continue
        if "*" in tok:
            n, v = tok.split("*")
            vals += [float(v.replace("D", "E"))] * int(n)
        else:
            vals.append(float(tok.replace("D", "E").replace("d", "e")))
    return vals


def extract(src_dir):
    text = open(src_dir + "/oprop.f90").read()
    stmts = _join_continuations(text)

    # ---- VOICON
    voi = {}
    for name in ["AV01", "AV51", "CG01", "CG51", "CFA01", "CFA51", "CFB01",
                 "CFB51", "CFC01", "CFC51", "CER01", "CER51"]:
        for s in stmts:
            if re.match(r"\s*DATA\s+" + name + r"\s*/", s):
                voi[name] = _data_values(s)
                break
        else:
            raise ValueError("VOICON array missing: " + name)
    tables = {
        "AVRAT": voi["AV01"] + voi["AV51"],
        "CGAUSS": voi["CG01"] + voi["CG51"],
        "CF1": voi["CFA01"] + voi["CFA51"],
        "CF2": voi["CFB01"] + voi["CFB51"],
        "CF3": voi["CFC01"] + voi["CFC51"],
        "CER": voi["CER01"] + voi["CER51"],
    }
    for k, v in tables.items():
        assert len(v) == 102, (k, len(v))

    # ---- Tdat
    tdat = None
    for s in stmts:
        if re.match(r"\s*data\s+Tdat\s*/", s, re.I):
            tdat = _data_values(s)
            break
    assert tdat is not None and len(tdat) == 119

    # ---- ISO_MAX and masses
    iso_max = None
    for s in stmts:
        if re.match(r"\s*DATA\s*\(ISO_MAX\(I\)", s, re.I):
            iso_max = [int(v) for v in _data_values(s)]
            break
    assert iso_max is not None
    smass = {}
    for s in stmts:
        m = re.match(r"\s*data\s*\(smass\((\d+),i\),i=1,(\d+)\)", s, re.I)
        if m:
            mol = int(m.group(1))
            vals = _data_values(s)
            assert len(vals) == int(m.group(2))
            smass[mol] = vals

    # ---- TIPS QofT per molecule
    qt = {}
    for mi, qname in enumerate(QT_NAMES, start=1):
        start = None
        for i, s in enumerate(stmts):
            if re.match(r"\s*Subroutine\s+" + qname + r"\b", s, re.I):
                start = i
                break
        assert start is not None, qname
        rows = {}
        xgj = None
        for s in stmts[start:]:
            if re.match(r"\s*end\s+subroutine", s, re.I):
                break
            m = re.match(r"\s*data\s*\(QofT\(\s*(\d+),J\),J=1,119\)", s, re.I)
            if m:
                rows[int(m.group(1))] = _data_values(s)
            if re.match(r"\s*data\s+xgj\s*/", s, re.I):
                xgj = _data_values(s)
        n = iso_max[mi - 1]
        for k in range(1, n + 1):
            assert k in rows and len(rows[k]) == 119, (qname, k)
        qt[mi] = [rows[k] for k in range(1, n + 1)]

    # ---- exact-Voigt build: AVC of voigt_init, armstrong's Chebyshev
    # coefficients C and BLOCK DATA ARMCOEFF's W, T (oprop_voigt.f90)
    vst = _join_continuations(open(src_dir + "/oprop_voigt.f90").read())
    voigt = {}
    for key, pat, n in [("AVC", r"\s*DATA\s+AVC\s*/", 102), ("ARM_C", r"\s*data\s+C\s*/", 34),
                        ("ARM_W", r"\s*DATA\s+W\s*/", 10), ("ARM_T", r"\s*DATA\s+T\s*/", 10)]:
        for st in vst:
            if re.match(pat, st, re.I):
                voigt[key] = _data_values(st)
                break
        assert key in voigt and len(voigt[key]) == n, (key, len(voigt.get(key, [])))
    tables["VOIGT"] = voigt
    return tables, tdat, iso_max, smass, qt


def emit(tables, tdat, iso_max, smass, qt, out):
    w = out.write
    w("/* Generated by gen_refdata.py from LBLRTM src/oprop.f90 DATA statements.\n"
      "   Numeric tables only: VOICON (oprop.f90:1993-2144), TIPS QofT for\n"
      "   molecules 1-7 (oprop.f90:6588-8231), Tdat (oprop.f90:6296-6316),\n"
      "   ISO_MAX and SMASS (oprop.f90:2697-2895); the exact-Voigt build's AVC,\n"
      "   armstrong C and ARMCOEFF W, T (oprop_voigt.f90:11797-11818,\n"
      "   12058-12068, 12136-12150).  Do not edit. */\n")
    w("#ifndef LBLX_REFDATA_H\n#define LBLX_REFDATA_H\n\n")
    w("#define LBLX_NMOL_TIPS %d\n#define LBLX_NT_TIPS 119\n#define LBLX_MAXISO 10\n\n" % NMOL)
    for k in ["AVRAT", "CGAUSS", "CF1", "CF2", "CF3", "CER"]:
        w("static const double LBLX_%s[102] = {\n" % k)
        v = tables[k]
        for i in range(0, 102, 6):
            w("  " + ", ".join(repr(x) for x in v[i:i + 6]) + ",\n")
        w("};\n")
    for k, v in tables["VOIGT"].items():
        w("static const double LBLX_VOIGT_%s[%d] = {\n" % (k, len(v)))
        for i in range(0, len(v), 6):
            w("  " + ", ".join(repr(x) for x in v[i:i + 6]) + ",\n")
        w("};\n")
    w("static const double LBLX_TDAT[119] = {\n")
    for i in range(0, 119, 8):
        w("  " + ", ".join(repr(x) for x in tdat[i:i + 8]) + ",\n")
    w("};\n")
    w("static const int LBLX_ISO_MAX[%d] = {%s};\n" %
      (NMOL, ", ".join(str(iso_max[m]) for m in range(NMOL))))
    w("static const double LBLX_SMASS[%d][10] = {\n" % NMOL)
    for m in range(1, NMOL + 1):
        v = smass[m] + [0.0] * (10 - len(smass[m]))
        w("  {" + ", ".join(repr(x) for x in v) + "},\n")
    w("};\n")
    w("static const double LBLX_QOFT[%d][10][119] = {\n" % NMOL)
    for m in range(1, NMOL + 1):
        w("  {\n")
        rows = qt[m]
        for r in range(10):
            v = rows[r] if r < len(rows) else [0.0] * 119
            w("    {")
            w(", ".join(repr(x) for x in v))
            w("},\n")
        w("  },\n")
    w("};\n\n#endif\n")


if __name__ == "__main__":
    t = extract(sys.argv[1])
    emit(*t, sys.stdout)
