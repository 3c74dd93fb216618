"""Extract the continuum tables of CONTNM from LBLRTM's sources (read as
text) and its MT_CKD data file, and write them as a C header.

Tables (physical data, not algorithm):
  MT_CKD H2O self/foreign reference absorption and self T-exponent on the
      10 cm-1 grid of data/absco-ref_wv-mt-ckd.nc (read by getData,
      src/read_module.f90:50-109; the classic netCDF file is read here with
      scipy.io.netcdf_file);
  CO2 continuum S(5003) on -4..10000 cm-1 (BLOCK DATA BFCO2,
      src/contnm.f90:1355-2493), the XFACCO2 scale on 2000-2998 cm-1
      (src/contnm.f90:57-167) and the band-head T exponents of FRNCO2
      (src/contnm.f90:1298-1305);
  O2 fundamental 1340-1850 cm-1 (BLOCK DATA bo2f, src/contnm.f90:7320-7383);
  N2 fundamental 1997.8-2901.6 cm-1 (BLOCK DATA bn2f,
      src/contnm.f90:2756-2912);
  N2 rotational (collision-induced) -10-350 cm-1 at 296 K and 220 K with
      the O2 scale factors (BLOCK DATA BN2T296, BN2T220,
      src/contnm.f90:2569-2666).

Usage:  python gen_contdata.py /path/to/lblrtm > lblx_contdata.h
The generated header is committed; nothing at run time reads the sources.
"""
import os
import re
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from gen_refdata import _join_continuations  # noqa: E402


def _vals(body):
    out = []
    for tok in body.split(","):
        tok = tok.strip()
        if not tok:
            continue
        if "*" in tok:
            n, v = tok.split("*")
            out += [float(v.replace("D", "E").replace("d", "e"))] * int(n)
        else:
            out.append(float(tok.replace("D", "E").replace("d", "e")))
    return out


def block_data(text, name):
    """Values of a BLOCK DATA's COMMON, in COMMON order: (scalars, arrays)."""
    m = re.search(r"BLOCK DATA\s+%s\b(.*?)end block data" % name, text, re.S | re.I)
    assert m, name
    common, data = None, {}
    for st in _join_continuations(m.group(1)):
        s = st.strip()
        if re.match(r"COMMON\s*/", s, re.I):
            lst = s.split("/", 2)[2]
            common = [(n, int(k) if k else 1)
                      for n, k in re.findall(r"(\w+)\s*(?:\(\s*(\d+)\s*\))?", lst)]
        elif re.match(r"DATA\b", s, re.I):
            m2 = re.match(r"DATA\s+(.*?)\s*/(.*)/\s*$", s, re.I | re.S)
            names = [x.strip() for x in m2.group(1).split(",")]
            vals = _vals(m2.group(2))
            if len(names) == 1:
                data[names[0]] = vals
            else:
                assert len(names) == len(vals), s
                for n, v in zip(names, vals):
                    data[n] = [v]
    head = [data[n][0] for n, k in common[:4]]
    arr = []
    for n, k in common[4:]:
        assert len(data[n]) == k, (name, n, len(data[n]), k)
        arr += data[n]
    return head, arr


def extract(root):
    text = open(os.path.join(root, "src", "contnm.f90")).read()
    stmts = _join_continuations(text)
    xfac = tdep = None
    for s in stmts:
        if re.match(r"\s*DATA\s+XFACCO2\s*/", s, re.I):
            xfac = _vals(re.search(r"/(.*)/\s*$", s, re.S).group(1))
        m = re.match(r"\s*data\s*\(tdep_bandhead\(i\),i=(\d+),(\d+)\)\s*/(.*)/\s*$", s, re.I | re.S)
        if m:
            tdep = (int(m.group(1)), _vals(m.group(3)))
            assert len(tdep[1]) == int(m.group(2)) - tdep[0] + 1
    assert xfac is not None and len(xfac) == 500
    assert tdep is not None
    co2_head, co2 = block_data(text, "BFCO2")
    o2_head, o2 = block_data(text, "bo2f")
    n2_head, n2 = block_data(text, "bn2f")
    n2r296_head, n2r296 = block_data(text, "BN2T296")
    n2r220_head, n2r220 = block_data(text, "BN2T220")
    assert n2r296_head == n2r220_head and int(n2r296_head[3]) == 73
    assert len(n2r296) == len(n2r220) == 146
    assert int(co2_head[3]) == len(co2) == 5003
    assert int(o2_head[3]) == 103 and len(o2) == 206
    assert int(n2_head[3]) == 228 and len(n2) == 684
    from scipy.io import netcdf_file
    f = netcdf_file(os.path.join(root, "data", "absco-ref_wv-mt-ckd.nc"), "r", mmap=False)
    nc = {k: f.variables[k].data.astype(float).tolist() for k in
          ("wavenumbers", "self_absco_ref", "for_absco_ref", "for_closure_absco_ref", "self_texp")}
    nc["ref_temp"] = float(f.variables["ref_temp"].data)
    nc["ref_press"] = float(f.variables["ref_press"].data)
    f.close()
    return dict(xfac=xfac, tdep=tdep, co2_head=co2_head, co2=co2, o2_head=o2_head, o2=o2,
                n2_head=n2_head, n2=n2, n2r_head=n2r296_head, n2r296=n2r296, n2r220=n2r220, nc=nc)


def _arr(w, name, v, per=6):
    w("static const double %s[%d] = {\n" % (name, len(v)))
    for i in range(0, len(v), per):
        w("  " + ", ".join(repr(x) for x in v[i:i + per]) + ",\n")
    w("};\n")


def emit(d, out):
    w = out.write
    w("/* Generated by gen_contdata.py from LBLRTM src/contnm.f90 DATA statements\n"
      "   (BFCO2, XFACCO2, FRNCO2 band-head exponents, bo2f, bn2f, BN2T296,\n"
      "   BN2T220) and\n"
      "   data/absco-ref_wv-mt-ckd.nc (MT_CKD H2O).  Do not edit. */\n")
    w("#ifndef LBLX_CONTDATA_H\n#define LBLX_CONTDATA_H\n\n")
    nc = d["nc"]
    w("#define LBLX_MTCKD_N %d\n" % len(nc["wavenumbers"]))
    w("static const double LBLX_MTCKD_REF_T = %r, LBLX_MTCKD_REF_P = %r;\n" %
      (nc["ref_temp"], nc["ref_press"]))
    _arr(w, "LBLX_MTCKD_WVN", nc["wavenumbers"])
    _arr(w, "LBLX_MTCKD_SELF", nc["self_absco_ref"])
    _arr(w, "LBLX_MTCKD_FOR", nc["for_absco_ref"])
    _arr(w, "LBLX_MTCKD_FORCL", nc["for_closure_absco_ref"])
    _arr(w, "LBLX_MTCKD_TEXP", nc["self_texp"])
    h = d["co2_head"]
    w("/* FCO2: V1, V2, DV, NPT */\nstatic const double LBLX_FCO2_HDR[4] = {%r, %r, %r, %r};\n" % tuple(h))
    _arr(w, "LBLX_FCO2", d["co2"])
    _arr(w, "LBLX_XFACCO2", d["xfac"])
    w("#define LBLX_CO2_TDEP_I0 %d\n" % d["tdep"][0])
    _arr(w, "LBLX_CO2_TDEP", d["tdep"][1])
    h = d["o2_head"]
    w("static const double LBLX_O2F_HDR[4] = {%r, %r, %r, %r};\n" % tuple(h))
    _arr(w, "LBLX_O2F_XO2", d["o2"][:103])
    _arr(w, "LBLX_O2F_XO2T", d["o2"][103:])
    h = d["n2_head"]
    w("static const double LBLX_N2F_HDR[4] = {%r, %r, %r, %r};\n" % tuple(h))
    _arr(w, "LBLX_N2F_272", d["n2"][:228])
    _arr(w, "LBLX_N2F_228", d["n2"][228:456])
    _arr(w, "LBLX_N2F_AH2O", d["n2"][456:])
    h = d["n2r_head"]
    w("static const double LBLX_N2R_HDR[4] = {%r, %r, %r, %r};\n" % tuple(h))
    _arr(w, "LBLX_N2R_C296", d["n2r296"][:73])
    _arr(w, "LBLX_N2R_SF296", d["n2r296"][73:])
    _arr(w, "LBLX_N2R_C220", d["n2r220"][:73])
    _arr(w, "LBLX_N2R_SF220", d["n2r220"][73:])
    w("\n#endif\n")


if __name__ == "__main__":
    emit(extract(sys.argv[1]), sys.stdout)
