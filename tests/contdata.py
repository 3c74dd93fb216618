"""Reads the continuum tables back from the committed C header
lblrtm_amd/csrc/lblx_contdata.h (for tests)."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "lblrtm_amd", "csrc", "lblx_contdata.h")


def tables():
    text = open(HDR).read()
    out = {}
    for m in re.finditer(r"static const double (\w+)\[(\d+)\] = \{(.*?)\};", text, re.S):
        v = np.array([float(x) for x in m.group(3).replace("\n", " ").split(",") if x.strip()])
        assert len(v) == int(m.group(2)), m.group(1)
        out[m.group(1)] = v
    for m in re.finditer(r"(\w+) = ([-+0-9.eE]+)[,;]", text):
        out[m.group(1)] = float(m.group(2))
    for m in re.finditer(r"#define (\w+) (\d+)", text):
        out[m.group(1)] = int(m.group(2))
    return out
