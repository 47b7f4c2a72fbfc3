#!/usr/bin/env python3
"""Regenerate base_matrices.h: the IEEE 802.16e QC-LDPC base model matrices.

The decoder and encoder need, per code rate, the Mb x 24 base model matrix
(circulant shift p(i, j) at z0 = 96, or -1 for an all-zero block) plus the
encoder constants D (shift of the top entry of the first parity column) and
X (base row of the middle, shift-0 entry of that column).  These are standard
data and must be identical to the tables the reference decodes with, so this
script recovers them from the reference's z = 96 expansion
(lib/matrices/wimax_2304_0_*.cc, generated by python_ldpc/ldpc_utils.py:316-387
from the base matrices at python_ldpc/ldpc_utils.py:11-90): row l*96 of each
expanded table holds, for every nonzero block b of base row l, the variable
b*96 + p(l, b).  The 5/6 matrix is therefore the variant the reference uses
(ldpc_utils.py:82-87), not the one printed in the 2012 standard.

Run from the repository root:  python wimax_ldpc_lib_amd/csrc/gen_base_matrices.py
"""
import os
import re
import sys

REF = os.environ.get("WIMAX_REF_ROOT", "/root/reference")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "base_matrices.h")

# reference file suffix per coderate enum value (include/enums.h:4-12)
SUFFIX = ["0_5", "0_66A", "0_66B", "0_75A", "0_75B", "0_83"]
NAMES = ["HALFRATE", "TWOTHIRDSA", "TWOTHIRDSB", "THREEQUARTERSA", "THREEQUARTERSB", "FIVESIXTHS"]


def parse(path):
    txt = open(path).read()
    d = int(re.search(r"_D\s*=\s*(\d+)", txt).group(1))
    x = int(re.search(r"_X\s*=\s*(\d+)", txt).group(1))
    body = txt[txt.index("= {") + 3:]
    rows = [[int(v) for v in r.split(",")] for r in re.findall(r"\{([^{}]*)\}", body)]
    return rows, d, x


def main():
    z = 96
    out = []
    out.append("// Generated by gen_base_matrices.py -- IEEE 802.16e base model matrices.")
    out.append("// Shift p(i, j) of each 96x96 circulant block (-1 = zero block); D and X are")
    out.append("// the dual-diagonal encoder constants at z0 = 96.  Do not edit by hand.")
    out.append("#pragma once")
    out.append("#include <stdint.h>")
    out.append("")
    out.append("// C sees plain const tables; C++ (host and HIP device code) sees constexpr")
    out.append("// ones so that kernels can fold shifts into immediates.")
    out.append("#ifdef __cplusplus")
    out.append("#define WL_BASE_QUAL static constexpr")
    out.append("#else")
    out.append("#define WL_BASE_QUAL static const")
    out.append("#endif")
    out.append("")
    out.append("#define WL_NB 24")
    out.append("#define WL_Z0 96")
    out.append("")
    meta = []
    for rate, suf in enumerate(SUFFIX):
        rows, d, x = parse(os.path.join(REF, "lib", "matrices", "wimax_2304_%s.cc" % suf))
        mb = len(rows) // z
        base = [[-1] * 24 for _ in range(mb)]
        for l in range(mb):
            for v in rows[l * z]:
                if v < 0:
                    continue
                base[l][v // z] = v % z
        out.append("WL_BASE_QUAL int8_t wl_base_%s[%d][24] = {" % (suf, mb))
        for l in range(mb):
            out.append("    {" + ", ".join("%d" % p for p in base[l]) + "},")
        out.append("};")
        meta.append((rate, suf, mb, d, x))
    out.append("")
    out.append("struct wl_base_info { const int8_t (*p)[24]; int mb; int d0; int x; };")
    out.append("WL_BASE_QUAL struct wl_base_info wl_base_table[6] = {")
    for rate, suf, mb, d, x in meta:
        out.append("    { wl_base_%s, %d, %d, %d },  // %s" % (suf, mb, d, x, NAMES[rate]))
    out.append("};")
    open(OUT, "w").write("\n".join(out) + "\n")
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
