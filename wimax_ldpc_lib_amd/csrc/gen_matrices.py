#!/usr/bin/env python3
"""Regenerate include/matrices.h and csrc/matrices_list.inc.

include/matrices.h is the public table header of the reference
(include/matrices.h:1-824, installed by include/CMakeLists.txt:4-8): for
every (rate, z) code an expanded parity-check table
`int16_t wimax_<n>_<rate>[M][C]` (ascending variable indices per check row,
-1 at the row tail) and the encoder constants `wimax_<n>_<rate>_D` and `_X`,
each behind its `Z_<z>` gate.  This library exports the same 114 arrays and
228 constants; it fills them at load time from the 802.16e base matrices
(csrc/matrices.cpp), so no generated table data is compiled in.
matrices_list.inc is the X-macro list matrices.cpp defines them from.

Run from the repository root:  python wimax_ldpc_lib_amd/csrc/gen_matrices.py
"""
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
# array-name suffix per coderate enum value (include/enums.h), base rows Mb
# and the largest check-row degree C of each rate
SUFFIX = ["0_5", "0_66A", "0_66B", "0_75A", "0_75B", "0_83"]
MB = [12, 8, 8, 6, 6, 4]
C = [7, 10, 11, 15, 15, 20]
ZS = list(range(24, 97, 4))


def main():
    h = ["/* matrices.h -- the expanded 802.16e parity-check tables, drop-in for the",
         " * reference's include/matrices.h.",
         " *",
         " * wimax_<n>_<rate>[M][C]: row m lists the variables of check m in ascending",
         " * order, -1 after the last one (n = 24 z, M = Mb z).  _D is the shift of the",
         " * top nonzero block of the first parity column at this z and _X the base row",
         " * of its middle (shift 0) block, the encoder constants.  libwimax_ldpc_amd.so",
         " * defines all of them and fills them when it is loaded.",
         " * Written by wimax_ldpc_lib_amd/csrc/gen_matrices.py. */",
         "#ifndef MATRICES_H", "#define MATRICES_H", "", "#include <stdint.h>", ""]
    h += ["#define Z_%d" % z for z in ZS]
    inc = ["// matrices_list.inc -- WL_MATRIX(name, rate, z, rows, cols) for every code of",
           "// include/matrices.h.  Written by gen_matrices.py."]
    for z in ZS:
        for r in range(6):
            name = "wimax_%d_%s" % (24 * z, SUFFIX[r])
            h += ["", "#ifdef Z_%d" % z,
                  "extern unsigned int %s_D;" % name,
                  "extern unsigned int %s_X;" % name,
                  "extern int16_t %s[%d][%d];" % (name, MB[r] * z, C[r]),
                  "#endif"]
            inc.append("WL_MATRIX(%s, %d, %d, %d, %d)" % (name, r, z, MB[r] * z, C[r]))
    h += ["", "#endif  // MATRICES_H", ""]
    with open(os.path.join(ROOT, "include", "matrices.h"), "w") as f:
        f.write("\n".join(h))
    with open(os.path.join(HERE, "matrices_list.inc"), "w") as f:
        f.write("\n".join(inc) + "\n")
    print("wrote include/matrices.h and matrices_list.inc:", len(inc) - 2, "codes")


if __name__ == "__main__":
    main()
