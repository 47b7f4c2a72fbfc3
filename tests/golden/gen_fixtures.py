"""Writes the committed golden fixtures from the reference's own compiled code
(oracle/_ref/libwimax_ref.so, built from the reference sources by
oracle/Makefile; this container only):

  encoder_ref.npz  one seeded info word and the reference encoder's codeword
                   for each (rate, z) with rate in {1/2, 2/3A, 2/3B, 3/4A, 3/4B,
                   5/6} and z in {24, 48, 76, 96} (3/4B: the reference's bytes,
                   which do not satisfy H)
  tables_ref.json  sha256 of every reference expanded table (int16 [M][C],
                   lib/matrices/*.cc) and its N, M, C, D, X, for all 114 codes

Run from the repository root:  python tests/golden/gen_fixtures.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def main():
    O.build()
    assert O.ref_available(), "reference build missing"
    out = {}
    rng = np.random.default_rng(20240607)
    for rate in range(6):
        for z in (24, 48, 76, 96):
            t = O.ref_table(rate, z)
            k = t["n"] - t["m"]
            info = rng.integers(0, 2, (2, k), dtype=np.uint8)
            info[1] = rng.integers(0, 256, k, dtype=np.uint8)
            out["info_%d_%d" % (rate, z)] = info
            out["cw_%d_%d" % (rate, z)] = O.ref_encode(rate, z, info)
    np.savez_compressed(os.path.join(HERE, "encoder_ref.npz"), **out)
    tabs = {}
    for rate in range(6):
        for z in O.Z_VALUES:
            t = O.ref_table(rate, z)
            tabs["%d_%d" % (rate, z)] = dict(
                sha256=hashlib.sha256(np.ascontiguousarray(t["table"], dtype="<i2").tobytes()).hexdigest(),
                n=t["n"], m=t["m"], c=t["cols"], d=t["d"], x=t["x"])
    with open(os.path.join(HERE, "tables_ref.json"), "w") as f:
        json.dump(tabs, f, indent=0, sort_keys=True)
    print("wrote", len(out) // 2, "encoder fixtures and", len(tabs), "table hashes")


if __name__ == "__main__":
    main()
