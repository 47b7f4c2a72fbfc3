"""Run-time quasi-cyclic codes, host side (no GPU): construction from a base
matrix, alist writing and reading in the reference's format
(python_ldpc/ldpc_utils.py:157-240, python_ldpc/ldpc_decoder.py:20-59),
lifting-size detection and the dual-diagonal encoder structure."""
import os

import numpy as np
import pytest

REF_ALIST = "/root/reference/alist"
SUFFIX = ["0_5", "0_66A", "0_66B", "0_75A", "0_75B", "0_83"]


def _wimax_base(W, rate, z):
    """Base matrix of a compiled 802.16e code, from its expanded table."""
    p = W.code_params(rate, z)
    t = W.ldpc(rate, z).checknode_array()
    base = np.full((p["mb"], 24), -1, np.int16)
    for l in range(p["mb"]):
        for v in t[l * z]:
            if v >= 0:
                base[l, v // z] = v % z
    return base


def test_base_matrix_round_trip_all_wimax_codes():
    """qc_code from the base matrix of every compiled code reproduces its
    table; the detected dual-diagonal constants equal the compiled ones for
    the five encodable rates, and 3/4B is reported without them."""
    import wimax_ldpc_lib_amd as W
    for rate in range(6):
        for z in range(24, 97, 4):
            base = _wimax_base(W, rate, z)
            q = W.qc_code(base=base, z=z)
            p = W.code_params(rate, z)
            assert (q.n, q.m, q.k, q.c, q.mb) == (p["n"], p["m"], p["k"], p["c"], p["mb"])
            assert np.array_equal(q.checknode_array(), W.ldpc(rate, z).checknode_array())
            assert np.array_equal(q.base(), base)
            if rate == 4:
                assert q.X == -1 and q.D == -1
            else:
                assert (q.X, q.D) == (p["x"], p["d"]), (rate, z)


@pytest.mark.skipif(not os.path.isdir(REF_ALIST), reason="reference alists not present")
def test_alist_writer_and_reader_match_reference_files():
    """Our alist text equals the reference's file byte for byte (all 114),
    and reading the reference's file finds z and the same base matrix."""
    import wimax_ldpc_lib_amd as W
    for rate in range(6):
        for z in range(24, 97, 4):
            n = 24 * z
            path = os.path.join(REF_ALIST, "wimax_%d_%s.alist" % (n, SUFFIX[rate]))
            ref = open(path).read()
            base = _wimax_base(W, rate, z)
            q = W.qc_code(base=base, z=z)
            assert q.alist() == ref, path
            q2 = W.qc_code(alist=ref)
            assert q2.z == z and np.array_equal(q2.base(), base)
            with pytest.raises(W.WLError):       # not circulant at half the lifting
                W.qc_code(alist=ref, z=z // 2)


def test_alist_round_trip_and_lifting_detection():
    import wimax_ldpc_lib_amd as W
    rng = np.random.default_rng(3)
    for z, nb, mb in ((27, 24, 12), (81, 24, 6), (10, 30, 9), (7, 12, 5)):
        base = np.where(rng.random((mb, nb)) < 0.35, rng.integers(0, z, (mb, nb)), -1).astype(np.int16)
        base[np.arange(mb), np.arange(mb) % nb] = rng.integers(0, z, mb)   # no empty base row
        q = W.qc_code(base=base, z=z)
        text = q.alist()
        q2 = W.qc_code(alist=text, z=z)
        assert np.array_equal(q2.base(), base)
        assert np.array_equal(q2.checknode_array(), q.checknode_array())
        q3 = W.qc_code(alist=text)           # detected: the largest lifting that fits
        assert q3.z % z == 0 or z % q3.z == 0
        assert np.array_equal(q3.checknode_array().ravel()[q3.checknode_array().ravel() >= 0].size,
                              q.checknode_array().ravel()[q.checknode_array().ravel() >= 0].size)


def test_non_quasi_cyclic_alist_decodes_row_by_row():
    """A code with no circulant structure reads as z = 1 (every check its own
    layer, the reference's row order)."""
    import wimax_ldpc_lib_amd as W
    rng = np.random.default_rng(5)
    n, m = 40, 20
    rows = [sorted(rng.choice(n, 5, replace=False) + 1) for _ in range(m)]
    cols = [[r + 1 for r in range(m) if c + 1 in rows[r]] for c in range(n)]
    wc, wr = max(len(c) for c in cols), 5
    lines = ["%d %d" % (n, m), "%d %d" % (wc, wr), " ".join(str(len(c)) for c in cols),
             " ".join("5" for _ in rows)]
    lines += [" ".join(str(v) for v in c + [0] * (wc - len(c))) for c in cols]
    lines += [" ".join(str(v) for v in r) for r in rows]
    q = W.qc_code(alist="\n".join(lines) + "\n")
    assert q.z == 1 and q.mb == m
    t = q.checknode_array()
    assert [list(t[r][t[r] >= 0] + 1) for r in range(m)] == [list(r) for r in rows]
    assert q.X == -1


@pytest.mark.parametrize("bad", ["shift", "empty", "square", "z"])
def test_invalid_codes_are_rejected(bad):
    import wimax_ldpc_lib_amd as W
    base = np.array([[1, 0, 2, -1], [0, 3, -1, 1]], np.int16)
    z = 4
    if bad == "shift":
        base[0, 0] = 4
    elif bad == "empty":
        base[1, :] = -1
    elif bad == "square":
        base = np.zeros((4, 4), np.int16)
    else:
        z = 1000
    with pytest.raises(W.WLError):
        W.qc_code(base=base, z=z)
    with pytest.raises(W.WLError):
        W.qc_code(alist="4 2\n1 1\n")


@pytest.mark.parametrize("text", [
    "4 2\n2 2000000000\n1 1 1 1\n2 2\n",                 # row weight beyond N
    "4 2\n9223372036854775807 2\n1 1 1 1\n2 2\n",        # column weight that overflows the size
    "4 2\n3 2\n1 1 1 1\n2 2\n",                          # column weight beyond M
    "4 2\n1 2\n1 1 1 1\n2 2\n1 2 3 4\n1 2\n3",          # check lists end early
])
def test_malformed_alist_headers_are_rejected(text):
    """Header weights are bounded by the code size and a short file is an
    error, not a read past the numbers."""
    import wimax_ldpc_lib_amd as W
    with pytest.raises(W.WLError, match="alist"):
        W.qc_code(alist=text)
