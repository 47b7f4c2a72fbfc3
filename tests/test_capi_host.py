"""CPU-side checks of the C-ABI library: it loads without a GPU, exports every
function include/*.h declares, and its host logic (code parameters, tables,
argument validation) matches the reference.  No compute calls here."""
import ctypes as C
import glob
import hashlib
import json
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        if 'extern "C"' not in txt:
            continue
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(wl_\w+)\s*\(", txt, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd import capi
    L = W.lib()
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(L, name), name
    assert declared == set(capi.SIGNATURES), declared ^ set(capi.SIGNATURES)
    # the data symbols of matrices.h: 114 tables and their D / X constants
    data = re.findall(r"^extern\s+[\w\s]+?\b(wimax_\w+)", open(os.path.join(ROOT, "include", "matrices.h")).read(),
                      flags=re.M)
    assert len(data) == 3 * 114
    for name in data:
        assert hasattr(L, name), name
    assert W.lib().wl_version().startswith(b"wimax_ldpc_amd")


def test_code_params_match_reference_hashes():
    import wimax_ldpc_lib_amd as W
    ref = json.load(open(os.path.join(GOLDEN, "tables_ref.json")))
    for key, r in ref.items():
        rate, z = map(int, key.split("_"))
        p = W.code_params(rate, z)
        assert (p["n"], p["m"], p["c"], p["d"], p["x"]) == (r["n"], r["m"], r["c"], r["d"], r["x"])
        t = W.ldpc(rate, z).checknode_array()
        assert hashlib.sha256(np.ascontiguousarray(t, dtype="<i2").tobytes()).hexdigest() == r["sha256"]


def test_matrices_header_tables_equal_reference(tmp_path):
    """include/matrices.h: a C++ caller that names the tables (wimax_2304_0_83,
    ...) and reads them through ldpc::m_checknode_array gets the reference's
    expanded tables and D / X for all 114 codes (include/matrices.h:1-824,
    lib/matrices/*.cc), and BER_runner() (include/BER_runner.h:51) builds."""
    import subprocess
    exe = os.path.join(os.path.dirname(GOLDEN), "cpp", "_bin", "matrices_check")
    assert os.path.exists(exe), "built by wimax_ldpc_lib_amd/build.py"
    out = tmp_path / "tables.bin"
    p = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    ref = json.load(open(os.path.join(GOLDEN, "tables_ref.json")))
    buf = out.read_bytes()
    pos, seen = 0, 0
    while pos < len(buf):
        rate, z, rows, cols, d, x = np.frombuffer(buf, dtype="<i4", count=6, offset=pos)
        pos += 24
        t = np.frombuffer(buf, dtype="<i2", count=rows * cols, offset=pos)
        pos += 2 * rows * cols
        r = ref["%d_%d" % (rate, z)]
        assert (rows, cols, d, x) == (r["m"], r["c"], r["d"], r["x"]), (rate, z)
        assert hashlib.sha256(t.tobytes()).hexdigest() == r["sha256"], (rate, z)
        seen += 1
    assert seen == 114


@pytest.mark.parametrize("rate,z", [(6, 96), (-1, 96), (0, 20), (0, 98), (0, 50), (5, 0)])
def test_invalid_rate_or_z_is_rejected(rate, z):
    """lib/ldpc.cc:46-57,1198-1203 throw; the C ABI returns WL_EINVAL with a message."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd import capi
    with pytest.raises(W.WLError) as e:
        W.code_params(rate, z)
    assert e.value.code == capi.WL_EINVAL
    h = C.c_void_p()
    assert W.lib().wl_decoder_create(rate, z, 10, 1, C.byref(h)) == capi.WL_EINVAL
    assert not h.value
    assert b"invalid" in W.lib().wl_last_error()
    assert W.lib().wl_encoder_create(rate, z, C.byref(h)) == capi.WL_EINVAL


def test_decoder_handle_host_logic():
    import wimax_ldpc_lib_amd as W
    dec = W.ldpc_decoder(5, 96, 10, early_stop=False)
    assert dec.frames_per_wg >= 1
    with pytest.raises(W.WLError):
        dec.frames_per_wg = 1000
    dec.frames_per_wg = 1
    assert dec.frames_per_wg == 1
    # compute_syndrome runs on the host (lib/ldpc_decoder.cc:67-104)
    from oracle import oracle as O
    code = O.Code(5, 96)
    rng = np.random.default_rng(0)
    hd = rng.integers(0, 2, code.n, dtype=np.int8)
    assert dec.compute_syndrome(hd, False) == code.syndrome(hd, False)
    assert dec.compute_syndrome(hd, True) == code.syndrome(hd, True)
    cw = code.encode(rng.integers(0, 2, (1, code.k), dtype=np.uint8))[0].astype(np.int8)
    assert dec.compute_syndrome(cw, False) == 0
    h = C.c_void_p()
    assert W.lib().wl_decoder_create(5, 96, 300, 0, C.byref(h)) != 0   # max_iter range
    dec.close()


def test_null_arguments_are_rejected_without_device_work():
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd import capi
    L = W.lib()
    assert L.wl_decode_f32_dev(None, None, 1, None, None, None, None, None, None) == capi.WL_EINVAL
    dec = W.ldpc_decoder(0, 24, 10)
    assert L.wl_decode_f32_dev(dec._h, None, 4, None, None, None, None, None, None) == capi.WL_EINVAL
    assert L.wl_decode_f32_dev(dec._h, None, 0, None, None, None, None, None, None) == capi.WL_OK
    assert L.wl_decode_learned_fallback_dev(dec._h, None, 4, 10, None, None, None, None, None, None,
                                            None) == capi.WL_EINVAL   # no offsets set
    assert L.wl_code_params(0, 24, None) == capi.WL_EINVAL


def test_header_flag_values_match_bindings():
    """The #define values of the C header equal the constants the Python
    bindings pass."""
    from wimax_ldpc_lib_amd import capi
    text = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                             "wimax_ldpc_amd.h")).read()
    defs = dict(re.findall(r"#define\s+(WL_FLAG_\w+)\s+(0x[0-9a-fA-F]+)u", text))
    assert defs, "no flags found"
    for name, val in defs.items():
        assert getattr(capi, name) == int(val, 16), name


def test_tuned_table_is_valid():
    """The autotuned launch table compiled into the library covers all 114
    codes once, with frames per workgroup the kernel accepts and each code
    shape only where it exists (0, 1 and 4 at any frame count; 2: z % 32 == 0
    at two frames per workgroup, 3: z = 32, 48, 64 at four)."""
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "wimax_ldpc_lib_amd",
                        "csrc", "tuned_table.inc")
    rows = [tuple(int(x) for x in re.findall(r"-?\d+", line)) for line in open(path) if line.startswith("{")]
    assert sorted((r, z) for r, z, _, _ in rows) == [(r, z) for r in range(6) for z in range(24, 97, 4)]
    for rate, z, f, v in rows:
        fmax = 512 // z
        while fmax > 1 and 512 + fmax * 24 * z * 4 > 65536:
            fmax -= 1
        assert 1 <= f <= fmax, (rate, z, f)
        assert v in (0, 1, 2, 3, 4), (rate, z, v)
        if v == 2:
            assert z % 32 == 0 and f == 2, (rate, z, f)
        if v == 3:
            assert z in (32, 48, 64) and f == 4, (rate, z, f)


def test_message_rule_flags_are_checked():
    """At most one of WL_FLAG_OMS / WL_FLAG_MS; mixed batches and run-time
    codes decode with normalised min-sum only (checked before any device
    call)."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd import capi
    h = C.c_void_p()
    assert W.lib().wl_decoder_create(5, 96, 10, capi.WL_FLAG_OMS | capi.WL_FLAG_MS, C.byref(h)) == capi.WL_EINVAL
    assert W.lib().wl_mixed_create(10, capi.WL_FLAG_MS, C.byref(h)) == capi.WL_ENOTSUP
