"""The C++ drop-in surface, run as a reference user would: the test_encoder /
test_decoder command-line tools (lib/encode_example.cc, lib/decode_example.cc)
with their file formats, the ldpc_decoder class under an OpenMP frame loop
(lib/decode_example.cc:118-135), and BER_runner (lib/BER_runner.cc:162-204)
against a restatement of its frame source decoded by the oracle."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

from tests.helpers import bits_equal_f32, noisy_llr

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "wimax_ldpc_lib_amd", "bin")
CHECK = os.path.join(ROOT, "tests", "cpp", "_bin", "drop_in_check")


def _run(args, env=None, timeout=120):
    for a in args[:1]:
        assert os.path.exists(a), "%s is not built (python -m wimax_ldpc_lib_amd.build)" % a
    e = dict(os.environ)
    e.update(env or {})
    print("running:", " ".join(args), flush=True)
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout, env=e)


def _k_of(rate, n):
    return {0: n // 2, 1: n * 2 // 3, 2: n * 2 // 3, 3: n * 3 // 4, 4: n * 3 // 4, 5: n * 5 // 6}[rate]


@pytest.mark.parametrize("rate,z,rounds,errors,threads", [(0, 96, 40, 25, 4), (5, 48, 30, 4, 3), (3, 24, 10, 0, 1)])
def test_cli_encode_then_decode(gpu, oracle_mod, tmp_path, rate, z, rounds, errors, threads):
    code = oracle_mod.Code(rate, z)
    n, k = code.n, code.k
    data_f, enc_f, dec_f = (str(tmp_path / x) for x in ("data.bin", "encoded.bin", "decoded.bin"))
    p = _run([os.path.join(BIN, "test_encoder"), str(rate), str(z), str(rounds), str(errors), data_f, enc_f],
             env={"WL_SEED": "12345"})
    assert p.returncode == 1, p.stdout + p.stderr      # the reference's success status
    assert len(re.findall(r"^Round\[\d+\] Rate \(Mbits/Sec\): ", p.stdout, re.M)) == rounds
    data = np.fromfile(data_f, dtype=np.uint8).reshape(rounds, k)
    enc = np.fromfile(enc_f, dtype=np.uint8).reshape(rounds, n)
    assert set(np.unique(data)) <= {0, 1} and len(np.unique(data, axis=0)) == rounds
    cw = code.encode(data)
    flips = (enc != cw).sum(axis=1)
    assert (flips <= errors).all() and ((flips % 2) == (errors % 2)).all()

    # decode the file: only whole groups of `threads` frames, first k bytes each
    p = _run([os.path.join(BIN, "test_decoder"), str(threads), str(rate), str(z), str(rounds), enc_f, dec_f])
    assert p.returncode == 1, p.stdout + p.stderr
    used = (rounds // threads) * threads
    out = np.fromfile(dec_f, dtype=np.uint8)
    assert out.size == used * k
    out = out.reshape(used, k)
    init = [int(x) for x in re.findall(r"^Initial Number of Errors \(\d+\): (\d+)$", p.stdout, re.M)]
    final = [int(x) for x in re.findall(r"^Final Number of Errors \(\d+\): (\d+)$", p.stdout, re.M)]
    assert len(init) == used and len(final) == used
    for f in range(used):
        r, hd, _ = oracle_mod.Code(rate, z).decode_i8(enc[f].astype(np.int8), 10, True)
        assert np.array_equal(out[f], hd[:k].astype(np.uint8)), f
        assert final[f] == r
        assert init[f] == code.syndrome(enc[f].astype(np.int8), False)
    if errors == 0:
        assert np.array_equal(out, data[:used])


def test_cli_argument_and_file_errors(gpu, tmp_path):
    p = _run([os.path.join(BIN, "test_decoder")])
    assert p.returncode == 0 and "Usage: ./test_decoder" in p.stdout
    p = _run([os.path.join(BIN, "test_encoder"), "0", "24"])
    assert p.returncode == 0 and "Usage: ./test_encoder" in p.stdout
    p = _run([os.path.join(BIN, "test_encoder"), "0", "24", "1", "576", str(tmp_path / "a"), str(tmp_path / "b")])
    assert p.returncode == 0 and "Num_errors greater than codeword length" in p.stdout
    short = tmp_path / "short.bin"
    short.write_bytes(b"\x01" * 100)
    p = _run([os.path.join(BIN, "test_decoder"), "1", "0", "24", "2", str(short), str(tmp_path / "d")])
    assert p.returncode == 0 and "File does not contain enough data" in p.stdout
    # a bad rate throws from the constructor, as the reference's does
    ok = tmp_path / "one.bin"
    ok.write_bytes(b"\x00" * 576)
    p = _run([os.path.join(BIN, "test_decoder"), "1", "7", "24", "1", str(ok), str(tmp_path / "d")])
    assert p.returncode != 0 and p.returncode != 1


def test_cxx_decoder_objects_under_openmp(gpu, oracle_mod, tmp_path):
    rate, z, F = 2, 64, 48
    code = oracle_mod.Code(rate, z)
    llr, _ = noisy_llr(code, F, 2.3, seed=77)
    llr[3] = 0.0
    llr[4, ::5] = -0.0
    path = str(tmp_path / "llr.f32")
    llr.tofile(path)
    p = _run([CHECK, "frames", str(rate), str(z), "10", "6", path, str(tmp_path / "out")])
    assert p.returncode == 0, p.stdout + p.stderr
    o = code.decode(llr, 10, True)
    for pre in ("", "b"):
        post = np.fromfile(str(tmp_path / ("out.%spost" % pre)), dtype=np.float32).reshape(F, code.n)
        dec = np.fromfile(str(tmp_path / ("out.%sdec" % pre)), dtype=np.int8).reshape(F, code.n)
        ret = np.fromfile(str(tmp_path / ("out.%sret" % pre)), dtype=np.uint8)
        assert bits_equal_f32(post, o["post"]), pre
        assert np.array_equal(dec, o["hd"]), pre
        assert np.array_equal(ret, o["ret"]), pre


@pytest.mark.parametrize("rate,z,ebn0,F", [(0, 24, 1.5, 300), (5, 96, 3.5, 120)])
def test_ber_runner_matches_restated_frame_source(gpu, oracle_mod, tmp_path, rate, z, ebn0, F):
    """run_iteration, run_batch and a split run_batch count the same errors,
    and those equal the oracle's decode of the restated frames."""
    out = str(tmp_path / "ber")
    p = _run([CHECK, "ber", str(rate), str(z), "10", repr(ebn0), str(F), out])
    assert p.returncode == 0, p.stdout + p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    code = oracle_mod.Code(rate, z)
    n, k = r["n"], r["k"]
    assert (n, k) == (code.n, _k_of(rate, code.n)) and k == code.k
    assert r["a_bits"] == r["b_bits"] == r["c_bits"] == F * k
    assert r["a_errors"] == r["b_errors"] == r["c_errors"]
    data = np.fromfile(out + ".data", dtype=np.uint8).reshape(F, k)
    noise = np.fromfile(out + ".noise", dtype=np.float32).reshape(F, n)
    cw = code.encode(data)
    # lib/BER_runner.cc:170-191 with its operand types: float noise amplitude,
    # float sample + (double sqrt(0.5) * float amplitude) * float normal
    ebn0_lin = 10.0 ** (ebn0 / 10.0)
    R = {0: 0.5, 1: 2.0 / 3.0, 2: 2.0 / 3.0, 3: 0.75, 5: 5.0 / 6.0}[rate]
    amp = np.float32((1.0 / np.sqrt(ebn0_lin)) * np.sqrt(1.0 / (np.float64(np.float32(R)) * 1)))
    ch = (cw.astype(np.float64) * 2 - 1 + (np.sqrt(0.5) * np.float64(amp)) * noise.astype(np.float64)).astype(np.float32)
    o = code.decode(ch, 10, True)
    errors = int((o["hd"].astype(np.uint8) != cw).sum())
    assert errors == r["a_errors"], (errors, r)
    assert 0 < errors < F * n // 20     # the operating point has some errors, not a broken channel
