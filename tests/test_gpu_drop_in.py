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


def _restated_channel(code, rate, noise, cw, ebn0):
    """lib/BER_runner.cc:170-191 with its operand types."""
    R = {0: 0.5, 1: 2.0 / 3.0, 2: 2.0 / 3.0, 3: 0.75, 5: 5.0 / 6.0}[rate]
    amp = np.float32((1.0 / np.sqrt(10.0 ** (ebn0 / 10.0))) * np.sqrt(1.0 / (np.float64(np.float32(R)) * 1)))
    return (cw.astype(np.float64) * 2 - 1
            + (np.sqrt(0.5) * np.float64(amp)) * noise.astype(np.float64)).astype(np.float32)


@pytest.mark.parametrize("rate,z,ebn0,max_frames,repeats,clean_runs",
                         [(0, 24, 2.5, 150, 6, True), (5, 48, 3.4, 120, 5, False)])
def test_run_until_error_matches_reference_loop(gpu, rate, z, ebn0, max_frames, repeats, clean_runs):
    """BER_runner::run_until_error (batched) stops where the reference's
    frame-by-frame loop stops, with the same counters, and leaves the frame
    stream at the same place."""
    p = _run([CHECK, "until", str(rate), str(z), repr(ebn0), str(max_frames), str(repeats)])
    assert p.returncode == 0, p.stdout + p.stderr
    r = json.loads(p.stdout.strip().splitlines()[-1])
    assert r["loop"] == r["until"], r
    assert r["tail_loop"] == r["tail_until"]
    flags = [x[1] for x in r["loop"]]
    assert sum(flags) > 0, r                   # some runs stop at an error
    if clean_runs:
        assert sum(flags) < len(flags), r      # and some run clean to max_frames


def test_ber_example_cli(gpu, oracle_mod, tmp_path):
    """BER_example's printed BER per point equals the oracle's decode of the
    restated frame stream with the reference's per-point sample sizes."""
    import math
    rate, z = 0, 24
    cwd = os.getcwd()
    os.chdir(str(tmp_path))
    try:
        p = _run([os.path.join(BIN, "BER_example"), "1", str(rate), str(z), "0", "2", "1"], timeout=300)
    finally:
        os.chdir(cwd)
    assert p.returncode == 1, p.stdout + p.stderr
    printed = [float(x) for x in re.findall(r"^BER: (\S+)$", p.stdout, re.M)]
    points = [float(x) for x in re.findall(r"^EbN0\(dB\): (\S+)$", p.stdout, re.M)]
    assert points == [0.0, 1.0, 2.0] and len(printed) == 3
    code = oracle_mod.Code(rate, z)
    n, k = code.n, code.k
    F = 4000
    q = _run([CHECK, "stream", str(rate), str(z), str(F), str(tmp_path / "s")])
    assert q.returncode == 0, q.stdout + q.stderr
    data = np.fromfile(str(tmp_path / "s.data"), dtype=np.uint8).reshape(F, k)
    noise = np.fromfile(str(tmp_path / "s.noise"), dtype=np.float32).reshape(F, n)
    cw = code.encode(data)
    pos = 0

    def errors(lo, hi, eb):
        o = code.decode(_restated_channel(code, rate, noise[lo:hi], cw[lo:hi], eb), 10, True)
        return int((o["hd"].astype(np.uint8) != cw[lo:hi]).sum())

    for eb, ber_printed in zip(points, printed):
        W = 46080 // n
        e = errors(pos, pos + W, eb)
        ber = e / (W * k)
        num_req = math.floor(3.912 / 1e-8) if ber == 0.0 else int(10 * math.floor(3.912 / ber))
        N2 = num_req // n
        assert pos + W + N2 <= F
        e += errors(pos + W, pos + W + N2, eb) if N2 else 0
        ber = e / ((W + N2) * k)
        assert "%e" % ber == "%e" % ber_printed, (eb, ber, ber_printed)
        pos += W + N2
    lines = open(str(tmp_path / "BER_LDPC_RATE0_N576_K288.csv")).read().strip().splitlines()
    assert lines[0].startswith("ebno_db,ber") and len(lines) == 4
    assert [float(l.split(",")[1]) for l in lines[1:]] == pytest.approx(printed, rel=1e-5)


def test_capacity_example_cli(gpu, tmp_path):
    """The Eb/N0 search at z = 24: one threshold per rate, ordered by rate,
    above the BPSK-AWGN limit and within a few dB of it."""
    cwd = os.getcwd()
    os.chdir(str(tmp_path))
    try:
        p = _run([os.path.join(BIN, "Capacity_example"), "1", "24"], timeout=600)
    finally:
        os.chdir(cwd)
    assert p.returncode == 1, p.stdout + p.stderr
    found = re.findall(r"^Rate: (\S+), EbNo: (\S+)$", p.stdout, re.M)
    assert [r for r, _ in found] == ["1/2", "2/3A", "2/3B", "3/4A", "5/6"]
    lines = open(str(tmp_path / "CAP_LDPC_N576.csv")).read().strip().splitlines()
    assert lines[0] == "rate,plot_rate,ebno_db,shannon_limit_db,bpsk_limit_db,frames_run" and len(lines) == 6
    rows = [l.split(",") for l in lines[1:]]
    eb = [float(r[2]) for r in rows]
    for r, (name, printed) in zip(rows, found):
        assert r[0] == name and abs(float(r[2]) - float(printed)) < 1e-6
        shannon, bpsk = float(r[3]), float(r[4])
        assert shannon < bpsk < float(r[2]) < bpsk + 6.0
    assert eb[0] < eb[1] and eb[0] < eb[2] and max(eb[1], eb[2]) < eb[4] and eb[3] < eb[4]
