"""GPU parity of the TDMP decoder against the CPU oracle, through the C ABI.

The oracle restates lib/ldpc_decoder.cc:122-262; every comparison here is
bit-exact: posterior float bits, packed hard decisions, decode()'s return
value, iteration counts and syndrome weights.
"""
import numpy as np
import pytest

from tests.helpers import Z_VALUES, bits_equal_f32, noisy_llr, pack_bits

pytestmark = pytest.mark.gpu


def _gpu_decode(torch, dec, llr_np, with_post=True):
    F, n = llr_np.shape
    llr = torch.from_numpy(llr_np).cuda()
    post = torch.empty_like(llr) if with_post else None
    hard = torch.empty((F, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    synd = torch.empty(F, dtype=torch.int16, device="cuda")
    dec.decode_dev(llr, post, hard, ret, iters, synd)
    torch.cuda.synchronize()
    out = dict(hard=hard.cpu().numpy(), ret=ret.cpu().numpy(), iters=iters.cpu().numpy(),
               synd=synd.cpu().numpy().view(np.uint16))
    if with_post:
        out["post"] = post.cpu().numpy()
    # the input buffer must be untouched
    assert bits_equal_f32(llr.cpu().numpy(), llr_np)
    return out


def _compare(g, o, label):
    assert np.array_equal(g["ret"], o["ret"]), label + " ret"
    assert np.array_equal(g["iters"], o["iters"]), label + " iters"
    assert np.array_equal(g["synd"], o["synd_w"]), label + " syndrome weight"
    assert np.array_equal(g["hard"], pack_bits(o["hd"])), label + " hard decisions"
    if "post" in g:
        bad = np.flatnonzero(g["post"].view(np.uint32) != o["post"].view(np.uint32))
        assert bad.size == 0, "%s posterior mismatch at %d positions, first %r: gpu %r oracle %r" % (
            label, bad.size, bad[:4], g["post"].ravel()[bad[:4]], o["post"].ravel()[bad[:4]])


def test_rate12_z24_1024_frames_early_stop(gpu, oracle_mod):
    """Rate 1/2, z=24, 1024 frames at 2.5 dB, max 10 iterations, early stop."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(0, 24)
    llr, _ = noisy_llr(code, 1024, 2.5, seed=11)
    o = code.decode(llr, max_iter=10, early_stop=True)
    dec = W.ldpc_decoder(0, 24, 10, early_stop=True)
    assert dec.small_batch >= 1024   # 256 compute units: this batch takes the small-batch kernel
    for small in (None, 0):
        dec.small_batch = small
        g = _gpu_decode(gpu, dec, llr)
        _compare(g, o, "rate 1/2 z=24 small_batch %r" % small)
    assert o["ret"].sum() >= 1 and len(np.unique(o["iters"])) > 3   # failures and varied stops


def test_rate56_z96_fixed10(gpu, oracle_mod):
    """Rate 5/6, z=96, exactly 10 iterations (no early stop), 512 frames at 4.0 dB."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(5, 96)
    llr, _ = noisy_llr(code, 512, 4.0, seed=12)
    o = code.decode(llr, max_iter=10, early_stop=False)
    dec = W.ldpc_decoder(5, 96, 10, early_stop=False)
    for small in (0, 1 << 30):   # the row-per-lane kernel, the small-batch kernel
        dec.small_batch = small
        _compare(_gpu_decode(gpu, dec, llr), o, "rate 5/6 z=96 small_batch %d" % small)


@pytest.mark.parametrize("rate", range(6))
def test_all_codes(gpu, oracle_mod, rate):
    """Every (rate, z), both stop modes, frames near the waterfall so that
    iteration counts vary (3/4B is decode-only: random LLRs, no codeword)."""
    import wimax_ldpc_lib_amd as W
    ebn0 = {0: 1.8, 1: 2.4, 2: 2.4, 3: 2.8, 4: 2.8, 5: 3.4}[rate]
    for z in Z_VALUES:
        code = oracle_mod.Code(rate, z)
        if rate == 4:
            rng = np.random.default_rng(100 + z)
            llr = rng.normal(1.0, 2.0, (24, code.n)).astype(np.float32)
        else:
            llr, _ = noisy_llr(code, 24, ebn0, seed=1000 * rate + z)
        for early in (True, False):
            o = code.decode(llr, max_iter=8, early_stop=early)
            dec = W.ldpc_decoder(rate, z, 8, early_stop=early)
            for small in (None, 0):   # the small-batch kernel (default at 24 frames), the row-per-lane one
                dec.small_batch = small
                _compare(_gpu_decode(gpu, dec, llr), o, "rate %d z %d early %d small_batch %r" % (rate, z, early, small))


@pytest.mark.parametrize("rate", range(6))
def test_every_code_in_every_any_frame_shape(gpu, oracle_mod, rate):
    """Every (rate, z) in the code shapes that take any frame count (0, 1,
    and 4: the hoisted shape compiled for 4 waves per SIMD), at the code's
    tuned frames per workgroup, both stop modes, an odd batch (a partly
    filled last workgroup), the row-per-lane kernel."""
    import wimax_ldpc_lib_amd as W
    ebn0 = {0: 1.8, 1: 2.4, 2: 2.4, 3: 2.8, 4: 2.8, 5: 3.4}[rate]
    for z in Z_VALUES:
        code = oracle_mod.Code(rate, z)
        if rate == 4:
            llr = np.random.default_rng(300 + z).normal(1.0, 2.0, (11, code.n)).astype(np.float32)
        else:
            llr, _ = noisy_llr(code, 11, ebn0, seed=5000 * rate + z)
        for early in (True, False):
            o = code.decode(llr, max_iter=7, early_stop=early)
            dec = W.ldpc_decoder(rate, z, 7, early_stop=early)
            dec.small_batch = 0
            f = dec.frames_per_wg
            for variant in (0, 1, 4):
                dec.variant = variant
                _compare(_gpu_decode(gpu, dec, llr), o,
                         "rate %d z %d early %d F %d variant %d" % (rate, z, early, f, variant))


@pytest.mark.parametrize("rate,z", [(0, 24), (5, 96), (3, 52), (1, 60)])
def test_frames_per_workgroup_and_ragged_batches(gpu, oracle_mod, rate, z):
    """Every frames-per-workgroup setting, in the code shapes that take any
    (0: hoisted slot addresses, 1: register-lean, 4: hoisted at 4 waves per
    SIMD), and batch sizes that leave partial workgroups."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(rate, z)
    llr, _ = noisy_llr(code, 37, 2.0 if rate < 3 else 3.0, seed=7 * z + rate)
    o = code.decode(llr, max_iter=6, early_stop=True)
    dec = W.ldpc_decoder(rate, z, 6, early_stop=True)
    dec.small_batch = 0   # the row-per-lane kernel at every batch size
    fmax = 1
    while True:
        try:
            dec.frames_per_wg = fmax + 1
            fmax += 1
        except W.WLError:
            break
    assert fmax >= 2
    for f in range(1, fmax + 1):
        dec.frames_per_wg = f
        for variant in (0, 1, 4):
            dec.variant = variant
            for nf in (1, f, 37):
                g = _gpu_decode(gpu, dec, np.ascontiguousarray(llr[:nf]))
                oo = {k: v[:nf] for k, v in o.items() if k != "seconds"}
                _compare(g, oo, "F=%d variant %d nframes=%d" % (f, variant, nf))


def test_edge_cases(gpu, oracle_mod):
    """Clean input (0 iterations, rx untouched), the max_iter = 0 quirk,
    |L| >= 10000 (the initialiser), signed zeros and all-zero input."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(5, 96)
    rng = np.random.default_rng(5)
    info = rng.integers(0, 2, (4, code.k), dtype=np.uint8)
    cw = code.encode(info)
    clean = np.where(cw == 1, 3.0, -3.0).astype(np.float32)
    huge = np.where(rng.random((4, code.n)) < 0.5, 20000.0, -15000.0).astype(np.float32)
    zeros = np.zeros((2, code.n), np.float32)
    negz = np.full((2, code.n), -0.0, np.float32)
    mixed = rng.normal(0, 1, (4, code.n)).astype(np.float32)
    mixed[:, ::7] = 0.0
    mixed[:, 3::11] = -0.0
    llr = np.concatenate([clean, huge, zeros, negz, mixed])
    for max_iter in (0, 1, 10):
        for early in (True, False):
            o = code.decode(llr, max_iter=max_iter, early_stop=early)
            dec = W.ldpc_decoder(5, 96, max_iter, early_stop=early)
            for small in (None, 0):
                dec.small_batch = small
                g = _gpu_decode(gpu, dec, llr)
                _compare(g, o, "edge max_iter %d early %d small_batch %r" % (max_iter, early, small))
                if early:
                    assert (g["iters"][:4] == 0).all() and bits_equal_f32(g["post"][:4], clean)
                    if max_iter == 0:
                        assert (g["ret"] == 0).all()   # lib/ldpc_decoder.cc:124,141 quirk


def test_host_api_semantics(gpu, oracle_mod):
    """decode(float* rx, int8_t* decoded): posterior in place, 0/1 bytes,
    return value; decode(int8_t*): the hard-input adapter."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(2, 48)
    llr, cw = noisy_llr(code, 3, 2.2, seed=3)
    dec = W.ldpc_decoder(2, 48, 10)
    for f in range(3):
        rx = llr[f].copy()
        decoded = np.zeros(code.n, np.int8)
        r = dec.decode(rx, decoded)
        o = code.decode(llr[f:f + 1], 10, True)
        assert r == o["ret"][0]
        assert bits_equal_f32(rx, o["post"][0])
        assert np.array_equal(decoded, o["hd"][0])
    bits = cw[0].astype(np.int8).copy()
    bits[::50] ^= 1
    keep = bits.copy()
    decoded = np.zeros(code.n, np.int8)
    r = dec.decode(bits, decoded)
    r2, hd2, _ = code.decode_i8(keep, 10, True)
    assert np.array_equal(bits, keep)
    assert r == r2 and np.array_equal(decoded, hd2)
    s = dec.compute_syndrome(decoded, False)
    assert s == code.syndrome(decoded, False)


def test_large_batch_properties(gpu, oracle_mod):
    """65536 frames of rate 5/6, z=96 at 4.0 dB, exactly 10 iterations:
    size-independent properties at full size (frames are independent, so
    any frame decoded alone equals its slot in the batch; every frame's
    syndrome weight is recomputed from its hard decisions) and a bit-exact
    oracle comparison on a spread sample of frames."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    code = oracle_mod.Code(5, 96)
    enc = W.ldpc_encoder(5, 96)
    F = 65536
    llr = torch.empty((F, code.n), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, code.n), dtype=torch.uint8, device="cuda")
    enc.channel_dev(1234, 0, F, 4.0, llr=llr, bits=bits)
    dec = W.ldpc_decoder(5, 96, 10, early_stop=False)
    hard = torch.empty((F, code.n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    synd = torch.empty(F, dtype=torch.int16, device="cuda")
    dec.decode_dev(llr, None, hard, ret, None, synd)
    torch.cuda.synchronize()
    hard_np = hard.cpu().numpy()
    sample = np.unique(np.concatenate([np.arange(0, F, 997), [F - 1]]))
    llr_s = llr[torch.from_numpy(sample).cuda()].cpu().numpy()
    o = code.decode(llr_s, 10, False)
    assert np.array_equal(hard_np[sample], pack_bits(o["hd"]))
    assert np.array_equal(ret.cpu().numpy()[sample], o["ret"])
    # frame errors against the transmitted codewords at 4.0 dB are rare
    cw_hard = np.packbits(bits.cpu().numpy(), axis=1, bitorder="little")
    fer = np.mean((cw_hard != hard_np).any(axis=1))
    assert fer < 1e-2
    H = code.dense_h().astype(np.uint8)
    hd = np.unpackbits(hard_np[:256], axis=1, bitorder="little")[:, :code.n]
    w = ((hd.astype(np.int64) @ H.T.astype(np.int64)) % 2).sum(axis=1)
    assert np.array_equal(w, synd.cpu().numpy()[:256])


@pytest.mark.parametrize("rate,z", [(5, 96), (0, 64), (3, 32), (4, 96)])
def test_two_halves_code_shape(gpu, oracle_mod, rate, z):
    """Code shape 2 (each wave holds 32 rows of two frames, per-wave compile-time
    slot addressing) equals the oracle in both stop modes, with a partial last
    workgroup (odd frame count); it needs two frames per workgroup, and another
    frames-per-workgroup value falls back to shape 1."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(rate, z)
    if rate == 4:
        llr = np.random.default_rng(z).normal(0.8, 2.0, (33, code.n)).astype(np.float32)
    else:
        llr, _ = noisy_llr(code, 33, {0: 1.8, 3: 2.8, 5: 3.4}[rate], seed=500 + z)
    for early in (True, False):
        o = code.decode(llr, max_iter=9, early_stop=early)
        dec = W.ldpc_decoder(rate, z, 9, early_stop=early)
        dec.small_batch = 0
        dec.variant = 0
        dec.frames_per_wg = 3
        with pytest.raises(W.WLError):
            dec.variant = 2
        dec.frames_per_wg = 2
        dec.variant = 2
        assert dec.variant == 2
        _compare(_gpu_decode(gpu, dec, llr), o, "shape 2 rate %d z %d early %d" % (rate, z, early))
        dec.frames_per_wg = 1
        assert dec.variant == 1
        fresh = W.ldpc_decoder(rate, z, 9, early_stop=early)
        dec.frames_per_wg = 0                     # back to the autotuned launch shape
        assert (dec.frames_per_wg, dec.variant) == (fresh.frames_per_wg, fresh.variant)
    with pytest.raises(W.WLError):
        W.ldpc_decoder(0, 24, 5).variant = 2      # z % 32 != 0: no such shape


@pytest.mark.parametrize("rate,z", [(5, 48), (0, 64), (2, 32), (4, 48)])
def test_quarters_code_shape(gpu, oracle_mod, rate, z):
    """Code shape 3 (each wave holds 16 rows of four frames, per-wave
    compile-time slot addressing) equals the oracle in both stop modes, with a
    partial last workgroup (35 frames = 8 x 4 + 3); it needs four frames per
    workgroup, and another value falls back to shape 1."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(rate, z)
    if rate == 4:
        llr = np.random.default_rng(z).normal(0.8, 2.0, (35, code.n)).astype(np.float32)
    else:
        llr, _ = noisy_llr(code, 35, {0: 1.8, 2: 2.4, 5: 3.4}[rate], seed=700 + z)
    for early in (True, False):
        o = code.decode(llr, max_iter=9, early_stop=early)
        dec = W.ldpc_decoder(rate, z, 9, early_stop=early)
        dec.small_batch = 0
        dec.variant = 0
        dec.frames_per_wg = 2
        with pytest.raises(W.WLError):
            dec.variant = 3
        dec.frames_per_wg = 4
        dec.variant = 3
        assert dec.variant == 3
        _compare(_gpu_decode(gpu, dec, llr), o, "shape 3 rate %d z %d early %d" % (rate, z, early))
        dec.frames_per_wg = 3
        assert dec.variant == 1
    with pytest.raises(W.WLError):
        W.ldpc_decoder(5, 96, 5).variant = 3      # more than four waves: no such shape


@pytest.mark.parametrize("rate,z", [(0, 24), (1, 60), (3, 32), (5, 96), (2, 92), (4, 36)])
def test_small_batch_kernel_matches_row_per_lane_kernel(gpu, oracle_mod, rate, z):
    """The small-batch kernel (one frame per workgroup, two lanes per row)
    forced onto a 3000-frame batch near the waterfall equals the row-per-lane
    kernel bit for bit in both stop modes (posteriors, hard decisions, status),
    and a spread sample equals the oracle."""
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(rate, z)
    F = 3000
    if rate == 4:
        llr = np.random.default_rng(z).normal(0.9, 2.0, (F, code.n)).astype(np.float32)
    else:
        llr, _ = noisy_llr(code, F, {0: 1.9, 1: 2.5, 2: 2.5, 3: 2.9, 5: 3.5}[rate], seed=900 + rate * 97 + z)
    sample = np.arange(0, F, 149)
    for early in (True, False):
        dec = W.ldpc_decoder(rate, z, 10, early_stop=early)
        dec.small_batch = 1 << 30
        a = _gpu_decode(gpu, dec, llr)
        dec.small_batch = 0
        b = _gpu_decode(gpu, dec, llr)
        for k in a:
            assert np.array_equal(a[k].view(np.uint8), b[k].view(np.uint8)), (k, early)
        o = code.decode(np.ascontiguousarray(llr[sample]), 10, early)
        _compare({k: v[sample] for k, v in a.items()}, o, "small-batch sample rate %d z %d early %d" % (rate, z, early))
        if early and rate != 4:   # 3/4B: random LLRs, no frame converges
            assert len(np.unique(a["iters"])) > 2


def test_small_batch_setting(gpu):
    """Default threshold per lifting size, 0 disables, None restores; the
    offset / plain min-sum rules report 0 (they never take that kernel)."""
    import wimax_ldpc_lib_amd as W
    d24, d96 = W.ldpc_decoder(0, 24, 10), W.ldpc_decoder(5, 96, 10)
    assert d24.small_batch > d96.small_batch > 0
    assert W.ldpc_decoder(0, 24, 10, early_stop=False).small_batch < d24.small_batch
    d24.small_batch = 0
    assert d24.small_batch == 0
    d24.small_batch = 77
    assert d24.small_batch == 77
    d24.small_batch = None
    assert d24.small_batch == W.ldpc_decoder(0, 24, 10).small_batch
    assert W.ldpc_decoder(0, 24, 10, rule="oms").small_batch == 0


def test_nan_input_stays_in_its_frame(gpu, oracle_mod):
    """A NaN LLR is outside the bit-exact contract: the reference is built
    with -ffast-math (lib/CMakeLists.txt:6), under which its comparisons with
    NaN are not defined, and the kernel's min3 / med3 selection does not skip
    NaN as the oracle's strict '<' does.  What is guaranteed: the decode of a
    batch with a NaN frame completes, the NaN posterior stays NaN, and every
    other frame of the batch (including the frames sharing the NaN frame's
    workgroup) equals the oracle bit for bit."""
    import wimax_ldpc_lib_amd as W
    for rate, z in ((5, 96), (0, 24), (3, 68)):
        code = oracle_mod.Code(rate, z)
        llr, _ = noisy_llr(code, 9, 2.5, seed=rate * 100 + z)
        bad = 4
        llr[bad, 17] = np.float32("nan")
        keep = np.array([f for f in range(9) if f != bad])
        for early in (True, False):
            o = code.decode(llr[keep], 10, early)
            dec = W.ldpc_decoder(rate, z, 10, early_stop=early)
            for small in (None, 0):
                dec.small_batch = small
                g = _gpu_decode(gpu, dec, llr)
                assert np.isnan(g["post"][bad, 17])
                gk = {k: v[keep] for k, v in g.items()}
                _compare(gk, o, "NaN batch rate %d z %d early %d small_batch %r" % (rate, z, early, small))


@pytest.mark.parametrize("rate, z", [(0, 24), (0, 96), (2, 56), (5, 32)])
def test_failures_only_in_late_layers(gpu, oracle_mod, rate, z):
    """The stop checks exit once a frame is known to fail (after layers 0, 1,
    3 and 7 in the row-per-lane kernel, after layer 0 in the small-batch
    kernel), and the fixed-iteration return value uses the same early-exit
    check; a frame whose only failing rows lie in the last base rows is the
    case where no early layer decides.  Strong noiseless LLRs of random
    codewords with one or two bits flipped in the last parity block column,
    which only the last two base rows touch: every kernel and stop mode,
    with and without the syndrome weight output, equals the oracle."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    code = oracle_mod.Code(rate, z)
    col0 = (24 - 1) * z   # first variable of the last block column
    rows = [m for m in range(code.m) if any(code.cn[m, j] == col0 for j in range(code.row_len[m]))]
    assert rows and min(rows) // z >= code.mb - 2, rows   # touched by the last two base rows only
    F = 67
    rng = np.random.default_rng(rate * 1000 + z)
    cw = code.encode(rng.integers(0, 2, (F, code.k), dtype=np.uint8))
    llr = np.where(cw == 1, 8.0, -8.0).astype(np.float32)
    for f in range(F):
        for t in {(3 * f) % z, (7 * f + 1) % z} if f % 3 == 0 else {(5 * f) % z}:
            llr[f, col0 + t] = -llr[f, col0 + t]
    for early, max_iter in ((True, 10), (True, 0), (False, 0), (False, 2)):
        o = code.decode(llr, max_iter, early)
        assert early or max_iter > 0 or (o["ret"] == 1).all()   # every frame fails its checks before decoding
        dec = W.ldpc_decoder(rate, z, max_iter, early_stop=early)
        for small in (None, 0):
            dec.small_batch = small
            label = "late-layer failures rate %d z %d early %d max_iter %d small_batch %r" % (
                rate, z, early, max_iter, small)
            _compare(_gpu_decode(torch, dec, llr), o, label)
            # return value without the syndrome weight output
            d_llr = torch.from_numpy(llr).cuda()
            hard = torch.empty((F, code.n // 8), dtype=torch.uint8, device="cuda")
            ret = torch.empty(F, dtype=torch.uint8, device="cuda")
            iters = torch.empty(F, dtype=torch.uint8, device="cuda")
            dec.decode_dev(d_llr, None, hard, ret, iters)
            torch.cuda.synchronize()
            assert np.array_equal(ret.cpu().numpy(), o["ret"]), label + " ret without synd"
            assert np.array_equal(iters.cpu().numpy(), o["iters"]), label + " iters without synd"


@pytest.mark.parametrize("rule", ["oms", "ms"])
@pytest.mark.parametrize("rate", range(6))
def test_offset_and_plain_min_sum_rules(gpu, oracle_mod, rule, rate):
    """The reference's other compile-time message rules
    (lib/ldpc_decoder.cc:20-21,225-238): offset min-sum (0.10, double) and
    plain min-sum (float), bit-exact against the oracle's restatement of
    each (itself pinned by an independent numpy decoder) for every rate at
    six lifting sizes, both stop modes, every frames-per-workgroup setting at
    one of them, and the special inputs of the edge-case test."""
    import wimax_ldpc_lib_amd as W
    variant = {"oms": oracle_mod.OMS, "ms": oracle_mod.MS}[rule]
    ebn0 = {0: 1.8, 1: 2.4, 2: 2.4, 3: 2.8, 4: 2.8, 5: 3.4}[rate]
    for z in (24, 32, 52, 68, 88, 96):
        code = oracle_mod.Code(rate, z)
        if rate == 4:
            llr = np.random.default_rng(200 + z).normal(1.0, 2.0, (24, code.n)).astype(np.float32)
        else:
            llr, _ = noisy_llr(code, 24, ebn0, seed=3000 * rate + z)
        for early in (True, False):
            o = code.decode(llr, max_iter=8, early_stop=early, variant=variant)
            dec = W.ldpc_decoder(rate, z, 8, early_stop=early, rule=rule)
            _compare(_gpu_decode(gpu, dec, llr), o, "%s rate %d z %d early %d" % (rule, rate, z, early))
    # frames per workgroup, partial workgroups
    code = oracle_mod.Code(rate, 48)
    llr, _ = noisy_llr(code, 37, ebn0, seed=77 + rate)
    o = code.decode(llr, max_iter=6, early_stop=True, variant=variant)
    dec = W.ldpc_decoder(rate, 48, 6, early_stop=True, rule=rule)
    f = 1
    while True:
        try:
            dec.frames_per_wg = f
        except W.WLError:
            break
        for nf in (1, 37):
            g = _gpu_decode(gpu, dec, np.ascontiguousarray(llr[:nf]))
            _compare(g, {k: v[:nf] for k, v in o.items() if k != "seconds"}, "%s F=%d nf=%d" % (rule, f, nf))
        f += 1
    assert f > 2
    # two frames per workgroup at z % 32 == 0 (the two-halves code shape),
    # an odd frame count (a half-filled last workgroup)
    for z in (32, 64, 96):
        code = oracle_mod.Code(rate, z)
        if rate == 4:
            llr = np.random.default_rng(400 + z).normal(1.0, 2.0, (37, code.n)).astype(np.float32)
        else:
            llr, _ = noisy_llr(code, 37, ebn0, seed=91 + 1000 * rate + z)
        for early in (True, False):
            o = code.decode(llr, max_iter=6, early_stop=early, variant=variant)
            dec = W.ldpc_decoder(rate, z, 6, early_stop=early, rule=rule)
            dec.frames_per_wg = 2
            _compare(_gpu_decode(gpu, dec, llr), o, "%s F=2 z %d early %d" % (rule, z, early))
    # special inputs: clean codewords, |L| >= 10000, signed zeros, max_iter 0
    if rate != 4:
        code = oracle_mod.Code(rate, 96)
        rng = np.random.default_rng(9)
        cw = code.encode(rng.integers(0, 2, (2, code.k), dtype=np.uint8))
        clean = np.where(cw == 1, 3.0, -3.0).astype(np.float32)
        huge = np.where(rng.random((2, code.n)) < 0.5, 20000.0, -15000.0).astype(np.float32)
        small = rng.normal(0, 0.05, (2, code.n)).astype(np.float32)   # |L| near the 0.10 offset
        small[:, ::5] = -0.0
        llr = np.concatenate([clean, huge, small])
        for max_iter in (0, 3):
            o = code.decode(llr, max_iter=max_iter, early_stop=True, variant=variant)
            dec = W.ldpc_decoder(rate, 96, max_iter, early_stop=True, rule=rule)
            _compare(_gpu_decode(gpu, dec, llr), o, "%s special max_iter %d" % (rule, max_iter))



@pytest.mark.parametrize("rate,z,F", [(5, 96, 4096), (0, 24, 64)])
def test_decode_recorded_in_a_graph(gpu, oracle_mod, rate, z, F):
    """wl_decode_f32_dev allocates nothing for device buffers, so a serving
    loop can record it into a HIP graph (torch.cuda.graph, side stream) and
    replay it on new inputs written into the captured buffer: three replays
    with eager work between them, each equal to an eager decode (the row-
    per-lane kernel at 4096 frames, the small-batch kernel at 64)."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    code = oracle_mod.Code(rate, z)
    dec = W.ldpc_decoder(rate, z, 10, early_stop=True)
    llr = torch.empty((F, code.n), dtype=torch.float32, device="cuda")
    post = torch.empty_like(llr)
    hard = torch.empty((F, code.n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    enc = W.ldpc_encoder(rate, z)
    enc.channel_dev(11, 0, F, 3.0, llr=llr)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        dec.decode_dev(llr, post, hard, ret, iters, stream=side.cuda_stream)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        dec.decode_dev(llr, post, hard, ret, iters, stream=side.cuda_stream)
    torch.cuda.synchronize()
    for seed in (21, 22, 23):
        enc.channel_dev(seed, 0, F, 2.0 + 0.5 * (seed - 21), llr=llr)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [t.clone() for t in (post, hard, ret, iters)]
        p2, h2 = torch.empty_like(post), torch.empty_like(hard)
        r2, i2 = torch.empty_like(ret), torch.empty_like(iters)
        dec.decode_dev(llr, p2, h2, r2, i2)
        torch.cuda.synchronize()
        assert torch.equal(got[0].view(torch.int32), p2.view(torch.int32)), seed
        assert torch.equal(got[1], h2) and torch.equal(got[2], r2) and torch.equal(got[3], i2), seed
    sample = llr[:8].cpu().numpy()
    o = code.decode(sample, 10, True)
    assert np.array_equal(hard[:8].cpu().numpy(), pack_bits(o["hd"]))
