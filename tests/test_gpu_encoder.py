"""GPU encoder and channel generator through the C ABI.

Encoder: bit-exact against the oracle (which is pinned to the reference's
own encoder build) and against the committed reference codeword fixtures.
Channel: exact info bits against a numpy Philox4x32-10, LLRs against a
float64 numpy Box-Muller within 1e-5 relative, and H c = 0.
"""
import os

import numpy as np
import pytest

from tests.helpers import Z_VALUES

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("rate", [0, 1, 2, 3, 5])
def test_encoder_matches_oracle(gpu, oracle_mod, rate):
    """Every z of the rate, through both encoder kernels: 97 frames (one full
    64-frame bit-sliced tile and a ragged one) and 33 frames (below one
    tile: the byte-per-frame kernel), device and host calls."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    for z in Z_VALUES:
        code = oracle_mod.Code(rate, z)
        rng = np.random.default_rng(rate * 100 + z)
        enc = W.ldpc_encoder(rate, z)
        for F in (97, 33):
            info = rng.integers(0, 2, (F, code.k), dtype=np.uint8)
            info[0, :] = rng.integers(0, 256, code.k, dtype=np.uint8)   # nonzero bytes mean 1, copied as is
            want = code.encode(info)
            d_info = torch.from_numpy(info).cuda()
            d_cw = torch.empty((F, code.n), dtype=torch.uint8, device="cuda")
            enc.encode_dev(d_info, d_cw)
            torch.cuda.synchronize()
            assert np.array_equal(d_cw.cpu().numpy(), want), (rate, z, F)
            host = np.zeros((F, code.n), np.uint8)
            enc.encode_data(info, host)
            assert np.array_equal(host, want), (rate, z, F)


@pytest.mark.parametrize("rate, z", [(0, 96), (3, 28), (5, 24), (4, 52), (1, 64), (2, 36)])
def test_encoder_tiles_and_pointer_alignment(gpu, oracle_mod, rate, z):
    """Compiled codes are encoded 64 frames per workgroup as bit-sliced words
    with 16-byte accesses (k % 16 == 0) or 8-byte ones (3/4 at z % 8 == 4);
    pointers not aligned to that take the byte-per-frame kernel.  Three full
    tiles and a ragged one, bytes of any value (nonzero means 1, copied as
    given), at pointer offsets 0, 8 and 1: every frame equals the oracle and
    no byte outside the output is written."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    code = oracle_mod.Code(rate, z)
    F = 3 * 64 + 17
    rng = np.random.default_rng(7000 + rate * 100 + z)
    info = rng.integers(0, 2, (F, code.k), dtype=np.uint8)
    wild = rng.random((F, code.k)) < 0.1
    info[wild] = rng.integers(0, 256, int(wild.sum()), dtype=np.uint8)
    want = code.encode(info)
    enc = W.ldpc_encoder(rate, z)
    for off in (0, 8, 1):
        buf_in = torch.zeros(F * code.k + 64, dtype=torch.uint8, device="cuda")
        buf_in[off:off + F * code.k] = torch.from_numpy(info.ravel()).cuda()
        buf_out = torch.full((F * code.n + 64,), 0xAB, dtype=torch.uint8, device="cuda")
        enc.encode_dev(buf_in[off:off + F * code.k], buf_out[off:off + F * code.n])
        torch.cuda.synchronize()
        out = buf_out.cpu().numpy()
        assert np.array_equal(out[off:off + F * code.n].reshape(F, code.n), want), (rate, z, off)
        assert (out[:off] == 0xAB).all() and (out[off + F * code.n:] == 0xAB).all(), (rate, z, off)


def test_encoder_matches_reference_fixtures(gpu):
    import wimax_ldpc_lib_amd as W
    fx = np.load(os.path.join(GOLDEN, "encoder_ref.npz"))
    for key in fx.files:
        if not key.startswith("info_"):
            continue
        _, rate, z = key.split("_")
        rate, z = int(rate), int(z)
        info = fx[key]
        want = fx["cw_%d_%d" % (rate, z)]
        enc = W.ldpc_encoder(rate, z)
        got = np.zeros_like(want)
        enc.encode_data(np.ascontiguousarray(info), got)
        assert np.array_equal(got, want), (rate, z)


def test_encoder_3_4b_matches_reference_behaviour(gpu, oracle_mod):
    """3/4B: the reference encoder runs but its codewords fail H (as the
    reference's own encoder does); the GPU reproduces those bytes exactly."""
    import wimax_ldpc_lib_amd as W
    for z in (96, 28):
        code = oracle_mod.Code(4, z)
        for F in (4, 70):   # the byte-per-frame and the bit-sliced kernel
            info = np.random.default_rng(z + F).integers(0, 2, (F, code.k), dtype=np.uint8)
            got = np.zeros((F, code.n), np.uint8)
            W.ldpc_encoder(4, z).encode_data(info, got)
            assert np.array_equal(got, code.encode(info)), (z, F)


def _philox(c, k0, k1):
    M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
    c = [np.uint64(x) for x in c]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    m32 = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0 = np.uint64(M0) * c[0]
        p1 = np.uint64(M1) * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & m32
        hi1, lo1 = p1 >> np.uint64(32), p1 & m32
        c = [(hi1 ^ c[1] ^ k0) & m32, lo1, (hi0 ^ c[3] ^ k1) & m32, lo0]
        k0 = (k0 + np.uint64(W0)) & m32
        k1 = (k1 + np.uint64(W1)) & m32
    return [int(x) for x in c]


def test_channel_generator(gpu, oracle_mod):
    import wimax_ldpc_lib_amd as W
    torch = gpu
    rate, z, F, seed, eb = 1, 48, 6, 0x1234567890ABCDEF, 2.0
    code = oracle_mod.Code(rate, z)
    enc = W.ldpc_encoder(rate, z)
    llr = torch.empty((F, code.n), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, code.n), dtype=torch.uint8, device="cuda")
    enc.channel_dev(seed, 5, F, eb, llr=llr, bits=bits)
    torch.cuda.synchronize()
    llr, bits = llr.cpu().numpy(), bits.cpu().numpy()
    k0, k1 = seed & 0xFFFFFFFF, seed >> 32
    R = code.k / code.n
    s2 = 1.0 / (2.0 * R * 10.0 ** (eb / 10.0))
    for f in range(F):
        g = 5 + f
        info = np.zeros(code.k, np.uint8)
        for blk in range((code.k + 127) // 128):
            w = _philox([g & 0xFFFFFFFF, g >> 32, blk, 0], k0, k1)
            for q in range(128):
                j = blk * 128 + q
                if j < code.k:
                    info[j] = (w[q >> 5] >> (q & 31)) & 1
        cw = code.encode(info[None])[0]
        assert np.array_equal(bits[f], cw)
        assert not ((code.dense_h().astype(np.int64) @ cw) % 2).any()
        want = np.zeros(code.n)
        for v2 in range(0, code.n // 2, 37):
            w = _philox([g & 0xFFFFFFFF, g >> 32, v2, 1], k0, k1)
            u1, u2 = (w[0] + 0.5) / 2 ** 32, (w[1] + 0.5) / 2 ** 32
            r = np.sqrt(-2 * np.log(u1))
            n0, n1 = r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)
            for v, nn in ((2 * v2, n0), (2 * v2 + 1, n1)):
                y = (-1.0 if cw[v] else 1.0) + np.sqrt(s2) * nn
                want[v] = -2.0 * y / s2
                assert abs(llr[f, v] - want[v]) <= 1e-5 * max(1.0, abs(want[v])), (f, v)


def test_channel_statistics(gpu, oracle_mod):
    """Per-frame Eb/N0 draws stay in range and the LLR moments match
    BPSK/AWGN: E[LLR | bit] = -/+ 2/sigma^2, Var = 4/sigma^2 (reference sign)."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    code = oracle_mod.Code(5, 96)
    enc = W.ldpc_encoder(5, 96)
    F = 2048
    llr = torch.empty((F, code.n), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, code.n), dtype=torch.uint8, device="cuda")
    eb = torch.empty(F, dtype=torch.float32, device="cuda")
    enc.channel_dev(77, 0, F, 1.0, 4.5, llr=llr, bits=bits, ebn0=eb)
    e = eb.cpu().numpy()
    assert e.min() >= 1.0 and e.max() < 4.5 and abs(e.mean() - 2.75) < 0.1
    enc.channel_dev(78, 0, F, 4.0, llr=llr, bits=bits)
    x, b = llr.cpu().numpy().ravel().astype(np.float64), bits.cpu().numpy().ravel()
    s2 = 1.0 / (2.0 * code.k / code.n * 10 ** 0.4)
    assert abs(b.mean() - 0.5) < 0.01
    m1, m0 = x[b == 1].mean(), x[b == 0].mean()
    assert abs(m1 - 2 / s2) < 0.01 * 2 / s2 and abs(m0 + 2 / s2) < 0.01 * 2 / s2
    assert abs(x[b == 1].var() - 4 / s2) < 0.02 * 4 / s2


@pytest.mark.parametrize("rate", range(6))
def test_standard_encoder(gpu, oracle_mod, rate):
    """wl_encoder_create_standard, the back-substitution encoder: at 3/4B its
    codewords satisfy H at every z and equal the textbook GF(2) solution of
    H_p p = H_u u (unique: the parity part is invertible); at the other five
    rates they equal the reference encoder's (the oracle's) bit for bit.  Its
    channel frames decode back to the codewords sent."""
    import wimax_ldpc_lib_amd as W
    from tests.test_oracle_pins import _gf2_parity
    torch = gpu
    for z in Z_VALUES:
        code = oracle_mod.Code(rate, z)
        rng = np.random.default_rng(900 + rate * 100 + z)
        enc = W.ldpc_encoder(rate, z, standard=True)
        for F in (17, 81):   # the byte-per-frame and the bit-sliced kernel
            info = rng.integers(0, 2, (F, code.k), dtype=np.uint8)
            d_info = torch.from_numpy(info).cuda()
            d_cw = torch.empty((F, code.n), dtype=torch.uint8, device="cuda")
            enc.encode_dev(d_info, d_cw)
            torch.cuda.synchronize()
            cw = d_cw.cpu().numpy()
            assert np.array_equal(cw[:, :code.k], info)
            assert (code.dense_h().astype(np.int64) @ cw.T.astype(np.int64) % 2 == 0).all(), (rate, z, F)
            if rate != 4:
                assert np.array_equal(cw, code.encode(info)), (rate, z, F)
            elif z in (24, 48):
                H = code.dense_h()
                for f in (0, 1, F - 1):
                    assert np.array_equal(cw[f, code.k:], _gf2_parity(H, code.k, info[f])), (z, F, f)
    # BPSK / AWGN frames of the standard encoder at 6 dB decode to what was sent
    z = 48
    enc = W.ldpc_encoder(rate, z, standard=True)
    F = 64
    llr = torch.empty((F, enc.n), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, enc.n), dtype=torch.uint8, device="cuda")
    enc.channel_dev(5, 0, F, 6.0, llr=llr, bits=bits)
    dec = W.ldpc_decoder(rate, z, 10)
    hard = torch.empty((F, enc.n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    dec.decode_dev(llr, None, hard, ret)
    torch.cuda.synchronize()
    assert (ret.cpu().numpy() == 0).all()
    assert np.array_equal(hard.cpu().numpy(), np.packbits(bits.cpu().numpy(), axis=1, bitorder="little"))
