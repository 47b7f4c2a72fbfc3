"""Run-time quasi-cyclic codes on the GPU: the generic decoder against the
oracle (the reference's table-driven decoder restated) and against the
compiled 802.16e kernels, and the encoder for dual-diagonal run-time codes."""
import numpy as np
import pytest

from tests.helpers import bits_equal_f32, noisy_llr, pack_bits

pytestmark = pytest.mark.gpu


def _wimax_base(W, rate, z):
    p = W.code_params(rate, z)
    t = W.ldpc(rate, z).checknode_array()
    base = np.full((p["mb"], 24), -1, np.int16)
    for l in range(p["mb"]):
        for v in t[l * z]:
            if v >= 0:
                base[l, v // z] = v % z
    return base


def _dual_diagonal_base(rng, nb, mb, z, density=0.4):
    """A random base matrix with the 802.16e / 802.11n parity part."""
    kb = nb - mb
    base = np.full((mb, nb), -1, np.int16)
    info = rng.random((mb, kb)) < density
    info[np.arange(mb), rng.integers(0, kb, mb)] = True
    base[:, :kb] = np.where(info, rng.integers(0, z, (mb, kb)), -1)
    x = int(rng.integers(1, mb - 1))
    p = int(rng.integers(1, z))
    base[0, kb] = p
    base[mb - 1, kb] = p
    base[x, kb] = 0
    for i in range(mb - 1):
        base[i, kb + 1 + i] = 0
        base[i + 1, kb + 1 + i] = 0
    return base, x, p


def _decode(torch, dec, llr):
    F, n = llr.shape
    d_llr = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    post = torch.empty_like(d_llr)
    hard = torch.empty((F, (n + 7) // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    synd = torch.empty(F, dtype=torch.int16, device="cuda")
    dec.decode_dev(d_llr, post, hard, ret, iters, synd)
    torch.cuda.synchronize()
    return dict(post=post.cpu().numpy(), hard=hard.cpu().numpy(), ret=ret.cpu().numpy(),
                iters=iters.cpu().numpy(), synd=synd.cpu().numpy().astype(np.uint16))


def _compare(g, o, label):
    assert np.array_equal(g["ret"], o["ret"]), label
    assert np.array_equal(g["iters"], o["iters"]), label
    assert np.array_equal(g["synd"], o["synd_w"]), label
    assert np.array_equal(g["hard"], pack_bits(o["hd"])), label
    assert bits_equal_f32(g["post"], o["post"]), label


@pytest.mark.parametrize("rate,z", [(0, 24), (5, 96), (2, 52), (3, 76), (4, 40), (1, 60)])
def test_generic_decoder_on_wimax_codes(gpu, oracle_mod, rate, z):
    """Read back from alist text, an 802.16e code decodes on the generic
    kernel exactly as on its compiled kernel and as the oracle does."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    q = W.qc_code(alist=W.qc_code(base=_wimax_base(W, rate, z), z=z).alist())
    assert q.z == z
    code = oracle_mod.Code(rate, z)
    if rate == 4:
        llr = np.random.default_rng(z).normal(0.8, 2.0, (40, code.n)).astype(np.float32)
    else:
        llr, _ = noisy_llr(code, 40, {0: 1.8, 1: 2.4, 2: 2.4, 3: 2.9, 5: 3.5}[rate], seed=rate * 100 + z)
    for early in (True, False):
        g = _decode(torch, q.decoder(8, early_stop=early), llr)
        o = code.decode(llr, 8, early)
        _compare(g, o, "rate %d z %d early %d" % (rate, z, early))
        c = _decode(torch, W.ldpc_decoder(rate, z, 8, early_stop=early), llr)
        for key in ("post", "hard", "ret", "iters", "synd"):
            assert np.array_equal(np.asarray(g[key]).view(np.uint8), np.asarray(c[key]).view(np.uint8)), key


@pytest.mark.parametrize("nb,mb,z", [(24, 12, 27), (24, 8, 54), (24, 4, 81), (30, 10, 13), (16, 6, 100)])
def test_random_dual_diagonal_codes(gpu, oracle_mod, nb, mb, z):
    """802.11n-sized and arbitrary codes: the generic encoder's codewords
    satisfy H c = 0 with the info bits as the systematic part, its channel
    frames decode bit-exactly against the oracle in both stop modes."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    rng = np.random.default_rng(nb * 1000 + z)
    base, x, p = _dual_diagonal_base(rng, nb, mb, z)
    q = W.qc_code(base=base, z=z)
    assert (q.X, q.D) == (x, p)
    code = oracle_mod.Code(None, z, base=base)
    H = code.dense_h().astype(np.int64)
    enc = q.encoder()
    F = 24
    info = rng.integers(0, 2, (F, q.k), dtype=np.uint8)
    cw = np.zeros((F, q.n), np.uint8)
    enc.encode_data(info, cw)
    assert np.array_equal(cw[:, :q.k], info)
    assert not ((cw.astype(np.int64) @ H.T) % 2).any()
    llr = torch.empty((F, q.n), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, q.n), dtype=torch.uint8, device="cuda")
    enc.channel_dev(99, 0, F, 1.0, 3.0, llr=llr, bits=bits)
    torch.cuda.synchronize()
    llr = llr.cpu().numpy()
    assert not ((bits.cpu().numpy().astype(np.int64) @ H.T) % 2).any()
    for early in (True, False):
        g = _decode(torch, q.decoder(10, early_stop=early), llr)
        o = code.decode(llr, 10, early)
        _compare(g, o, "nb %d mb %d z %d early %d" % (nb, mb, z, early))
    assert len(np.unique(o["iters"])) > 1 or o["ret"].any()


def test_odd_length_code_and_host_path(gpu, oracle_mod):
    """n not a multiple of 8 (77): packed decisions are padded per frame, and
    the host decode() returns the oracle's posteriors and 0/1 bytes."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    rng = np.random.default_rng(11)
    base, _, _ = _dual_diagonal_base(rng, 11, 5, 7)
    q = W.qc_code(base=base, z=7)
    code = oracle_mod.Code(None, 7, base=base)
    llr = rng.normal(0.5, 1.5, (9, q.n)).astype(np.float32)
    g = _decode(torch, q.decoder(6), llr)
    o = code.decode(llr, 6, True)
    _compare(g, o, "n=77")
    dec = q.decoder(6)
    rx = llr.copy()
    decoded = np.zeros((9, q.n), np.int8)
    ret = np.zeros(9, np.uint8)
    dec.decode_batch_host(rx, decoded, ret)
    assert bits_equal_f32(rx, o["post"]) and np.array_equal(decoded, o["hd"]) and np.array_equal(ret, o["ret"])
    assert dec.compute_syndrome(decoded[0], False) == code.syndrome(decoded[0], False)


def test_non_quasi_cyclic_code(gpu, oracle_mod):
    """z = 1: every check its own layer, in the reference's row order."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    rng = np.random.default_rng(7)
    n, m = 48, 24
    base = np.full((m, n), -1, np.int16)
    for r in range(m):
        base[r, rng.choice(n, 6, replace=False)] = 0
    q = W.qc_code(base=base, z=1)
    code = oracle_mod.Code(None, 1, base=base)
    llr = rng.normal(0.3, 1.2, (16, n)).astype(np.float32)
    for early in (True, False):
        _compare(_decode(torch, q.decoder(7, early_stop=early), llr), code.decode(llr, 7, early), "z=1")
    with pytest.raises(W.WLError):
        q.encoder()
