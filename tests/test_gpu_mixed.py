"""Mixed-(rate, z) batches: on-device bucketing + one captured graph of
per-code persistent kernels, checked bit-exact against per-code decoding."""
import numpy as np
import pytest

from tests.helpers import bits_equal_f32

pytestmark = pytest.mark.gpu


def _setup(torch, F, seed, codes, lo=1.0, hi=4.5):
    from wimax_ldpc_lib_amd.mixed import MixedDecoder
    rng = np.random.default_rng(seed)
    cfg_np = rng.choice(np.array(codes, dtype=np.uint8), F)
    cfg = torch.from_numpy(cfg_np).cuda()
    md = MixedDecoder(max_iter=10, early_stop=True)
    off = torch.empty(F, dtype=torch.int64, device="cuda")
    md.offsets_dev(cfg, off)
    n_of = np.array([24 * (24 + 4 * (int(c) % 19)) for c in cfg_np], dtype=np.int64)
    want = np.concatenate([[0], np.cumsum(n_of)[:-1]])
    assert np.array_equal(off.cpu().numpy(), want)
    total = int(n_of.sum())
    llr = torch.empty(total, dtype=torch.float32, device="cuda")
    bits = torch.empty(total, dtype=torch.uint8, device="cuda")
    md.channel_dev(seed, cfg, off, lo, hi, llr, bits)
    torch.cuda.synchronize()
    return md, cfg_np, cfg, off, want, n_of, llr, bits


def _mixed_decode(torch, md, cfg, off, llr):
    F = cfg.numel()
    total = llr.numel()
    post = torch.empty(total, dtype=torch.float32, device="cuda")
    hard = torch.empty(total // 8, dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    synd = torch.empty(F, dtype=torch.int16, device="cuda")
    md.decode_dev(cfg, off, llr, post, hard, ret, iters, synd)
    torch.cuda.synchronize()
    return post, hard, ret, iters, synd


def test_mixed_channel_matches_per_code_channel(gpu):
    import wimax_ldpc_lib_amd as W
    torch = gpu
    codes = [0 * 19 + 0, 1 * 19 + 9, 5 * 19 + 18, 3 * 19 + 7]
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, 64, 5, codes, 1.0, 4.5)
    llr_np, bits_np = llr.cpu().numpy(), bits.cpu().numpy()
    for c in codes:
        rate, z = int(c) // 19, 24 + 4 * (int(c) % 19)
        enc = W.ldpc_encoder(rate, z)
        sel = np.flatnonzero(cfg_np == c)
        G = int(sel.max()) + 1
        l1 = torch.empty((G, enc.n), dtype=torch.float32, device="cuda")
        b1 = torch.empty((G, enc.n), dtype=torch.uint8, device="cuda")
        enc.channel_dev(5, 0, G, 1.0, 4.5, llr=l1, bits=b1)
        l1, b1 = l1.cpu().numpy(), b1.cpu().numpy()
        for g in sel:
            o = offs[g]
            assert bits_equal_f32(llr_np[o:o + enc.n], l1[g])
            assert np.array_equal(bits_np[o:o + enc.n], b1[g])


def _check_against_per_code(torch, cfg_np, offs, llr, outs):
    """Every frame of a mixed decode equals the per-code decoder's result."""
    import wimax_ldpc_lib_amd as W
    post, hard, ret, iters, synd = outs
    post, hard = post.cpu().numpy(), hard.cpu().numpy()
    ret, iters, synd = ret.cpu().numpy(), iters.cpu().numpy(), synd.cpu().numpy()
    llr_np = llr.cpu().numpy()
    for c in np.unique(cfg_np):
        rate, z = int(c) // 19, 24 + 4 * (int(c) % 19)
        sel = np.flatnonzero(cfg_np == c)
        n = 24 * z
        frames = np.stack([llr_np[offs[g]:offs[g] + n] for g in sel])
        dec = W.ldpc_decoder(rate, z, 10, early_stop=True)
        d_llr = torch.from_numpy(frames).cuda()
        p1 = torch.empty_like(d_llr)
        h1 = torch.empty((len(sel), n // 8), dtype=torch.uint8, device="cuda")
        r1 = torch.empty(len(sel), dtype=torch.uint8, device="cuda")
        i1 = torch.empty(len(sel), dtype=torch.uint8, device="cuda")
        s1 = torch.empty(len(sel), dtype=torch.int16, device="cuda")
        dec.decode_dev(d_llr, p1, h1, r1, i1, s1)
        torch.cuda.synchronize()
        p1, h1 = p1.cpu().numpy(), h1.cpu().numpy()
        for q, g in enumerate(sel):
            o = offs[g]
            assert bits_equal_f32(post[o:o + n], p1[q]), (c, g)
            assert np.array_equal(hard[o // 8:(o + n) // 8], h1[q]), (c, g)
        assert np.array_equal(ret[sel], r1.cpu().numpy())
        assert np.array_equal(iters[sel], i1.cpu().numpy())
        assert np.array_equal(synd[sel], s1.cpu().numpy())


def test_mixed_decode_matches_per_code(gpu):
    """All 114 codes in one batch (3/4B frames carry the reference encoder's
    bytes, which fail H: they exercise non-converging decodes)."""
    torch = gpu
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, 6000, 11, list(range(114)))
    outs = _mixed_decode(torch, md, cfg, off, llr)
    assert md.invalid_count() == 0
    _check_against_per_code(torch, cfg_np, offs, llr, outs)


@pytest.mark.parametrize("nframes", [1, 3, 17])
def test_mixed_tiny_batches_match_per_code(gpu, nframes):
    """Batches smaller than one tile per code: most codes and lifting-size
    groups are empty, and the few tiles are partly filled."""
    torch = gpu
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, nframes, 40 + nframes, list(range(114)))
    outs = _mixed_decode(torch, md, cfg, off, llr)
    assert md.invalid_count() == 0
    _check_against_per_code(torch, cfg_np, offs, llr, outs)


def test_mixed_graph_replays(gpu):
    """The captured graph is replayed on new contents of the same buffers;
    every replay equals a freshly captured decode."""
    from wimax_ldpc_lib_amd.mixed import MixedDecoder
    torch = gpu
    codes = [c for c in range(114) if c // 19 != 4]
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, 3000, 21, codes)
    outs = []
    for rep in range(3):
        md.channel_dev(100 + rep, cfg, off, 1.0, 4.5, llr)
        torch.cuda.synchronize()
        got = [t.cpu().numpy() for t in _mixed_decode(torch, md, cfg, off, llr)]
        fresh = MixedDecoder(max_iter=10, early_stop=True)
        want = [t.cpu().numpy() for t in _mixed_decode(torch, fresh, cfg, off, llr)]
        for a, b in zip(got, want):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        outs.append(got[1])
    assert not np.array_equal(outs[0], outs[1])


def test_mixed_invalid_and_disabled_codes_are_skipped(gpu):
    torch = gpu
    codes = [5 * 19 + 18, 1 * 19 + 18]
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, 100, 31, codes)
    md.set_codes([5 * 19 + 18])
    total = llr.numel()
    post = torch.full((total,), 7.0, dtype=torch.float32, device="cuda")
    ret = torch.full((100,), 9, dtype=torch.uint8, device="cuda")
    md.decode_dev(cfg, off, llr, post, None, ret)
    torch.cuda.synchronize()
    skipped = cfg_np != 5 * 19 + 18
    assert md.invalid_count() == skipped.sum()
    r = ret.cpu().numpy()
    assert (r[skipped] == 9).all() and (r[~skipped] <= 1).all()
    p = post.cpu().numpy()
    g = int(np.flatnonzero(skipped)[0])
    assert (p[offs[g]:offs[g] + int(n_of[g])] == 7.0).all()


def test_mixed_reliability_order_leaves_results_unchanged(gpu):
    """Frames arranged by reliability inside the code buckets (the early-stop
    default) decode exactly as without the arrangement, including frames whose
    reliability estimate falls off either end of its range (all-zero LLRs,
    tiny and huge magnitudes), and those frames equal a per-code decode."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd.mixed import MixedDecoder
    torch = gpu
    codes = [0 * 19 + 0, 1 * 19 + 9, 2 * 19 + 3, 3 * 19 + 12, 5 * 19 + 18]
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, 4000, 41, codes)
    x = llr.cpu().numpy()
    edits = {0: 0.0, 1: 1e-6, 2: 1e30}
    for g, scale in edits.items():
        o, n = offs[g], int(n_of[g])
        x[o:o + n] *= np.float32(scale)
    llr = torch.from_numpy(x).cuda()
    got = [t.cpu().numpy() for t in _mixed_decode(torch, md, cfg, off, llr)]
    plain = MixedDecoder(max_iter=10, early_stop=True, ordered=False)
    want = [t.cpu().numpy() for t in _mixed_decode(torch, plain, cfg, off, llr)]
    for a, b in zip(got, want):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    post, hard, ret, iters, synd = got
    for g in edits:
        c = int(cfg_np[g])
        rate, z = c // 19, 24 + 4 * (c % 19)
        n, o = 24 * z, offs[g]
        dec = W.ldpc_decoder(rate, z, 10, early_stop=True)
        d = torch.from_numpy(np.ascontiguousarray(x[o:o + n]).reshape(1, n)).cuda()
        p1 = torch.empty_like(d)
        h1 = torch.empty((1, n // 8), dtype=torch.uint8, device="cuda")
        r1 = torch.empty(1, dtype=torch.uint8, device="cuda")
        i1 = torch.empty(1, dtype=torch.uint8, device="cuda")
        dec.decode_dev(d, p1, h1, r1, i1)
        torch.cuda.synchronize()
        assert np.array_equal(post[o:o + n].view(np.uint32), p1.cpu().numpy()[0].view(np.uint32)), g
        assert np.array_equal(hard[o // 8:(o + n) // 8], h1.cpu().numpy()[0]), g
        assert ret[g] == r1.cpu().numpy()[0] and iters[g] == i1.cpu().numpy()[0], g


def test_mixed_group_frames_per_workgroup_leave_results_unchanged(gpu):
    """Every frames-per-workgroup value of the group kernels (1, the largest
    that fits, and values between) decodes bit-identically; values outside
    1 .. largest are rejected, 0 restores the compiled default."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd.mixed import MixedDecoder
    torch = gpu
    codes = [r * 19 + zi for r in (0, 1, 2, 3, 5) for zi in (0, 1, 8, 18)]
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, 5000, 43, codes)
    want = [t.cpu().numpy() for t in _mixed_decode(torch, md, cfg, off, llr)]
    for z in (24, 28, 56, 96):
        default = md.launch(0, z)[0]
        fmax = 1
        while True:
            try:
                md.set_group_frames_per_wg(z, fmax + 1)
                fmax += 1
            except W.WLError:
                break
        assert fmax >= default >= 1
        with pytest.raises(W.WLError):
            md.set_group_frames_per_wg(z, -1)
        for f in sorted({1, 2, (fmax + 1) // 2, fmax}):
            md.set_group_frames_per_wg(z, f)
            assert md.launch(3, z)[0] == f
            got = [t.cpu().numpy() for t in _mixed_decode(torch, md, cfg, off, llr)]
            for a, b in zip(got, want):
                assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (z, f)
        md.set_group_frames_per_wg(z, 0)
        assert md.launch(0, z)[0] == default
    fresh = MixedDecoder(max_iter=10, early_stop=True)
    assert all(fresh.launch(0, z)[0] == md.launch(0, z)[0] for z in range(24, 97, 4))


def _gather_index(torch, offs, sel, n):
    """[len(sel), n] element indices of frames sel of a ragged buffer (device)."""
    base = torch.from_numpy(offs[sel].astype(np.int64)).cuda()
    return base[:, None] + torch.arange(n, device="cuda", dtype=torch.int64)[None, :]


def test_mixed_full_size_262144_frames(gpu):
    """The mixed workload at its full size: 262144 frames, rate uniform over
    the five encoder rates and z over the 19 lifting sizes (95 codes), Eb/N0
    uniform in [1, 4.5] dB per frame, max 10 iterations, early stop; one
    captured graph.  Checked: every frame is in exactly one code bucket, its
    own; every frame equals the per-code decoder (posteriors, hard decisions,
    return value, iterations, syndrome weight); a second replay into
    sentinel-filled outputs is identical; a spread sample equals the oracle.
    (A group kernel that kept its tile table in LDS decoded a few frames of
    each such batch from the next code's bucket at random; batches of a few
    thousand frames, one tile per workgroup, did not show it.)"""
    import wimax_ldpc_lib_amd as W
    from oracle import oracle as O
    torch = gpu
    F = 262144
    codes = [r * 19 + zi for r in (0, 1, 2, 3, 5) for zi in range(19)]
    md, cfg_np, cfg, off, offs, n_of, llr, bits = _setup(torch, F, 31337, codes)
    total = llr.numel()

    def run():
        post = torch.full((total,), float("nan"), dtype=torch.float32, device="cuda")
        hard = torch.full((total // 8,), 0xAB, dtype=torch.uint8, device="cuda")
        ret = torch.full((F,), 0xAB, dtype=torch.uint8, device="cuda")
        iters = torch.full((F,), 0xAB, dtype=torch.uint8, device="cuda")
        synd = torch.full((F,), -1, dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        md.decode_dev(cfg, off, llr, post, hard, ret, iters, synd)
        torch.cuda.synchronize()
        return post, hard, ret, iters, synd

    first = run()
    # (1) buckets: the bucket of code c holds exactly the frames of code c
    idx = torch.empty(F, dtype=torch.int32, device="cuda")
    b = md.bucket_list(idx).astype(np.int64)
    idx_np = idx.cpu().numpy().astype(np.int64)
    assert b[:, 1].sum() == F and len(np.unique(idx_np)) == F
    for c in range(114):
        got = np.sort(idx_np[b[c, 0]:b[c, 0] + b[c, 1]])
        assert np.array_equal(got, np.flatnonzero(cfg_np == c)), c
    # (2) every frame against the per-code decoder, on the device
    post, hard, ret, iters, synd = first
    for c in codes:
        rate, z = c // 19, 24 + 4 * (c % 19)
        n = 24 * z
        sel = np.flatnonzero(cfg_np == c)
        gi = _gather_index(torch, offs, sel, n)
        hi = _gather_index(torch, offs // 8, sel, n // 8)
        dec = W.ldpc_decoder(rate, z, 10, early_stop=True)
        frames = llr[gi].contiguous()
        p1 = torch.empty_like(frames)
        h1 = torch.empty((len(sel), n // 8), dtype=torch.uint8, device="cuda")
        r1 = torch.empty(len(sel), dtype=torch.uint8, device="cuda")
        i1 = torch.empty(len(sel), dtype=torch.uint8, device="cuda")
        s1 = torch.empty(len(sel), dtype=torch.int16, device="cuda")
        dec.decode_dev(frames, p1, h1, r1, i1, s1)
        torch.cuda.synchronize()
        s = torch.from_numpy(sel).cuda()
        assert torch.equal(post[gi].view(torch.int32), p1.view(torch.int32)), c
        assert torch.equal(hard[hi], h1), c
        assert torch.equal(ret[s], r1) and torch.equal(iters[s], i1) and torch.equal(synd[s], s1), c
    # (3) a second replay, outputs re-filled with sentinels
    second = run()
    for a, b2 in zip(first, second):
        assert torch.equal(a.view(torch.uint8), b2.view(torch.uint8))
    # (4) a spread sample against the oracle: two frames of every code
    llr_np = llr.cpu().numpy()
    post_np, hard_np = post.cpu().numpy(), hard.cpu().numpy()
    ret_np, it_np, sy_np = ret.cpu().numpy(), iters.cpu().numpy(), synd.cpu().numpy()
    rng = np.random.default_rng(7)
    for c in codes:
        rate, z = c // 19, 24 + 4 * (c % 19)
        n = 24 * z
        sel = rng.choice(np.flatnonzero(cfg_np == c), 2, replace=False)
        o = O.Code(rate, z).decode(np.stack([llr_np[offs[g]:offs[g] + n] for g in sel]), 10, True)
        for q, g in enumerate(sel):
            assert bits_equal_f32(post_np[offs[g]:offs[g] + n], o["post"][q]), (c, g)
            assert np.array_equal(hard_np[offs[g] // 8:(offs[g] + n) // 8],
                                  np.packbits(o["hd"][q].astype(np.uint8), bitorder="little")), (c, g)
            assert ret_np[g] == o["ret"][q] and it_np[g] == o["iters"][q] and sy_np[g] == o["synd_w"][q], (c, g)
