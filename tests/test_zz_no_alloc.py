"""wl_decode_learned_fallback_noalloc_dev (static device scratch, no
allocation) against wl_decode_learned_fallback_dev, through the C ABI: two
calls in a row, the second smaller, on the persistent re-queue path (z = 96)
and on the compaction path (z = 48); the frame limit; and the entry point
recorded into a HIP graph and replayed."""
import ctypes as C

import numpy as np
import pytest

from tests.helpers import noisy_llr

pytestmark = pytest.mark.gpu


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _run(torch, L, fn, dec, llr):
    F, n = llr.shape
    post = torch.full_like(llr, float("nan"))
    hard = torch.full((F, n // 8), 0xAB, dtype=torch.uint8, device="cuda")
    ret = torch.full((F,), 0xAB, dtype=torch.uint8, device="cuda")
    iters = torch.full((F,), 0xAB, dtype=torch.uint8, device="cuda")
    synd = torch.full((F,), -1, dtype=torch.int16, device="cuda")
    fb = torch.full((F,), 0xAB, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(L, fn)(dec._h, _ptr(llr), F, 10, _ptr(post), _ptr(hard), _ptr(ret), _ptr(iters), _ptr(synd),
                        _ptr(fb), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, L.wl_last_error()
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (post, hard, ret, iters, synd, fb)]


@pytest.mark.parametrize("z", [96, 48])
def test_noalloc_matches_allocating_entry_point(gpu, oracle_mod, z):
    import wimax_ldpc_lib_amd as W
    torch = gpu
    L = W.lib()
    code = oracle_mod.Code(1, z)
    dec = W.ldpc_decoder(1, z, 5, early_stop=True)
    dec.set_offsets(np.random.default_rng(z).uniform(0.1, 0.5, (5, code.mb, code.c)).astype(np.float32))
    for nf, seed in ((300, 1), (77, 2)):
        good, _ = noisy_llr(code, nf - nf // 4, 2.2, seed=seed)
        noise = np.random.default_rng(seed + 10).normal(0, 1, (nf // 4, code.n)).astype(np.float32)
        llr = torch.from_numpy(np.concatenate([good, noise])).cuda()
        want = _run(torch, L, "wl_decode_learned_fallback_dev", dec, llr)
        got = _run(torch, L, "wl_decode_learned_fallback_noalloc_dev", dec, llr)
        assert want[5].sum() >= nf // 4          # some frames fell back
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
        for a, b in zip(got[1:], want[1:]):
            assert np.array_equal(a, b)


def test_noalloc_frame_limit(gpu):
    """More than WL_NOALLOC_MAX_FRAMES (1 << 20) frames: WL_EINVAL, before
    any device access."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd import capi
    L = W.lib()
    dec = W.ldpc_decoder(1, 96, 5, early_stop=True)
    dec.set_offsets(np.full((5, 8, 10), 0.15, np.float32))
    dummy = C.c_void_p(16)
    rc = L.wl_decode_learned_fallback_noalloc_dev(dec._h, dummy, (1 << 20) + 1, 10, None, None, dummy, None, None,
                                                  None, None)
    assert rc == capi.WL_EINVAL and b"allocation-free limit" in L.wl_last_error()


@pytest.mark.parametrize("z", [96, 48])
def test_noalloc_under_graph_capture(gpu, oracle_mod, z):
    """The allocation-free entry point recorded into a HIP graph on a side
    stream (torch.cuda.graph) and replayed on three new inputs written into
    the captured input buffer, with eager work and synchronisations between
    the replays: every replay equals the allocating entry point on the same
    input (persistent re-queue path at z = 96, compaction path at z = 48)."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    L = W.lib()
    code = oracle_mod.Code(1, z)
    dec = W.ldpc_decoder(1, z, 5, early_stop=True)
    dec.set_offsets(np.random.default_rng(z + 1).uniform(0.1, 0.5, (5, code.mb, code.c)).astype(np.float32))
    F, n = 120, code.n

    def inputs(seed):
        good, _ = noisy_llr(code, F - F // 4, 2.2, seed=seed)
        noise = np.random.default_rng(seed + 10).normal(0, 1, (F // 4, n)).astype(np.float32)
        return torch.from_numpy(np.concatenate([good, noise])).cuda()

    llr = inputs(100)
    post = torch.empty_like(llr)
    hard = torch.empty((F, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    synd = torch.empty(F, dtype=torch.int16, device="cuda")
    fb = torch.empty(F, dtype=torch.uint8, device="cuda")
    side = torch.cuda.Stream()

    def call():
        rc = L.wl_decode_learned_fallback_noalloc_dev(dec._h, _ptr(llr), F, 10, _ptr(post), _ptr(hard), _ptr(ret),
                                                      _ptr(iters), _ptr(synd), _ptr(fb), C.c_void_p(side.cuda_stream))
        assert rc == 0, L.wl_last_error()

    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        call()   # the first call on this device resolves the scratch address, outside the capture
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        call()
    torch.cuda.synchronize()
    for seed in (1, 2, 3):
        new = inputs(seed)
        llr.copy_(new)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [t.cpu().numpy() for t in (post, hard, ret, iters, synd, fb)]
        want = _run(torch, L, "wl_decode_learned_fallback_dev", dec, new)
        assert want[5].sum() >= F // 4, seed
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32)), seed
        for a, b in zip(got[1:], want[1:]):
            assert np.array_equal(a, b), seed
