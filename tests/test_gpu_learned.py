"""Learned offset min-sum on the GPU: inference kernel vs the oracle (bit
exact), forward/backward kernels vs the oracle's exact gradient, training,
checkpoint/resume, fallback, and the gain over the untrained network."""
import numpy as np
import pytest

from tests.helpers import bits_equal_f32, noisy_llr, pack_bits

pytestmark = pytest.mark.gpu


def _decode(torch, dec, llr_np):
    F, n = llr_np.shape
    llr = torch.from_numpy(llr_np).cuda()
    post = torch.empty_like(llr)
    hard = torch.empty((F, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    synd = torch.empty(F, dtype=torch.int16, device="cuda")
    dec.decode_dev(llr, post, hard, ret, iters, synd)
    torch.cuda.synchronize()
    return dict(post=post.cpu().numpy(), hard=hard.cpu().numpy(), ret=ret.cpu().numpy(),
                iters=iters.cpu().numpy(), synd=synd.cpu().numpy().view(np.uint16))


@pytest.mark.parametrize("rate,z", [(1, 96), (0, 24), (5, 96), (3, 40)])
def test_learned_inference_matches_oracle(gpu, oracle_mod, rate, z):
    import wimax_ldpc_lib_amd as W
    code = oracle_mod.Code(rate, z)
    llr, _ = noisy_llr(code, 40, 2.2 if rate < 3 else 3.2, seed=31 + rate)
    off = np.random.default_rng(rate).uniform(0.0, 0.6, (5, code.mb, code.c)).astype(np.float32)
    for early in (True, False):
        dec = W.ldpc_decoder(rate, z, 7, early_stop=early)    # iterations past T reuse offsets[T-1]
        dec.set_offsets(off)
        g = _decode(gpu, dec, llr)
        o = code.decode(llr, 7, early, variant=oracle_mod.LEARNED, offsets=off)
        assert bits_equal_f32(g["post"], o["post"])
        assert np.array_equal(g["hard"], pack_bits(o["hd"]))
        assert np.array_equal(g["ret"], o["ret"]) and np.array_equal(g["iters"], o["iters"])
        assert np.array_equal(g["synd"], o["synd_w"])


@pytest.mark.parametrize("rate,z", [(1, 24), (0, 24), (5, 96)])
def test_gradients_match_oracle(gpu, oracle_mod, rate, z):
    """GPU forward + backward against the oracle's exact backward: same loss
    and final posteriors, gradients within 1e-3 relative (float summation
    order differs)."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd.learned import Trainer
    torch = gpu
    code = oracle_mod.Code(rate, z)
    llr, cw = noisy_llr(code, 48, 2.0 if rate < 3 else 3.0, seed=41)
    T = 4
    off = np.random.default_rng(42).uniform(0.05, 0.4, (T, code.mb, code.c)).astype(np.float32)
    tr = Trainer(rate, z, T=T, batch=64, seed=1)
    tr.set_offsets(off)
    d_llr = torch.from_numpy(llr).cuda()
    d_bits = torch.from_numpy(cw).cuda()
    post = torch.empty_like(d_llr)
    loss, g = tr.grad(d_llr, d_bits, post)
    o_loss, o_post, o_g = code.learned_fwd_bwd(off, llr, cw)
    assert bits_equal_f32(post.cpu().numpy(), o_post)
    assert abs(loss - o_loss) <= 1e-4 * abs(o_loss)
    scale = np.abs(o_g).max()
    assert scale > 0
    assert np.abs(g - o_g).max() <= 1e-3 * scale, np.abs(g - o_g).max() / scale


def test_training_reduces_loss_and_resumes(gpu, oracle_mod, tmp_path):
    from wimax_ldpc_lib_amd.learned import Trainer
    tr = Trainer(1, 24, T=5, batch=512, seed=7, lr=1e-2)
    l0 = tr.step(1, 1.0, 4.0)
    l1 = tr.step(300, 1.0, 4.0)
    assert tr.global_step() == 301
    assert l1 < l0
    ck = tmp_path / "ckpt.bin"
    tr.save(ck)
    a = tr.offsets()
    la = tr.step(20, 1.0, 4.0)
    tr2 = Trainer(1, 24, T=5, batch=512, seed=999, lr=1e-2)
    tr2.load(ck)
    assert tr2.global_step() == 301 and np.array_equal(tr2.offsets(), a)
    lb = tr2.step(20, 1.0, 4.0)
    assert la == lb and np.array_equal(tr.offsets(), tr2.offsets())   # resumed run is identical


def test_fallback_reroutes_failed_frames(gpu, oracle_mod):
    """Frames the learned pass cannot clean (pure noise) are decoded again by
    plain normalised min-sum from their input LLRs; clean ones are not."""
    import wimax_ldpc_lib_amd as W
    torch = gpu
    code = oracle_mod.Code(1, 96)
    good, _ = noisy_llr(code, 16, 4.0, seed=51)
    noise = np.random.default_rng(52).normal(0, 1, (16, code.n)).astype(np.float32)
    llr = np.concatenate([good, noise])
    off = np.full((5, code.mb, code.c), 0.15, np.float32)
    dec = W.ldpc_decoder(1, 96, 5, early_stop=True)
    dec.set_offsets(off)
    F = llr.shape[0]
    d_llr = torch.from_numpy(llr).cuda()
    post = torch.empty_like(d_llr)
    hard = torch.empty((F, code.n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    fb = torch.empty(F, dtype=torch.uint8, device="cuda")
    dec.decode_learned_fallback_dev(d_llr, 10, post, hard, ret, iters, None, fb)
    torch.cuda.synchronize()
    fb = fb.cpu().numpy()
    o_l = code.decode(llr, 5, True, variant=oracle_mod.LEARNED, offsets=off)
    o_n = code.decode(llr, 10, True)
    assert np.array_equal(fb, o_l["ret"]) and fb[16:].all() and not fb[:16].any()
    want_post = np.where(fb[:, None] == 1, o_n["post"], o_l["post"])
    want_hd = np.where(fb[:, None] == 1, o_n["hd"], o_l["hd"])
    assert bits_equal_f32(post.cpu().numpy(), want_post)
    assert np.array_equal(hard.cpu().numpy(), pack_bits(want_hd))
    assert np.array_equal(ret.cpu().numpy(), np.where(fb == 1, o_n["ret"], 0))


def test_trained_offsets_beat_untrained(gpu, oracle_mod):
    """A short training run improves frame error rate of the 5-iteration
    learned decoder over its untrained start (all offsets 0.15) on fresh
    frames."""
    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd.learned import Trainer
    torch = gpu
    rate, z, T = 1, 24, 5
    tr = Trainer(rate, z, T=T, batch=1024, seed=3, lr=1e-2, init_offset=0.15)
    untrained = tr.offsets()
    tr.step(400, 1.0, 4.0)
    trained = tr.offsets()
    enc = W.ldpc_encoder(rate, z)
    F = 65536
    llr = torch.empty((F, 24 * z), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, 24 * z), dtype=torch.uint8, device="cuda")
    enc.channel_dev(987654, 10 ** 9, F, 2.5, llr=llr, bits=bits)
    ref = np.packbits(bits.cpu().numpy(), axis=1, bitorder="little")
    fer = {}
    for name, off in (("untrained", untrained), ("trained", trained)):
        dec = W.ldpc_decoder(rate, z, T, early_stop=False)
        dec.set_offsets(off)
        hard = torch.empty((F, 3 * z), dtype=torch.uint8, device="cuda")
        dec.decode_dev(llr, None, hard)
        fer[name] = float(np.mean((hard.cpu().numpy() != ref).any(axis=1)))
    assert fer["trained"] < 0.9 * fer["untrained"], fer
