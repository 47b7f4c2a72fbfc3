"""Measurements of the decoder workloads other than the headline bench.

  mixed  262144 frames, rate uniform over {1/2, 2/3A, 2/3B, 3/4A, 5/6}, z
         uniform over the 19 values, Eb/N0 uniform in [1.0, 4.5] dB, max 10
         iterations with early stop, decoded as one mixed batch (one captured
         graph of per-code persistent kernels); compared with decoding each
         code's frames by its own launch, bit for bit
  fallback 524288 frames, the five encoder rates at z = 96, Eb/N0 uniform in
         [1.5, 5.0] dB; learned offsets per rate (trained here if missing),
         learned 5 iterations with in-kernel re-queue of syndrome failures to
         normalised min-sum (10 iterations), against plain normalised
         min-sum (10 iterations, with early stop and at exactly 10) on the
         same frames: info Gbit/s, fallback fraction, BER / FER per rate

    python -m wimax_ldpc_lib_amd.workloads mixed|fallback|groups [--out file.json]

  groups per lifting size, the mixed graph over the five rates of that z
         against the sum of the five per-code decodes at full occupancy, and
         that per-code sum scaled to the mixed workload's draw

(the small rate 1/2, z = 24 workload, measured against the CPU decoder, is
bench.py --workload small)
"""
import argparse
import json
import os
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(PKG))
RATES5 = [0, 1, 2, 3, 5]
K_OF = {0: 12, 1: 16, 2: 16, 3: 18, 5: 20}   # info blocks per rate


def log(*a):
    print(*a, flush=True)


def timed(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


def mixed_inputs(torch, md, np, F, codes, seed, lo, hi):
    rng = np.random.default_rng(seed)
    cfg_np = rng.choice(np.array(codes, dtype=np.uint8), F)
    cfg = torch.from_numpy(cfg_np).cuda()
    off = torch.empty(F, dtype=torch.int64, device="cuda")
    md.offsets_dev(cfg, off)
    n_of = 24 * (24 + 4 * (cfg_np.astype(np.int64) % 19))
    total = int(n_of.sum())
    llr = torch.empty(total, dtype=torch.float32, device="cuda")
    bits = torch.empty(total, dtype=torch.uint8, device="cuda")
    md.channel_dev(seed, cfg, off, lo, hi, llr, bits)
    torch.cuda.synchronize()
    return cfg_np, cfg, off, llr, bits, n_of


def mixed(a, torch, W, np):
    from wimax_ldpc_lib_amd.mixed import MixedDecoder, code_id
    F = a.frames or 262144
    codes = [code_id(r, z) for r in RATES5 for z in range(24, 97, 4)]
    md = MixedDecoder(max_iter=10, early_stop=True, codes=codes)
    cfg_np, cfg, off, llr, bits, n_of = mixed_inputs(torch, md, np, F, codes, a.seed, 1.0, 4.5)
    total = llr.numel()
    k_of = np.array([K_OF[int(c) // 19] * (24 + 4 * (int(c) % 19)) for c in cfg_np], dtype=np.int64)
    info_bits = int(k_of.sum())
    post = torch.empty_like(llr)
    hard = torch.empty(total // 8, dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    ms = timed(torch, lambda: md.decode_dev(cfg, off, llr, None, hard, ret, iters), a.reps)
    log("mixed graph: %.3f ms, %.2f info Gbit/s" % (ms, info_bits / (ms * 1e-3) / 1e9))
    hard_m, ret_m, it_m = hard.cpu().numpy(), ret.cpu().numpy(), iters.cpu().numpy()
    # per-code launches over gathered frames, and the comparison
    llr_np = llr.cpu().numpy()
    offs = off.cpu().numpy()
    per = []
    same = True
    tot_ms = 0.0
    for c in codes:
        sel = np.flatnonzero(cfg_np == c)
        if not len(sel):
            continue
        rate, z = c // 19, 24 + 4 * (c % 19)
        n = 24 * z
        frames = torch.from_numpy(np.stack([llr_np[offs[g]:offs[g] + n] for g in sel])).cuda()
        dec = W.ldpc_decoder(rate, z, 10, early_stop=True)
        h1 = torch.empty((len(sel), n // 8), dtype=torch.uint8, device="cuda")
        r1 = torch.empty(len(sel), dtype=torch.uint8, device="cuda")
        i1 = torch.empty(len(sel), dtype=torch.uint8, device="cuda")
        tot_ms += timed(torch, lambda: dec.decode_dev(frames, None, h1, r1, i1), 3)
        h1 = h1.cpu().numpy()
        for q, g in enumerate(sel):
            if not np.array_equal(hard_m[offs[g] // 8:(offs[g] + n) // 8], h1[q]):
                same = False
        same = same and np.array_equal(ret_m[sel], r1.cpu().numpy()) and np.array_equal(it_m[sel], i1.cpu().numpy())
        per.append(len(sel))
    out = {"frames": F, "codes": len(per), "info_bits": info_bits, "mixed_ms": ms,
           "mixed_info_gbps": info_bits / (ms * 1e-3) / 1e9, "per_code_launches_ms_sum": tot_ms,
           "per_code_info_gbps": info_bits / (tot_ms * 1e-3) / 1e9, "bit_exact_vs_per_code": bool(same),
           "avg_iters": float(it_m.mean()), "syndrome_fail_rate": float(ret_m.mean()),
           "frames_per_code_min_max": [int(min(per)), int(max(per))]}
    log(json.dumps(out))
    return out


def groups(a, torch, W, np):
    """Per lifting size z: the mixed graph over the five rates of that z
    alone (5 x F frames, rate uniform, Eb/N0 uniform in [1.0, 4.5] dB, 10
    iterations with early stop) against the sum of the five per-code decodes
    of F frames each, every code alone on its own reliability-sorted frames
    (mean |LLR| of the first 256 LLRs) at its tuned shape: the cost of
    grouping at a size where every launch fills the chip.  The sum of the
    per-code times scaled to the mixed workload's draw (262144 frames over
    95 codes) is the full-occupancy figure the mixed graph is compared with."""
    from wimax_ldpc_lib_amd.mixed import MixedDecoder, code_id
    F = a.frames or 65536
    zs = list(range(24, 97, 4))
    draw = np.random.default_rng(a.seed).choice(
        np.array([code_id(r, z) for r in RATES5 for z in zs], dtype=np.uint8), 262144)
    rows, bound = [], 0.0
    for z in zs:
        per = 0.0
        for r in RATES5:
            enc = W.ldpc_encoder(r, z)
            llr = torch.empty((F, enc.n), dtype=torch.float32, device="cuda")
            enc.channel_dev(a.seed + 100 * r + z, 0, F, 1.0, 4.5, llr=llr)
            llr = llr.index_select(0, torch.argsort(llr[:, :256].abs().mean(dim=1))).contiguous()
            hard = torch.empty((F, enc.n // 8), dtype=torch.uint8, device="cuda")
            dec = W.ldpc_decoder(r, z, 10, early_stop=True)
            dec.small_batch = 0
            ms = timed(torch, lambda: dec.decode_dev(llr, None, hard), a.reps)
            per += ms
            bound += ms * int((draw == code_id(r, z)).sum()) / F
            del llr, hard
        codes = [code_id(r, z) for r in RATES5]
        md = MixedDecoder(max_iter=10, early_stop=True, codes=codes)
        cfg_np, cfg, off, llr, bits, n_of = mixed_inputs(torch, md, np, 5 * F, codes, a.seed + z, 1.0, 4.5)
        hard = torch.empty(llr.numel() // 8, dtype=torch.uint8, device="cuda")
        ret = torch.empty(5 * F, dtype=torch.uint8, device="cuda")
        g = timed(torch, lambda: md.decode_dev(cfg, off, llr, None, hard, ret), a.reps)
        md.close()
        del llr, bits, hard
        rows.append({"z": z, "group_graph_ms": g, "per_code_sum_ms": per, "ratio": g / per})
        log("z %2d: group graph %.3f ms, per-code sum %.3f ms, ratio %.3f" % (z, g, per, g / per))
    out = {"frames_per_code": F, "per_z": rows,
           "group_total_ms": sum(r["group_graph_ms"] for r in rows),
           "per_code_total_ms": sum(r["per_code_sum_ms"] for r in rows),
           "per_code_full_occupancy_ms_for_the_mixed_draw": bound}
    out["ratio"] = out["group_total_ms"] / out["per_code_total_ms"]
    log(json.dumps({k: v for k, v in out.items() if k != "per_z"}))
    return out


def offsets_for(rate, a, np, log):
    path = os.path.join(PKG, "trained", "offsets_r%d_z96_T5.json" % rate)
    if os.path.exists(path):
        return np.array(json.load(open(path))["offsets"], dtype=np.float32), path
    from wimax_ldpc_lib_amd.learned import Trainer
    tr = Trainer(rate, 96, T=5, batch=4096, seed=a.seed + rate, lr=1e-3, init_offset=0.15)
    t0 = time.time()
    for i in range(a.train_steps // 1000):
        loss = tr.step(1000, 1.0, 4.0)
        log("rate %d step %d loss %.5f (%.0f s)" % (rate, (i + 1) * 1000, loss, time.time() - t0))
    o = tr.offsets()
    if a.save_offsets:
        json.dump({"rate": rate, "z": 96, "T": 5, "offsets": o.tolist(),
                   "training": {"steps": a.train_steps, "batch": 4096, "lr": 1e-3, "init_offset": 0.15,
                                "ebn0_range_db": [1.0, 4.0], "seconds": time.time() - t0}}, open(path, "w"))
    return o, path


def fallback(a, torch, W, np):
    from wimax_ldpc_lib_amd.mixed import MixedDecoder, code_id
    F = a.frames or 524288
    codes = [code_id(r, 96) for r in RATES5]
    md_l = MixedDecoder(max_iter=5, early_stop=True, codes=codes)
    md_p = MixedDecoder(max_iter=10, early_stop=True, codes=codes)
    srcs = {}
    for r in RATES5:
        o, src = offsets_for(r, a, np, log)
        md_l.set_offsets(r, o)
        srcs[r] = os.path.basename(src)
    cfg_np, cfg, off, llr, bits, n_of = mixed_inputs(torch, md_l, np, F, codes, a.seed + 5, 1.5, 5.0)
    k_of = np.array([K_OF[int(c) // 19] * 96 for c in cfg_np], dtype=np.int64)
    info_bits = int(k_of.sum())
    hard = torch.empty(F * 288, dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    fb = torch.empty(F, dtype=torch.uint8, device="cuda")
    ms_l = timed(torch, lambda: md_l.decode_learned_fallback_dev(cfg, off, llr, 10, None, hard, ret, iters, None, fb),
                 a.reps)
    res = {}
    bits2 = bits.view(F, 2304)
    weights = (1 << torch.arange(8, dtype=torch.int32, device="cuda")).view(1, 1, 8)
    ref = (bits2.view(F, 288, 8).int() * weights).sum(dim=2).to(torch.uint8)
    popc = torch.tensor([bin(i).count("1") for i in range(256)], dtype=torch.int32, device="cuda")

    def per_rate(hard_t):
        x = hard_t.view(F, 288) ^ ref
        out = {}
        for r in RATES5:
            sel = torch.from_numpy(np.flatnonzero(cfg_np == code_id(r, 96))).cuda()
            kb = K_OF[r] * 96 // 8
            e = popc[x[sel, :kb].long()].sum(dim=1)
            out[str(r)] = {"frames": int(sel.numel()), "ber": float(e.sum()) / (sel.numel() * K_OF[r] * 96),
                           "fer": float((e > 0).float().mean())}
        return out
    res["learned_fallback"] = {"ms": ms_l, "info_gbps": info_bits / (ms_l * 1e-3) / 1e9,
                               "fallback_fraction": float(fb.float().mean()),
                               "fallback_fraction_per_rate": {str(r): float(fb[torch.from_numpy(
                                   np.flatnonzero(cfg_np == code_id(r, 96))).cuda()].float().mean()) for r in RATES5},
                               "avg_iters": float(iters.float().mean()), "per_rate": per_rate(hard)}
    ms_p = timed(torch, lambda: md_p.decode_dev(cfg, off, llr, None, hard, ret, iters), a.reps)
    res["nms10"] = {"ms": ms_p, "info_gbps": info_bits / (ms_p * 1e-3) / 1e9, "avg_iters": float(iters.float().mean()),
                    "per_rate": per_rate(hard)}
    # plain decoding at exactly 10 iterations (the headline's mode)
    md_f = MixedDecoder(max_iter=10, early_stop=False, codes=codes)
    ms_f = timed(torch, lambda: md_f.decode_dev(cfg, off, llr, None, hard, ret, iters), a.reps)
    res["nms10_fixed"] = {"ms": ms_f, "info_gbps": info_bits / (ms_f * 1e-3) / 1e9, "per_rate": per_rate(hard)}
    out = {"frames": F, "info_bits": info_bits, "offsets": srcs, **res,
           "speedup_learned_fallback_vs_nms10": ms_p / ms_l,
           "speedup_learned_fallback_vs_nms10_fixed": ms_f / ms_l}
    log(json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["mixed", "fallback", "groups"])
    ap.add_argument("--frames", type=int, default=0)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--seed", type=int, default=31337)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--save-offsets", action="store_true")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import numpy as np
    import torch

    import wimax_ldpc_lib_amd as W
    out = {"mixed": mixed, "fallback": fallback, "groups": groups}[a.which](a, torch, W, np)
    if a.out:
        json.dump(out, open(a.out, "w"), indent=1)


if __name__ == "__main__":
    main()
