"""Train and evaluate the learned offset min-sum decoder (GPU).

Training: rate 2/3A, z = 96 by default, T = 5 unrolled iterations, one offset
per nonzero base-matrix block per iteration initialised to 0.15, batches of
4096 fresh GPU-generated frames (seeded random bits, encoder, BPSK/AWGN with
Eb/N0 uniform in [1, 4] dB, or --train-lo / --train-hi), 20000 Adam steps at
lr 1e-3, mean binary cross-entropy on the coded bits (of the final
posteriors, or with --per-iteration the mean over every unrolled
iteration's).

Evaluation: fresh frames (another seed) at Eb/N0 = 1.0 .. 4.0 dB in 0.25 dB
steps, `--eval-frames` per point, decoded by
  learned-5  the trained offsets, 5 iterations
  nms-5      the reference's normalised min-sum (0.815), 5 iterations
  nms-10     the same, 10 iterations (the reference's MAX_ITER)
  ms-10      plain min-sum, 10 iterations (the reference's compile-time
             alternative, lib/ldpc_decoder.cc:21,233-238)
  oms-10     offset min-sum with the reference's offset 0.10, 10 iterations
             (lib/ldpc_decoder.cc:20,225-232)
  learned-5+fb  learned-5, frames failing its syndrome check decoded again
             from their channel LLRs by nms-10 (the fallback)
all with the reference's syndrome early stop.  BER counts info-bit errors
over the k systematic bits; FER counts frames with any info-bit error.  The
Eb/N0 at BER 1e-5 is interpolated log-linearly between the straddling points.
Throughput (info Gbit/s) of learned-5, learned-5+fb, nms-10 and ms-10 is timed on
65536 frames at `--speed-ebn0` and at nms-10's BER-1e-5 point, with early
stop and with fixed iterations, after a second of warm-up, each leg the
fastest of three interleaved sweeps.

    python -m wimax_ldpc_lib_amd.learned_eval --out results.json [--steps 20000]
"""
import argparse
import json
import math
import os
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(PKG))

POPC = None


def popcount_rows(torch, x):
    global POPC
    if POPC is None or POPC.device != x.device:
        POPC = torch.tensor([bin(i).count("1") for i in range(256)], dtype=torch.int32, device=x.device)
    return POPC[x.long()].sum(dim=1)


def threshold_db(points, key, target=1e-5):
    """Eb/N0 where BER crosses target (log-linear), or None."""
    pts = [(p["ebn0"], p[key]["ber"]) for p in points]
    for (e0, b0), (e1, b1) in zip(pts, pts[1:]):
        if b0 >= target > b1:
            if b1 <= 0:
                return e1   # no errors observed at e1: bound
            f = (math.log10(b0) - math.log10(target)) / (math.log10(b0) - math.log10(b1))
            return e0 + f * (e1 - e0)
    return None


def evaluate(torch, W, rate, z, offsets, ebn0s, frames, chunk, seed, T, log):
    enc = W.ldpc_encoder(rate, z)
    n, k = enc.n, enc.k
    decs = {}
    d = W.ldpc_decoder(rate, z, T, early_stop=True)
    d.set_offsets(offsets)
    decs["learned-%d" % T] = d
    decs["nms-%d" % T] = W.ldpc_decoder(rate, z, T, early_stop=True)
    decs["nms-10"] = W.ldpc_decoder(rate, z, 10, early_stop=True)
    decs["ms-10"] = W.ldpc_decoder(rate, z, 10, early_stop=True, rule="ms")
    decs["oms-10"] = W.ldpc_decoder(rate, z, 10, early_stop=True, rule="oms")
    decs["learned-%d+fb" % T] = d
    llr = torch.empty((chunk, n), dtype=torch.float32, device="cuda")
    bits = torch.empty((chunk, n), dtype=torch.uint8, device="cuda")
    hard = torch.empty((chunk, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(chunk, dtype=torch.uint8, device="cuda")
    iters = torch.empty(chunk, dtype=torch.uint8, device="cuda")
    fb = torch.empty(chunk, dtype=torch.uint8, device="cuda")
    weights = (1 << torch.arange(8, dtype=torch.int32, device="cuda")).view(1, 1, 8)   # LSB-first packing
    points = []
    for pi, eb in enumerate(ebn0s):
        acc = {name: dict(bit_err=0, frame_err=0, coded_err=0, iters=0, fail=0, fallback=0) for name in decs}
        done = 0
        while done < frames:
            m = min(chunk, frames - done)
            enc.channel_dev(seed, pi * (1 << 32) + done, m, eb, llr=llr[:m], bits=bits[:m])
            ref = (bits[:m].view(m, n // 8, 8).int() * weights).sum(dim=2).to(torch.uint8)
            for name, dec in decs.items():
                a = acc[name]
                if name.endswith("+fb"):
                    dec.decode_learned_fallback_dev(llr[:m], 10, None, hard[:m], ret[:m], iters[:m],
                                                    fallback=fb[:m])
                    a["fallback"] += int(fb[:m].int().sum())
                else:
                    dec.decode_dev(llr[:m], None, hard[:m], ret[:m], iters[:m])
                x = hard[:m] ^ ref
                ie = popcount_rows(torch, x[:, :k // 8])
                ce = popcount_rows(torch, x)
                a["bit_err"] += int(ie.sum())
                a["frame_err"] += int((ie > 0).sum())
                a["coded_err"] += int(ce.sum())
                a["iters"] += int(iters[:m].int().sum())
                a["fail"] += int(ret[:m].int().sum())
            done += m
        p = {"ebn0": eb, "frames": frames}
        for name, a in acc.items():
            p[name] = {"ber": a["bit_err"] / (frames * k), "fer": a["frame_err"] / frames,
                       "ber_runner_def": a["coded_err"] / (frames * k),
                       "avg_iters": a["iters"] / frames, "syndrome_fail_rate": a["fail"] / frames}
            if name.endswith("+fb"):
                p[name]["fallback_fraction"] = a["fallback"] / frames
        points.append(p)
        log("Eb/N0 %.2f dB: " % eb + "  ".join("%s BER %.3g FER %.3g it %.2f" % (
            nm, p[nm]["ber"], p[nm]["fer"], p[nm]["avg_iters"]) for nm in decs))
    return points


def speed(torch, W, rate, z, offsets, T, eb, frames, reps=10, rounds=3):
    enc = W.ldpc_encoder(rate, z)
    n, k = enc.n, enc.k
    llr = torch.empty((frames, n), dtype=torch.float32, device="cuda")
    enc.channel_dev(77, 0, frames, eb, llr=llr)
    hard = torch.empty((frames, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(frames, dtype=torch.uint8, device="cuda")
    fb = torch.empty(frames, dtype=torch.uint8, device="cuda")
    # about a second of decoding first, so every timed leg runs at the
    # clocks a sustained load settles at (the first legs otherwise read slow)
    warm = W.ldpc_decoder(rate, z, 10, early_stop=False)
    for _ in range(400):
        warm.decode_dev(llr, None, hard, ret)
    torch.cuda.synchronize()
    # every leg is timed in each of `rounds` interleaved sweeps and keeps its
    # fastest (clock drift then spreads over all legs)
    legs = []
    for early in (False, True):
        for name, it, off in (("learned-%d" % T, T, offsets), ("learned-%d+fb" % T, T, offsets),
                              ("nms-10", 10, None), ("ms-10", 10, None)):
            d = W.ldpc_decoder(rate, z, it, early_stop=early, rule="ms" if name == "ms-10" else "nms")
            if off is not None:
                d.set_offsets(off)
            legs.append(("%s%s" % (name, " early-stop" if early else " fixed"), d, name.endswith("+fb")))
    best = {}
    fb_frac = {}
    for _ in range(rounds):
        for key, d, with_fb in legs:
            if with_fb:
                def run():
                    d.decode_learned_fallback_dev(llr, 10, None, hard, ret, fallback=fb)
            else:
                def run():
                    d.decode_dev(llr, None, hard, ret)
            run()
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(reps):
                run()
            e.record()
            torch.cuda.synchronize()
            ms = s.elapsed_time(e) / reps
            best[key] = min(ms, best.get(key, ms))
            if with_fb:
                fb_frac[key] = float(fb.float().mean())
    out = {"ebn0": eb}
    for key, _, with_fb in legs:
        r = {"ms": best[key], "info_gbps": frames * k / (best[key] * 1e-3) / 1e9}
        if with_fb:
            r["fallback_fraction"] = fb_frac[key]
        out[key] = r
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=int, default=1)
    ap.add_argument("--z", type=int, default=96)
    ap.add_argument("--T", type=int, default=5)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init", type=float, default=0.15)
    ap.add_argument("--train-lo", type=float, default=1.0)
    ap.add_argument("--train-hi", type=float, default=4.0)
    ap.add_argument("--per-iteration", action="store_true",
                    help="loss: mean of every unrolled iteration's cross-entropy (default: the final one)")
    ap.add_argument("--seed", type=int, default=20240101)
    ap.add_argument("--eval-frames", type=int, default=1 << 20)
    ap.add_argument("--eval-lo", type=float, default=1.0)
    ap.add_argument("--eval-hi", type=float, default=4.0)
    ap.add_argument("--eval-step", type=float, default=0.25)
    ap.add_argument("--chunk", type=int, default=65536)
    ap.add_argument("--speed-ebn0", type=float, default=3.0)
    ap.add_argument("--offsets-in", default=None, help="skip training, evaluate these offsets (json)")
    ap.add_argument("--offsets-out", default=None)
    ap.add_argument("--checkpoint", default=None)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import numpy as np
    import torch

    import wimax_ldpc_lib_amd as W
    from wimax_ldpc_lib_amd.learned import Trainer

    def log(*x):
        print(*x, flush=True)

    result = {"rate": a.rate, "z": a.z, "T": a.T}
    if a.offsets_in:
        offsets = np.array(json.load(open(a.offsets_in))["offsets"], dtype=np.float32)
    else:
        tr = Trainer(a.rate, a.z, T=a.T, batch=a.batch, seed=a.seed, lr=a.lr, init_offset=a.init)
        tr.set_loss(a.per_iteration)
        t0 = time.time()
        losses = []
        done = 0
        while done < a.steps:
            m = min(1000, a.steps - done)
            loss = tr.step(m, a.train_lo, a.train_hi)
            done += m
            losses.append((done, loss))
            log("step %6d  loss %.6f  (%.1f s)" % (done, loss, time.time() - t0))
        train_s = time.time() - t0
        offsets = tr.offsets()
        if a.checkpoint:
            tr.save(a.checkpoint)
        result["training"] = {"steps": a.steps, "batch": a.batch, "lr": a.lr, "init_offset": a.init,
                              "ebn0_range_db": [a.train_lo, a.train_hi], "per_iteration_loss": a.per_iteration,
                              "seconds": train_s,
                              "ms_per_step": train_s * 1e3 / max(1, a.steps), "loss_curve": losses}
        if a.offsets_out:
            json.dump({"rate": a.rate, "z": a.z, "T": a.T, "offsets": offsets.tolist(),
                       "training": result["training"]}, open(a.offsets_out, "w"))
    ebs = [round(a.eval_lo + i * a.eval_step, 4)
           for i in range(int(round((a.eval_hi - a.eval_lo) / a.eval_step)) + 1)]
    t0 = time.time()
    pts = evaluate(torch, W, a.rate, a.z, offsets, ebs, a.eval_frames, a.chunk, a.seed + 1, a.T, log)
    result["eval"] = {"frames_per_point": a.eval_frames, "points": pts, "seconds": time.time() - t0}
    th = {name: threshold_db(pts, name) for name in pts[0] if name not in ("ebn0", "frames")}
    result["ebn0_at_ber_1e-5"] = th
    sp = speed(torch, W, a.rate, a.z, offsets, a.T, a.speed_ebn0, a.chunk)
    result["speed"] = sp
    L, LF, N = th["learned-%d" % a.T], th["learned-%d+fb" % a.T], th["nms-10"]

    def gap(x, y):
        return (x - y) if (x is not None and y is not None) else None

    def ratio(sp, name, mode, base="nms-10"):
        return sp["%s %s" % (name, mode)]["info_gbps"] / sp["%s %s" % (base, mode)]["info_gbps"]

    lt, lf = "learned-%d" % a.T, "learned-%d+fb" % a.T
    result["summary"] = {
        "gap_db_learned_vs_nms10": gap(L, N),
        "gap_db_learned_vs_ms10": gap(L, th["ms-10"]),
        "gap_db_learned_vs_oms10": gap(L, th["oms-10"]),
        "gap_db_learned_fb_vs_nms10": gap(LF, N),
        "speedup_fixed": ratio(sp, lt, "fixed"),
        "speedup_early_stop": ratio(sp, lt, "early-stop"),
        "speedup_fixed_vs_ms10": ratio(sp, lt, "fixed", "ms-10"),
        "speedup_early_stop_vs_ms10": ratio(sp, lt, "early-stop", "ms-10"),
    }
    if N is not None:
        op = speed(torch, W, a.rate, a.z, offsets, a.T, round(N, 3), a.chunk)
        result["speed_at_nms10_ber_1e-5"] = op
        result["summary"].update({
            "op_ebn0_db": op["ebn0"],
            "op_speedup_learned_fixed": ratio(op, lt, "fixed"),
            "op_speedup_learned_fb_fixed": ratio(op, lf, "fixed"),
            "op_speedup_learned_fb_early_stop": ratio(op, lf, "early-stop"),
            "op_fallback_fraction": op["%s early-stop" % lf]["fallback_fraction"],
        })
    log(json.dumps(result["summary"]), json.dumps(th))
    if a.out:
        json.dump(result, open(a.out, "w"), indent=1)


if __name__ == "__main__":
    main()
