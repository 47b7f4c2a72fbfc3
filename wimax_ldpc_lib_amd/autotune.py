"""Autotuner of the decoder's launch shape per (rate, z).

For every code it times the normalised min-sum kernel (exactly `--iters`
iterations, no early stop, so the work is the same for every shape) over
each code-shape variant and every frames-per-workgroup value the kernel
accepts (shapes 0, 1 and 4 at any; shape 2, where z % 32 == 0, at two
frames per workgroup only; shape 3, at z = 32, 48, 64, at four),
on a batch of seeded channel frames holding about 150 Mbit of
codeword bits, in three interleaved sweeps (each shape keeps its fastest),
and keeps the fastest shape.  Every shape is checked to give the same hard
decisions as the first one.  Results go to

  wimax_ldpc_lib_amd/csrc/tuned_table.inc   compiled into the library
  wimax_ldpc_lib_amd/tuned.json             the measured times

    python -m wimax_ldpc_lib_amd.autotune [--rates 0,1,...] [--zs 24,...] [--merge tuned.json]

With --groups it tunes the mixed-batch decoder instead: the frames per
workgroup of its one persistent kernel per lifting size z, by coordinate
descent over the z values on the mixed workload itself (262144 frames, the
five encoder rates x 19 z uniform, Eb/N0 uniform in [1.0, 4.5] dB, 10
iterations with early stop, the whole captured graph timed), every value
kept only if it is faster than the current one by more than --min-gain and
every run checked to give the same hard decisions.  Results go to

  wimax_ldpc_lib_amd/csrc/group_tuned.inc   compiled into the library
  wimax_ldpc_lib_amd/group_tuned.json       the measured times

    python -m wimax_ldpc_lib_amd.autotune --groups [--frames 262144] [--passes 2]
"""
import argparse
import json
import os
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(PKG))


def tune_one(torch, W, rate, z, iters, reps, bits_target, rounds=3):
    enc = W.ldpc_encoder(rate, z)
    dec = W.ldpc_decoder(rate, z, iters, early_stop=False)
    n = dec.n
    F = max(1024, (bits_target // n) // 64 * 64)
    llr = torch.empty((F, n), dtype=torch.float32, device="cuda")
    enc.channel_dev(4242 + rate * 100 + z, 0, F, 3.0, llr=llr)
    hard = torch.empty((F, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    ref = None
    results = []
    dec.variant = 0   # any frames-per-workgroup value is allowed with shapes 0, 1 and 4
    fmax = 1
    while True:
        try:
            dec.frames_per_wg = fmax + 1
            fmax += 1
        except W.WLError:
            break
    shapes = [(v, f) for v in (0, 1, 4) for f in range(1, fmax + 1)]
    if z % 32 == 0 and fmax >= 2:
        shapes.append((2, 2))   # two-halves lane mapping: two frames per workgroup only
    if z % 16 == 0 and z <= 64 and fmax >= 4:
        shapes.append((3, 4))   # quarters lane mapping: four frames per workgroup only
    # Every shape is timed in each of `rounds` sweeps and keeps its fastest
    # sweep: interleaving the shapes spreads clock drift over all of them,
    # and the minimum discards sweeps that drift slowed.
    times = {}
    for _ in range(rounds):
        for variant, f in shapes:
            dec.variant = 0
            dec.frames_per_wg = f
            dec.variant = variant
            dec.decode_dev(llr, None, hard, ret)
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(reps):
                dec.decode_dev(llr, None, hard, ret)
            e.record()
            torch.cuda.synchronize()
            ms = s.elapsed_time(e) / reps
            times[(variant, f)] = min(ms, times.get((variant, f), ms))
            h = hard.cpu()
            if ref is None:
                ref = h
            elif not torch.equal(ref, h):
                raise RuntimeError("rate %d z %d variant %d F %d: hard decisions differ" % (rate, z, variant, f))
    for (variant, f), ms in times.items():
        results.append({"variant": variant, "f": f, "ms": ms,
                        "info_gbps": F * dec.k / (ms * 1e-3) / 1e9})
    best = min(results, key=lambda r: r["ms"])
    return {"rate": rate, "z": z, "frames": F, "iters": iters, "best": best, "all": results}


ZS = list(range(24, 97, 4))
RATES5 = [0, 1, 2, 3, 5]
K_OF = {0: 12, 1: 16, 2: 16, 3: 18, 5: 20}   # info blocks per rate


def tune_groups(torch, np, frames, reps, rounds, passes, min_gain, seed=31337):
    from wimax_ldpc_lib_amd.mixed import MixedDecoder, code_id
    import wimax_ldpc_lib_amd as W
    codes = [code_id(r, z) for r in RATES5 for z in ZS]
    md = MixedDecoder(max_iter=10, early_stop=True, codes=codes)
    rng = np.random.default_rng(seed)
    cfg_np = rng.choice(np.array(codes, dtype=np.uint8), frames)
    cfg = torch.from_numpy(cfg_np).cuda()
    off = torch.empty(frames, dtype=torch.int64, device="cuda")
    md.offsets_dev(cfg, off)
    n_of = 24 * (24 + 4 * (cfg_np.astype(np.int64) % 19))
    info_bits = int(sum(K_OF[int(c) // 19] * (24 + 4 * (int(c) % 19)) for c in cfg_np))
    total = int(n_of.sum())
    llr = torch.empty(total, dtype=torch.float32, device="cuda")
    md.channel_dev(seed, cfg, off, 1.0, 4.5, llr)
    hard = torch.empty(total // 8, dtype=torch.uint8, device="cuda")
    ret = torch.empty(frames, dtype=torch.uint8, device="cuda")
    iters = torch.empty(frames, dtype=torch.uint8, device="cuda")

    def run():
        md.decode_dev(cfg, off, llr, None, hard, ret, iters)

    ref = []

    def measure():
        run()   # captures the graph after a change of shape
        torch.cuda.synchronize()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(reps):
            run()
        e.record()
        torch.cuda.synchronize()
        if not ref:
            ref.append(hard.clone())
        elif not torch.equal(ref[0], hard):
            raise RuntimeError("hard decisions differ between group shapes")
        return s.elapsed_time(e) / reps

    fmax, cur = {}, {}
    for z in ZS:
        cur[z] = md.launch(0, z)[0]
        f = 1
        while True:
            try:
                md.set_group_frames_per_wg(z, f + 1)
                f += 1
            except W.WLError:
                break
        fmax[z] = f
        md.set_group_frames_per_wg(z, cur[z])
    default = dict(cur)
    log = []
    for p in range(passes):
        for z in ZS:
            times = {}
            for _ in range(rounds):
                for f in range(1, fmax[z] + 1):
                    md.set_group_frames_per_wg(z, f)
                    ms = measure()
                    times[f] = min(ms, times.get(f, ms))
            best = min(times, key=times.get)
            keep = best if times[best] < times[cur[z]] * (1.0 - min_gain) else cur[z]
            print("pass %d z %2d: F %2d -> %2d  (%.3f ms at %d, %.3f ms at %d)"
                  % (p, z, cur[z], keep, times[cur[z]], cur[z], times[best], best), flush=True)
            log.append({"pass": p, "z": z, "ms": {str(f): t for f, t in sorted(times.items())}, "kept": keep})
            cur[z] = keep
            md.set_group_frames_per_wg(z, keep)
    # the default and the tuned shapes, interleaved
    t_def, t_new = [], []
    for _ in range(rounds + 2):
        for z in ZS:
            md.set_group_frames_per_wg(z, default[z])
        t_def.append(measure())
        for z in ZS:
            md.set_group_frames_per_wg(z, cur[z])
        t_new.append(measure())
    out = {"frames": frames, "info_bits": info_bits, "default_f": default, "tuned_f": cur, "fmax": fmax,
           "default_ms": min(t_def), "tuned_ms": min(t_new),
           "default_info_gbps": info_bits / (min(t_def) * 1e-3) / 1e9,
           "tuned_info_gbps": info_bits / (min(t_new) * 1e-3) / 1e9, "sweeps": log}
    print("default %.3f ms (%.2f info Gbit/s), tuned %.3f ms (%.2f info Gbit/s)"
          % (out["default_ms"], out["default_info_gbps"], out["tuned_ms"], out["tuned_info_gbps"]), flush=True)
    return out


def write_group_outputs(out, out_dir):
    os.makedirs(os.path.join(out_dir, "csrc"), exist_ok=True)
    with open(os.path.join(out_dir, "group_tuned.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    lines = ["// group_tuned.inc -- {z, frames per workgroup} of the mixed-batch group kernels, written by",
             "// wimax_ldpc_lib_amd/autotune.py --groups on MI355X (%d frames: %.3f ms at the per-code"
             % (out["frames"], out["default_ms"]),
             "// table's most common F, %.3f ms at these); generated file." % out["tuned_ms"]]
    for z in ZS:
        lines.append("{%d, %d}," % (z, out["tuned_f"][z] if z in out["tuned_f"] else out["tuned_f"][str(z)]))
    with open(os.path.join(out_dir, "csrc", "group_tuned.inc"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--groups", action="store_true", help="tune the mixed-batch group kernels")
    ap.add_argument("--frames", type=int, default=262144, help="--groups: frames of the mixed batch")
    ap.add_argument("--passes", type=int, default=2, help="--groups: coordinate-descent passes over z")
    ap.add_argument("--min-gain", type=float, default=0.002,
                    help="--groups: relative gain a new F needs over the current one")
    ap.add_argument("--rates", default="0,1,2,3,4,5")
    ap.add_argument("--zs", default=",".join(str(z) for z in range(24, 97, 4)))
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=3, help="sweeps over all shapes; each keeps its fastest")
    ap.add_argument("--mbit", type=float, default=150.0)
    ap.add_argument("--out-dir", default=PKG)
    ap.add_argument("--from-json", default=None,
                    help="skip measuring: write the table from a tuned.json of an earlier run")
    ap.add_argument("--merge", default=None,
                    help="tuned.json of an earlier run: codes measured now replace their entries, "
                         "the others are kept")
    a = ap.parse_args()
    if a.from_json and a.groups:
        write_group_outputs(json.load(open(a.from_json)), a.out_dir)
        return
    if a.from_json:
        write_outputs(json.load(open(a.from_json)), a.out_dir, a.iters)
        return
    import torch

    import wimax_ldpc_lib_amd as W
    if a.groups:
        import numpy as np
        out = tune_groups(torch, np, a.frames, a.reps, a.rounds, a.passes, a.min_gain)
        write_group_outputs(out, a.out_dir)
        return
    rates = [int(x) for x in a.rates.split(",")]
    zs = [int(x) for x in a.zs.split(",")]
    out = []
    t0 = time.time()
    for rate in rates:
        for z in zs:
            r = tune_one(torch, W, rate, z, a.iters, a.reps, int(a.mbit * 1e6), a.rounds)
            out.append(r)
            b = r["best"]
            print("rate %d z %2d: variant %d F %2d  %.3f ms  %.2f info Gbit/s  (%.0f s)"
                  % (rate, z, b["variant"], b["f"], b["ms"], b["info_gbps"], time.time() - t0), flush=True)
    if a.merge:
        new = {(r["rate"], r["z"]): r for r in out}
        old = json.load(open(a.merge))
        out = [new.pop((r["rate"], r["z"]), r) for r in old] + list(new.values())
        out.sort(key=lambda r: (r["rate"], r["z"]))
    write_outputs(out, a.out_dir, a.iters)


def write_outputs(out, out_dir, iters):
    os.makedirs(os.path.join(out_dir, "csrc"), exist_ok=True)
    with open(os.path.join(out_dir, "tuned.json"), "w") as fh:
        json.dump(out, fh, indent=0)
    lines = ["// (rate, z, frames_per_wg, variant): fastest launch shape measured on MI355X by",
             "// wimax_ldpc_lib_amd/autotune.py (%d-iteration normalised min-sum); generated file." % iters]
    for r in out:
        lines.append("{%d, %d, %d, %d}," % (r["rate"], r["z"], r["best"]["f"], r["best"]["variant"]))
    with open(os.path.join(out_dir, "csrc", "tuned_table.inc"), "w") as fh:
        fh.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
