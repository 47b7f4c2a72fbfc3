"""Autotuner of the decoder's launch shape per (rate, z).

For every code it times the normalised min-sum kernel (exactly `--iters`
iterations, no early stop, so the work is the same for every shape) over
each code-shape variant and every frames-per-workgroup value the kernel
accepts (shape 2, where z % 32 == 0, at two frames per workgroup only;
shape 3, at z = 32, 48, 64, at four),
on a batch of seeded channel frames holding about 150 Mbit of
codeword bits, in three interleaved sweeps (each shape keeps its fastest),
and keeps the fastest shape.  Every shape is checked to give the same hard
decisions as the first one.  Results go to

  wimax_ldpc_lib_amd/csrc/tuned_table.inc   compiled into the library
  wimax_ldpc_lib_amd/tuned.json             the measured times

    python -m wimax_ldpc_lib_amd.autotune [--rates 0,1,...] [--zs 24,...] [--merge tuned.json]
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
    dec.variant = 0   # any frames-per-workgroup value is allowed with shapes 0 and 1
    fmax = 1
    while True:
        try:
            dec.frames_per_wg = fmax + 1
            fmax += 1
        except W.WLError:
            break
    shapes = [(v, f) for v in (0, 1) for f in range(1, fmax + 1)]
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


def main():
    ap = argparse.ArgumentParser()
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
    if a.from_json:
        write_outputs(json.load(open(a.from_json)), a.out_dir, a.iters)
        return
    import torch

    import wimax_ldpc_lib_amd as W
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
