#!/usr/bin/env python3
"""Throughput benchmark of the gfx950 TDMP decoder (one JSON line on stdout).

Workload: rate 5/6, z = 96 (n = 2304, k = 1920), 65536 seeded BPSK/AWGN
frames at Eb/N0 = 4.0 dB generated on the GPU, FP32 LLRs, normalised
min-sum with exactly 10 iterations (no early stop).  One step = one decode
of the whole batch (LLRs resident in HBM -> packed hard decisions and
per-frame status in HBM).  The CPU baseline is the oracle (the C restatement
of the reference decoder, OpenMP over frames as lib/decode_example.cc does)
on the same LLR buffer, timed on this host's cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 20] [--no-cpu]
    python bench.py --workload small   # rate 1/2, z = 24, 1024 frames at 2.5 dB,
                                       # max 10 iterations with early stop, vs the CPU decoder
    python bench.py --workload encode  # rate 1/2, z = 96 encoder (the reference's
                                       # test_encoder configuration), vs the CPU encoder

The CPU baselines run with OMP_PROC_BIND=close and OMP_PLACES=cores (set
here before any OpenMP runtime starts, unless the caller set them), on every
core this process may use; the CPU model is reported with them.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# thread placement of the CPU baselines (read when an OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# CPUs this process may use, read before an OpenMP runtime pins this thread
USABLE_CPUS = len(os.sched_getaffinity(0))


def cpu_model():
    """Model name of the first CPU in /proc/cpuinfo, with its family / model /
    stepping numbers (which name the generation where the model name is
    generic, as on virtual machines)."""
    fields = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if fields:
                    break   # end of the first CPU's block
                continue
            key, val = (x.strip() for x in line.split(":", 1))
            fields.setdefault(key, val)
    except OSError:
        pass
    name = fields.get("model name", "unknown")
    ids = [("family", fields.get("cpu family")), ("model", fields.get("model")), ("stepping", fields.get("stepping"))]
    ids = ", ".join("%s %s" % (k, v) for k, v in ids if v)
    return "%s (%s)" % (name, ids) if ids else name


def cpu_env(threads):
    return {"cpu_model": cpu_model(), "threads": threads, "usable_cpus": USABLE_CPUS,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES")}


def median_seconds(fn, runs=5):
    """Median of `runs` timed calls of fn() -> seconds (after one warm-up call)."""
    fn()
    secs = sorted(fn() for _ in range(runs))
    return secs[len(secs) // 2], secs


RATE, Z, FRAMES, EBN0, ITERS, SEED = 5, 96, 65536, 4.0, 10, 20260101
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--frames", type=int, default=FRAMES)
    ap.add_argument("--frames-per-wg", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=FRAMES,
                    help="frames of the same buffer decoded by the CPU baseline")
    ap.add_argument("--workload", choices=["headline", "small", "encode"], default="headline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; the decoder shards frames across GPUs "
              "with no exchange, so run one process per GPU (--gpus 1).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import wimax_ldpc_lib_amd as W
    if not torch.cuda.is_available() or W.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.workload == "small":
        return small(args, np, torch, W)
    if args.workload == "encode":
        return encode(args, np, torch, W)

    F = args.frames
    enc = W.ldpc_encoder(RATE, Z)
    dec = W.ldpc_decoder(RATE, Z, ITERS, early_stop=False)
    if args.frames_per_wg:
        dec.frames_per_wg = args.frames_per_wg
    n, k = dec.n, dec.k
    stream = torch.cuda.current_stream()
    llr = torch.empty((F, n), dtype=torch.float32, device="cuda")
    bits = torch.empty((F, n), dtype=torch.uint8, device="cuda")
    enc.channel_dev(SEED, 0, F, EBN0, llr=llr, bits=bits)
    hard = torch.empty((F, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    log("frames %d, frames/workgroup %d" % (F, dec.frames_per_wg))

    for _ in range(args.warmup):
        dec.decode_dev(llr, None, hard, ret)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        dec.decode_dev(llr, None, hard, ret)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_step = wall * 1e3 / args.steps
    kavg = sum(kern_ms) / len(kern_ms)
    value = F * k / (ms_step * 1e-3) / 1e9
    bytes_per_frame = 4 * n + n // 8
    achieved = F * bytes_per_frame / (kavg * 1e-3) / 1e9
    hard_np = hard.cpu().numpy()
    ber = float(np.mean(np.unpackbits(hard_np[:, :k // 8], axis=1, bitorder="little")
                        != bits[:, :k].cpu().numpy()))
    log("kernel avg %.3f ms (min %.3f), step %.3f ms, %.2f info Gbit/s, info BER %.3g"
        % (kavg, min(kern_ms), ms_step, value, ber))

    # HBM bytes and VALU issue of this kernel from the committed rocprofv3
    # counter runs (FETCH_SIZE x2 and WRITE_SIZE x1 after copy calibration)
    traffic, valu, lds = None, None, None
    pmc = os.path.join(ROOT, "profiles", "tdmp_r5z96_f65536_traffic.json")
    if os.path.exists(pmc):
        prof = json.load(open(pmc))
        traffic = prof.get("hbm_bytes_per_launch")
        if prof.get("valu_wave_insts_per_launch"):
            # issue rate over this run's kernel time: wave instructions per CU per cycle
            # at the 2.4 GHz peak engine clock (peak 1.0: 4 SIMDs x 1/4 per cycle)
            ipc = prof["valu_wave_insts_per_launch"] / (prof["num_cu"] * 2.4e9 * kavg * 1e-3)
            valu = {"wave_insts_per_launch": prof["valu_wave_insts_per_launch"],
                    "issue_per_cu_cycle_at_2.4GHz": ipc,
                    "issue_frac_counters": prof.get("valu_issue_frac"),
                    "note": "the kernel is bound by VALU issue, not HBM"}
        if prof.get("lds_bytes_per_launch_est"):
            # LDS peak: 128 B/clk/CU x 256 CUs x 2.4 GHz
            lds_tbs = prof["lds_bytes_per_launch_est"] / (kavg * 1e-3) / 1e12
            lds = {"bytes_per_launch": prof["lds_bytes_per_launch_est"], "achieved_tbs": lds_tbs,
                   "peak_tbs": 78.6, "frac": lds_tbs / 78.6,
                   "bank_conflict_cycles": prof.get("lds_bank_conflict_cycles")}

    # The same decode through the host-pointer entry point (decode()'s
    # semantics: host LLRs in, posteriors back into them, 0/1 bytes out):
    # PCIe-inclusive rate, reported beside `value`, never as it.
    rx_host = llr.cpu().numpy()
    dec_host = np.empty((F, n), dtype=np.int8)
    ret_host = np.empty(F, dtype=np.uint8)
    dec.decode_batch_host(rx_host.copy(), dec_host, ret_host)
    rx_run = rx_host.copy()
    t0 = time.perf_counter()
    dec.decode_batch_host(rx_run, dec_host, ret_host)
    host_s = time.perf_counter() - t0
    host_same = bool(np.array_equal(np.packbits(dec_host.view(np.uint8), axis=1, bitorder="little"), hard_np))
    pcie = {"value": F * k / host_s / 1e9, "unit": "info Gbit/s", "seconds": host_s,
            "path": "wl_decode_f32_host: pageable H2D of LLRs, decode, D2H of posteriors and "
                    "0/1 bytes", "same_hard_decisions": host_same}
    log("host-pointer decode (PCIe-inclusive): %.3f s, %.2f info Gbit/s, same decisions: %s"
        % (host_s, pcie["value"], host_same))
    del rx_host, rx_run, dec_host

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        code = O.Code(RATE, Z)
        S = min(args.cpu_frames, F)
        sample = llr[:S].cpu().numpy()
        threads = O.omp_max_threads()
        res = code.decode(sample, ITERS, False, nthreads=threads)   # also the bit-exactness check
        same = bool(np.array_equal(np.packbits(res["hd"], axis=1, bitorder="little"), hard_np[:S]))
        fixed_s, fixed_all = median_seconds(lambda: code.decode(sample, ITERS, False, nthreads=threads)["seconds"])
        es_s, es_all = median_seconds(lambda: code.decode(sample, ITERS, True, nthreads=threads)["seconds"])
        cpu_gbs = S * k / fixed_s / 1e9
        cpu_es = S * k / es_s / 1e9
        log("cpu baseline (%d threads, %s): fixed-10 %.4f info Gbit/s (median %.2f s of %s), early-stop %.4f; "
            "GPU == CPU hard decisions on %d frames: %s"
            % (threads, cpu_model(), cpu_gbs, fixed_s, ["%.2f" % x for x in fixed_all], cpu_es, S, same))
        cpu = {"value": cpu_gbs, "unit": "info Gbit/s", "cores": threads, "kind": "port",
               "sample": "%d of the same frames, exactly 10 iterations, median of 5 timed runs after a warm-up "
                         "(early stop as shipped: %.4f info Gbit/s); OpenMP frame loop" % (S, cpu_es),
               "seconds": fixed_s, "runs_s": fixed_all, "early_stop_runs_s": es_all, "bit_exact_vs_gpu": same,
               **cpu_env(threads)}

    out = {
        "metric": "info Gbit/s decoded (rate 5/6, z=96, 10 it)",
        "value": value,
        "unit": "info Gbit/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "TDMP NMS decode, rate 5/6, z=96, n=2304, k=1920, %d frames, "
                               "BPSK/AWGN Eb/N0 4.0 dB, exactly 10 iterations" % F,
                   "frames": F, "frames_per_wg": dec.frames_per_wg, "iterations": ITERS},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                     "kernel_ms": kavg, "bytes_per_frame": bytes_per_frame, "valu": valu, "lds": lds},
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "info_ber": ber,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


def small(args, np, torch, W):
    """Rate 1/2, z = 24 (n = 576), 1024 seeded frames at Eb/N0 2.5 dB, at most
    10 iterations with the reference's early stop: GPU decode (posteriors,
    hard decisions, status) against the CPU decoder on the same frames."""
    F, rate, z, eb = 1024, 0, 24, 2.5
    enc = W.ldpc_encoder(rate, z)
    dec = W.ldpc_decoder(rate, z, ITERS, early_stop=True)
    n, k = dec.n, dec.k
    llr = torch.empty((F, n), dtype=torch.float32, device="cuda")
    enc.channel_dev(SEED, 0, F, eb, llr=llr)
    post = torch.empty_like(llr)
    hard = torch.empty((F, n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(F, dtype=torch.uint8, device="cuda")
    iters = torch.empty(F, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        dec.decode_dev(llr, post, hard, ret, iters)
    torch.cuda.synchronize()
    steps = max(args.steps, 20)
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record(stream)
    for _ in range(steps):
        dec.decode_dev(llr, post, hard, ret, iters)
    e.record(stream)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    ms = s.elapsed_time(e) / steps
    value = F * k / (wall_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        code = O.Code(rate, z)
        threads = O.omp_max_threads()
        sample = llr.cpu().numpy()
        code.decode(sample[:64], ITERS, True, nthreads=threads)
        secs = []
        for _ in range(5):
            o = code.decode(sample, ITERS, True, nthreads=threads)
            secs.append(o["seconds"])
        sec = sorted(secs)[len(secs) // 2]
        same = (np.array_equal(post.cpu().numpy().view(np.uint32), o["post"].view(np.uint32)) and
                np.array_equal(hard.cpu().numpy(), np.packbits(o["hd"], axis=1, bitorder="little")) and
                np.array_equal(iters.cpu().numpy(), o["iters"]) and np.array_equal(ret.cpu().numpy(), o["ret"]))
        cpu = {"value": F * k / sec / 1e9, "unit": "info Gbit/s", "cores": threads, "kind": "port",
               "sample": "the same 1024 frames, early stop as shipped; median of 5", "seconds": sec,
               "bit_exact_vs_gpu": bool(same), **cpu_env(threads)}
    out = {"metric": "info Gbit/s decoded (rate 1/2, z=24, early stop)", "value": value, "unit": "info Gbit/s",
           "n_gpus": 1, "steps": steps, "warmup": args.warmup, "ms_per_step": wall_ms, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": "TDMP NMS decode, rate 1/2, z=24, n=576, k=288, 1024 frames, BPSK/AWGN "
                                  "Eb/N0 2.5 dB, at most 10 iterations, early stop", "frames": F},
           "kernel_ms": ms, "avg_iters": float(iters.float().mean()), "cpu_baseline": cpu}
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


def encode(args, np, torch, W):
    """Rate 1/2, z = 96 (n = 2304, k = 1152): the reference's test_encoder
    configuration (README: `test_encoder 0 96 500 100`, 171.904706 coded
    Mbit/s on an i7-4771, lib/encode_example.cc:131-158).  One step encodes
    the whole batch: byte-per-bit info words in HBM -> byte-per-bit codewords
    in HBM (the reference's format), k + n bytes per frame."""
    rate, z = 0, 96
    F = 524288
    enc = W.ldpc_encoder(rate, z)
    n, k = enc.n, enc.k
    g = torch.Generator(device="cuda")
    g.manual_seed(SEED)
    info = torch.randint(0, 2, (F, k), dtype=torch.uint8, device="cuda", generator=g)
    cw = torch.empty((F, n), dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        enc.encode_dev(info, cw)
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        enc.encode_dev(info, cw)
        ends[i].record(stream)
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    kern = [a.elapsed_time(b) for a, b in zip(starts, ends)]
    kavg = sum(kern) / len(kern)
    value = F * n / (ms_step * 1e-3) / 1e9
    bytes_per_frame = k + n
    achieved = F * bytes_per_frame / (kavg * 1e-3) / 1e9
    log("encode: kernel avg %.3f ms, step %.3f ms, %.1f coded Gbit/s, %.0f GB/s" % (kavg, ms_step, value, achieved))
    # HBM bytes of this kernel from the committed rocprofv3 counter runs
    # (FETCH_SIZE x2 and WRITE_SIZE x1 after copy calibration)
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "encode_r12z96_f524288_traffic.json")
    if os.path.exists(pmc):
        traffic = json.load(open(pmc)).get("hbm_bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        code = O.Code(rate, z)
        threads = O.omp_max_threads()
        S = 65536
        sample = info[:S].cpu().numpy()
        ref, _ = code.encode_batch(sample[:4096], threads)
        same = bool(np.array_equal(ref, cw[:4096].cpu().numpy()))
        sec, runs = median_seconds(lambda: code.encode_batch(sample, threads)[1])
        cpu = {"value": S * n / sec / 1e9, "unit": "coded Gbit/s", "cores": threads, "kind": "port",
               "sample": "the first %d info words of the batch, median of 5 timed runs after a warm-up; OpenMP "
                         "frame loop over the oracle's restatement of ldpc_encoder::encode_data" % S,
               "seconds": sec, "runs_s": runs, "bit_exact_vs_gpu": same, **cpu_env(threads)}
        log("cpu encoder (%d threads): %.3f coded Gbit/s, GPU == CPU on 4096 frames: %s" % (threads, cpu["value"], same))
    out = {"metric": "coded Gbit/s encoded (rate 1/2, z=96)", "value": value, "unit": "coded Gbit/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": value / 0.171904706, "dtype": "u8", "data": "synthetic",
           "config": {"workload": "dual-diagonal encoder, rate 1/2, z=96, n=2304, k=1152, %d frames, one byte per "
                                  "bit in and out" % F, "frames": F},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                        "frac": achieved / PEAK_HBM_GBS, "traffic": traffic, "kernel_ms": kavg,
                        "bytes_per_frame": bytes_per_frame},
           "cpu_baseline": cpu,
           "vs_baseline_note": "the reference's README figure, 171.904706 coded Mbit/s, i7-4771 (4 threads)"}
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
