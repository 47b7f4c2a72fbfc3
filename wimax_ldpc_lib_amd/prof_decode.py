"""Short decode workload for hardware-counter runs (rocprofv3 --pmc).

Runs the 65536-frame rate 5/6, z=96, fixed-10-iteration decode a few times
and then a plain device copy of a known size, whose FETCH_SIZE / WRITE_SIZE
readings calibrate the counters (bytes the copy must move: 2 x its size).

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o run -- python3 -m wimax_ldpc_lib_amd.prof_decode
"""
import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=int, default=5)
    ap.add_argument("--z", type=int, default=96)
    ap.add_argument("--frames", type=int, default=65536)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--frames-per-wg", type=int, default=0)
    ap.add_argument("--copy-mib", type=int, default=512)
    a = ap.parse_args()
    import torch

    import wimax_ldpc_lib_amd as W
    enc = W.ldpc_encoder(a.rate, a.z)
    dec = W.ldpc_decoder(a.rate, a.z, a.iters, early_stop=False)
    if a.frames_per_wg:
        dec.frames_per_wg = a.frames_per_wg
    llr = torch.empty((a.frames, dec.n), dtype=torch.float32, device="cuda")
    enc.channel_dev(7, 0, a.frames, 4.0, llr=llr)
    hard = torch.empty((a.frames, dec.n // 8), dtype=torch.uint8, device="cuda")
    ret = torch.empty(a.frames, dtype=torch.uint8, device="cuda")
    for _ in range(a.reps):
        dec.decode_dev(llr, None, hard, ret)
    torch.cuda.synchronize()
    src = torch.ones(a.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    print("decode: %d frames x %d reps (bytes/frame algorithmic %d); copy: %d MiB x 2"
          % (a.frames, a.reps, 4 * dec.n + dec.n // 8, a.copy_mib), flush=True)


if __name__ == "__main__":
    main()
