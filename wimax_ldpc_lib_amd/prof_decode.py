"""Short decode (or encode) workload for hardware-counter runs (rocprofv3 --pmc).

Runs the 65536-frame rate 5/6, z=96, fixed-10-iteration decode a few times
(--encode: the encoder of bench.py --workload encode, 524288 frames of rate
1/2, z=96) and then a plain device copy of a known size, whose FETCH_SIZE /
WRITE_SIZE readings calibrate the counters (bytes the copy must move: 2 x
its size).

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o run -- python3 -m wimax_ldpc_lib_amd.prof_decode [--encode]
"""
import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def copy(torch, mib):
    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=int, default=5)
    ap.add_argument("--z", type=int, default=96)
    ap.add_argument("--frames", type=int, default=65536)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--frames-per-wg", type=int, default=0)
    ap.add_argument("--copy-mib", type=int, default=512)
    ap.add_argument("--encode", action="store_true")
    a = ap.parse_args()
    import torch

    import wimax_ldpc_lib_amd as W
    if a.encode:
        enc = W.ldpc_encoder(0, 96)
        F = 524288
        info = torch.randint(0, 2, (F, enc.k), dtype=torch.uint8, device="cuda")
        cw = torch.empty((F, enc.n), dtype=torch.uint8, device="cuda")
        for _ in range(a.reps):
            enc.encode_dev(info, cw)
        torch.cuda.synchronize()
        copy(torch, a.copy_mib)
        print("encode: %d frames x %d reps (bytes/frame algorithmic %d); copy: %d MiB x 2"
              % (F, a.reps, enc.k + enc.n, a.copy_mib), flush=True)
        return
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
    copy(torch, a.copy_mib)
    print("decode: %d frames x %d reps (bytes/frame algorithmic %d); copy: %d MiB x 2"
          % (a.frames, a.reps, 4 * dec.n + dec.n // 8, a.copy_mib), flush=True)


if __name__ == "__main__":
    main()
