"""Shared helpers for the test-suite: seeded frames and oracle comparisons."""
import numpy as np

Z_VALUES = list(range(24, 97, 4))


def noisy_llr(code, nframes, ebn0_db, seed):
    """Seeded BPSK/AWGN frames of random codewords (oracle encoder), returned
    as float32 LLRs in the reference convention (>= 0 slices to 1), plus the
    codewords."""
    rng = np.random.default_rng(seed)
    info = rng.integers(0, 2, (nframes, code.k), dtype=np.uint8)
    cw = code.encode(info)
    R = code.k / code.n
    s2 = 1.0 / (2.0 * R * 10.0 ** (ebn0_db / 10.0))
    y = (1.0 - 2.0 * cw.astype(np.float64)) + rng.normal(0.0, np.sqrt(s2), cw.shape)
    llr = (-(2.0 * y / s2)).astype(np.float32)
    return llr, cw


def pack_bits(hd):
    return np.packbits(np.asarray(hd, dtype=np.uint8), axis=1, bitorder="little")


def bits_equal_f32(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))
