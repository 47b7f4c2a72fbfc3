"""Mixed-(rate, z) batches over wl_mixed_* of the C ABI."""
import ctypes as C

from .capi import WL_FLAG_EARLY_STOP, WL_FLAG_MIXED_UNORDERED, check, lib
from .codec import _ptr, _stream

NUM_Z = 19


def code_id(rate, z):
    return int(rate) * NUM_Z + (int(z) - 24) // 4


class MixedDecoder:
    def __init__(self, max_iter=10, early_stop=True, codes=None, ordered=True):
        """ordered=False: frames of a code bucket are not arranged by
        reliability (WL_FLAG_MIXED_UNORDERED); results are the same."""
        h = C.c_void_p()
        flags = (WL_FLAG_EARLY_STOP if early_stop else 0) | (0 if ordered else WL_FLAG_MIXED_UNORDERED)
        check(lib().wl_mixed_create(int(max_iter), flags, C.byref(h)))
        self._h = h
        if codes is not None:
            self.set_codes(codes)

    def close(self):
        if getattr(self, "_h", None):
            lib().wl_mixed_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_codes(self, codes):
        arr = (C.c_int * len(codes))(*[int(c) for c in codes])
        check(lib().wl_mixed_set_codes(self._h, arr if len(codes) else None, len(codes)))

    def launch(self, rate, z):
        f, v, g = C.c_int(), C.c_int(), C.c_int()
        check(lib().wl_mixed_get_launch(self._h, int(rate), int(z), C.byref(f), C.byref(v), C.byref(g)))
        return f.value, v.value, g.value

    def set_group_frames_per_wg(self, z, frames_per_wg):
        """Frames per workgroup of the kernel of lifting size z (0: the tuned default)."""
        check(lib().wl_mixed_set_group_frames_per_wg(self._h, int(z), int(frames_per_wg)))

    def offsets_dev(self, cfg, off, stream=None):
        check(lib().wl_mixed_offsets_dev(self._h, _ptr(cfg), cfg.numel(), _ptr(off), _stream(stream)))

    def channel_dev(self, seed, cfg, off, ebn0_lo, ebn0_hi, llr, bits=None, ebn0=None, stream=None):
        check(lib().wl_mixed_channel_dev(self._h, int(seed), _ptr(cfg), _ptr(off), cfg.numel(), float(ebn0_lo),
                                         float(ebn0_hi), _ptr(llr), _ptr(bits), _ptr(ebn0), _stream(stream)))

    def decode_dev(self, cfg, off, llr, post=None, hard=None, ret=None, iters=None, synd=None, stream=None):
        check(lib().wl_mixed_decode_dev(self._h, _ptr(cfg), _ptr(off), _ptr(llr), cfg.numel(), _ptr(post),
                                        _ptr(hard), _ptr(ret), _ptr(iters), _ptr(synd), _stream(stream)))

    def set_offsets(self, rate, offsets):
        import numpy as np
        if offsets is None:
            check(lib().wl_mixed_set_offsets(self._h, int(rate), None, 0))
            return
        o = np.ascontiguousarray(offsets, dtype=np.float32)
        check(lib().wl_mixed_set_offsets(self._h, int(rate), _ptr(o), o.shape[0]))

    def decode_learned_fallback_dev(self, cfg, off, llr, fallback_iters, post=None, hard=None, ret=None,
                                    iters=None, synd=None, fallback=None, stream=None):
        check(lib().wl_mixed_decode_learned_fallback_dev(self._h, _ptr(cfg), _ptr(off), _ptr(llr), cfg.numel(),
                                                         int(fallback_iters), _ptr(post), _ptr(hard), _ptr(ret),
                                                         _ptr(iters), _ptr(synd), _ptr(fallback),
                                                         _stream(stream)))

    def bucket_list(self, idx=None):
        """{first slot, count} per code id ([114, 2] numpy) of the last
        decode; with idx (a uint32/int32 device tensor of the batch size) also
        the frame ids in slot order."""
        import numpy as np
        b = np.zeros((114, 2), dtype=np.uint32)
        check(lib().wl_mixed_bucket_list(self._h, _ptr(b), _ptr(idx), 0 if idx is None else idx.numel()))
        return b

    def invalid_count(self):
        v = C.c_uint64()
        check(lib().wl_mixed_invalid_count(self._h, C.byref(v)))
        return v.value
