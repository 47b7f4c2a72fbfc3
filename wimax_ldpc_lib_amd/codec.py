"""Python host mirror of the reference's codec classes over the C ABI.

ldpc_decoder / ldpc_encoder follow include/ldpc_decoder.h and
include/ldpc_encoder.h of the reference: same constructor arguments, decode()
returns what the reference's decode returns and replaces rx by the posterior,
encode_data() fills the codeword.  The *_dev methods take torch tensors on the
GPU (the device memory of this library's callers) and enqueue on the current
torch stream; the host methods take numpy arrays.
"""
import ctypes as C

import numpy as np

from . import capi
from .capi import check, lib

HALFRATE, TWOTHIRDSA, TWOTHIRDSB, THREEQUARTERSA, THREEQUARTERSB, FIVESIXTHS = range(6)


def _ptr(t):
    """Raw pointer of a torch tensor / numpy array / None."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        assert t.flags.c_contiguous
        return t.ctypes.data_as(C.c_void_p)
    assert t.is_contiguous()
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is not None:
        return C.c_void_p(stream)
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class ldpc:
    """Code parameters of (rate, z) -- include/ldpc.h:35-61."""

    def __init__(self, rate, z_factor, num_threads=1):
        self._set_params(capi.code_params(int(rate), int(z_factor)))
        self.num_threads = num_threads

    def _set_params(self, p):
        self.rate, self.z = p["rate"], p["z"]
        self.n, self.m, self.k, self.c = p["n"], p["m"], p["k"], p["c"]
        self.mb, self.D, self.X = p["mb"], p["d"], p["x"]

    def checknode_array(self):
        t = np.zeros((self.m, self.c), dtype=np.int16)
        if getattr(self, "qc", None) is not None:
            check(lib().wl_qc_table(self.qc._h, _ptr(t), t.size))
        else:
            check(lib().wl_code_table(self.rate, self.z, _ptr(t), t.size))
        return t


class qc_code:
    """A quasi-cyclic code given at run time: qc_code(base=[mb][nb] shifts,
    z=lifting) or qc_code(alist=text, z=0 to detect the lifting size)."""

    def __init__(self, base=None, z=None, alist=None):
        h = C.c_void_p()
        if alist is not None:
            text = alist.encode() if isinstance(alist, str) else bytes(alist)
            check(lib().wl_qc_from_alist(text, len(text), int(z or 0), C.byref(h)))
        else:
            b = np.ascontiguousarray(base, dtype=np.int16)
            check(lib().wl_qc_create(b.shape[1], b.shape[0], int(z), _ptr(b), C.byref(h)))
        self._h = h
        ci = capi._CodeInfo()
        check(lib().wl_qc_params(h, C.byref(ci)))
        self.params = {f: getattr(ci, f) for f, _ in capi._CodeInfo._fields_}
        self.z, self.n, self.m, self.k = ci.z, ci.n, ci.m, ci.k
        self.mb, self.c, self.D, self.X = ci.mb, ci.c, ci.d, ci.x

    def close(self):
        if getattr(self, "_h", None):
            lib().wl_qc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def base(self):
        nb, mb = C.c_int(), C.c_int()
        check(lib().wl_qc_base(self._h, C.byref(nb), C.byref(mb), None, 0))
        b = np.zeros((mb.value, nb.value), dtype=np.int16)
        check(lib().wl_qc_base(self._h, None, None, _ptr(b), b.size))
        return b

    def checknode_array(self):
        t = np.zeros((self.m, self.c), dtype=np.int16)
        check(lib().wl_qc_table(self._h, _ptr(t), t.size))
        return t

    def alist(self):
        n = C.c_size_t()
        check(lib().wl_qc_alist(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        check(lib().wl_qc_alist(self._h, buf, n.value + 1, C.byref(n)))
        return buf.value.decode()

    def decoder(self, max_iter, early_stop=True):
        return ldpc_decoder.from_qc(self, max_iter, early_stop)

    def encoder(self):
        return ldpc_encoder.from_qc(self)


class ldpc_decoder(ldpc):
    """include/ldpc_decoder.h:12-34 over wl_decoder."""

    RULES = {"nms": 0, "oms": capi.WL_FLAG_OMS, "ms": capi.WL_FLAG_MS}

    def __init__(self, rate, z_factor, max_iter, num_threads=1, early_stop=True, rule="nms"):
        """rule: the reference's compile-time message rule
        (lib/ldpc_decoder.cc:17-21,215-238): "nms" normalised min-sum (as
        shipped), "oms" offset min-sum, "ms" plain min-sum."""
        super().__init__(rate, z_factor, num_threads)
        h = C.c_void_p()
        flags = (capi.WL_FLAG_EARLY_STOP if early_stop else 0) | self.RULES[rule]
        check(lib().wl_decoder_create(int(rate), int(z_factor), int(max_iter), flags, C.byref(h)))
        self._h = h
        self.max_iter = max_iter
        self.early_stop = early_stop

    @classmethod
    def from_qc(cls, qc, max_iter, early_stop=True):
        """Decoder of a run-time quasi-cyclic code (qc_code)."""
        self = cls.__new__(cls)
        self._set_params(qc.params)
        self.num_threads = 1
        self.qc = qc
        h = C.c_void_p()
        flags = capi.WL_FLAG_EARLY_STOP if early_stop else 0
        check(lib().wl_decoder_create_qc(qc._h, int(max_iter), flags, C.byref(h)))
        self._h = h
        self.max_iter = max_iter
        self.early_stop = early_stop
        return self

    def close(self):
        if getattr(self, "_h", None):
            lib().wl_decoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def frames_per_wg(self):
        v = C.c_int()
        check(lib().wl_decoder_get_frames_per_wg(self._h, C.byref(v)))
        return v.value

    @frames_per_wg.setter
    def frames_per_wg(self, f):
        check(lib().wl_decoder_set_frames_per_wg(self._h, int(f)))

    @property
    def variant(self):
        v = C.c_int()
        check(lib().wl_decoder_get_variant(self._h, C.byref(v)))
        return v.value

    @variant.setter
    def variant(self, v):
        check(lib().wl_decoder_set_variant(self._h, int(v)))

    @property
    def small_batch(self):
        """Contiguous batches of at most this many frames use the small-batch
        kernel (one frame per workgroup, two lanes per row); 0 = never."""
        v = C.c_size_t()
        check(lib().wl_decoder_get_small_batch(self._h, C.byref(v)))
        return v.value

    @small_batch.setter
    def small_batch(self, n):
        """None: the default for this device; 0: never."""
        check(lib().wl_decoder_set_small_batch(self._h, capi.WL_SMALL_BATCH_AUTO if n is None else int(n)))

    def set_offsets(self, offsets):
        """Learned offsets [T][mb][c] (numpy float32), or None for plain NMS."""
        if offsets is None:
            check(lib().wl_decoder_set_offsets(self._h, None, 0))
            return
        o = np.ascontiguousarray(offsets, dtype=np.float32)
        assert o.ndim == 3 and o.shape[1:] == (self.mb, self.c), o.shape
        check(lib().wl_decoder_set_offsets(self._h, _ptr(o), o.shape[0]))

    def compute_syndrome(self, rx_codeword, exit_early):
        hd = np.ascontiguousarray(rx_codeword, dtype=np.int8)
        out = C.c_uint()
        check(lib().wl_compute_syndrome_host(self._h, _ptr(hd), int(bool(exit_early)), C.byref(out)))
        return out.value

    def decode(self, rx_codeword, decoded):
        """decode(float32 rx[n] or int8 bits[n], int8 decoded[n]) -> 0/1, in place."""
        ret = np.zeros(1, dtype=np.uint8)
        if rx_codeword.dtype == np.int8:
            check(lib().wl_decode_i8_host(self._h, _ptr(rx_codeword), 1, _ptr(decoded), _ptr(ret), None))
        else:
            assert rx_codeword.dtype == np.float32
            check(lib().wl_decode_f32_host(self._h, _ptr(rx_codeword), 1, _ptr(decoded), _ptr(ret), None))
        return int(ret[0])

    def decode_batch_host(self, rx, decoded, ret, iters=None):
        """Host batch: rx [F][n] float32 replaced by posteriors."""
        F = rx.shape[0]
        check(lib().wl_decode_f32_host(self._h, _ptr(rx), F, _ptr(decoded), _ptr(ret), _ptr(iters)))

    def decode_dev(self, llr, post=None, hard=None, ret=None, iters=None, synd=None, stream=None):
        """Device batch: llr [F][n] float32 CUDA tensor."""
        F = llr.shape[0] if llr.dim() == 2 else llr.numel() // self.n
        check(lib().wl_decode_f32_dev(self._h, _ptr(llr), F, _ptr(post), _ptr(hard), _ptr(ret),
                                      _ptr(iters), _ptr(synd), _stream(stream)))

    def decode_learned_fallback_dev(self, llr, fallback_iters, post=None, hard=None, ret=None, iters=None,
                                    synd=None, fallback=None, stream=None):
        F = llr.shape[0]
        check(lib().wl_decode_learned_fallback_dev(self._h, _ptr(llr), F, int(fallback_iters), _ptr(post),
                                                   _ptr(hard), _ptr(ret), _ptr(iters), _ptr(synd),
                                                   _ptr(fallback), _stream(stream)))

    def decode_learned_fallback_noalloc_dev(self, llr, fallback_iters, post=None, hard=None, ret=None, iters=None,
                                            synd=None, fallback=None, stream=None):
        """decode_learned_fallback_dev without any device allocation (static
        scratch; at most 1 << 20 frames per call; calls that overlap on one
        GPU race)."""
        F = llr.shape[0]
        check(lib().wl_decode_learned_fallback_noalloc_dev(self._h, _ptr(llr), F, int(fallback_iters), _ptr(post),
                                                           _ptr(hard), _ptr(ret), _ptr(iters), _ptr(synd),
                                                           _ptr(fallback), _stream(stream)))


class ldpc_encoder(ldpc):
    """include/ldpc_encoder.h:10-27 over wl_encoder."""

    def __init__(self, rate, z_factor, num_threads=1, standard=False):
        """standard=True: the general back-substitution encoder, valid for
        3/4B too (wl_encoder_create_standard); the default is the
        reference's encoder, whose 3/4B codewords fail H."""
        super().__init__(rate, z_factor, num_threads)
        h = C.c_void_p()
        fn = lib().wl_encoder_create_standard if standard else lib().wl_encoder_create
        check(fn(int(rate), int(z_factor), C.byref(h)))
        self._h = h

    @classmethod
    def from_qc(cls, qc):
        """Encoder of a run-time quasi-cyclic code with the dual-diagonal parity part."""
        self = cls.__new__(cls)
        self._set_params(qc.params)
        self.num_threads = 1
        self.qc = qc
        h = C.c_void_p()
        check(lib().wl_encoder_create_qc(qc._h, C.byref(h)))
        self._h = h
        return self

    def close(self):
        if getattr(self, "_h", None):
            lib().wl_encoder_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def encode_data(self, infoword, codeword):
        F = infoword.size // self.k
        check(lib().wl_encode_u8_host(self._h, _ptr(infoword), F, _ptr(codeword)))

    def encode_dev(self, info, cw, stream=None):
        F = info.numel() // self.k
        check(lib().wl_encode_u8_dev(self._h, _ptr(info), F, _ptr(cw), _stream(stream)))

    def channel_dev(self, seed, frame_offset, nframes, ebn0_lo, ebn0_hi=None, llr=None, bits=None,
                    ebn0=None, stream=None):
        hi = ebn0_lo if ebn0_hi is None else ebn0_hi
        check(lib().wl_channel_dev(self._h, int(seed), int(frame_offset), int(nframes), float(ebn0_lo),
                                   float(hi), _ptr(llr), _ptr(bits), _ptr(ebn0), _stream(stream)))


def unpack_hard(hard, n):
    """Packed hard decisions [F][n/8] (uint8, numpy) -> [F][n] int8 0/1."""
    h = np.asarray(hard, dtype=np.uint8)
    return np.unpackbits(h.reshape(h.shape[0], -1), axis=1, bitorder="little")[:, :n].astype(np.int8)
