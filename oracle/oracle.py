"""ctypes view of liboracle.so (the C restatement in oracle.c) and of the
reference build in oracle/_ref -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, to check results against.  Function-level reference citations are in oracle.c.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libwimax_ref.so")

NMS, OMS, MS, LEARNED = 0, 1, 2, 3
RATE_NAMES = ["1/2", "2/3A", "2/3B", "3/4A", "3/4B", "5/6"]
Z_VALUES = list(range(24, 97, 4))


class _Code(C.Structure):
    _fields_ = [("rate", C.c_int), ("z", C.c_int), ("n", C.c_int), ("m", C.c_int),
                ("k", C.c_int), ("c", C.c_int), ("mb", C.c_int), ("d", C.c_int),
                ("x", C.c_int), ("cn", C.POINTER(C.c_int16)), ("row_len", C.POINTER(C.c_int)),
                ("blk", C.POINTER(C.c_int)), ("shift", C.POINTER(C.c_int)),
                ("deg", C.POINTER(C.c_int))]


class _Cfg(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("early_stop", C.c_int), ("variant", C.c_int),
                ("offsets", C.POINTER(C.c_float)), ("n_offset_iters", C.c_int)]


def build(ref_root=None):
    """Compile liboracle.so (and oracle/_ref when the reference is present)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    ref = ref_root or "/root/reference"
    if os.path.isdir(os.path.join(ref, "lib")):
        subprocess.check_call(["make", "-s", "-j8", "-C", HERE, "ref", "REF=" + ref])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.or_code_init.argtypes = [C.POINTER(_Code), C.c_int, C.c_int]
        L.or_code_init_qc.argtypes = [C.POINTER(_Code), C.c_int, C.c_int, C.c_int, P]
        L.or_code_free.argtypes = [C.POINTER(_Code)]
        L.or_syndrome.argtypes = [C.POINTER(_Code), P, C.c_int]
        L.or_syndrome.restype = C.c_uint
        L.or_decode_f32.argtypes = [C.POINTER(_Code), C.POINTER(_Cfg), P, P, P, C.POINTER(C.c_int)]
        L.or_decode_f32.restype = C.c_uint
        L.or_decode_i8.argtypes = [C.POINTER(_Code), C.POINTER(_Cfg), P, P, P, C.POINTER(C.c_int)]
        L.or_decode_i8.restype = C.c_uint
        L.or_encode.argtypes = [C.POINTER(_Code), P, P]
        L.or_encode_batch.argtypes = [C.POINTER(_Code), P, P, C.c_size_t, C.c_int]
        L.or_encode_batch.restype = C.c_double
        L.or_decode_batch.argtypes = [C.POINTER(_Code), C.POINTER(_Cfg), P, C.c_size_t, P, P, P, P, C.c_int]
        L.or_decode_batch.restype = C.c_double
        L.or_learned_fwd_bwd.argtypes = [C.POINTER(_Code), P, C.c_int, P, P, C.c_size_t, P, P, C.c_int]
        L.or_learned_fwd_bwd.restype = C.c_double
        L.or_omp_max_threads.restype = C.c_int
        L.or_row_update.argtypes = [C.c_int, C.c_int, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Code:
    """One (rate, z) code: sizes, the expanded [m][c] table and layer descriptors."""

    def __init__(self, rate, z, base=None):
        """rate, z: a compiled 802.16e code; rate None with base: the
        quasi-cyclic code of base matrix base [mb][nb] (shifts, -1 = zero) at z."""
        self._c = _Code()
        if base is not None:
            base = np.ascontiguousarray(base, dtype=np.int16)
            rc = lib().or_code_init_qc(C.byref(self._c), base.shape[1], base.shape[0], z, _p(base))
        else:
            rc = lib().or_code_init(C.byref(self._c), rate, z)
        if rc != 0:
            raise ValueError("invalid code %r/%r" % (rate, z))
        c = self._c
        rate = c.rate
        self.rate, self.z, self.n, self.m, self.k = rate, z, c.n, c.m, c.k
        self.c, self.mb, self.d, self.x = c.c, c.mb, c.d, c.x
        self.cn = np.ctypeslib.as_array(c.cn, shape=(c.m, c.c)).copy()
        self.row_len = np.ctypeslib.as_array(c.row_len, shape=(c.m,)).copy()
        self.blk = np.ctypeslib.as_array(c.blk, shape=(c.mb, c.c)).copy()
        self.shift = np.ctypeslib.as_array(c.shift, shape=(c.mb, c.c)).copy()
        self.deg = np.ctypeslib.as_array(c.deg, shape=(c.mb,)).copy()

    def __del__(self):
        try:
            lib().or_code_free(C.byref(self._c))
        except Exception:
            pass

    def dense_h(self):
        H = np.zeros((self.m, self.n), dtype=np.uint8)
        for mm in range(self.m):
            for j in range(self.row_len[mm]):
                H[mm, self.cn[mm, j]] = 1
        return H

    def syndrome(self, hd, early_exit=False):
        hd = np.ascontiguousarray(hd, dtype=np.int8)
        return lib().or_syndrome(C.byref(self._c), _p(hd), int(early_exit))

    def encode(self, info):
        """info [F][k] uint8 -> codewords [F][n] uint8."""
        info = np.ascontiguousarray(info, dtype=np.uint8).reshape(-1, self.k)
        cw = np.zeros((info.shape[0], self.n), dtype=np.uint8)
        for f in range(info.shape[0]):
            lib().or_encode(C.byref(self._c), _p(info[f]), _p(cw[f]))
        return cw

    def encode_batch(self, info, nthreads=0):
        """info [F][k] uint8 -> (codewords [F][n], seconds of the OpenMP frame loop)."""
        info = np.ascontiguousarray(info, dtype=np.uint8).reshape(-1, self.k)
        cw = np.zeros((info.shape[0], self.n), dtype=np.uint8)
        secs = lib().or_encode_batch(C.byref(self._c), _p(info), _p(cw), info.shape[0], nthreads)
        return cw, secs

    def _cfg(self, max_iter, early_stop, variant, offsets):
        cfg = _Cfg(max_iter, int(early_stop), variant, None, 0)
        keep = None
        if variant == LEARNED:
            keep = np.ascontiguousarray(offsets, dtype=np.float32)
            assert keep.ndim == 3 and keep.shape[1:] == (self.mb, self.c)
            cfg.offsets = keep.ctypes.data_as(C.POINTER(C.c_float))
            cfg.n_offset_iters = keep.shape[0]
        return cfg, keep

    def decode(self, llr, max_iter=10, early_stop=True, variant=NMS, offsets=None, nthreads=0):
        """Batch decode; llr [F][n] float32 (reference convention).  Returns
        dict(post, hd, ret, iters, synd_w, seconds)."""
        post = np.array(llr, dtype=np.float32, copy=True).reshape(-1, self.n)
        F = post.shape[0]
        hd = np.zeros((F, self.n), dtype=np.int8)
        ret = np.zeros(F, dtype=np.uint8)
        iters = np.zeros(F, dtype=np.uint8)
        sw = np.zeros(F, dtype=np.uint16)
        cfg, keep = self._cfg(max_iter, early_stop, variant, offsets)
        secs = lib().or_decode_batch(C.byref(self._c), C.byref(cfg), _p(post), F, _p(hd), _p(ret),
                                     _p(iters), _p(sw), nthreads)
        del keep
        return dict(post=post, hd=hd, ret=ret, iters=iters, synd_w=sw, seconds=secs)

    def decode_i8(self, bits, max_iter=10, early_stop=True):
        bits = np.ascontiguousarray(bits, dtype=np.int8).reshape(self.n)
        hd = np.zeros(self.n, dtype=np.int8)
        lmn = np.zeros(self.m * self.c, dtype=np.float32)
        it = C.c_int(0)
        cfg, _ = self._cfg(max_iter, early_stop, NMS, None)
        r = lib().or_decode_i8(C.byref(self._c), C.byref(cfg), _p(bits), _p(hd), _p(lmn), C.byref(it))
        return r, hd, it.value

    def learned_fwd_bwd(self, offsets, llr, bits, per_iteration=False):
        offsets = np.ascontiguousarray(offsets, dtype=np.float32)
        T = offsets.shape[0]
        llr = np.ascontiguousarray(llr, dtype=np.float32).reshape(-1, self.n)
        bits = np.ascontiguousarray(bits, dtype=np.uint8).reshape(-1, self.n)
        post = np.zeros_like(llr)
        g = np.zeros_like(offsets)
        loss = lib().or_learned_fwd_bwd(C.byref(self._c), _p(offsets), T, _p(llr), _p(bits),
                                        llr.shape[0], _p(post), _p(g), int(bool(per_iteration)))
        return loss, post, g


def row_update(variant, values, msg=None, beta=None):
    """One check row: returns (posteriors, new messages) for the row's
    variable values and previous messages."""
    x = np.array(values, dtype=np.float32, copy=True)
    m = np.zeros_like(x) if msg is None else np.array(msg, dtype=np.float32, copy=True)
    b = None if beta is None else np.ascontiguousarray(beta, dtype=np.float32)
    lib().or_row_update(variant, x.size, _p(x), _p(m), _p(b))
    return x, m


def omp_max_threads():
    return lib().or_omp_max_threads()


# ---------------- reference build (oracle/_ref), this container only -----------

_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        P = C.c_void_p
        I = C.POINTER(C.c_int)
        R.ref_table_info.argtypes = [C.c_int, C.c_int, I, I, I, I, I, I]
        R.ref_table_copy.argtypes = [C.c_int, C.c_int, P]
        R.ref_encode.argtypes = [C.c_int, C.c_int, C.c_int, P, P, C.c_int]
        _ref = R
    return _ref


def ref_table(rate, z):
    """The reference's own expanded table and constants for (rate, z)."""
    v = [C.c_int() for _ in range(6)]
    rc = ref_lib().ref_table_info(rate, z, *[C.byref(x) for x in v])
    if rc != 0:
        raise ValueError("reference rejects rate/z %r/%r" % (rate, z))
    n, m, rows, cols, d, x = [x.value for x in v]
    t = np.zeros((rows, cols), dtype=np.int16)
    ref_lib().ref_table_copy(rate, z, _p(t))
    return dict(n=n, m=m, rows=rows, cols=cols, d=d, x=x, table=t)


def ref_encode(rate, z, info, nthreads=1):
    info = np.ascontiguousarray(info, dtype=np.uint8)
    t = ref_table(rate, z)
    k = t["n"] - t["m"]
    info = info.reshape(-1, k)
    cw = np.zeros((info.shape[0], t["n"]), dtype=np.uint8)
    rc = ref_lib().ref_encode(rate, z, nthreads, _p(info), _p(cw), info.shape[0])
    assert rc == 0
    return cw
