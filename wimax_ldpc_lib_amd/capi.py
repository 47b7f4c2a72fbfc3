"""ctypes binding of the C ABI in include/wimax_ldpc_amd.h.

Loads wimax_ldpc_lib_amd/lib/libwimax_ldpc_amd.so (built in-tree by
wimax_ldpc_lib_amd/build.py).  There is no fallback: if the library is
missing or a call fails, an exception is raised.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libwimax_ldpc_amd.so")

WL_OK, WL_EINVAL, WL_ENOMEM, WL_EDEVICE, WL_ENOTSUP = 0, -22, -12, -5, -95
WL_FLAG_EARLY_STOP = 0x1
WL_FLAG_MIXED_UNORDERED = 0x2
WL_FLAG_OMS = 0x4
WL_FLAG_MS = 0x8
WL_SMALL_BATCH_AUTO = C.c_size_t(-1).value

# every function the header declares: name -> (restype, argtypes)
_P = C.c_void_p
_SZ = C.c_size_t
SIGNATURES = {
    "wl_last_error": (C.c_char_p, []),
    "wl_version": (C.c_char_p, []),
    "wl_device_count": (C.c_int, []),
    "wl_code_params": (C.c_int, [C.c_int, C.c_int, _P]),
    "wl_code_table": (C.c_int, [C.c_int, C.c_int, _P, _SZ]),
    "wl_decoder_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_uint, C.POINTER(_P)]),
    "wl_decoder_destroy": (C.c_int, [_P]),
    "wl_decoder_set_frames_per_wg": (C.c_int, [_P, C.c_int]),
    "wl_decoder_get_frames_per_wg": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "wl_decoder_set_variant": (C.c_int, [_P, C.c_int]),
    "wl_decoder_get_variant": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "wl_decoder_set_small_batch": (C.c_int, [_P, _SZ]),
    "wl_decoder_get_small_batch": (C.c_int, [_P, C.POINTER(_SZ)]),
    "wl_tuned_launch": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "wl_decode_f32_dev": (C.c_int, [_P, _P, _SZ, _P, _P, _P, _P, _P, _P]),
    "wl_decode_f32_host": (C.c_int, [_P, _P, _SZ, _P, _P, _P]),
    "wl_decode_i8_host": (C.c_int, [_P, _P, _SZ, _P, _P, _P]),
    "wl_compute_syndrome_host": (C.c_int, [_P, _P, C.c_int, C.POINTER(C.c_uint)]),
    "wl_decoder_set_offsets": (C.c_int, [_P, _P, C.c_int]),
    "wl_decode_learned_fallback_dev": (C.c_int, [_P, _P, _SZ, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
    "wl_decode_learned_fallback_noalloc_dev": (C.c_int, [_P, _P, _SZ, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
    "wl_mixed_create": (C.c_int, [C.c_int, C.c_uint, C.POINTER(_P)]),
    "wl_mixed_destroy": (C.c_int, [_P]),
    "wl_mixed_set_codes": (C.c_int, [_P, _P, C.c_int]),
    "wl_mixed_get_launch": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "wl_mixed_set_group_frames_per_wg": (C.c_int, [_P, C.c_int, C.c_int]),
    "wl_mixed_offsets_dev": (C.c_int, [_P, _P, _SZ, _P, _P]),
    "wl_mixed_decode_dev": (C.c_int, [_P, _P, _P, _P, _SZ, _P, _P, _P, _P, _P, _P]),
    "wl_mixed_set_offsets": (C.c_int, [_P, C.c_int, _P, C.c_int]),
    "wl_mixed_decode_learned_fallback_dev": (C.c_int, [_P, _P, _P, _P, _SZ, C.c_int, _P, _P, _P, _P, _P, _P, _P]),
    "wl_mixed_invalid_count": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "wl_mixed_bucket_list": (C.c_int, [_P, _P, _P, _SZ]),
    "wl_mixed_channel_dev": (C.c_int, [_P, C.c_uint64, _P, _P, _SZ, C.c_float, C.c_float, _P, _P, _P, _P]),
    "wl_trainer_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_float, C.c_float,
                                     C.POINTER(_P)]),
    "wl_trainer_destroy": (C.c_int, [_P]),
    "wl_trainer_set_hparams": (C.c_int, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "wl_trainer_set_loss": (C.c_int, [_P, C.c_int]),
    "wl_trainer_step": (C.c_int, [_P, C.c_int, C.c_float, C.c_float, C.POINTER(C.c_float)]),
    "wl_trainer_num_params": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "wl_trainer_get_offsets": (C.c_int, [_P, _P]),
    "wl_trainer_set_offsets": (C.c_int, [_P, _P]),
    "wl_trainer_get_step": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "wl_trainer_save": (C.c_int, [_P, C.c_char_p]),
    "wl_trainer_load": (C.c_int, [_P, C.c_char_p]),
    "wl_trainer_grad_dev": (C.c_int, [_P, _P, _P, _SZ, _P, _P, C.POINTER(C.c_double)]),
    "wl_encoder_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(_P)]),
    "wl_encoder_destroy": (C.c_int, [_P]),
    "wl_encoder_create_standard": (C.c_int, [C.c_int, C.c_int, C.POINTER(_P)]),
    "wl_encode_u8_dev": (C.c_int, [_P, _P, _SZ, _P, _P]),
    "wl_encode_u8_host": (C.c_int, [_P, _P, _SZ, _P]),
    "wl_channel_dev": (C.c_int, [_P, C.c_uint64, C.c_uint64, _SZ, C.c_float, C.c_float, _P, _P, _P, _P]),
    "wl_qc_create": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, C.POINTER(_P)]),
    "wl_qc_from_alist": (C.c_int, [C.c_char_p, _SZ, C.c_int, C.POINTER(_P)]),
    "wl_qc_destroy": (C.c_int, [_P]),
    "wl_qc_params": (C.c_int, [_P, _P]),
    "wl_qc_base": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int), _P, _SZ]),
    "wl_qc_table": (C.c_int, [_P, _P, _SZ]),
    "wl_qc_alist": (C.c_int, [_P, _P, _SZ, C.POINTER(_SZ)]),
    "wl_decoder_create_qc": (C.c_int, [_P, C.c_int, C.c_uint, C.POINTER(_P)]),
    "wl_encoder_create_qc": (C.c_int, [_P, C.POINTER(_P)]),
}


class WLError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("wimax_ldpc_amd error %d: %s" % (code, msg))
        self.code = code


class _CodeInfo(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("rate", "z", "n", "m", "k", "c", "mb", "d", "x")]


_lib = None


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libwimax_ldpc_amd.so is not built (%s); run "
                              "python -m wimax_ldpc_lib_amd.build" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != WL_OK:
        raise WLError(rc, lib().wl_last_error().decode())
    return rc


def code_params(rate, z):
    ci = _CodeInfo()
    check(lib().wl_code_params(rate, z, C.byref(ci)))
    return {f: getattr(ci, f) for f, _ in _CodeInfo._fields_}


def device_count():
    return lib().wl_device_count()
