"""Learned offset min-sum: trainer binding (wl_trainer_* of the C ABI).

    tr = Trainer(rate=1, z=96, T=5, batch=4096, seed=1, lr=1e-3, init_offset=0.15)
    tr.step(20000, 1.0, 4.0)           # GPU training, one hipGraph launch per step
    offsets = tr.offsets()             # [T][mb][c] float32 for ldpc_decoder.set_offsets
"""
import ctypes as C

import numpy as np

from .capi import check, lib
from .codec import _ptr


class Trainer:
    def __init__(self, rate, z, T=5, batch=4096, seed=1, lr=1e-3, init_offset=0.15):
        h = C.c_void_p()
        check(lib().wl_trainer_create(int(rate), int(z), int(T), int(batch), int(seed), float(lr),
                                      float(init_offset), C.byref(h)))
        self._h = h
        t, mb, c = C.c_int(), C.c_int(), C.c_int()
        check(lib().wl_trainer_num_params(h, C.byref(t), C.byref(mb), C.byref(c)))
        self.shape = (t.value, mb.value, c.value)
        self.rate, self.z, self.batch = rate, z, batch

    def close(self):
        if getattr(self, "_h", None):
            lib().wl_trainer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, nsteps, ebn0_lo, ebn0_hi):
        loss = C.c_float()
        check(lib().wl_trainer_step(self._h, int(nsteps), float(ebn0_lo), float(ebn0_hi), C.byref(loss)))
        return loss.value

    def offsets(self):
        o = np.zeros(self.shape, np.float32)
        check(lib().wl_trainer_get_offsets(self._h, _ptr(o)))
        return o

    def set_offsets(self, o):
        o = np.ascontiguousarray(o, dtype=np.float32).reshape(self.shape)
        check(lib().wl_trainer_set_offsets(self._h, _ptr(o)))

    def set_hparams(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        check(lib().wl_trainer_set_hparams(self._h, float(lr), float(beta1), float(beta2), float(eps)))

    def set_loss(self, per_iteration):
        """True: mean over the T iterations of each iteration's cross-entropy;
        False: the final posteriors' cross-entropy (default)."""
        check(lib().wl_trainer_set_loss(self._h, 1 if per_iteration else 0))

    def global_step(self):
        s = C.c_uint64()
        check(lib().wl_trainer_get_step(self._h, C.byref(s)))
        return s.value

    def save(self, path):
        check(lib().wl_trainer_save(self._h, str(path).encode()))

    def load(self, path):
        check(lib().wl_trainer_load(self._h, str(path).encode()))

    def grad(self, llr, bits, post=None):
        """(loss summed over frames and bits, gradient of the mean loss) on device frames."""
        g = np.zeros(self.shape, np.float32)
        loss = C.c_double()
        check(lib().wl_trainer_grad_dev(self._h, _ptr(llr), _ptr(bits), llr.shape[0], _ptr(post), _ptr(g),
                                        C.byref(loss)))
        return loss.value, g
