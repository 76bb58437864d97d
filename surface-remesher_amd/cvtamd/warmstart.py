"""Host API of the learned warm start (cvt_unet_* of include/cvt.h)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import call, lib
from .engine import _ptr

IN_C = 32
WEIGHTS = Path(__file__).resolve().parent.parent / "weights" / "warmstart_r256.bin"


class UNet:
    def __init__(self, R: int = 256, max_batch: int = 1, seed: int = 0):
        h = C.c_void_p()
        call("cvt_unet_create", C.byref(h), int(R), int(max_batch), C.c_ulonglong(seed))
        self._h, self.R, self.max_batch = h, int(R), int(max_batch)

    def close(self):
        if getattr(self, "_h", None):
            lib().cvt_unet_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def num_params(self) -> int:
        return call("cvt_unet_num_params", self._h)

    def set_stream(self, handle):
        call("cvt_unet_set_stream", self._h, C.c_void_p(handle or 0))

    def params(self, out=None):
        out = np.empty(self.num_params, np.float32) if out is None else out
        p, k = _ptr(out)
        call("cvt_unet_get_params", self._h, p)
        return out

    def set_params(self, params):
        p, k = _ptr(params, np.float32)
        call("cvt_unet_set_params", self._h, p)

    def grads(self, out=None):
        out = np.empty(self.num_params, np.float32) if out is None else out
        p, k = _ptr(out)
        call("cvt_unet_get_grads", self._h, p)
        return out

    def prepare(self, slot, density, xy, dst=None):
        d, k1 = _ptr(density, np.float32)
        n = int(density.shape[0])
        S = int(xy.shape[0]) if not hasattr(xy, "numel") else xy.numel() // 2
        x, k2 = _ptr(xy, np.float32)
        o, k3 = _ptr(dst)
        call("cvt_unet_prepare", self._h, int(slot), d, n, x, S, o)

    def set_input(self, x, nb):
        p, k = _ptr(x)
        call("cvt_unet_set_input", self._h, p, int(nb))

    def forward(self, nb):
        call("cvt_unet_forward", self._h, int(nb))

    def field(self, nb, out=None):
        out = np.empty((nb, self.R, self.R, 2), np.float32) if out is None else out
        p, k = _ptr(out)
        call("cvt_unet_get_field", self._h, p, int(nb))
        return out

    def loss_backward(self, nb, uv, target, S, sync=True):
        pu, k1 = _ptr(uv, np.float32)
        pt, k2 = _ptr(target, np.float32)
        loss = C.c_float()
        call("cvt_unet_loss_backward", self._h, int(nb), pu, pt, int(S),
             C.byref(loss) if sync else None)
        return loss.value if sync else None

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        call("cvt_unet_adam", self._h, float(lr), float(beta1), float(beta2), float(eps))

    def sample(self, slot, uv):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.empty_like(uv)
        call("cvt_unet_sample", self._h, int(slot), uv.ctypes.data_as(C.c_void_p), uv.shape[0],
             out.ctypes.data_as(C.c_void_p))
        return out

    def warm_start(self, engine):
        """Moves the engine's sites by the network's displacement field, on
        the engine's stream (cvt_unet_warm_start)."""
        call("cvt_unet_warm_start", self._h, engine._h)

    def save(self, path):
        call("cvt_unet_save", self._h, str(path).encode())

    def load(self, path):
        call("cvt_unet_load", self._h, str(path).encode())

    def synchronize(self):
        call("cvt_unet_synchronize", self._h)
