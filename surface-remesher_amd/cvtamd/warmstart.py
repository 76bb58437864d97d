"""Host API of the learned warm start (cvt_unet_* of include/cvt.h) and the
checked warm start: plain Lloyd decides whether the network's sites are used."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from ._lib import call, lib
from .engine import _out_ptr, _ptr

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
        call("cvt_unet_get_params", self._h, _out_ptr(out, np.float32, self.num_params))
        return out

    def set_params(self, params):
        p, k = _ptr(params, np.float32)
        call("cvt_unet_set_params", self._h, p)

    def grads(self, out=None):
        out = np.empty(self.num_params, np.float32) if out is None else out
        call("cvt_unet_get_grads", self._h, _out_ptr(out, np.float32, self.num_params))
        return out

    def prepare(self, slot, density, xy, dst=None):
        d, k1 = _ptr(density, np.float32)
        n = int(density.shape[0])
        S = int(xy.shape[0]) if not hasattr(xy, "numel") else xy.numel() // 2
        x, k2 = _ptr(xy, np.float32)
        o = None if dst is None else _out_ptr(dst, np.uint16, self.R * self.R * 32,
                                                ('bfloat16', 'int16'))
        call("cvt_unet_prepare", self._h, int(slot), d, n, x, S, o)

    def set_input(self, x, nb):
        p, k = _ptr(x)
        call("cvt_unet_set_input", self._h, p, int(nb))

    def forward(self, nb):
        call("cvt_unet_forward", self._h, int(nb))

    def field(self, nb, out=None):
        out = np.empty((nb, self.R, self.R, 2), np.float32) if out is None else out
        call("cvt_unet_get_field", self._h, _out_ptr(out, np.float32, nb * self.R * self.R * 2),
             int(nb))
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


TRIAL = 10


def warm_start_checked(net: UNet, engine, xy0, trial: int = TRIAL) -> dict:
    """Warm start with plain Lloyd as the arbiter and the fallback.

    From the initial sites xy0 (device or host, pixel units), run `trial`
    plain Lloyd iterations and keep their sites; then, from xy0 again, move
    the sites by the network (cvt_unet_warm_start) and run `trial`
    iterations.  The warm-started run continues if its energy is finite and
    not above the plain one's; otherwise the engine goes back to the plain
    run's sites.  Either way the engine is left after `trial` iterations of
    the chosen run, with its history cleared.  Returns {"warm": bool,
    "e_plain", "e_warm": the energies at the last trial iteration,
    "trial_energies": the chosen run's per-iteration energies}.
    """
    import torch
    engine.set_sites(xy0)
    engine.clear_history()
    engine.lloyd(trial)
    h_plain = engine.history()[:, 0].copy()
    dev = xy0.device if hasattr(xy0, "device") else torch.device("cuda")
    kept = torch.empty((engine.num_sites, 2), dtype=torch.float32, device=dev)
    engine.sites(out=kept)
    engine.set_sites(xy0)
    net.warm_start(engine)
    engine.clear_history()
    engine.lloyd(trial)
    h_warm = engine.history()[:, 0].copy()
    engine.clear_history()
    e_plain, e_warm = float(h_plain[-1]), float(h_warm[-1])
    warm = bool(np.isfinite(e_warm) and e_warm <= e_plain)
    if not warm:
        engine.set_sites(kept)
    return {"warm": warm, "e_plain": e_plain, "e_warm": e_warm,
            "trial_energies": h_warm if warm else h_plain}
