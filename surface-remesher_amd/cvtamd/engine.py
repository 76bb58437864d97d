"""Python host API over the C ABI of include/cvt.h (one grid per Engine).

Inputs may be numpy arrays (host) or torch tensors (host or the engine's
device); outputs come back as numpy arrays unless a destination tensor is
given.  The Python layer holds no compute: every call goes to
libcvt_hip.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LABEL_NONE, call, lib


def _torch_ready(a):
    """A torch device tensor may still be written by work queued on the
    caller's current torch stream, which the library's streams do not wait
    for: wait for that stream before the library reads or writes it."""
    if hasattr(a, "data_ptr") and getattr(a, "is_cuda", False):
        import torch
        torch.cuda.current_stream(a.device).synchronize()


def _ptr(a, dtype=None):
    """(pointer, keep-alive) for an input: a numpy array (converted to a
    contiguous copy of dtype when needed) or a contiguous torch tensor of
    that dtype."""
    if a is None:
        return None, None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        if dtype is not None:
            got, want = str(a.dtype).replace("torch.", ""), np.dtype(dtype).name
            if got != want and not (got == "bool" and want == "uint8"):
                raise TypeError(f"tensor dtype {a.dtype}, expected {want}")
        _torch_ready(a)
        return C.c_void_p(a.data_ptr()), a
    arr = np.ascontiguousarray(a, dtype=dtype)
    return arr.ctypes.data_as(C.c_void_p), arr


def _out_ptr(a, dtype, count, torch_dtypes=()):
    """Pointer to an output the library writes in place: a numpy array or
    torch tensor of exactly dtype (or, for a tensor, one of torch_dtypes,
    e.g. bfloat16 for raw 16-bit words), C-contiguous, writable and holding
    exactly count elements.  Never a copy (a copy would receive the results
    and the caller's array would be left unchanged): anything else raises."""
    want = np.dtype(dtype)
    if hasattr(a, "data_ptr"):
        got = str(a.dtype).replace("torch.", "")
        if got != want.name and got not in torch_dtypes:
            raise TypeError(f"out tensor dtype {a.dtype}, expected {want.name}")
        if not a.is_contiguous():
            raise ValueError("out tensor must be contiguous")
        if a.numel() != count:
            raise ValueError(f"out tensor holds {a.numel()} values, expected {count}")
        _torch_ready(a)
        return C.c_void_p(a.data_ptr())
    if not isinstance(a, np.ndarray):
        raise TypeError("out must be a numpy array or a torch tensor")
    if a.dtype != want:
        raise TypeError(f"out array dtype {a.dtype}, expected {want.name}")
    if not a.flags["C_CONTIGUOUS"] or not a.flags["WRITEABLE"]:
        raise ValueError("out array must be C-contiguous and writable")
    if a.size != count:
        raise ValueError(f"out array holds {a.size} values, expected {count}")
    return a.ctypes.data_as(C.c_void_p)


class Engine:
    """A CVT engine context for a width x height grid and up to max_sites sites."""

    def __init__(self, width: int, height: int | None = None, max_sites: int = 0):
        height = width if height is None else height
        h = C.c_void_p()
        call("cvt_create", C.byref(h), int(width), int(height), int(max_sites))
        self._h = h
        self.width, self.height, self.max_sites = int(width), int(height), int(max_sites)

    # -- lifetime --------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().cvt_destroy(self._h)
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

    # -- inputs ----------------------------------------------------------
    def set_stream(self, stream_handle: int | None):
        call("cvt_set_stream", self._h, C.c_void_p(stream_handle or 0))

    def set_density(self, density):
        if hasattr(density, "numel"):
            assert density.numel() == self.width * self.height
        else:
            density = np.ascontiguousarray(density, np.float32)
            assert density.size == self.width * self.height, density.shape
        p, keep = _ptr(density, np.float32)
        call("cvt_set_density", self._h, p)

    def set_sites(self, xy, fixed=None):
        if hasattr(xy, "numel"):
            n = xy.numel() // 2
        else:
            xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
            n = xy.shape[0]
        p, k1 = _ptr(xy, np.float32)
        f, k2 = _ptr(fixed, np.uint8)
        call("cvt_set_sites", self._h, p, f, int(n))

    # -- queries ---------------------------------------------------------
    @property
    def num_sites(self) -> int:
        return call("cvt_num_sites", self._h)

    def sites(self, out=None):
        """(num_sites, 2) float32 sites; out: an array or tensor of exactly
        num_sites * 2 float32 values to write instead."""
        n = self.num_sites
        if out is None:
            out = np.empty((n, 2), np.float32)
        call("cvt_get_sites", self._h, _out_ptr(out, np.float32, 2 * n))
        return out

    def voronoi(self):
        call("cvt_voronoi", self._h)

    def labels(self, out=None):
        """(height, width) int32 labels; out: an array or tensor of exactly
        width * height int32 values to write instead."""
        if out is None:
            out = np.empty((self.height, self.width), np.int32)
        call("cvt_get_labels", self._h, _out_ptr(out, np.int32, self.width * self.height))
        return out

    def moments(self):
        n = self.num_sites
        out = np.zeros((max(n, 1), 4), np.float64)
        call("cvt_moments", self._h, _out_ptr(out, np.float64, out.size))
        return out[:n]

    def lloyd(self, iterations: int, omega: float = 1.0):
        call("cvt_lloyd", self._h, int(iterations), float(omega))

    def lloyd_until(self, max_iter: int, rel_tol: float = 0.0, window: int = 0,
                    target: float = -np.inf, omega: float = 1.0) -> int:
        """Lloyd steps that stop on the device (cvt_lloyd_until): after the
        first iteration with energy <= target, or whose energy fell by less
        than rel_tol * E over the last `window` iterations of the call.
        Returns the number of iterations run."""
        n = C.c_int(0)
        call("cvt_lloyd_until", self._h, float(omega), float(rel_tol), int(window),
             float(target), int(max_iter), C.byref(n))
        return n.value

    def history(self):
        n = call("cvt_history_length", self._h)
        out = np.zeros((max(n, 1), 2), np.float64)
        call("cvt_get_history", self._h, _out_ptr(out, np.float64, out.size), n)
        return out[:n]

    def clear_history(self):
        call("cvt_clear_history", self._h)

    def energy(self):
        e = C.c_double()
        g = C.c_double()
        call("cvt_energy", self._h, C.byref(e), C.byref(g))
        return e.value, g.value

    def profile(self, iterations: int = 5, omega: float = 1.0) -> dict:
        """Eager Lloyd steps timed per phase with HIP events (ms); jfa_pass is
        one dense JFA pass (passes 1 .. P-1, mean), jfa_first the first pass
        (from the seed map)."""
        ms = np.zeros(8, np.float64)
        call("cvt_profile", self._h, int(iterations), float(omega), _out_ptr(ms, np.float64, 8))
        return {"jfa_pass": ms[0], "jfa_first": ms[2], "moments": ms[1], "update": ms[3],
                "jfa": ms[4], "step": ms[5], "jfa_passes": int(ms[6]), "tile": int(ms[7])}

    def synchronize(self):
        call("cvt_synchronize", self._h)

    @property
    def jfa_shape(self) -> tuple:
        """JFA launch shape (runs_mask, blocks_per_cu).  Set: runs_mask -1 =
        automatic, else bit k = runs at step 2^k (bands where clear);
        blocks_per_cu 0 = the built-in grid (cvt_set_jfa_shape).  Read: the
        steps that take runs under the current shape, and blocks_per_cu."""
        mask, bpc = C.c_int(), C.c_int()
        call("cvt_get_jfa_shape", self._h, C.byref(mask), C.byref(bpc))
        return mask.value, bpc.value

    @jfa_shape.setter
    def jfa_shape(self, shape):
        mask, bpc = shape
        call("cvt_set_jfa_shape", self._h, int(mask), int(bpc))

    @property
    def tile(self) -> int:
        return call("cvt_get_tile", self._h)

    @tile.setter
    def tile(self, t: int):
        call("cvt_set_tile", self._h, int(t))


__all__ = ["Engine", "LABEL_NONE"]


class Batch:
    """Many patches advanced together (cvt_batch_* of include/cvt.h)."""

    def __init__(self, sizes, max_sites):
        """sizes: [(width, height)], max_sites: [int] per patch."""
        w = np.array([s[0] for s in sizes], np.int32)
        h = np.array([s[1] for s in sizes], np.int32)
        m = np.array(max_sites, np.int32)
        handle = C.c_void_p()
        call("cvt_batch_create", C.byref(handle), len(sizes), w.ctypes.data_as(C.c_void_p),
             h.ctypes.data_as(C.c_void_p), m.ctypes.data_as(C.c_void_p))
        self._h = handle
        self.sizes = [(int(a), int(b)) for a, b in sizes]
        self._S = [0] * len(sizes)

    def close(self):
        if getattr(self, "_h", None):
            lib().cvt_batch_destroy(self._h)
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

    def __len__(self):
        return len(self.sizes)

    def set_stream(self, stream_handle):
        call("cvt_batch_set_stream", self._h, C.c_void_p(stream_handle or 0))

    def set_density(self, p, density):
        pd, keep = _ptr(density, np.float32)
        call("cvt_batch_set_density", self._h, int(p), pd)

    def set_sites(self, p, xy, fixed=None):
        if hasattr(xy, "numel"):
            n = xy.numel() // 2
        else:
            xy = np.ascontiguousarray(xy, np.float32).reshape(-1, 2)
            n = xy.shape[0]
        px, k1 = _ptr(xy, np.float32)
        pf, k2 = _ptr(fixed, np.uint8)
        call("cvt_batch_set_sites", self._h, int(p), px, pf, int(n))
        self._S[p] = int(n)

    def sites(self, p):
        out = np.empty((self._S[p], 2), np.float32)
        call("cvt_batch_get_sites", self._h, int(p), out.ctypes.data_as(C.c_void_p))
        return out

    def lloyd(self, iterations, omega=1.0):
        call("cvt_batch_lloyd", self._h, int(iterations), float(omega))

    def voronoi(self):
        call("cvt_batch_voronoi", self._h)

    def labels(self, p):
        w, h = self.sizes[p]
        out = np.empty((h, w), np.int32)
        call("cvt_batch_get_labels", self._h, int(p), out.ctypes.data_as(C.c_void_p))
        return out

    def history(self, p):
        n = call("cvt_batch_history", self._h, int(p), None, 0)
        out = np.zeros((max(n, 1), 2), np.float64)
        k = call("cvt_batch_history", self._h, int(p), out.ctypes.data_as(C.c_void_p), n)
        return out[:k]

    def set_tile(self, p, tile):
        call("cvt_batch_set_tile", self._h, int(p), int(tile))

    def tile(self, p):
        return call("cvt_batch_get_tile", self._h, int(p))

    def synchronize(self):
        call("cvt_batch_synchronize", self._h)
