"""Oracle (test infrastructure only): numpy front end of ``fft_ref.c``.

Applies the C restatement of the DFT along one axis of an N-d array.
Computation is always FP64 (complex128 / float64).  Only the test-suite,
``__graft_entry__.smoke()`` and bench.py's CPU baseline may use it.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liboracle_fft.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def _load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        lib = ctypes.CDLL(str(_LIB_PATH))
        dp = ctypes.POINTER(ctypes.c_double)
        for name in ("fft_c2c", "dft_ld_c2c"):
            f = getattr(lib, name)
            f.argtypes = [ctypes.c_long, ctypes.c_long, dp, dp, ctypes.c_int]
            f.restype = None
        for name in ("fft_r2c", "fft_c2r"):
            f = getattr(lib, name)
            f.argtypes = [ctypes.c_long, ctypes.c_long, dp, dp]
            f.restype = None
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _lines(x, axis):
    y = np.ascontiguousarray(np.moveaxis(x, axis, -1))
    return y


def c2c(x: np.ndarray, axis: int, sign: int, direct: bool = False) -> np.ndarray:
    """Unscaled C2C along ``axis`` (sign -1 forward, +1 backward)."""
    lib = _load()
    y = _lines(np.asarray(x, dtype=np.complex128), axis)
    n = y.shape[-1]
    out = np.empty_like(y)
    if y.size:
        fn = lib.dft_ld_c2c if direct else lib.fft_c2c
        fn(n, y.size // n, _ptr(y), _ptr(out), int(sign))
    return np.moveaxis(out, -1, axis)


def r2c(x: np.ndarray, axis: int) -> np.ndarray:
    lib = _load()
    y = _lines(np.asarray(x, dtype=np.float64), axis)
    n = y.shape[-1]
    out = np.empty(y.shape[:-1] + (n // 2 + 1,), dtype=np.complex128)
    if y.size:
        lib.fft_r2c(n, y.size // n, _ptr(y), _ptr(out))
    return np.moveaxis(out, -1, axis)


def c2r(x: np.ndarray, n: int, axis: int) -> np.ndarray:
    lib = _load()
    y = _lines(np.asarray(x, dtype=np.complex128), axis)
    if y.shape[-1] != n // 2 + 1:
        raise ValueError("C2R input extent must be n//2+1")
    out = np.empty(y.shape[:-1] + (n,), dtype=np.float64)
    if y.size:
        lib.fft_c2r(n, y.size // y.shape[-1], _ptr(y), _ptr(out))
    return np.moveaxis(out, -1, axis)


def norm_factor(norm: str, direction: str, n_total: int) -> float:
    """KokkosFFT::Impl::normalize scaling (called at SlabPlan.hpp:1347-1357)."""
    if norm == "none":
        return 1.0
    if norm == "ortho":
        return 1.0 / np.sqrt(float(n_total))
    if norm == "backward":
        return 1.0 if direction == "forward" else 1.0 / n_total
    if norm == "forward":
        return 1.0 / n_total if direction == "forward" else 1.0
    raise ValueError(norm)


def fftn(x, axes, real_input=False, norm="backward"):
    """Global forward transform over ``axes`` (R2C on ``axes[-1]`` when the
    input is real), then normalisation -- the plan-level semantics."""
    axes = list(axes)
    if real_input:
        y = r2c(x, axes[-1])
    else:
        y = c2c(x, axes[-1], -1)
    for a in reversed(axes[:-1]):
        y = c2c(y, a, -1)
    n_total = int(np.prod([x.shape[a] for a in axes]))
    return y * norm_factor(norm, "forward", n_total)


def ifftn(x, axes, n_last=None, real_output=False, norm="backward"):
    """Global backward transform over ``axes`` (C2R on ``axes[-1]`` when the
    output is real; ``n_last`` is its real length)."""
    axes = list(axes)
    y = np.asarray(x, dtype=np.complex128)
    for a in axes[:-1]:
        y = c2c(y, a, +1)
    if real_output:
        y = c2r(y, n_last, axes[-1])
        lens = [x.shape[a] for a in axes[:-1]] + [n_last]
    else:
        y = c2c(y, axes[-1], +1)
        lens = [x.shape[a] for a in axes]
    n_total = int(np.prod(lens))
    return y * norm_factor(norm, "backward", n_total)
