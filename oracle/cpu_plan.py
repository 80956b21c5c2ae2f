"""Oracle (test infrastructure only): OpenMP CPU executor of the reference's
slab/pencil block sequence (``cpu_plan.c``), the timed CPU baseline of
bench.py.  Only the test-suite and bench.py's ``cpu_baseline`` leg use it.

``CpuPlan`` walks the same block list as ``plan_exec.OraclePlan`` (the
planner's ``block_analyses`` per rank), on preallocated per-rank buffers as
the reference's internal plans hold them (``SlabPlan.hpp:882-961``,
``PencilPlan.hpp:1082-1188``): FFT blocks transform along the innermost
``fft_dim`` axes of the mapped local arrays (out of place on the first axis
pass when the input is the caller's array), Transpose blocks run
pack -> all-to-all -> unpack inside each (sub-)communicator
(``TransBlock.hpp:76-107``), a final local transpose returns to the caller's
layout and the normalisation is a separate sweep (``SlabPlan.hpp:1344-1358``).
C2C, complex128, LayoutRight.  Pinned against ``OraclePlan`` in
tests/test_oracle.py.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

from . import fft as F
from . import plan_exec as PE
from . import planner as PL

_HERE = Path(__file__).resolve().parent
_lib = None


def _has_avx2():
    try:
        flags = Path("/proc/cpuinfo").read_text().split()
    except OSError:
        return False
    return "avx2" in flags and "fma" in flags


def load():
    global _lib
    if _lib is None:
        name = "liboracle_cpuplan_avx2.so" if _has_avx2() else "liboracle_cpuplan.so"
        path = _HERE / "_build" / name
        if not path.exists():
            F.build()
        lib = ctypes.CDLL(str(path))
        vpp = ctypes.POINTER(ctypes.c_void_p)
        lp = ctypes.POINTER(ctypes.c_long)
        ip = ctypes.POINTER(ctypes.c_int)
        lib.cpu_fft_axis.argtypes = [ctypes.c_int, vpp, vpp, lp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.cpu_fft_axis.restype = ctypes.c_int
        for fn in (lib.cpu_pack, lib.cpu_unpack):
            fn.argtypes = [ctypes.c_int, vpp, lp, ctypes.c_int, vpp, lp, ip, ctypes.c_int]
            fn.restype = None
        lib.cpu_all2all.argtypes = [ctypes.c_int, ctypes.c_int, ip, vpp, vpp, ctypes.c_long]
        lib.cpu_transpose.argtypes = [ctypes.c_int, vpp, vpp, lp, ctypes.c_int, ip]
        lib.cpu_scale.argtypes = [ctypes.c_int, vpp, lp, ctypes.c_double]
        lib.cpu_copy.argtypes = [ctypes.c_int, vpp, vpp, lp]
        lib.cpu_zero.argtypes = [ctypes.c_int, vpp, lp]
        lib.cpu_set_threads.argtypes = [ctypes.c_int]
        lib.cpu_get_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def set_threads(n: int):
    load().cpu_set_threads(int(n))


def get_threads() -> int:
    return int(load().cpu_get_threads())


def _ptrs(arrs):
    return (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])


def _longs(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_long * len(vals))(*vals)


def _ints(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int * len(vals))(*vals)


def fft_axis(arrs, axis, sign, outs=None):
    """Unscaled C2C along ``axis`` of every array of ``arrs`` (complex128, C order)."""
    outs = arrs if outs is None else outs
    d = arrs[0].ndim
    ext = [e for a in arrs for e in a.shape]
    rc = load().cpu_fft_axis(len(arrs), _ptrs(arrs), _ptrs(outs), _longs(ext), d, int(axis), int(sign))
    if rc != 0:
        raise ValueError("transform axis length differs between ranks")
    return outs


class CpuPlan:
    """C2C plan over P virtual ranks executing the reference's block list with
    the OpenMP kernels of ``cpu_plan.c``."""

    def __init__(self, gin, axes, in_topology, out_topology, norm="backward"):
        self.meta = PE.OraclePlan(gin, gin, axes, in_topology, out_topology, False, norm)
        m = self.meta
        if m.kind not in (PL.SLAB, PL.PENCIL):
            raise NotImplementedError("CpuPlan runs slab and pencil plans")
        self.P = m.P
        self.dim = m.dim
        n_total = int(np.prod(m.fft_extents))
        self.fwd_scale = F.norm_factor(norm, "forward", n_total)
        self.bwd_scale = F.norm_factor(norm, "backward", n_total)
        # per-rank work arrays (two), send and receive buffers, sized for the
        # largest mapped local array / (P, e...) buffer of the block list
        big = max(int(np.prod(e)) for e in m.in_local + m.out_local)
        for a in m.analyses:
            for b in a.blocks:
                big = max(big, int(np.prod(b.in_extents)), int(np.prod(b.out_extents)))
        buf = max(int(np.prod(b.buffer_extents)) for a in m.analyses for b in a.blocks
                  if b.block_type == PL.TRANSPOSE)
        self.work = [[np.empty(big, np.complex128) for _ in range(self.P)] for _ in range(2)]
        self.send = [np.empty(buf, np.complex128) for _ in range(self.P)]
        self.recv = [np.empty(buf, np.complex128) for _ in range(self.P)]
        lib = load()
        for group in self.work + [self.send, self.recv]:
            lib.cpu_zero(self.P, _ptrs(group), _longs([a.size for a in group]))

    # ------------------------------------------------------------- helpers
    def _views(self, wi, shapes):
        return [self.work[wi][r][:int(np.prod(s))].reshape(s) for r, s in enumerate(shapes)]

    def _transpose(self, st, perm, wi):
        shapes = [[a.shape[p] for p in perm] for a in st]
        dst = self._views(wi, shapes)
        load().cpu_transpose(self.P, _ptrs(st), _ptrs(dst), _longs([e for a in st for e in a.shape]),
                             self.dim, _ints(perm))
        return dst

    def _fft_block(self, st, st_wi, fft_dim, sign):
        """Transforms along the innermost fft_dim axes (innermost first forward,
        outermost first backward); returns (arrays, their work index)."""
        inner = list(range(self.dim - fft_dim, self.dim))
        order = list(reversed(inner)) if sign < 0 else inner
        for ax in order:
            if st_wi is None:  # the caller's array: first pass out of place
                dst = self._views(0, [a.shape for a in st])
                fft_axis(st, ax, sign, dst)
                st, st_wi = dst, 0
            else:
                fft_axis(st, ax, sign)
        return st, st_wi

    def _exchange(self, st, st_wi, bi, forward):
        lib = load()
        m = self.meta
        blocks = [a.blocks[bi] for a in m.analyses]
        b0 = blocks[0]
        groups = m.groups(b0.comm_axis)
        gsize = len(groups[0])
        bext = list(b0.buffer_extents)
        if any(list(b.buffer_extents) != bext for b in blocks) or bext[0] != gsize:
            raise RuntimeError("buffer extents differ between ranks or from the communicator size")
        smaps = [(b.in_map if forward else b.out_map) for b in blocks]
        sax = b0.in_axis if forward else b0.out_axis
        dmaps = [(b.out_map if forward else b.in_map) for b in blocks]
        dax = b0.out_axis if forward else b0.in_axis
        dshapes = [list(b.out_extents if forward else b.in_extents) for b in blocks]
        nb = int(np.prod(bext))
        sends = [s[:nb] for s in self.send]
        recvs = [s[:nb] for s in self.recv]
        lib.cpu_pack(self.P, _ptrs(st), _longs([e for a in st for e in a.shape]), self.dim, _ptrs(sends),
                     _longs(bext), _ints([v for mp in smaps for v in mp]), int(sax))
        lib.cpu_all2all(len(groups), gsize, _ints([r for g in groups for r in g]), _ptrs(sends), _ptrs(recvs),
                        nb // gsize)
        wi = 1 if st_wi == 0 else 0
        dst = self._views(wi, dshapes)
        lib.cpu_unpack(self.P, _ptrs(dst), _longs([e for s in dshapes for e in s]), self.dim, _ptrs(recvs),
                       _longs(bext), _ints([v for mp in dmaps for v in mp]), int(dax))
        return dst, wi

    def _finish(self, st, cur, outs, scale):
        lib = load()
        ident = PL.identity(self.dim)
        if cur != ident:
            perm = PE._perm_between(cur, ident)
            lib.cpu_transpose(self.P, _ptrs(st), _ptrs(outs), _longs([e for a in st for e in a.shape]),
                              self.dim, _ints(perm))
        else:
            lib.cpu_copy(self.P, _ptrs(st), _ptrs(outs), _longs([a.size for a in st]))
        if scale != 1.0:
            lib.cpu_scale(self.P, _ptrs(outs), _longs([a.size for a in outs]), scale)

    # ------------------------------------------------------------- execute
    def forward(self, ins, outs):
        m = self.meta
        ident = PL.identity(self.dim)
        st, st_wi, cur = list(ins), None, ident
        for bi, b0 in enumerate(m.blocks):
            if b0.block_type == PL.FFT:
                if cur != b0.in_map:
                    wi = 1 if st_wi == 0 else 0
                    st, st_wi = self._transpose(st, PE._perm_between(cur, b0.in_map), wi), wi
                st, st_wi = self._fft_block(st, st_wi, b0.fft_dim, -1)
                cur = b0.in_map
            else:
                st, st_wi = self._exchange(st, st_wi, bi, True)
                cur = b0.out_map
        self._finish(st, cur, outs, self.fwd_scale)
        return outs

    def backward(self, outs, ins):
        m = self.meta
        ident = PL.identity(self.dim)
        st, st_wi, cur = list(outs), None, ident
        for bi in range(len(m.blocks) - 1, -1, -1):
            b0 = m.blocks[bi]
            if b0.block_type == PL.FFT:
                if cur != b0.out_map:
                    wi = 1 if st_wi == 0 else 0
                    st, st_wi = self._transpose(st, PE._perm_between(cur, b0.out_map), wi), wi
                st, st_wi = self._fft_block(st, st_wi, b0.fft_dim, +1)
                cur = b0.in_map
            else:
                st, st_wi = self._exchange(st, st_wi, bi, False)
                cur = b0.in_map
        self._finish(st, cur, ins, self.bwd_scale)
        return ins


def cores_available() -> int:
    return len(os.sched_getaffinity(0))
