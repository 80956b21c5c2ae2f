"""Python host mirror of ``KokkosFFT::Distributed`` over the C ABI of libdfft.so.

The reference's API (``KokkosFFT_Distributed_Plan.hpp:124-238``)::

    Plan plan(exec, in, out, axes, in_topology, out_topology, comm, norm);
    execute(plan, in, out, KokkosFFT::Direction::forward);
    plan.name();

becomes, for P virtual ranks held in one process on one GPU::

    plan = Plan(in_views, out_views, axes, in_topology, out_topology, norm="backward")
    execute(plan, in_views, out_views, Direction.forward)
    plan.name()

where ``in_views`` / ``out_views`` are lists of P device tensors (one local
array per virtual rank, shaped like the reference's per-rank Kokkos::View).
``Topology`` carries the LayoutRight/LayoutLeft tag of the reference's
``Topology<T, N, Layout>`` (``KokkosFFT_Distributed_Types.hpp:174-428``).

Device memory and streams come from torch; every transform runs in the HIP
kernels of libdfft.so -- there is no fallback path.
"""

from __future__ import annotations

import ctypes
import enum
import os
from pathlib import Path
from typing import List, Sequence

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libdfft.so"

SUCCESS = 0
ERROR_INVALID_TOPOLOGY = 1
ERROR_INVALID_EXTENTS = 2
ERROR_MISMATCH = 3
ERROR_ALIAS = 4
ERROR_UNSUPPORTED = 5
ERROR_HIP = 6
ERROR_INVALID_ARGUMENT = 7
ERROR_OUT_OF_MEMORY = 8

MAX_RANK = 4
MAX_VRANKS = 64

C2C, R2C = 0, 1
FP32, FP64 = 0, 1
LAYOUT_RIGHT, LAYOUT_LEFT = 0, 1
NORMS = {"backward": 0, "forward": 1, "ortho": 2, "none": 3}
EXEC_FUSED, EXEC_BLOCKWALK, EXEC_VENDOR = 0, 1, 2


class Direction(enum.IntEnum):
    forward = 0
    backward = 1


class DFFTError(RuntimeError):
    """Raised for every non-zero status; ``code`` is the C ABI status."""

    def __init__(self, code, msg):
        super().__init__(f"[dfft status {code}] {msg}")
        self.code = code


class LayoutRight:  # tags mirroring Kokkos::LayoutRight / LayoutLeft
    value = LAYOUT_RIGHT


class LayoutLeft:
    value = LAYOUT_LEFT


class Topology(tuple):
    """``Topology<std::size_t, N, Layout>``: a tuple with a layout tag."""

    def __new__(cls, values, layout=LayoutRight):
        obj = super().__new__(cls, [int(v) for v in values])
        obj.layout = layout
        return obj

    def array(self):
        return tuple(self)


def _as_topology(t):
    return t if isinstance(t, Topology) else Topology(t)


class _Desc(ctypes.Structure):
    _fields_ = [
        ("rank", ctypes.c_int),
        ("fft_rank", ctypes.c_int),
        ("axes", ctypes.c_int * 3),
        ("transform", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("layout", ctypes.c_int),
        ("in_extents", ctypes.c_size_t * MAX_RANK),
        ("out_extents", ctypes.c_size_t * MAX_RANK),
        ("in_topology", ctypes.c_size_t * MAX_RANK),
        ("out_topology", ctypes.c_size_t * MAX_RANK),
        ("in_topology_layout", ctypes.c_int),
        ("out_topology_layout", ctypes.c_int),
        ("norm", ctypes.c_int),
        ("nranks", ctypes.c_int),
        ("exec_mode", ctypes.c_int),
    ]


class BlockInfo(ctypes.Structure):
    """POD mirror of ``BlockInfo<DIM>`` (Types.hpp:61-154)."""

    _fields_ = [
        ("in_topology", ctypes.c_size_t * MAX_RANK),
        ("out_topology", ctypes.c_size_t * MAX_RANK),
        ("in_extents", ctypes.c_size_t * MAX_RANK),
        ("out_extents", ctypes.c_size_t * MAX_RANK),
        ("buffer_extents", ctypes.c_size_t * (MAX_RANK + 1)),
        ("in_map", ctypes.c_size_t * MAX_RANK),
        ("out_map", ctypes.c_size_t * MAX_RANK),
        ("in_axis", ctypes.c_size_t),
        ("out_axis", ctypes.c_size_t),
        ("fft_dim", ctypes.c_size_t),
        ("comm_axis", ctypes.c_size_t),
        ("block_type", ctypes.c_int),
        ("block_idx", ctypes.c_size_t),
    ]

    def as_dict(self, rank):
        def v(a, n=rank):
            return [int(a[i]) for i in range(n)]
        return {
            "in_topology": v(self.in_topology), "out_topology": v(self.out_topology),
            "in_extents": v(self.in_extents), "out_extents": v(self.out_extents),
            "buffer_extents": v(self.buffer_extents, rank + 1), "in_map": v(self.in_map),
            "out_map": v(self.out_map), "in_axis": int(self.in_axis), "out_axis": int(self.out_axis),
            "fft_dim": int(self.fft_dim), "comm_axis": int(self.comm_axis),
            "block_type": "FFT" if self.block_type == 1 else "Transpose",
            "block_idx": int(self.block_idx),
        }


class NsDesc(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("px", ctypes.c_int),
        ("py", ctypes.c_int),
        ("lx", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("form", ctypes.c_int),
    ]


EXPORTED = [
    "dfft_last_error", "dfft_version", "dfft_plan_create", "dfft_plan_destroy", "dfft_plan_name",
    "dfft_plan_local_extents", "dfft_plan_num_blocks", "dfft_plan_block_info",
    "dfft_plan_max_buffer_size", "dfft_plan_workspace_bytes", "dfft_execute", "dfft_execute_adjoint",
    "dfft_plan_num_passes",
    "dfft_execute_timed", "dfft_plan_set_option", "dfft_plan_get_option", "dfft_plan_autotune", "dfft_pack",
    "dfft_unpack", "dfft_all2all", "dfft_exchange", "dfft_get_all_slab_topologies", "dfft_rank_to_coord",
    "dfft_compute_local_extents_and_starts",
    "dfft_get_all_pencil_topologies", "dfft_block_analyses", "dfft_group_create", "dfft_group_destroy",
    "dfft_group_execute", "dfft_group_execute_noalloc", "dfft_group_num_launches", "dfft_group_set_option",
    "dfft_group_get_option",
    "dfft_group_autotune", "dfft_ns_create", "dfft_ns_destroy",
    "dfft_ns_local_extents", "dfft_ns_init_taylor_green", "dfft_ns_set_physical", "dfft_ns_set_state",
    "dfft_ns_get_state", "dfft_ns_step", "dfft_ns_autotune", "dfft_ns_set_option", "dfft_ns_get_option", "dfft_ns_energy", "dfft_ns_velocity",
    "dfft_ns_vorticity",
    "dfft_ns_energy_spectrum", "dfft_ns_diag", "dfft_ns_rhs", "dfft_ns_rhs_tangent",
    "dfft_ns_rhs_adjoint", "dfft_ns_gradient",
]

_lib = None


def load_library(path: os.PathLike | None = None):
    """Load libdfft.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(str(p))
    sz = ctypes.c_size_t
    szp = ctypes.POINTER(sz)
    vpp = ctypes.POINTER(ctypes.c_void_p)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "dfft_last_error": ([], ctypes.c_char_p),
        "dfft_version": ([ip, ip], ctypes.c_int),
        "dfft_plan_create": ([ctypes.POINTER(_Desc), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "dfft_plan_destroy": ([ctypes.c_void_p], ctypes.c_int),
        "dfft_plan_name": ([ctypes.c_void_p, ctypes.c_char_p, sz], ctypes.c_int),
        "dfft_plan_local_extents": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, szp, szp], ctypes.c_int),
        "dfft_plan_num_blocks": ([ctypes.c_void_p, ip], ctypes.c_int),
        "dfft_plan_block_info": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(BlockInfo)],
                                 ctypes.c_int),
        "dfft_plan_max_buffer_size": ([ctypes.c_void_p, szp], ctypes.c_int),
        "dfft_plan_workspace_bytes": ([ctypes.c_void_p, szp], ctypes.c_int),
        "dfft_execute": ([ctypes.c_void_p, vpp, vpp, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_execute_adjoint": ([ctypes.c_void_p, vpp, vpp, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_plan_num_passes": ([ctypes.c_void_p, ctypes.c_int, ip], ctypes.c_int),
        "dfft_execute_timed": ([ctypes.c_void_p, vpp, vpp, ctypes.c_int, ctypes.c_void_p,
                                ctypes.POINTER(ctypes.c_float), ctypes.c_int, ip], ctypes.c_int),
        "dfft_plan_set_option": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_long], ctypes.c_int),
        "dfft_plan_get_option": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "dfft_plan_autotune": ([ctypes.c_void_p, vpp, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_pack": ([ctypes.c_void_p, szp, szp, ctypes.c_int, ctypes.c_void_p, szp, ctypes.c_int,
                       ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_unpack": ([ctypes.c_void_p, szp, ctypes.c_int, ctypes.c_void_p, szp, szp, ctypes.c_int,
                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_all2all": ([vpp, vpp, ctypes.c_int, sz, ctypes.c_void_p], ctypes.c_int),
        "dfft_exchange": ([vpp, vpp, ctypes.c_int, szp, szp, ctypes.c_int, szp, szp, ctypes.c_int, szp,
                           ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_get_all_slab_topologies": ([ctypes.c_int, szp, szp, ctypes.c_int, ip, szp, ip], ctypes.c_int),
        "dfft_get_all_pencil_topologies": ([ctypes.c_int, szp, ctypes.c_int, szp, ctypes.c_int, ctypes.c_int,
                                            ip, szp, szp, szp, ip], ctypes.c_int),
        "dfft_rank_to_coord": ([ctypes.c_int, szp, ctypes.c_int, ctypes.c_int, szp], ctypes.c_int),
        "dfft_compute_local_extents_and_starts": ([ctypes.c_int, szp, szp, ctypes.c_int, ctypes.c_int, szp, szp],
                                                  ctypes.c_int),
        "dfft_block_analyses": ([ctypes.POINTER(_Desc), ctypes.c_int, ctypes.POINTER(BlockInfo), ip, szp],
                                ctypes.c_int),
        "dfft_group_create": ([ctypes.c_int, ctypes.POINTER(_Desc), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "dfft_group_destroy": ([ctypes.c_void_p], ctypes.c_int),
        "dfft_group_execute": ([ctypes.c_void_p, vpp, vpp, ctypes.c_int, ctypes.c_int, ctypes.c_void_p],
                               ctypes.c_int),
        "dfft_group_execute_noalloc": ([ctypes.c_void_p, vpp, vpp, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_group_num_launches": ([ctypes.c_void_p, ctypes.c_int, ip], ctypes.c_int),
        "dfft_group_set_option": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_long], ctypes.c_int),
        "dfft_group_get_option": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "dfft_group_autotune": ([ctypes.c_void_p, vpp, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_create": ([ctypes.POINTER(NsDesc), ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "dfft_ns_destroy": ([ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_local_extents": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, szp, szp], ctypes.c_int),
        "dfft_ns_init_taylor_green": ([ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_set_physical": ([ctypes.c_void_p, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_set_state": ([ctypes.c_void_p, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_get_state": ([ctypes.c_void_p, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_step": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_autotune": ([ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_set_option": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_long], ctypes.c_int),
        "dfft_ns_get_option": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_long)], ctypes.c_int),
        "dfft_ns_energy": ([ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_velocity": ([ctypes.c_void_p, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_vorticity": ([ctypes.c_void_p, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_energy_spectrum": ([ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                     ctypes.POINTER(ctypes.c_double), ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_diag": ([ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                          ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_rhs": ([ctypes.c_void_p, vpp, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_rhs_tangent": ([ctypes.c_void_p, vpp, vpp, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_rhs_adjoint": ([ctypes.c_void_p, vpp, vpp, vpp, ctypes.c_void_p], ctypes.c_int),
        "dfft_ns_gradient": ([ctypes.c_void_p, ctypes.c_int, ctypes.c_int, vpp, vpp,
                              ctypes.POINTER(ctypes.c_double), ctypes.c_void_p], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def _check(status):
    if status != SUCCESS:
        msg = _lib.dfft_last_error().decode() if _lib is not None else ""
        raise DFFTError(status, msg)


def _sizes(values, n=None):
    vals = list(values)
    if n is not None:
        vals = vals + [0] * (n - len(vals))
    return (ctypes.c_size_t * len(vals))(*[int(v) for v in vals])


def _ints(values):
    vals = list(values)
    return (ctypes.c_int * len(vals))(*[int(v) for v in vals])


def _ptrs(tensors):
    arr = (ctypes.c_void_p * len(tensors))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr() if (t is not None and t.numel() > 0) else None
    return arr


def _stream_handle(stream):
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


# ---------------------------------------------------------------------------
# Host-only planner helpers (no device needed)
# ---------------------------------------------------------------------------

def get_all_slab_topologies(in_topology, out_topology, axes):
    lib = load_library()
    rank = len(in_topology)
    cap = ctypes.c_int(16)
    buf = (ctypes.c_size_t * (16 * rank))()
    _check(lib.dfft_get_all_slab_topologies(rank, _sizes(in_topology), _sizes(out_topology), len(axes),
                                            _ints(axes), buf, ctypes.byref(cap)))
    return [[int(buf[i * rank + k]) for k in range(rank)] for i in range(cap.value)]


def get_all_pencil_topologies(in_topology, out_topology, axes):
    lib = load_library()
    tin, tout = _as_topology(in_topology), _as_topology(out_topology)
    rank = len(tin)
    cap = ctypes.c_int(16)
    buf = (ctypes.c_size_t * (16 * rank))()
    tax = (ctypes.c_size_t * 16)()
    lay = (ctypes.c_size_t * 16)()
    _check(lib.dfft_get_all_pencil_topologies(rank, _sizes(tin), tin.layout.value, _sizes(tout),
                                              tout.layout.value, len(axes), _ints(axes), buf, tax, lay,
                                              ctypes.byref(cap)))
    n = cap.value
    topos = [[int(buf[i * rank + k]) for k in range(rank)] for i in range(n)]
    return topos, [int(tax[i]) for i in range(n - 1)], [int(lay[i]) for i in range(n)]


def _make_desc(gin, gout, axes, in_topology, out_topology, real, fp64, norm, exec_mode, layout=LAYOUT_RIGHT):
    tin, tout = _as_topology(in_topology), _as_topology(out_topology)
    d = _Desc()
    d.rank = len(gin)
    d.fft_rank = len(axes)
    for i, a in enumerate(axes):
        d.axes[i] = int(a)
    d.transform = R2C if real else C2C
    d.precision = FP64 if fp64 else FP32
    d.layout = layout
    for i in range(len(gin)):
        d.in_extents[i] = int(gin[i])
        d.out_extents[i] = int(gout[i])
        d.in_topology[i] = int(tin[i])
        d.out_topology[i] = int(tout[i])
    d.in_topology_layout = tin.layout.value
    d.out_topology_layout = tout.layout.value
    d.norm = NORMS[norm] if isinstance(norm, str) else int(norm)
    n = 1
    for t in tin:
        n *= int(t)
    d.nranks = n
    d.exec_mode = exec_mode
    return d


def block_analyses(gin, gout, axes, in_topology, out_topology, rank, real=False, layout=LayoutRight):
    """Block list of one virtual rank (BlockAnalyses.hpp:415-506) for views of ``layout``."""
    lib = load_library()
    d = _make_desc(gin, gout, axes, in_topology, out_topology, real, True, "backward", EXEC_FUSED,
                   _layout_value(layout))
    infos = (BlockInfo * 16)()
    nb = ctypes.c_int(16)
    mx = ctypes.c_size_t(0)
    _check(lib.dfft_block_analyses(ctypes.byref(d), int(rank), infos, ctypes.byref(nb), ctypes.byref(mx)))
    return [infos[i].as_dict(len(gin)) for i in range(nb.value)], int(mx.value)


def rank_to_coord(topology, rank):
    """MPI_Extents.hpp:492-529 (``dfft_rank_to_coord``)."""
    t = _as_topology(topology)
    lib = load_library()
    c = (ctypes.c_size_t * len(t))()
    _check(lib.dfft_rank_to_coord(len(t), _sizes(t), t.layout.value, int(rank), c))
    return [int(v) for v in c]


def compute_local_extents_and_starts(extents, topology, rank):
    """MPI_Extents.hpp:546-577 for virtual rank ``rank``
    (``dfft_compute_local_extents_and_starts``)."""
    t = _as_topology(topology)
    if len(extents) != len(t):
        raise DFFTError(ERROR_INVALID_ARGUMENT, "extents and topology must have the same rank")
    lib = load_library()
    e = (ctypes.c_size_t * len(t))()
    st = (ctypes.c_size_t * len(t))()
    _check(lib.dfft_compute_local_extents_and_starts(len(t), _sizes(extents), _sizes(t), t.layout.value, int(rank),
                                                     e, st))
    return [int(v) for v in e], [int(v) for v in st]


def _prod(v):
    out = 1
    for x in v:
        out *= int(x)
    return out


# ---------------------------------------------------------------------------
# Plan / execute
# ---------------------------------------------------------------------------

class Plan:
    """``KokkosFFT::Distributed::Plan`` for P virtual ranks on one GPU."""

    def __init__(self, in_views: Sequence, out_views: Sequence, axes, in_topology, out_topology,
                 norm="backward", exec_mode=EXEC_FUSED, global_in_extents=None, global_out_extents=None,
                 layout=LayoutRight):
        """``layout``: LayoutRight (row-major local arrays, the default) or
        LayoutLeft (column-major: a view of extents (e0, e1, e2) is a tensor
        with strides (1, e0, e0*e1), e.g. ``torch.empty(e[::-1]).permute(2, 1, 0)``)
        -- the reference's view layout template parameter."""
        import torch
        lib = load_library()
        self._lib = lib
        self._plan = ctypes.c_void_p()
        tin, tout = _as_topology(in_topology), _as_topology(out_topology)
        self.in_topology, self.out_topology = tin, tout
        self.nranks = _prod(tin)
        if len(in_views) != self.nranks or len(out_views) != self.nranks:
            raise DFFTError(ERROR_INVALID_TOPOLOGY, "topology size must be identical to the number of views")
        in_dtype, out_dtype = in_views[0].dtype, out_views[0].dtype
        if not out_dtype.is_complex:
            raise DFFTError(ERROR_INVALID_ARGUMENT, "Plan: the output type must be complex")
        real = not in_dtype.is_complex
        fp64 = out_dtype == torch.complex128
        want_in = (torch.float64 if fp64 else torch.float32) if real else out_dtype
        if in_dtype != want_in:
            raise DFFTError(ERROR_MISMATCH, f"Plan: input type {in_dtype} does not match output type {out_dtype}")
        self.in_dtype, self.out_dtype = in_dtype, out_dtype
        self.device = _device_of(list(in_views) + list(out_views))
        self.real, self.fp64 = real, fp64
        rank = in_views[0].dim()
        self.rank = rank
        gin = global_in_extents or _global_extents([list(v.shape) for v in in_views], tin)
        gout = global_out_extents or _global_extents([list(v.shape) for v in out_views], tout)
        self.global_in_extents, self.global_out_extents = list(gin), list(gout)
        self.axes = [int(a) for a in axes]
        self.layout = layout.value if hasattr(layout, "value") else int(layout)
        d = _make_desc(gin, gout, self.axes, tin, tout, real, fp64, norm, exec_mode, self.layout)
        _check(lib.dfft_plan_create(ctypes.byref(d), ctypes.byref(self._plan)))
        self.in_extents = [list(v.shape) for v in in_views]
        self.out_extents = [list(v.shape) for v in out_views]
        for r in range(self.nranks):
            e_in, _ = self.local_extents(r, 0)
            e_out, _ = self.local_extents(r, 1)
            if e_in != self.in_extents[r] or e_out != self.out_extents[r]:
                self.destroy()
                raise DFFTError(ERROR_INVALID_EXTENTS,
                                f"local extents of rank {r} do not match the topology: "
                                f"in {self.in_extents[r]} vs {e_in}, out {self.out_extents[r]} vs {e_out}")

    def name(self) -> str:
        buf = ctypes.create_string_buffer(64)
        _check(self._lib.dfft_plan_name(self._plan, buf, 64))
        return buf.value.decode()

    def local_extents(self, vrank, which):
        e = (ctypes.c_size_t * MAX_RANK)()
        s = (ctypes.c_size_t * MAX_RANK)()
        _check(self._lib.dfft_plan_local_extents(self._plan, int(vrank), int(which), e, s))
        return [int(e[i]) for i in range(self.rank)], [int(s[i]) for i in range(self.rank)]

    def blocks(self, vrank=0):
        nb = ctypes.c_int()
        _check(self._lib.dfft_plan_num_blocks(self._plan, ctypes.byref(nb)))
        out = []
        for i in range(nb.value):
            info = BlockInfo()
            _check(self._lib.dfft_plan_block_info(self._plan, int(vrank), i, ctypes.byref(info)))
            out.append(info.as_dict(self.rank))
        return out

    def max_buffer_size(self):
        v = ctypes.c_size_t()
        _check(self._lib.dfft_plan_max_buffer_size(self._plan, ctypes.byref(v)))
        return int(v.value)

    def workspace_bytes(self):
        v = ctypes.c_size_t()
        _check(self._lib.dfft_plan_workspace_bytes(self._plan, ctypes.byref(v)))
        return int(v.value)

    def _execute(self, ins, outs, direction, stream=None):
        _check(self._lib.dfft_execute(self._plan, _ptrs(ins), _ptrs(outs), int(direction),
                                      _stream_handle(stream)))

    def check_views(self, ins, outs, direction):
        """The reference's ``good()`` (InternalPlan.hpp:109-142) for the arrays
        that ``execute(direction)`` reads (``ins``) and writes (``outs``):
        one per rank, with the plan's extents, element types, device and
        contiguity.  Raises DFFTError(ERROR_MISMATCH) otherwise."""
        d = Direction(direction)
        fwd = d == Direction.forward
        shapes_in, shapes_out = (self.in_extents, self.out_extents) if fwd else (self.out_extents, self.in_extents)
        dt_in, dt_out = (self.in_dtype, self.out_dtype) if fwd else (self.out_dtype, self.in_dtype)
        if len(ins) != self.nranks or len(outs) != self.nranks:
            raise DFFTError(ERROR_MISMATCH, f"expected {self.nranks} input and output views, "
                            f"got {len(ins)} and {len(outs)}")
        for r in range(self.nranks):
            for v, shape, dt, side in ((ins[r], shapes_in[r], dt_in, "input"), (outs[r], shapes_out[r], dt_out, "output")):
                if list(v.shape) != list(shape):
                    raise DFFTError(ERROR_MISMATCH, f"extents of {side} View for plan and execution are not "
                                    f"identical: View {list(v.shape)}, Plan {list(shape)}")
                if v.dtype != dt:
                    raise DFFTError(ERROR_MISMATCH, f"{side} View of rank {r} has type {v.dtype}, the plan "
                                    f"needs {dt}")
                if v.numel():
                    if not v.is_cuda or v.device.index != self.device:
                        raise DFFTError(ERROR_MISMATCH, f"{side} View of rank {r} is on {v.device}, the plan "
                                        f"runs on cuda:{self.device}")
                    if not _is_layout_contiguous(v, self.layout):
                        raise DFFTError(ERROR_INVALID_ARGUMENT, f"views must be contiguous in the plan's layout "
                                        f"({'LayoutLeft' if self.layout == LAYOUT_LEFT else 'LayoutRight'})")

    def execute_adjoint(self, ins, outs, direction, stream=None):
        """Adjoint of ``execute(direction)``: ``ins`` are shaped like that direction's outputs."""
        d = Direction(direction)
        self.check_views(ins, outs, Direction.backward if d == Direction.forward else Direction.forward)
        _check(self._lib.dfft_execute_adjoint(self._plan, _ptrs(ins), _ptrs(outs), int(direction),
                                              _stream_handle(stream)))

    def num_passes(self, direction):
        v = ctypes.c_int()
        _check(self._lib.dfft_plan_num_passes(self._plan, int(direction), ctypes.byref(v)))
        return int(v.value)

    def execute_timed(self, ins, outs, direction, stream=None):
        """Execute with HIP events around every launch; returns ms per launch."""
        self.check_views(ins, outs, direction)
        cap = 64
        buf = (ctypes.c_float * cap)()
        n = ctypes.c_int()
        _check(self._lib.dfft_execute_timed(self._plan, _ptrs(ins), _ptrs(outs), int(direction),
                                            _stream_handle(stream), buf, cap, ctypes.byref(n)))
        return [float(buf[i]) for i in range(min(n.value, cap))]

    def set_option(self, key, value):
        _check(self._lib.dfft_plan_set_option(self._plan, key.encode(), int(value)))

    def get_option(self, key):
        v = ctypes.c_long()
        _check(self._lib.dfft_plan_get_option(self._plan, key.encode(), ctypes.byref(v)))
        return int(v.value)

    def options(self):
        """Current launch shapes of every pass and the chunk sizes."""
        out = {"fwd.chunk": self.get_option("fwd.chunk"), "bwd.chunk": self.get_option("bwd.chunk")}
        for d, dirn in (("fwd", Direction.forward), ("bwd", Direction.backward)):
            for i in range(max(0, self.num_passes(dirn))):
                for f in ("lpb", "vpt", "lf", "nt", "pair", "xcd", "swz", "split", "dual", "lin"):
                    out[f"{d}.{i}.{f}"] = self.get_option(f"{d}.{i}.{f}")
        return out

    def autotune(self, in_views, out_views, stream=None):
        """Measure launch shapes on these arrays (their contents are overwritten)."""
        self.check_views(in_views, out_views, Direction.forward)
        _check(self._lib.dfft_plan_autotune(self._plan, _ptrs(in_views), _ptrs(out_views),
                                            _stream_handle(stream)))

    def destroy(self):
        if getattr(self, "_plan", None) and self._plan.value:
            self._lib.dfft_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Group:
    """Independent single-GPU transforms of possibly different extents run
    together: each line pass of all members is one grouped kernel launch.
    ``ins[i]`` / ``outs[i]`` are member i's arrays; ``axes`` is one axes list
    for all members or one per member (default: every axis)."""

    def __init__(self, ins: Sequence, outs: Sequence, axes=None, norm="backward"):
        import torch
        lib = load_library()
        self._lib = lib
        self._h = ctypes.c_void_p()
        if len(ins) != len(outs) or not ins:
            raise DFFTError(ERROR_INVALID_ARGUMENT, "need one output per input")
        n = len(ins)
        if axes is None:
            axes = [list(range(t.dim())) for t in ins]
        elif not isinstance(axes[0], (list, tuple)):
            axes = [list(axes)] * n
        descs = (_Desc * n)()
        self.in_shapes = [list(t.shape) for t in ins]
        self.out_shapes = [list(t.shape) for t in outs]
        self.in_dtypes = [t.dtype for t in ins]
        self.out_dtypes = [t.dtype for t in outs]
        self.device = _device_of(list(ins) + list(outs))
        for i, (a, b) in enumerate(zip(ins, outs)):
            if not b.dtype.is_complex:
                raise DFFTError(ERROR_INVALID_ARGUMENT, "Group: the output type must be complex")
            real = not a.dtype.is_complex
            fp64 = b.dtype == torch.complex128
            want_in = (torch.float64 if fp64 else torch.float32) if real else b.dtype
            if a.dtype != want_in:
                raise DFFTError(ERROR_MISMATCH, f"Group member {i}: input type {a.dtype} does not match "
                                f"output type {b.dtype}")
            ones = [1] * a.dim()
            descs[i] = _make_desc(list(a.shape), list(b.shape), axes[i], ones, ones, real, fp64, norm, EXEC_FUSED)
        _check(lib.dfft_group_create(n, descs, ctypes.byref(self._h)))
        self.count = n

    def check_views(self, ins, outs, direction):
        """Member arrays of ``execute(direction)``: count, extents, types,
        device and contiguity as at creation (raises ERROR_MISMATCH)."""
        fwd = Direction(direction) == Direction.forward
        if len(ins) != self.count or len(outs) != self.count:
            raise DFFTError(ERROR_MISMATCH, f"expected {self.count} member arrays per side")
        for i in range(self.count):
            si, so = (self.in_shapes[i], self.out_shapes[i]) if fwd else (self.out_shapes[i], self.in_shapes[i])
            di, do = (self.in_dtypes[i], self.out_dtypes[i]) if fwd else (self.out_dtypes[i], self.in_dtypes[i])
            for v, shape, dt, side in ((ins[i], si, di, "input"), (outs[i], so, do, "output")):
                if list(v.shape) != shape or v.dtype != dt:
                    raise DFFTError(ERROR_MISMATCH, f"Group member {i} {side}: {list(v.shape)} {v.dtype}, "
                                    f"the group needs {shape} {dt}")
                if v.numel() and (not v.is_cuda or v.device.index != self.device or not v.is_contiguous()):
                    raise DFFTError(ERROR_MISMATCH, f"Group member {i} {side} must be a contiguous tensor on "
                                    f"cuda:{self.device}")

    def execute(self, ins, outs, direction, use_graph=False, stream=None):
        self.check_views(ins, outs, direction)
        _check(self._lib.dfft_group_execute(self._h, _ptrs(ins), _ptrs(outs), int(direction), int(bool(use_graph)),
                                            _stream_handle(stream)))

    def execute_noalloc(self, ins, outs, direction, stream=None):
        """``execute`` without device allocation: the pass tables go into the
        library's static table area (4 MiB per GPU, shared by every call on
        the device; overlapping calls race)."""
        self.check_views(ins, outs, direction)
        _check(self._lib.dfft_group_execute_noalloc(self._h, _ptrs(ins), _ptrs(outs), int(direction),
                                                    _stream_handle(stream)))

    def num_launches(self, direction):
        v = ctypes.c_int()
        _check(self._lib.dfft_group_num_launches(self._h, int(direction), ctypes.byref(v)))
        return int(v.value)

    def set_option(self, key, value):
        _check(self._lib.dfft_group_set_option(self._h, key.encode(), int(value)))

    def get_option(self, key):
        v = ctypes.c_long()
        _check(self._lib.dfft_group_get_option(self._h, key.encode(), ctypes.byref(v)))
        return int(v.value)

    def autotune(self, ins, outs, stream=None):
        """Time LDS caps and grid sizes on these arrays (their contents are overwritten)."""
        self.check_views(ins, outs, Direction.forward)
        _check(self._lib.dfft_group_autotune(self._h, _ptrs(ins), _ptrs(outs), _stream_handle(stream)))

    def destroy(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.dfft_group_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _device_of(tensors):
    """CUDA device index shared by the non-empty tensors (raises otherwise)."""
    dev = None
    for t in tensors:
        if t is None or t.numel() == 0:
            continue
        if not t.is_cuda:
            raise DFFTError(ERROR_MISMATCH, f"views must be device tensors, got one on {t.device}")
        if dev is None:
            dev = t.device.index
        elif t.device.index != dev:
            raise DFFTError(ERROR_MISMATCH, f"views on cuda:{dev} and {t.device}")
    if dev is None:
        import torch
        dev = torch.cuda.current_device()
    return dev


def _is_layout_contiguous(t, layout):
    """True if ``t`` is dense in LayoutRight (C) or LayoutLeft (Fortran) order."""
    if layout == LAYOUT_RIGHT:
        return t.is_contiguous()
    expect = 1
    for size, stride in zip(t.shape, t.stride()):
        if size != 1 and stride != expect:
            return False
        expect *= size
    return True


def _global_extents(local_shapes, topology):
    """compute_global_extents (MPI_Extents.hpp:114-129) over the virtual ranks."""
    t = _as_topology(topology)
    rank = len(t)
    out = []
    for a in range(rank):
        if t[a] == 1:
            out.append(int(local_shapes[0][a]))
            continue
        total = 0
        for c in range(t[a]):
            # the first rank with coordinate c along a and zero elsewhere
            coord = [0] * rank
            coord[a] = c
            r = _coord_to_rank(coord, t)
            total += int(local_shapes[r][a])
        out.append(total)
    return out


def _coord_to_rank(coord, t):
    r = 0
    if t.layout is LayoutRight:
        for i in range(len(t)):
            r = r * t[i] + coord[i]
    else:
        for i in range(len(t) - 1, -1, -1):
            r = r * t[i] + coord[i]
    return r


def execute(plan: Plan, in_views, out_views, direction, stream=None):
    """``execute(plan, in, out, direction)`` (Plan.hpp:212-238)."""
    try:
        d = Direction(direction)
    except ValueError:
        raise DFFTError(ERROR_INVALID_ARGUMENT, "Invalid FFT direction specified.") from None
    if d == Direction.forward:
        if not out_views[0].dtype.is_complex:
            raise DFFTError(ERROR_MISMATCH, "Forward FFT operation requires complex output type.")
    else:
        if not in_views[0].dtype.is_complex:
            raise DFFTError(ERROR_MISMATCH, "Backward FFT operation requires complex input type.")
    plan.check_views(in_views, out_views, d)
    plan._execute(in_views, out_views, d, stream)


# ---------------------------------------------------------------------------
# Exchange layer
# ---------------------------------------------------------------------------

def _layout_value(layout):
    return layout.value if hasattr(layout, "value") else int(layout)


def _check_exchange_views(view, buffer, layout):
    for what, t in (("view", view), ("buffer", buffer)):
        if t.numel() and not _is_layout_contiguous(t, layout):
            raise DFFTError(ERROR_INVALID_ARGUMENT, f"the {what} must be contiguous in the "
                            f"{'LayoutLeft' if layout == LAYOUT_LEFT else 'LayoutRight'} order")
        if t.numel() and not t.is_cuda:
            raise DFFTError(ERROR_INVALID_ARGUMENT, f"the {what} must be a device array")
    if view.dtype != buffer.dtype:
        raise DFFTError(ERROR_MISMATCH, "view and buffer value types differ")


def pack(src, buffer, src_map, axis, stream=None, layout=LayoutRight):
    """``pack(exec, src, dst, src_map, axis)`` (PackUnpack.hpp:439-460).

    ``layout``: the views' layout.  LayoutRight buffers are (P, e_0..e_{D-1});
    LayoutLeft ones (e_0..e_{D-1}, P), both given by their logical extents
    (a LayoutLeft tensor is column-major: strides increase along the axes)."""
    lib = load_library()
    lay = _layout_value(layout)
    _check_exchange_views(src, buffer, lay)
    rank = src.dim()
    nprocs = int(buffer.shape[-1] if lay == LAYOUT_LEFT else buffer.shape[0])
    _check(lib.dfft_pack(ctypes.c_void_p(src.data_ptr() if src.numel() else None), _sizes(src.shape),
                         _sizes(src_map), rank, ctypes.c_void_p(buffer.data_ptr() if buffer.numel() else None),
                         _sizes(buffer.shape), nprocs, int(axis), lay, src.element_size(),
                         _stream_handle(stream)))


def unpack(buffer, dst, dst_map, axis, stream=None, layout=LayoutRight):
    """``unpack(exec, src, dst, dst_map, axis)`` (PackUnpack.hpp:478-499)."""
    lib = load_library()
    lay = _layout_value(layout)
    _check_exchange_views(dst, buffer, lay)
    rank = dst.dim()
    nprocs = int(buffer.shape[-1] if lay == LAYOUT_LEFT else buffer.shape[0])
    _check(lib.dfft_unpack(ctypes.c_void_p(buffer.data_ptr() if buffer.numel() else None), _sizes(buffer.shape),
                           nprocs, ctypes.c_void_p(dst.data_ptr() if dst.numel() else None),
                           _sizes(dst.shape), _sizes(dst_map), rank, int(axis), lay, dst.element_size(),
                           _stream_handle(stream)))


def all2all(sends, recvs, stream=None):
    """``all2all(send, recv, comm)`` (MPI_All2All.hpp:23-72) over P virtual ranks."""
    lib = load_library()
    n = len(sends)
    for s in list(sends) + list(recvs):
        if int(s.shape[0]) != n:
            raise DFFTError(ERROR_MISMATCH,
                            "Extent of dimension to be transposed must be equal to the number of processes")
    block = sends[0].numel() // n * sends[0].element_size()
    _check(lib.dfft_all2all(_ptrs(sends), _ptrs(recvs), n, block, _stream_handle(stream)))


def exchange(ins, outs, global_extents, in_topology, in_map, out_topology, out_map, stream=None):
    """TransBlock (TransBlock.hpp:76-107) for all virtual ranks as one remap."""
    lib = load_library()
    tin, tout = _as_topology(in_topology), _as_topology(out_topology)
    rank = len(global_extents)
    _check(lib.dfft_exchange(_ptrs(ins), _ptrs(outs), rank, _sizes(global_extents), _sizes(tin),
                             tin.layout.value, _sizes(in_map), _sizes(tout), tout.layout.value, _sizes(out_map),
                             ins[0].element_size(), _stream_handle(stream)))
