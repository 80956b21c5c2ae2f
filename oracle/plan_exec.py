"""Oracle (test infrastructure only): virtual-rank CPU executor of the
reference's distributed plans.

``OraclePlan`` restates ``SlabPlan``/``PencilPlan``/``SharedPlan``
(``KokkosFFT_Distributed_SlabPlan.hpp:810-1363``,
``KokkosFFT_Distributed_PencilPlan.hpp:993-1696``) for P virtual ranks held
in one process: the block list comes from ``planner.block_analyses`` for every
rank, FFT blocks use ``oracle.fft`` on the innermost ``fft_dim`` axes of each
rank's mapped local array, Transpose blocks run ``pack`` -> ``all2all`` ->
``unpack`` inside the rank's (sub-)communicator exactly like
``TransBlock::operator()`` (``KokkosFFT_Distributed_TransBlock.hpp:76-107``),
and the plan finishes with ``normalize`` over ``compute_fft_extents``.

Pencil sub-communicators follow ``create_pencil_communicators``
(``PencilPlan.hpp:32-71``): a 2-D row-major Cartesian grid of the non-one
factors of the input topology; communicator axis 0 groups ranks with equal
second coordinate (ordered by the first), axis 1 the converse.
"""

from __future__ import annotations

import numpy as np

from . import fft as F
from . import packunpack as PU
from . import planner as PL


def _perm_between(cur_map, new_map):
    """Axes permutation taking an array mapped by cur_map to one mapped by new_map."""
    return [list(cur_map).index(a) for a in new_map]


class OraclePlan:
    def __init__(self, gin, gout, axes, in_topology, out_topology, real_input=False,
                 norm="backward", layout=PL.LAYOUT_RIGHT, in_layout=PL.LAYOUT_RIGHT,
                 out_layout=PL.LAYOUT_RIGHT):
        # ``layout``: the views' layout (the reference's LayoutType template
        # parameter): it selects the block analyses' maps and buffer shapes
        # and the storage axes the FFT blocks transform; values are indexed
        # alike for both layouts
        self.layout = layout
        self.gin = list(gin)
        self.gout = list(gout)
        self.dim = len(gin)
        self.axes = PL.convert_negative_axes(axes, self.dim)
        self.in_topology = list(in_topology)
        self.out_topology = list(out_topology)
        self.in_layout = in_layout
        self.out_layout = out_layout
        self.real_input = real_input
        self.norm = norm
        if PL.total_size(in_topology) != PL.total_size(out_topology):
            raise PL.PlannerError("in_topology and out_topology must have the same size.")
        self.P = PL.total_size(in_topology)
        PL.are_valid_extents(self.gin, self.gout, self.axes, real_input, False)
        self.kind = PL.get_common_topology_type(self.in_topology, self.out_topology)
        if self.kind == PL.EMPTY:
            raise PL.PlannerError("Empty topology is not supported for FFT plan.")
        if self.kind == PL.BRICK:
            raise PL.PlannerError("Brick topology is not supported for FFT plan.")
        if self.kind not in (PL.SHARED, PL.SLAB, PL.PENCIL):
            raise PL.PlannerError("Unsupported topology for FFT plan.")
        self.name = {PL.SHARED: "SharedPlan", PL.SLAB: "SlabPlan", PL.PENCIL: "PencilPlan"}[self.kind]
        ident = PL.identity(self.dim)
        self.in_local = [PL.compute_local_extents(self.gin, self.in_topology, ident, r, in_layout)
                         for r in range(self.P)]
        self.out_local = [PL.compute_local_extents(self.gout, self.out_topology, ident, r, out_layout)
                          for r in range(self.P)]
        self.fft_extents = PL.compute_fft_extents(self.gin, self.gout, self.axes)
        self.analyses = None
        if self.kind in (PL.SLAB, PL.PENCIL):
            self.analyses = [PL.block_analyses(self.kind, real_input, layout,
                                               self.in_local[r], self.gin, self.gout,
                                               self.in_topology, self.out_topology, self.axes, r,
                                               in_layout, out_layout)
                             for r in range(self.P)]
            self.max_buffer_size = max(a.max_buffer_size for a in self.analyses)
        if self.kind == PL.PENCIL:
            dims = [t for t in self.in_topology if t > 1]
            self.cart = [(r // dims[1], r % dims[1]) for r in range(self.P)]
            self.cart_dims = dims

    # ----------------------------------------------------------------- groups
    def groups(self, comm_axis):
        if self.kind == PL.SLAB:
            return [list(range(self.P))]
        keep = comm_axis
        other = 1 - comm_axis
        out = {}
        for r, c in enumerate(self.cart):
            out.setdefault(c[other], []).append((c[keep], r))
        return [[r for _, r in sorted(v)] for _, v in sorted(out.items())]

    @property
    def blocks(self):
        return self.analyses[0].blocks

    # ---------------------------------------------------------------- execute
    def _exchange(self, state, bi, forward):
        """TransBlock::operator() for every (sub-)communicator."""
        new_state = list(state)
        for grp in self.groups(self.blocks[bi].comm_axis):
            sends = []
            for r in grp:
                b = self.analyses[r].blocks[bi]
                smap, sax = (b.in_map, b.in_axis) if forward else (b.out_map, b.out_axis)
                pdim = 0 if self.layout == PL.LAYOUT_RIGHT else -1
                if len(grp) != b.buffer_extents[pdim]:
                    raise RuntimeError("communicator size differs from the buffer's rank extent")
                sends.append(PU.pack(state[r], b.buffer_extents, smap, sax, self.layout))
            recvs = PU.all2all(sends, self.layout)
            for r, rb in zip(grp, recvs):
                b = self.analyses[r].blocks[bi]
                dmap, dax = (b.out_map, b.out_axis) if forward else (b.in_map, b.in_axis)
                dshape = b.out_extents if forward else b.in_extents
                new_state[r] = PU.unpack(rb, dshape, dmap, dax, layout=self.layout)
        return new_state

    def _fft_block(self, a, b, forward):
        """kokkos-fft DynPlan over the block's fft_dim innermost storage axes:
        the last ones of a LayoutRight array, the first ones of a LayoutLeft
        array (listed here so that the R2C axis -- the contiguous one -- is
        last)."""
        d = self.dim
        if self.layout == PL.LAYOUT_RIGHT:
            inner = list(range(d - b.fft_dim, d))
        else:
            inner = list(range(b.fft_dim - 1, -1, -1))
        is_r2c = self.real_input and b.block_idx == 0
        if forward:
            if is_r2c:
                a = F.r2c(a, inner[-1])
            else:
                a = F.c2c(a, inner[-1], -1)
            for ax in reversed(inner[:-1]):
                a = F.c2c(a, ax, -1)
        else:
            for ax in inner[:-1]:
                a = F.c2c(a, ax, +1)
            if is_r2c:
                a = F.c2r(a, b.in_extents[inner[-1]], inner[-1])
            else:
                a = F.c2c(a, inner[-1], +1)
        return a

    def forward(self, ins):
        ins = [np.asarray(x) for x in ins]
        for r in range(self.P):
            if list(ins[r].shape) != self.in_local[r]:
                raise RuntimeError("extents of input View for plan and execution are not identical")
        n_total = int(np.prod(self.fft_extents))
        scale = F.norm_factor(self.norm, "forward", n_total)
        if self.kind == PL.SHARED:
            return [F.fftn(ins[0], self.axes, self.real_input, self.norm)]
        ident = PL.identity(self.dim)
        state = [x.astype(np.float64 if self.real_input else np.complex128) for x in ins]
        cur = ident
        for bi, b0 in enumerate(self.blocks):
            if b0.block_type == PL.FFT:
                for r in range(self.P):
                    b = self.analyses[r].blocks[bi]
                    a = state[r]
                    if cur != b.in_map:
                        a = PU.transpose(a, _perm_between(cur, b.in_map))
                    state[r] = self._fft_block(a, b, True)
                cur = b0.in_map
            else:
                state = self._exchange(state, bi, True)
                cur = b0.out_map
        if cur != ident:
            state = [PU.transpose(a, _perm_between(cur, ident)) for a in state]
        return [a * scale for a in state]

    def backward(self, outs):
        outs = [np.asarray(x, dtype=np.complex128) for x in outs]
        for r in range(self.P):
            if list(outs[r].shape) != self.out_local[r]:
                raise RuntimeError("extents of output View for plan and execution are not identical")
        n_total = int(np.prod(self.fft_extents))
        scale = F.norm_factor(self.norm, "backward", n_total)
        if self.kind == PL.SHARED:
            return [F.ifftn(outs[0], self.axes, self.gin[self.axes[-1]], self.real_input, self.norm)]
        ident = PL.identity(self.dim)
        state = list(outs)
        cur = ident
        for bi in range(len(self.blocks) - 1, -1, -1):
            b0 = self.blocks[bi]
            if b0.block_type == PL.FFT:
                for r in range(self.P):
                    b = self.analyses[r].blocks[bi]
                    a = state[r]
                    if cur != b.out_map:
                        a = PU.transpose(a, _perm_between(cur, b.out_map))
                    state[r] = self._fft_block(a, b, False)
                cur = b0.in_map
            else:
                state = self._exchange(state, bi, False)
                cur = b0.in_map
        if cur != ident:
            state = [PU.transpose(a, _perm_between(cur, ident)) for a in state]
        return [a * scale for a in state]


# ------------------------------------------------------------------ helpers

def scatter(global_array, gext, topology, layout=PL.LAYOUT_RIGHT):
    """Split a global array into the per-rank local blocks of ``topology``
    (compute_local_extents_and_starts, MPI_Extents.hpp:546-577)."""
    out = []
    for r in range(PL.total_size(topology)):
        ext, st = PL.compute_local_extents_and_starts(gext, topology, r, layout)
        out.append(np.ascontiguousarray(global_array[tuple(slice(s, s + e) for s, e in zip(st, ext))]))
    return out


def gather(locals_, gext, topology, dtype, layout=PL.LAYOUT_RIGHT):
    g = np.zeros(gext, dtype=dtype)
    for r, a in enumerate(locals_):
        ext, st = PL.compute_local_extents_and_starts(gext, topology, r, layout)
        g[tuple(slice(s, s + e) for s, e in zip(st, ext))] = a
    return g
