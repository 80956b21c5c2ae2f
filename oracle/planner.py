"""Oracle (test infrastructure only): CPU restatement of the reference planner.

Only the test-suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker.  The product
planner lives in the HIP/C++ library and is checked against this module and
against the golden tables transcribed from the reference's own unit tests.

Every function restates one function of the reference (cited file:line).
All arithmetic is integer; topologies, extents and maps are python lists.

Semantics of helpers that live in the absent kokkos-fft submodule
(``get_map_axes``, ``get_index``, ``compute_mapped_extents``) are inferred from
their call sites and pinned by the reference's planner golden tables
(``Test_Topologies.cpp``, ``Test_BlockAnalyses.cpp``, ``Test_Mapping.cpp``).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

LAYOUT_RIGHT = "right"
LAYOUT_LEFT = "left"


class PlannerError(RuntimeError):
    """Mirrors the std::runtime_error thrown by KOKKOSFFT_THROW_IF."""


# ---------------------------------------------------------------------------
# kokkos-fft helper semantics (inferred; pinned by golden tables)
# ---------------------------------------------------------------------------

def total_size(values: Sequence[int]) -> int:
    out = 1
    for v in values:
        out *= int(v)
    return out


def get_index(values: Sequence[int], value: int) -> int:
    """Index of the first occurrence; raises when absent (KokkosFFT::Impl::get_index)."""
    for i, v in enumerate(values):
        if v == value:
            return i
    raise PlannerError("value is not included in the container")


def convert_negative_axes(axes: Sequence[int], rank: int) -> List[int]:
    out = []
    for a in axes:
        a = int(a)
        if a < -rank or a >= rank:
            raise PlannerError("axis out of range")
        out.append(a + rank if a < 0 else a)
    return out


def get_map_axes(layout: str, rank: int, axes: Sequence[int]) -> Tuple[List[int], List[int]]:
    """Map that brings the FFT axes innermost (kokkos-fft get_map_axes).

    LayoutRight: [non-FFT axes ascending] + [axes in given order].
    LayoutLeft:  [axes reversed] + [non-FFT axes ascending].
    """
    ax = convert_negative_axes(axes, rank)
    if layout == LAYOUT_RIGHT:
        m = [i for i in range(rank) if i not in ax] + list(ax)
    else:
        m = list(reversed(ax)) + [i for i in range(rank) if i not in ax]
    inv = [m.index(i) for i in range(rank)]
    return m, inv


def compute_mapped_extents(extents: Sequence[int], m: Sequence[int]) -> List[int]:
    return [int(extents[m[i]]) for i in range(len(m))]


def identity(rank: int) -> List[int]:
    return list(range(rank))


# ---------------------------------------------------------------------------
# ContainerAnalyses  (KokkosFFT_Distributed_ContainerAnalyses.hpp)
# ---------------------------------------------------------------------------

def count_non_ones(values):  # :18-28
    return sum(1 for v in values if v != 1)


def extract_different_indices(a, b):  # :37-54
    if len(a) != len(b):
        raise PlannerError("Containers must have the same size.")
    return [i for i in range(len(a)) if a[i] != b[i]]


def extract_different_value_set(a, b):  # :63-78
    if len(a) != len(b):
        raise PlannerError("Containers must have the same size.")
    s = set()
    for i in range(len(a)):
        s.add(a[i])
        s.add(b[i])
    return s


def extract_non_one_indices(a, b):  # :89-104
    if len(a) != len(b):
        raise PlannerError("Containers must have the same size.")
    return [i for i in range(len(a)) if a[i] != 1 or b[i] != 1]


def extract_non_one_values(a):  # :113-125
    return [v for v in a if v != 1]


def has_identical_non_ones(non_ones):  # :133-145
    if count_non_ones(non_ones) < 2:
        return False
    return len(non_ones) == 2 and len(set(non_ones)) == 1


def swap_elements(arr, i, j):  # :157-164
    out = list(arr)
    out[i], out[j] = out[j], out[i]
    return out


def merge_topology(in_topology, out_topology):  # :181-226
    if total_size(in_topology) != total_size(out_topology):
        raise PlannerError("Input and output topologies must have the same size.")
    if total_size(in_topology) == 1:
        return list(in_topology)
    diff = extract_different_indices(in_topology, out_topology)
    if len(diff) != 2:
        raise PlannerError("Input and output topologies must differ exactly two positions")
    return [max(a, b) for a, b in zip(in_topology, out_topology)]


def diff_topology(in_topology, out_topology):  # :243-257
    if total_size(in_topology) == 1 and total_size(out_topology) == 1:
        return 1
    diff = extract_different_indices(in_topology, out_topology)
    if len(diff) != 1:
        raise PlannerError("Input and output topologies must differ at exactly one position.")
    d = diff[0]
    return max(in_topology[d], out_topology[d])


# ---------------------------------------------------------------------------
# Topology types  (KokkosFFT_Distributed_Topologies.hpp:25-89)
# ---------------------------------------------------------------------------

EMPTY, SHARED, SLAB, PENCIL, BRICK, INVALID = "Empty", "Shared", "Slab", "Pencil", "Brick", "Invalid"


def to_topology_type(topology):  # :25-40
    if total_size(topology) == 0:
        return EMPTY
    n = count_non_ones(topology)
    return {0: SHARED, 1: SLAB, 2: PENCIL, 3: BRICK}.get(n, INVALID)


def are_specified_topologies(ttype, *topologies):  # :48-58
    return all(to_topology_type(t) == ttype for t in topologies)


def get_common_topology_type(*topologies):  # :65-89
    if any(to_topology_type(t) == EMPTY for t in topologies):
        return EMPTY
    for t in (SHARED, SLAB, PENCIL, BRICK):
        if are_specified_topologies(t, *topologies):
            return t
    return INVALID


def slab_in_out_axes(in_topology, out_topology):  # :108-132
    if total_size(in_topology) != total_size(out_topology):
        raise PlannerError("Input and output topologies must have the same size.")
    if not are_specified_topologies(SLAB, in_topology, out_topology):
        raise PlannerError("Input and output topologies must be slab topologies.")
    in_axis = out_axis = 0
    for i in range(len(in_topology)):
        if in_topology[i] > 1 and out_topology[i] == 1:
            out_axis = i
        if in_topology[i] == 1 and out_topology[i] > 1:
            in_axis = i
    return in_axis, out_axis


def pencil_in_out_axes(in_topology, out_topology):  # :150-175
    if total_size(in_topology) == 1 or total_size(out_topology) == 1:
        raise PlannerError("Input and output topologies must have at least one non-trivial dimension.")
    if total_size(in_topology) != total_size(out_topology):
        raise PlannerError("Input and output topologies must have the same size.")
    in_axis = out_axis = 0
    for i in range(len(in_topology)):
        if in_topology[i] != out_topology[i]:
            if in_topology[i] == 1:
                in_axis = i
            if out_topology[i] == 1:
                out_axis = i
    return in_axis, out_axis


def propose_mid_array(in_t, out_t):  # :190-242
    diff_indices = extract_different_indices(in_t, out_t)
    diff_value_set = extract_different_value_set(in_t, out_t)
    diff_non_one = extract_non_one_indices(in_t, out_t)
    if len(diff_non_one) < 3:
        raise PlannerError("The total number of non-one elements either in Input and output topologies must be three.")
    if len(diff_indices) < 3 and len(diff_value_set) == 3:
        raise PlannerError("Input and output topologies must differ exactly three positions.")
    dim = len(in_t)
    in_trim = [0] * dim
    out_trim = [0] * dim
    for d in diff_indices:
        in_trim[d] = in_t[d]
        out_trim[d] = out_t[d]
    idx_one_in = get_index(in_trim, 1)
    idx_one_out = get_index(out_trim, 1)
    for i in range(len(diff_non_one)):
        for j in range(i + 1, len(diff_non_one)):
            mid = swap_elements(in_t, diff_non_one[i], diff_non_one[j])
            idx_one_mid = get_index(mid, 1)
            mid_in = extract_different_indices(mid, in_t)
            mid_out = extract_different_indices(mid, out_t)
            if len(mid_in) == 2 and len(mid_out) == 2 and not (
                idx_one_mid == idx_one_in or idx_one_mid == idx_one_out
            ):
                a0, a1 = mid[mid_in[0]], mid[mid_in[1]]
                b0, b1 = mid[mid_out[0]], mid[mid_out[1]]
                if (a0 == 1 or a1 == 1) and (b0 == 1 or b1 == 1):
                    return mid
    return list(out_t)


def decompose_axes(topologies, axes):  # :258-331
    dim = len(topologies[0])
    ax = convert_negative_axes(axes, dim)
    axes_reversed = list(reversed(ax))
    all_axes = []
    for topo in topologies:
        ready = []
        for a in axes_reversed:
            if topo[a] > 1:
                break
            ready.append(a)
        all_axes.append(list(reversed(ready)))
        for a in ready:
            if a in axes_reversed:
                axes_reversed.remove(a)
    if sum(len(r) for r in all_axes) != len(axes):
        raise PlannerError("Axes are not decomposed correctly")
    return all_axes


def compute_trans_axis(in_t, out_t, first_non_one):  # :351-399
    in_non = extract_non_one_values(in_t)
    out_non = extract_non_one_values(out_t)
    if len(in_non) != 2 or len(out_non) != 2:
        raise PlannerError("Input and output topologies must have exactly two non-one elements.")
    if has_identical_non_ones(in_non) or has_identical_non_ones(out_non):
        raise PlannerError("Input and output topologies must not have identical non-one elements.")
    diff = extract_different_indices(in_t, out_t)
    if len(diff) != 2:
        raise PlannerError("Input and output topologies must differ exactly two positions")
    exchange = 0
    for d in diff:
        if in_t[d] > 1:
            exchange = in_t[d]
            break
    return int(not (exchange == first_non_one))


def get_all_slab_topologies(in_topology, out_topology, axes):  # :420-618
    dim = len(in_topology)
    fft_dim = len(axes)
    if not (1 <= fft_dim <= 3) or dim < 2 or dim < fft_dim:
        raise PlannerError("invalid dimensions")
    if not are_specified_topologies(SLAB, in_topology, out_topology):
        raise PlannerError("Input and output topologies must be slab topologies.")
    in_topology = list(in_topology)
    out_topology = list(out_topology)
    axes = list(axes)
    topologies = [in_topology]
    p = total_size(in_topology)

    def add_at(axis):
        t = [1] * dim
        t[axis] = p
        if topologies[-1] != t:
            topologies.append(t)

    def add_at_first_one(topo):
        for i in range(dim):
            if topo[i] == 1:
                add_at(i)
                break

    def finalize():
        if topologies[-1] != out_topology:
            topologies.append(out_topology)
        return topologies

    if dim == 2 and fft_dim == 2:
        if in_topology == out_topology:
            add_at_first_one(in_topology)
        else:
            if in_topology[axes[-1]] != 1:
                add_at_first_one(in_topology)
                if topologies[-1] != in_topology:
                    topologies.append(in_topology)
        return finalize()

    if dim == 3 and fft_dim == 3:
        if in_topology == out_topology:
            if in_topology[axes[-1]] != 1:
                add_at(axes[0])
            else:
                add_at(axes[-1])
        else:
            if in_topology[axes[-1]] != 1:
                add_at(axes[0])
                if out_topology[axes[0]] != 1:
                    add_at_first_one(topologies[-1])
            else:
                if in_topology[axes[1]] != 1:
                    add_at(axes[-1])
                else:
                    if out_topology[axes[0]] != 1:
                        add_at_first_one(topologies[-1])
        return finalize()

    axes_reversed = list(axes)

    def is_ready(topo):
        return all(topo[a] <= 1 for a in axes_reversed)

    if is_ready(in_topology) or is_ready(out_topology):
        return finalize()

    if fft_dim == 1:
        for i in range(dim):
            if i not in axes_reversed:
                add_at(i)
                break
        return finalize()
    elif dim > 3 and fft_dim == 3:
        sorted_axes = sorted(axes)
        axes_reversed.reverse()
        ready = []
        for a in axes_reversed:
            if in_topology[a] > 1:
                break
            ready.append(a)
        if len(ready) == 0:
            add_at(axes_reversed[-1])
            for a in axes_reversed:
                if topologies[-1][a] > 1:
                    break
                ready.append(a)
        for a in ready:
            if a in axes_reversed:
                axes_reversed.remove(a)
        if not is_ready(out_topology):
            for a in sorted_axes:
                if a not in axes_reversed:
                    add_at(a)
                    break
        return finalize()
    else:
        axes_reversed.reverse()
        ready = []
        for a in axes_reversed:
            if in_topology[a] > 1:
                break
            ready.append(a)
        for a in ready:
            if a in axes_reversed:
                axes_reversed.remove(a)
        if not is_ready(out_topology):
            for i in range(dim):
                if i not in axes_reversed:
                    add_at(i)
                    break
        return finalize()


def get_all_pencil_topologies(in_topology, out_topology, axes,
                              in_layout=LAYOUT_RIGHT, out_layout=LAYOUT_RIGHT):
    """Returns (topologies, trans_axes, layouts) -- Topologies.hpp:642-811."""
    dim = len(in_topology)
    fft_dim = len(axes)
    if not (1 <= fft_dim <= 3) or dim < 3 or dim < fft_dim:
        raise PlannerError("invalid dimensions")
    if not are_specified_topologies(PENCIL, in_topology, out_topology):
        raise PlannerError("Input and output topologies must be pencil topologies.")
    axes_reversed = list(axes)
    non_ones = extract_non_one_values(in_topology)
    same = has_identical_non_ones(non_ones)
    in_tmp = list(in_topology)
    out_tmp = list(out_topology)
    if same:
        dummies = [2, 3]
        c = 0
        for i in range(dim):
            if in_tmp[i] > 1:
                in_tmp[i] = dummies[c]
                c += 1
        c = 0
        if in_layout != out_layout:
            dummies.reverse()
        for i in range(dim):
            if out_tmp[i] > 1:
                out_tmp[i] = dummies[c]
                c += 1
    nz = extract_non_one_values(in_tmp)
    first_non_one = nz[0] if in_layout == LAYOUT_RIGHT else nz[1]

    def to_original(topologies, trans_axes, layouts):
        if same:
            v = non_ones[0]
            orig = [[v if x > 1 else x for x in t] for t in topologies]
            return orig, trans_axes, layouts
        return topologies, trans_axes, layouts

    def get_layout(topo):
        return int(extract_non_one_values(topo)[0] == first_non_one)

    topologies = [in_tmp]
    trans_axes: List[int] = []
    layouts = [get_layout(in_tmp)]

    def add(topo):
        if topologies[-1] != topo:
            trans_axes.append(compute_trans_axis(topologies[-1], topo, first_non_one))
            topologies.append(list(topo))
            layouts.append(get_layout(topo))

    def try_add_mid():
        try:
            mid = propose_mid_array(topologies[-1], out_tmp)
            add(mid)
        except PlannerError:
            pass

    def finalize():
        add(out_tmp)
        return to_original(topologies, trans_axes, layouts)

    def is_ready(topo, cur_axes):
        return all(topo[a] == 1 for a in cur_axes)

    if is_ready(in_tmp, axes_reversed) or is_ready(out_tmp, axes_reversed):
        try_add_mid()
        return finalize()

    axes_reversed.reverse()
    shuffled = list(in_tmp)
    last_axis = axes_reversed[0]
    first_axis = axes_reversed[-1]
    first_dim = in_tmp[last_axis]
    last_dim = out_tmp[first_axis]
    if first_dim == 1:
        axes_reversed.pop(0)
    if last_dim == 1 and axes_reversed:
        axes_reversed.pop()

    for axis in axes_reversed:
        swap_idx = 0
        if in_tmp == out_tmp:
            for idx in range(dim):
                if shuffled[idx] == 1 and idx != axis:
                    swap_idx = idx
                    break
        else:
            for d in extract_non_one_indices(in_tmp, out_tmp):
                if shuffled[d] == 1 and d != axis:
                    swap_idx = d
                    break
        shuffled = swap_elements(shuffled, axis, swap_idx)
        add(shuffled)

    if topologies[-1] == out_tmp:
        return to_original(topologies, trans_axes, layouts)
    try_add_mid()
    return finalize()


# ---------------------------------------------------------------------------
# Mapping / Extents  (Mapping.hpp:39-110, Extents.hpp:35-137)
# ---------------------------------------------------------------------------

def permute_map_by_axes(layout, src_map, axes):  # Mapping.hpp:39-85
    if isinstance(axes, int):
        axes = [axes]
    axes = list(axes)
    m = []
    if layout == LAYOUT_RIGHT:
        for s in src_map:
            if s not in axes:
                m.append(s)
        m.extend(axes)
    else:
        rev = list(reversed(axes))
        m.extend(rev)
        for s in src_map:
            if s not in rev:
                m.append(s)
    return m[: len(src_map)]


def compute_buffer_extents(layout, extents, in_topology, out_topology):  # Extents.hpp:35-60
    merged = merge_topology(in_topology, out_topology)
    p0 = diff_topology(merged, in_topology)
    body = [(extents[i] - 1) // merged[i] + 1 for i in range(len(extents))]
    if layout == LAYOUT_RIGHT:
        return [p0] + body
    return body + [p0]


def compute_fft_extents(gin, gout, axes):  # Extents.hpp:116-137
    return [max(gin[a], gout[a]) for a in axes]


# ---------------------------------------------------------------------------
# MPI_Extents  (KokkosFFT_Distributed_MPI_Extents.hpp)
# ---------------------------------------------------------------------------

def rank_to_coord(topology, rank, layout=LAYOUT_RIGHT):  # :492-529
    if rank >= total_size(topology):
        raise PlannerError("rank must be less than topology size.")
    coord = [0] * len(topology)
    r = rank
    order = range(len(topology) - 1, -1, -1) if layout == LAYOUT_RIGHT else range(len(topology))
    for i in order:
        coord[i] = r % topology[i]
        r //= topology[i]
    return coord


def compute_local_extents(extents, topology, m, rank, layout=LAYOUT_RIGHT):  # :195-254
    local = list(extents)
    coords = rank_to_coord(topology, rank, layout)
    for i in range(len(extents)):
        if topology[i] > 1:
            n, t = extents[i], topology[i]
            q, r = divmod(n, t)
            local[i] = q + 1 if coords[i] < r else q
    return [local[m[i]] for i in range(len(extents))]


def compute_local_starts(local_extents_of_all_ranks, topology, rank, layout=LAYOUT_RIGHT):
    """accumulate_extents (:55-107) with the Allgather replaced by the table of
    every rank's local extents (index = rank)."""
    dim = len(topology)
    total = total_size(topology)
    coords = rank_to_coord(topology, rank, layout)
    acc = [0] * dim
    if layout == LAYOUT_RIGHT:
        stride = total
        for i in range(dim):
            stride //= topology[i]
            acc[i] = sum(local_extents_of_all_ranks[stride * j][i] for j in range(coords[i]))
    else:
        stride = 1
        for i in range(dim):
            acc[i] = sum(local_extents_of_all_ranks[stride * j][i] for j in range(coords[i]))
            stride *= topology[i]
    return acc


def compute_local_extents_and_starts(extents, topology, rank, layout=LAYOUT_RIGHT):  # :546-577
    ident = identity(len(extents))
    allext = [compute_local_extents(extents, topology, ident, r, layout)
              for r in range(total_size(topology))]
    return allext[rank], compute_local_starts(allext, topology, rank, layout)


def compute_global_extents(local_extents_of_all_ranks, topology, rank=0, layout=LAYOUT_RIGHT):
    """compute_global_extents (:114-129) with the Allgather replaced by the table."""
    coords = list(topology)
    dim = len(topology)
    total = total_size(topology)
    acc = [0] * dim
    if layout == LAYOUT_RIGHT:
        stride = total
        for i in range(dim):
            stride //= topology[i]
            acc[i] = sum(local_extents_of_all_ranks[stride * j][i] for j in range(coords[i]))
    else:
        stride = 1
        for i in range(dim):
            acc[i] = sum(local_extents_of_all_ranks[stride * j][i] for j in range(coords[i]))
            stride *= topology[i]
    mine = local_extents_of_all_ranks[rank]
    for i in range(dim):
        if topology[i] == 1:
            acc[i] = mine[i]
    return acc


def extent_after_transform(n, is_r2c):
    return n // 2 + 1 if is_r2c else n


def are_valid_extents(gin, gout, axes, is_real_in, is_real_out, layout=LAYOUT_RIGHT):  # :322-417
    rank = len(gin)
    if is_real_in and is_real_out:
        raise PlannerError("real to real transform is not supported")
    ax = convert_negative_axes(axes, rank)
    if len(set(ax)) != len(ax):
        raise PlannerError("input axes are not valid for the view")
    inner = 0 if layout == LAYOUT_LEFT else rank - 1
    m, _ = get_map_axes(layout, rank, ax)
    ine = compute_mapped_extents(gin, m)
    oute = compute_mapped_extents(gout, m)
    for i in range(rank):
        if i == inner:
            continue
        if ine[i] != oute[i]:
            raise PlannerError("input and output extents must be the same except for the transform axis")
    if not is_real_in and not is_real_out and ine[inner] != oute[inner]:
        raise PlannerError("input and output extents must be the same for C2C transform")
    if is_real_in and oute[inner] != ine[inner] // 2 + 1:
        raise PlannerError("For R2C, the 'output extent' of transform must be equal to 'input extent'/2 + 1")
    if is_real_out and ine[inner] != oute[inner] // 2 + 1:
        raise PlannerError("For C2R, the 'input extent' of transform must be equal to 'output extent' / 2 + 1")
    return True


# ---------------------------------------------------------------------------
# Block analyses  (KokkosFFT_Distributed_BlockAnalyses.hpp)
# ---------------------------------------------------------------------------

TRANSPOSE, FFT = "Transpose", "FFT"


@dataclass
class BlockInfo:  # Types.hpp:61-154
    in_topology: List[int] = field(default_factory=list)
    out_topology: List[int] = field(default_factory=list)
    in_extents: List[int] = field(default_factory=list)
    out_extents: List[int] = field(default_factory=list)
    buffer_extents: List[int] = field(default_factory=list)
    in_map: List[int] = field(default_factory=list)
    out_map: List[int] = field(default_factory=list)
    in_axis: int = 0
    out_axis: int = 0
    fft_dim: int = 0
    comm_axis: int = 0
    block_type: str = FFT
    block_idx: int = 0


def count_blocks(blocks, btype):  # :21-29
    return sum(1 for b in blocks if b.block_type == btype)


def count_fft_dims(blocks):  # :36-46
    return sum(b.fft_dim for b in blocks if b.block_type == FFT)


def propose_fft_block(m, in_extents, out_extents, fft_dim, block_idx):  # :57-74
    b = BlockInfo(block_type=FFT, fft_dim=fft_dim, in_extents=list(in_extents),
                  out_extents=list(out_extents), in_map=list(m), out_map=list(m),
                  block_idx=block_idx)
    b.in_topology = []
    b.out_topology = []
    dim = len(m)
    b.in_topology = [0] * dim
    b.out_topology = [0] * dim
    b.buffer_extents = [0] * (dim + 1)
    return b, total_size(b.out_extents) * 2


def propose_transpose_block(layout, m, in_topology, out_topology, in_extents,
                            global_extents, axes, rank, is_layout_right, is_last,
                            in_axis, out_axis, comm_axis, block_idx, size_factor):  # :95-138
    dim = len(m)
    b = BlockInfo(block_type=TRANSPOSE, in_map=list(m), in_axis=in_axis, out_axis=out_axis,
                  comm_axis=comm_axis, in_topology=list(in_topology),
                  out_topology=list(out_topology), in_extents=list(in_extents),
                  block_idx=block_idx)
    if is_last:
        b.out_map = identity(dim)
    else:
        b.out_map = (permute_map_by_axes(layout, b.in_map, out_axis) if len(axes) == 0
                     else permute_map_by_axes(layout, b.in_map, axes))
    b.out_extents = compute_local_extents(global_extents, out_topology, b.out_map, rank,
                                          LAYOUT_RIGHT if is_layout_right else LAYOUT_LEFT)
    b.buffer_extents = compute_buffer_extents(layout, global_extents, in_topology, out_topology)
    mx = max(total_size(b.in_extents), total_size(b.buffer_extents), total_size(b.out_extents))
    return b, mx * size_factor


def _propose_block(kind, is_real_value, layout, fft_total_dim, in_extents, gin, gout,
                   blocks, all_topologies, all_trans_axes, all_layouts, all_axes,
                   m, rank, block_idx):
    """Slab (:158-293) and pencil (:311-395) propose_block share this body."""
    dim = len(in_extents)

    def transpose_block(blocks_, size_factor):
        is_last = block_idx == len(all_topologies) - 1
        nb_trans = count_blocks(blocks_, TRANSPOSE)
        cur = all_topologies[nb_trans]
        nxt = all_topologies[nb_trans + 1]
        if kind == SLAB:
            in_axis, out_axis = slab_in_out_axes(cur, nxt)
            comm_axis = 0
            is_lr = True
        else:
            in_axis, out_axis = pencil_in_out_axes(cur, nxt)
            comm_axis = all_trans_axes[nb_trans]
            is_lr = bool(all_layouts[nb_trans + 1])
        next_axes = [] if is_last else list(all_axes[block_idx + 1])
        if not blocks_:
            return propose_transpose_block(layout, identity(dim), cur, nxt, in_extents, gin,
                                           next_axes, rank, is_lr, False, in_axis, out_axis,
                                           comm_axis, 0, size_factor)
        prev = blocks_[-1]
        return propose_transpose_block(layout, prev.out_map, cur, nxt, prev.out_extents, gout,
                                       next_axes, rank, is_lr, is_last, in_axis, out_axis,
                                       comm_axis, nb_trans, 2)

    out = []
    axes = all_axes[block_idx]
    fft_dim = len(axes)
    if fft_dim == 0:
        out.append(transpose_block(blocks, 1 if is_real_value else 2))
        return out
    if kind == SLAB:
        lay = LAYOUT_RIGHT
    else:
        lay = LAYOUT_RIGHT if all_layouts[block_idx] else LAYOUT_LEFT
    is_last_fft = (count_fft_dims(blocks) + fft_dim) == fft_total_dim
    if not blocks:
        topo = all_topologies[block_idx]
        fin = compute_mapped_extents(in_extents, m)
        fout = compute_local_extents(gout, topo, m, rank, lay)
        t = propose_fft_block(m, fin, fout, fft_dim, 0)
        out.append(t)
        if not is_last_fft:
            out.append(transpose_block(blocks + [t[0]], 2))
        return out
    nb_fft = count_blocks(blocks, FFT)
    in_map = blocks[-1].out_map
    fin = blocks[-1].out_extents
    if nb_fft == 0:
        topo = blocks[-1].out_topology
        fout = compute_local_extents(gout, topo, in_map, rank, lay)
        t = propose_fft_block(in_map, fin, fout, fft_dim, 0)
    else:
        t = propose_fft_block(in_map, fin, fin, fft_dim, nb_fft)
    out.append(t)
    if not is_last_fft:
        out.append(transpose_block(blocks + [t[0]], 2))
    return out


@dataclass
class BlockAnalyses:
    blocks: List[BlockInfo]
    max_buffer_size: int  # local (this rank) maximum, in scalar units
    topologies: List[List[int]]
    trans_axes: List[int]
    layouts: List[int]
    all_axes: List[List[int]]


def block_analyses(kind, is_real_value, layout, in_extents, gin, gout, in_topology,
                   out_topology, axes, rank, in_layout=LAYOUT_RIGHT, out_layout=LAYOUT_RIGHT):
    """Slab/PencilBlockAnalysesInternal (:415-444, :476-506) for one rank.

    The MPI_Allreduce(MAX) of the buffer size is left to the caller
    (``global_max_buffer_size``)."""
    dim = len(in_extents)
    m, _ = get_map_axes(layout, dim, axes)
    if kind == SLAB:
        topos = get_all_slab_topologies(in_topology, out_topology, axes)
        tax, lays = [], []
    else:
        topos, tax, lays = get_all_pencil_topologies(in_topology, out_topology, axes,
                                                     in_layout, out_layout)
    all_axes = decompose_axes(topos, axes)
    blocks: List[BlockInfo] = []
    sizes = []
    for itopo in range(len(topos)):
        for b, sz in _propose_block(kind, is_real_value, layout, len(axes), in_extents, gin,
                                    gout, blocks, topos, tax, lays, all_axes, m, rank, itopo):
            blocks.append(b)
            sizes.append(sz)
    return BlockAnalyses(blocks, max(sizes), topos, tax, lays, all_axes)
