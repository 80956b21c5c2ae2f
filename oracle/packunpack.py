"""Oracle (test infrastructure only): CPU restatement of the reference's data
movement on the hot path -- bin mapping, pack, all-to-all, unpack, local
transpose.  Only the test-suite, ``__graft_entry__.smoke()`` and the CPU
baseline of ``bench.py`` may use it, and only as the checker.

Arrays are numpy arrays indexed like the reference's views (the storage
order of a LayoutLeft view does not change its index semantics).  A local
array "mapped by m" stores logical axis ``m[k]`` at array axis ``k``.
LayoutRight buffers are ``(P, e_0, ..., e_{D-1})`` and LayoutLeft buffers
``(e_0, ..., e_{D-1}, P)``, with the ``e_i`` in *logical* axis order, as in
``KokkosFFT_Distributed_PackUnpack.hpp:176-178, 217-233`` and
``KokkosFFT_Distributed_Extents.hpp:44-60``.
"""

from __future__ import annotations

import itertools

import numpy as np

OUT_OF_BOUNDS = -1  # PackUnpack.hpp:13-21
LAYOUT_RIGHT, LAYOUT_LEFT = "right", "left"  # the planner's layout tags


def _right_shape(buffer_shape, layout):
    """(P, e...) shape of a buffer given in its layout's order."""
    buffer_shape = list(buffer_shape)
    return buffer_shape if layout == LAYOUT_RIGHT else [buffer_shape[-1]] + buffer_shape[:-1]


def bin_mapping(n: int, nbins: int, ibin: int):
    """(start, length) of bin ``ibin`` of [0, n) -- PackUnpack.hpp:40-46."""
    base, rem = divmod(n, nbins)
    length = base + (1 if ibin < rem else 0)
    start = ibin * base + min(ibin, rem)
    return start, length


def merge_indices(idx: int, start: int, extent: int, axis: int, merged_axis: int) -> int:
    """PackUnpack.hpp:62-73."""
    if axis == merged_axis:
        return OUT_OF_BOUNDS if idx >= extent else idx + start
    return idx


def _logical_view(a: np.ndarray, m):
    """View of a mapped array with axes in logical order (no copy)."""
    inv = [list(m).index(i) for i in range(len(m))]
    return a.transpose(inv)


def pack(src: np.ndarray, buffer_shape, src_map, axis: int, layout=LAYOUT_RIGHT) -> np.ndarray:
    """Pack::PackInternal (PackUnpack.hpp:217-233); the process index is the
    buffer's first index (LayoutRight) or last (LayoutLeft, :176-178).

    ``merged_size`` = src.extent(index of ``axis`` in ``src_map``) (:446-447).
    Vectorised; ``pack_literal`` is the element-wise form it is tested against.
    """
    if layout != LAYOUT_RIGHT:
        return np.moveaxis(pack(src, _right_shape(buffer_shape, layout), src_map, axis), 0, -1)
    nprocs = buffer_shape[0]
    body = list(buffer_shape[1:])
    dst = np.zeros(buffer_shape, dtype=src.dtype)
    lg = _logical_view(src, src_map)
    merged = src.shape[list(src_map).index(axis)]
    for p in range(nprocs):
        start, length = bin_mapping(merged, nprocs, p)
        sl_dst = [p]
        sl_src = []
        for i in range(len(body)):
            if i == axis:
                n = max(0, min(length, body[i], lg.shape[i] - start))
                sl_dst.append(slice(0, n))
                sl_src.append(slice(start, start + n))
            else:
                n = min(body[i], lg.shape[i])
                sl_dst.append(slice(0, n))
                sl_src.append(slice(0, n))
        dst[tuple(sl_dst)] = lg[tuple(sl_src)]
    return dst


def unpack(buf: np.ndarray, dst_shape, dst_map, axis: int, dst: np.ndarray | None = None,
           layout=LAYOUT_RIGHT) -> np.ndarray:
    """Unpack::UnpackInternal (PackUnpack.hpp:385-417); LayoutLeft buffers
    carry the process index last.

    Entries of ``dst`` that no buffer element maps to are left untouched
    (zero-initialised when ``dst`` is not given)."""
    if layout != LAYOUT_RIGHT:
        return unpack(np.moveaxis(buf, -1, 0), dst_shape, dst_map, axis, dst)
    if dst is None:
        dst = np.zeros(dst_shape, dtype=buf.dtype)
    nprocs = buf.shape[0]
    body = list(buf.shape[1:])
    lg = _logical_view(dst, dst_map)
    merged = dst.shape[list(dst_map).index(axis)]
    for p in range(nprocs):
        start, length = bin_mapping(merged, nprocs, p)
        sl_buf = [p]
        sl_dst = []
        for i in range(len(body)):
            if i == axis:
                n = max(0, min(length, body[i], lg.shape[i] - start))
                sl_buf.append(slice(0, n))
                sl_dst.append(slice(start, start + n))
            else:
                n = min(body[i], lg.shape[i])
                sl_buf.append(slice(0, n))
                sl_dst.append(slice(0, n))
        lg[tuple(sl_dst)] = buf[tuple(sl_buf)]
    return dst


def pack_literal(src, buffer_shape, src_map, axis, layout=LAYOUT_RIGHT):
    """Element-by-element statement of Pack (PackUnpack.hpp:217-233); the
    LayoutLeft branch (:236-247) takes p from the last buffer index."""
    dst = np.zeros(buffer_shape, dtype=src.dtype)
    left = layout != LAYOUT_RIGHT
    nprocs = buffer_shape[-1] if left else buffer_shape[0]
    d = len(src_map)
    merged = src.shape[list(src_map).index(axis)]
    for idx in itertools.product(*[range(n) for n in buffer_shape]):
        p = idx[-1] if left else idx[0]
        body = idx[:-1] if left else idx[1:]
        start, extent = bin_mapping(merged, nprocs, p)
        l = [merge_indices(body[i], start, extent, i, axis) for i in range(d)]
        ok = True
        for k in range(d):
            v = l[src_map[k]]
            if v == OUT_OF_BOUNDS or v >= src.shape[k]:
                ok = False
                break
        if ok:
            dst[idx] = src[tuple(l[src_map[k]] for k in range(d))]
    return dst


def unpack_literal(buf, dst_shape, dst_map, axis, layout=LAYOUT_RIGHT):
    """Element-by-element statement of Unpack (PackUnpack.hpp:385-417)."""
    dst = np.zeros(dst_shape, dtype=buf.dtype)
    left = layout != LAYOUT_RIGHT
    nprocs = buf.shape[-1] if left else buf.shape[0]
    d = len(dst_map)
    merged = dst_shape[list(dst_map).index(axis)]
    for idx in itertools.product(*[range(n) for n in buf.shape]):
        p = idx[-1] if left else idx[0]
        body = idx[:-1] if left else idx[1:]
        start, extent = bin_mapping(merged, nprocs, p)
        l = [merge_indices(body[i], start, extent, i, axis) for i in range(d)]
        ok = True
        for k in range(d):
            v = l[dst_map[k]]
            if v == OUT_OF_BOUNDS or v >= dst_shape[k]:
                ok = False
                break
        if ok:
            dst[tuple(l[dst_map[k]] for k in range(d))] = buf[idx]
    return dst


def all2all(sends, layout=LAYOUT_RIGHT):
    """MPI_Alltoall over the process dimension of the buffers -- the first
    (LayoutRight) or last (LayoutLeft) one (MPI_All2All.hpp:44-57): rank r's
    block p lands at rank p as block r."""
    nprocs = len(sends)
    pdim = 0 if layout == LAYOUT_RIGHT else -1
    for s in sends:
        if s.shape[pdim] != nprocs:
            raise RuntimeError("Extent of dimension to be transposed must be equal to the number of processes")
    if layout != LAYOUT_RIGHT:
        return [np.stack([sends[r][..., p] for r in range(nprocs)], axis=-1) for p in range(nprocs)]
    return [np.stack([sends[r][p] for r in range(nprocs)]) for p in range(nprocs)]


def transpose(a: np.ndarray, m) -> np.ndarray:
    """kokkos-fft Impl::transpose: out.extent(k) = in.extent(m[k]),
    out(j) = in(i) with i[m[k]] = j_k (numpy transpose(axes=m))."""
    return np.ascontiguousarray(a.transpose(list(m)))
