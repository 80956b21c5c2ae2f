"""Expand ``topology_tables.json`` (symbolic, transcribed from the reference's
``Test_Topologies.cpp``) into concrete cases for a given ``nprocs``."""

import json
from pathlib import Path

_TABLES = json.loads(Path(__file__).with_name("topology_tables.json").read_text())


def _resolve(fn, name, nprocs):
    defs, consts = fn["defs"], fn["consts"]
    v = defs[name]
    if isinstance(v, str):  # ref_topoK alias
        v = defs[v]
    out = []
    for x in v:
        if isinstance(x, int):
            out.append(x)
        elif x == "nprocs":
            out.append(nprocs)
        elif x in consts:
            out.append(consts[x])
        else:
            out.append(int(x))
    return out


def slab_cases(nprocs):
    """Yield (function, in_topo, out_topo, axes, expected_topologies)."""
    for fn in _TABLES["functions"]:
        if "slab" not in fn["function"]:
            continue
        for table in fn["tables"]:
            for c in table["cases"]:
                tin, tout, ax, ref = c
                yield (fn["function"], _resolve(fn, tin, nprocs), _resolve(fn, tout, nprocs),
                       _resolve(fn, ax, nprocs), [_resolve(fn, r, nprocs) for r in ref])


def pencil_cases(nprocs):
    """Yield (function, in_topo, in_layout, out_topo, out_layout, axes,
    expected_trans_axes, expected_layouts, expected_topologies)."""
    for fn in _TABLES["functions"]:
        if "pencil" not in fn["function"]:
            continue
        for table in fn["tables"]:
            for c in table["cases"]:
                tin, tout, ax, ref_axes, ref_layouts, ref = c
                yield (fn["function"], _resolve(fn, tin, nprocs), fn["layouts"][tin],
                       _resolve(fn, tout, nprocs), fn["layouts"][tout],
                       _resolve(fn, ax, nprocs), [int(a) for a in ref_axes],
                       [int(a) for a in ref_layouts], [_resolve(fn, r, nprocs) for r in ref])
