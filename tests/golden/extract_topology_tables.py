"""Transcribe the planner golden tables of the reference's unit tests into JSON.

Reads ``distributed/unit_test/Test_Topologies.cpp`` of the reference checkout
as text (nothing is compiled or run) and writes ``topology_tables.json``: for
every ``test_get_all_slab_topologies*`` / ``test_get_all_pencil_topologies*``
function, the symbolic topology/axes definitions (``nprocs`` left symbolic)
and each case ``(in, out, axes[, trans_axes, layouts], expected topologies)``.

Usage:  python tests/golden/extract_topology_tables.py /root/reference
"""

import json
import re
import sys
from pathlib import Path


def split_top(s):
    """Split a comma list at brace depth 0."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def parse_brace_list(s):
    """'{a, b, {c}}' -> python structure of strings."""
    s = s.strip()
    m = re.match(r"^[A-Za-z_0-9:]*\s*\{(.*)\}$", s, re.S)
    if not m:
        return s
    return [parse_brace_list(x) for x in split_top(m.group(1))]


def function_bodies(text, prefix):
    for m in re.finditer(r"void (" + prefix + r"\w*)\(std::size_t nprocs\) \{", text):
        start = m.end()
        depth, i = 1, start
        while depth:
            if text[i] == "{":
                depth += 1
            elif text[i] == "}":
                depth -= 1
            i += 1
        yield m.group(1), text[start:i - 1]


def parse_function(name, body):
    defs = {}
    layouts = {}
    # declarations "topo_r_type topo0{...}, topo1{...};"
    for decl in re.finditer(r"\b(topo_type|topo_r_type|topo_l_type|axes_type)\s+([^;]*?);", body, re.S):
        kind, rest = decl.group(1), decl.group(2)
        if "=" in rest and "{" not in rest.split("=")[0]:
            continue
        for item in split_top(rest):
            mm = re.match(r"(\w+)\s*\{(.*)\}$", item.strip(), re.S)
            if not mm:
                continue
            vals = [v.strip() for v in mm.group(2).split(",")]
            defs[mm.group(1)] = [int(v) if v.isdigit() else v for v in vals]
            if kind != "axes_type":
                layouts[mm.group(1)] = "left" if kind == "topo_l_type" else "right"
    for mm in re.finditer(r"(ref_topo\d+)\s*=\s*(topo\d+)\.array\(\)", body):
        defs[mm.group(1)] = mm.group(2)
    consts = {}
    for mm in re.finditer(r"std::size_t\s+(np\d)\s*=\s*(\d+);", body):
        consts[mm.group(1)] = int(mm.group(2))
    tables = []
    for mm in re.finditer(r"std::vector<(\w+)>\s+(\w+test_cases)\s*=\s*\{(.*?)\};", body, re.S):
        cases = []
        for c in split_top(mm.group(3)):
            cases.append(parse_brace_list(c))
        tables.append({"tuple_type": mm.group(1), "name": mm.group(2), "cases": cases})
    return {"function": name, "defs": defs, "layouts": layouts, "consts": consts,
            "tables": tables}


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = (ref / "distributed/unit_test/Test_Topologies.cpp").read_text()
    out = []
    for prefix in ("test_get_all_slab_topologies", "test_get_all_pencil_topologies"):
        for name, body in function_bodies(src, prefix):
            out.append(parse_function(name, body))
    dst = Path(__file__).with_name("topology_tables.json")
    dst.write_text(json.dumps({"source": "distributed/unit_test/Test_Topologies.cpp",
                               "nprocs_values": [2, 3, 4, 5, 6], "functions": out}, indent=1))
    print(dst, sum(len(t["cases"]) for f in out for t in f["tables"]), "cases")


if __name__ == "__main__":
    main()
