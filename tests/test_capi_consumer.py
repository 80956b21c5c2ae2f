"""A compiled C++ consumer of ``include/dfft.h`` and the C++ mirror
``include/dfft.hpp`` (``tests/capi_consumer.cpp``, built by ``tests/Makefile``
from ``__graft_entry__.build()``).

* layout: every struct of the C ABI has the size and field offsets that the
  ctypes declarations in ``dfft.py`` assume;
* host: invalid plan descriptions return the status codes that mirror the
  reference's throw sites (``KokkosFFT_Distributed_Plan.hpp:62-112``,
  ``KokkosFFT_Distributed_MPI_Extents.hpp:322-417``), and the mirror's host
  helpers agree with the planner;
* gpu: slab and pencil round trips through ``dfft::Plan`` / ``dfft::execute``
  on hipMalloc'd views, the execute-time extent check (``good()``,
  ``KokkosFFT_Distributed_InternalPlan.hpp:109-142``) and the C ABI's
  allocation guard.
"""

import json
import subprocess
from pathlib import Path

import pytest

import dfft

ROOT = Path(__file__).resolve().parent.parent
BUILD = ROOT / "tests" / "_build"


def _consumer(name):
    exe = BUILD / name
    if not exe.exists():
        # host variant: cheap to build here; the GPU variant comes from build()
        if name == "capi_consumer":
            subprocess.run(["make", "-s", "-C", str(ROOT / "tests"), str(exe)], check=True)
        else:
            pytest.fail(f"{exe} is missing: run __graft_entry__.build() first")
    return exe


def _run(name, mode):
    out = subprocess.run([str(_consumer(name)), mode], check=True, capture_output=True, text=True, timeout=120)
    return json.loads(out.stdout)


def test_struct_layout_matches_ctypes():
    got = _run("capi_consumer", "layout")
    for cname, cls in (("dfft_desc", dfft._Desc), ("dfft_block_info", dfft.BlockInfo),
                       ("dfft_ns_desc", dfft.NsDesc)):
        import ctypes
        assert got[f"sizeof.{cname}"] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[f"{cname}.{field}"] == getattr(cls, field).offset, (cname, field)


def test_invalid_plans_return_reference_status_codes():
    got = _run("capi_consumer", "host")
    expect = {
        "topology_sizes_differ": dfft.ERROR_INVALID_TOPOLOGY,     # Plan.hpp:65-66
        "topology_size_not_nranks": dfft.ERROR_INVALID_TOPOLOGY,  # Plan.hpp:67-68
        "c2c_extents_differ": dfft.ERROR_INVALID_EXTENTS,         # MPI_Extents.hpp:393-396
        "r2c_out_not_half": dfft.ERROR_INVALID_EXTENTS,           # MPI_Extents.hpp:399-406
        "brick": dfft.ERROR_UNSUPPORTED,                          # Plan.hpp:107-108
        "empty": dfft.ERROR_UNSUPPORTED,                          # Plan.hpp:105-106
        "slab_to_pencil": dfft.ERROR_INVALID_TOPOLOGY,            # Plan.hpp:109-111 (no common type)
        "repeated_axis": dfft.ERROR_INVALID_EXTENTS,              # MPI_Extents.hpp:339-340
        "pencil_rank2": dfft.ERROR_UNSUPPORTED,                   # Plan.hpp:101-103
        "fft_rank_4": dfft.ERROR_INVALID_ARGUMENT,                # Plan.hpp:137 (static_assert)
        "bad_norm": dfft.ERROR_INVALID_ARGUMENT,
        "extent_too_large": dfft.ERROR_UNSUPPORTED,               # no counterpart: int32 element indices
    }
    for k, v in expect.items():
        assert got[k] == v, (k, got[k])


def test_host_helpers_of_the_cpp_mirror():
    got = _run("capi_consumer", "host")
    assert got["slab_topologies"] == [[4, 1, 1], [1, 4, 1]]
    # 129 over 2 ranks: {65, 64} (MPI_Extents.hpp:211-221)
    assert got["local_extents_129_r1"] == [64, 64, 10]
    assert got["local_starts_129_r1"] == [65, 0, 0]
    # LayoutLeft numbering: rank 4 of (2,3,1) is (0, 2, 0) (MPI_Extents.hpp:507-512)
    assert got["coord_left_231_r4"] == dfft.rank_to_coord(dfft.Topology([2, 3, 1], dfft.LayoutLeft), 4)
    assert got["rank_to_coord_out_of_range_throws"] is True
    # config A block list: FFT2, Transpose with buffer (2, 32, 32, 64), FFT1
    assert got["config_a_blocks"] == 3
    assert got["config_a_block1_buffer"] == [2, 32, 32, 64]
    assert got["config_a_max_buffer"] == 2 * 32 * 32 * 64 * 2


@pytest.mark.gpu
def test_cpp_mirror_round_trips_and_checks_on_gpu():
    print("[gpu] capi_consumer gpu: slab C2C 16x12x10 P=2, pencil R2C 16^3 2x2", flush=True)
    got = _run("capi_consumer_gpu", "gpu")
    assert got["slab_name"] == "SlabPlan"
    assert got["pencil_name"] == "PencilPlan"
    assert got["slab_c2c_roundtrip"] < 1e-12
    assert got["pencil_r2c_roundtrip"] < 1e-12
    assert got["good_mismatch_code"] == dfft.ERROR_MISMATCH
    assert got["short_buffer_code"] == dfft.ERROR_MISMATCH
