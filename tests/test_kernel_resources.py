"""Resource use of the built line-pass kernels (CPU test on libdfft.so).

Reads the gfx950 code object that hipcc embedded in libdfft.so (the
.hip_fatbin offload bundle) and its AMDGPU metadata note, and checks the
private (scratch) memory of every kernel.  The FFT kernels keep their lines
in registers and LDS; a kernel that needs more than a few spill slots of
private memory per lane is a code-generation regression (for instance the
compiler copying the 1.6 KB kernel argument into private memory when it
cannot address it in place, which made a 1024-point pass ten times slower).
"""

import struct
from pathlib import Path

import pytest

msgpack = pytest.importorskip("msgpack")

LIB = Path(__file__).resolve().parents[1] / "distributed-fft-for-kokkos_amd" / "libdfft.so"
NT_AMDGPU_METADATA = 32


def _sections(elf):
    shoff = struct.unpack_from("<Q", elf, 0x28)[0]
    shentsize, shnum, shstrndx = struct.unpack_from("<HHH", elf, 0x3A)
    raw = []
    for i in range(shnum):
        name, typ, _flags, _addr, offset, size = struct.unpack_from("<IIQQQQ", elf, shoff + i * shentsize)
        raw.append((name, typ, offset, size))
    stroff = raw[shstrndx][2]

    def nm(n):
        return elf[stroff + n: elf.index(b"\0", stroff + n)].decode()
    return [(nm(n), typ, off, size) for n, typ, off, size in raw]


def _code_objects(lib_bytes, arch="gfx950"):
    """The gfx950 code objects of every offload bundle in .hip_fatbin (one
    bundle per HIP translation unit)."""
    fat = [s for s in _sections(lib_bytes) if s[0] == ".hip_fatbin"]
    assert fat, "libdfft.so has no .hip_fatbin section"
    _, _, off, size = fat[0]
    blob = lib_bytes[off: off + size]
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    assert blob.startswith(magic), "unexpected offload bundle format"
    found = []
    start = 0
    while True:
        start = blob.find(magic, start)
        if start < 0:
            break
        (count,) = struct.unpack_from("<Q", blob, start + len(magic))
        pos = start + len(magic) + 8
        for _ in range(count):
            eoff, esize, tlen = struct.unpack_from("<QQQ", blob, pos)
            triple = blob[pos + 24: pos + 24 + tlen].decode()
            pos += 24 + tlen
            if triple.endswith(arch):
                found.append(blob[start + eoff: start + eoff + esize])
        start += len(magic)
    assert found, f"no {arch} code object in the bundles"
    return found


def _kernels(co):
    for name, typ, off, size in _sections(co):
        if typ != 7:  # SHT_NOTE
            continue
        p = off
        while p < off + size:
            namesz, descsz, ntype = struct.unpack_from("<III", co, p)
            nname = co[p + 12: p + 12 + namesz].rstrip(b"\0")
            dpos = p + 12 + ((namesz + 3) & ~3)
            if nname == b"AMDGPU" and ntype == NT_AMDGPU_METADATA:
                meta = msgpack.unpackb(co[dpos: dpos + descsz], raw=False)
                return meta["amdhsa.kernels"]
            p = dpos + ((descsz + 3) & ~3)
    raise AssertionError("no AMDGPU metadata note in the code object")


@pytest.fixture(scope="module")
def kernels():
    if not LIB.exists():
        pytest.skip("libdfft.so is not built")
    return [k for co in _code_objects(LIB.read_bytes()) for k in _kernels(co)]


def test_metadata_lists_the_line_kernels(kernels):
    """Every kernel family is present, and each kernel is instantiated in one
    translation unit only (the line kernels are spread over several)."""
    names = [k[".name"] for k in kernels]
    assert len(names) == len(set(names)), "a kernel is instantiated in more than one unit"
    assert any("fft_line_fast" in n for n in names)
    assert any("fft_pair_fast" in n for n in names)
    assert any("fft_group_pass" in n for n in names)


def test_no_private_memory_in_power_of_two_line_kernels(kernels):
    """Fast kernels of lengths 256..2048 (the large-array passes) use no more
    than a few spill slots; any kernel stays far below a copy of its argument."""
    big = []
    for k in kernels:
        name, priv = k[".name"], k[".private_segment_fixed_size"]
        if ("fft_line_fast" in name or "fft_pair_fast" in name) and any(
                f", {n}," in name or f"Li{n}E" in name for n in (256, 512, 1024, 2048)):
            if priv > 64:
                big.append((name, priv))
        elif priv > 512:
            big.append((name, priv))
    assert not big, f"kernels with private memory: {big}"


OBJDUMP = Path("/opt/rocm/llvm/bin/llvm-objdump")


def test_no_flat_memory_instructions():
    """Every global access of the kernels compiles to global_* instructions.
    A flat access (an address the compiler cannot place in global memory, e.g.
    a pointer read from a table in device memory) also counts against the LDS
    wait counter, so each LDS wait of the radix stages would wait for the line
    loads in flight too (the grouped kernel ran 12 % slower that way)."""
    import subprocess
    import tempfile
    if not LIB.exists():
        pytest.skip("libdfft.so is not built")
    if not OBJDUMP.exists():
        pytest.skip("llvm-objdump is not installed")
    flat = {}
    with tempfile.TemporaryDirectory() as tmp:
        for i, co in enumerate(_code_objects(LIB.read_bytes())):
            path = Path(tmp) / f"co{i}.o"
            path.write_bytes(co)
            out = subprocess.run([str(OBJDUMP), "-d", "--no-show-raw-insn", str(path)], check=True,
                                 capture_output=True, text=True).stdout
            func = None
            for line in out.splitlines():
                if line.endswith(">:"):
                    func = line[line.find("<") + 1: -2]
                elif "flat_load" in line or "flat_store" in line or "flat_atomic" in line:
                    flat[func] = flat.get(func, 0) + 1
    assert not flat, f"flat memory instructions: {sorted(flat.items())[:10]}"
