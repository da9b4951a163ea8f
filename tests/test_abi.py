"""The C ABI: libbmx_hip.so loads and exports every entry point include/bmx.h declares.

No compute calls here (CPU-only environment); the GPU tests exercise them.
"""
import os
import re
import subprocess

from biomedisa_amd import _hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "bmx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(bmx_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_path():
    syms = declared_symbols()
    for need in ("bmx_open", "bmx_seeds", "bmx_walk", "bmx_argmax", "bmx_smooth_argmax",
                 "bmx_uncertainty", "bmx_plane_flags", "bmx_last_error"):
        assert need in syms


def test_library_exports_every_declared_symbol():
    lib = _hip.load()
    nm = subprocess.run(["nm", "-D", "--defined-only", _hip.LIB_PATH], capture_output=True,
                        text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(bmx_\w+)", nm))
    for s in declared_symbols():
        assert s in exported, s
        getattr(lib, s)


def test_bindings_cover_the_header():
    assert set(declared_symbols()) == set(_hip.SIGNATURES)


def test_last_error_starts_empty_and_version():
    lib = _hip.load()
    assert lib.bmx_version() >= 1
    assert isinstance(lib.bmx_last_error(), bytes)


def test_struct_layouts_match_the_header(tmp_path):
    """bmx_seed, bmx_walk_opts and bmx_volume as the bindings lay them out: sizes
    and field offsets from a C program compiled against include/bmx.h."""
    import ctypes
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "bmx.h"\n'
        '#define F(T, m) printf(#T "." #m " %zu\\n", offsetof(T, m))\n'
        'int main(void) {\n'
        ' printf("seed %zu opts %zu volume %zu\\n", sizeof(bmx_seed), sizeof(bmx_walk_opts), sizeof(bmx_volume));\n'
        ' F(bmx_seed, var); F(bmx_seed, reserved);\n'
        ' F(bmx_walk_opts, seed_overrides); F(bmx_walk_opts, tile_shape); F(bmx_walk_opts, graph_rounds);\n'
        ' F(bmx_volume, Z); F(bmx_volume, nlabels); F(bmx_volume, hits_offset); F(bmx_volume, seed_offset);\n'
        ' F(bmx_volume, nseeds);\n'
        ' return 0; }\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = dict(zip(out[0].split()[0::2], map(int, out[0].split()[1::2])))
    assert sizes == {"seed": _hip.SEED_DTYPE.itemsize, "opts": ctypes.sizeof(_hip.WalkOpts),
                     "volume": _hip.VOLUME_DTYPE.itemsize}
    offs = dict(line.split() for line in out[1:] if line)
    assert int(offs["bmx_seed.var"]) == _hip.SEED_DTYPE.fields["var"][1]
    assert int(offs["bmx_seed.reserved"]) == _hip.SEED_DTYPE.fields["reserved"][1]
    for f in ("seed_overrides", "tile_shape", "graph_rounds"):
        assert int(offs["bmx_walk_opts." + f]) == getattr(_hip.WalkOpts, f).offset
    for f in ("Z", "nlabels", "hits_offset", "seed_offset", "nseeds"):
        assert int(offs["bmx_volume." + f]) == _hip.VOLUME_DTYPE.fields[f][1]
