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
