"""CPU checks of the C ABI: the library loads, exports every symbol include/olx.h
declares, answers host-only queries, and refuses to run without a GPU."""
import ctypes
import os
import re

import pytest

from openol_amd import _lib

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "olx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(olx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers them all
    assert set(names) <= set(_lib.SIGNATURES) | {"olx_last_error"}


def test_host_only_entry_points():
    L = _lib.load()
    assert L.olx_version() >= 10000
    # AsmPropCheck / FresnelPropCheck are pure host arithmetic
    assert L.olx_asm_prop_check(1024, 1024, 8e-6, 8e-6, 532e-9, 0.1) == 1
    assert L.olx_asm_prop_check(1024, 1024, 8e-6, 8e-6, 532e-9, 10.0) == 0
    assert L.olx_fresnel_prop_check(64, 64, 8e-6, 8e-6, 532e-9, 1.0) == 1
    assert abs(L.olx_wrp_plane_z(64, 64, 8e-6, 532e-9, 0.15, 8e-6, ctypes.c_float(0.9), 0)
               - 0.13499999046325684) < 1e-15
    assert L.olx_layer_depth(0.1234, 0.10, 0.20, 64) == pytest.approx(0.10 + 14.5 * 0.1 / 64)
    assert L.olx_layer_depth(0.25, 0.10, 0.20, 64) == pytest.approx(0.10 + 63.5 * 0.1 / 64)


def test_null_context_is_an_error_not_a_crash():
    L = _lib.load()
    rc = L.olx_asm_c64(None, None, 16, 16, 8e-6, 8e-6, 532e-9, 0.1, 0)
    assert rc == _lib.EINVAL
    assert "context" in _lib.last_error()


def test_ops_refuse_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.OlxError):
        _lib.ctx()


def test_host_logic_matches_oracle():
    import olref
    L = _lib.load()
    for (ny, nx, d, z) in [(64, 64, 8e-6, 0.01), (64, 64, 8e-6, 0.05), (128, 32, 4e-6, 0.002)]:
        assert bool(L.olx_asm_prop_check(ny, nx, d, d, 532e-9, z)) == olref.asm_prop_check(ny, nx, d, d, 532e-9, z)
        assert bool(L.olx_fresnel_prop_check(ny, nx, d, d, 532e-9, z)) == olref.fresnel_prop_check(ny, nx, d, d, 532e-9, z)
    for s, sq in [(0.9, 0), (-1.0, 0), (-1.0, 1), (0.5, 1)]:
        assert L.olx_wrp_plane_z(512, 256, 8e-6, 532e-9, 0.15, 8e-6, ctypes.c_float(s), sq) == \
            olref.wrp_plane_z(512, 256, 8e-6, 532e-9, 0.15, s, square=bool(sq))


def test_dropin_headers_compile():
    """The C++ drop-in headers are valid standalone C++17 (no HIP headers needed)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    for prog in ("dropin_test.cpp", "ingest_test.cpp"):
        src = os.path.join(ROOT, "tests", "cpp", prog)
        r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                            "-I" + os.path.join(ROOT, "include"), src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
