"""GPU checks of the float-PREC_T compatibility mode (OLX_PREC_F32): the reference's
tests compile with typedef float PREC_T (tests/para.h:8), which rounds the ASM phase
(up to ~1e6 rad) and the CGH angle to float.  The mode must reproduce the oracle's
emulation of that arithmetic (OLREF_PREC_F32) -- relative L2 <= 1e-5 for ASM, <= 1e-4
for CGH -- while staying far from the FP64-phase default, and the drop-in must select
it for PREC_T = float when OL_PREC_T_FLOAT_COMPAT is defined."""
import os
import subprocess

import numpy as np
import pytest
import torch

import olref
from openol_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
LAM = 532e-9


def ol():
    import openol_amd
    return openol_amd


@pytest.mark.parametrize("shape,z,bl", [((64, 64), 0.01, False), ((1024, 1024), 0.10, False),
                                        ((2048, 1024), 0.10, True), ((270, 480), 0.05, False)])
def test_asm_prec_f32_matches_oracle(shape, z, bl):
    u0 = synth.random_field(*shape, seed=95)
    u = torch.from_numpy(u0).cuda()
    ol().AsmProp(u, *shape, 8e-6, 8e-6, LAM, z, bandlimit=bl, prec_f32=True)
    fl = olref.PREC_F32 | (olref.BANDLIMIT if bl else 0)
    ref = olref.asm_prop(u0, 8e-6, 8e-6, LAM, z, fl)
    got = u.cpu().numpy()
    err = olref.rel_l2(got, ref)
    d64 = olref.rel_l2(got, olref.asm_prop(u0.astype(np.complex128), 8e-6, 8e-6, LAM, z,
                                           olref.BANDLIMIT if bl else 0))
    print(f"ASM PREC_T=float {shape}: rel L2 vs float oracle {err:.3e}, vs FP64 {d64:.3e}")
    assert err < 1e-5
    if z >= 0.05:
        assert d64 > 1e-4  # the float phase is far from FP64, as in the reference


def test_cgh_prec_f32_matches_oracle():
    ny, nx, n = 96, 128, 500
    pts = synth.random_points(n, seed=96, xmax=nx * 8e-6, ymax=ny * 8e-6, zmin=0.10, zmax=0.20)
    o = ol().Object(pts["x"], pts["y"], pts["z"])
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().Cgh(o, U, ny, nx, 8e-6, 8e-6, LAM, prec_f32=True)
    ref = olref.cgh(pts["x"], pts["y"], pts["z"], ny, nx, 8e-6, 8e-6, LAM, flags=olref.PREC_F32)
    assert olref.rel_l2(U.cpu().numpy(), ref) < 1e-4
    with pytest.raises(ValueError):
        ol().Cgh(o, U, ny, nx, 8e-6, 8e-6, LAM, exact_r=True, prec_f32=True)


def test_dropin_float_prec_compat(tmp_path):
    exe = os.path.join(ROOT, "tests", "cpp", "compat_test")
    assert os.path.exists(exe), "build() compiles tests/cpp/compat_test"
    n, m = 128, 200
    u0 = synth.random_field(n, n, seed=97)
    pts = synth.random_points(m, seed=98, xmax=n * 8e-6, ymax=n * 8e-6, zmin=0.10, zmax=0.20)
    xyz = np.stack([pts["x"], pts["y"], pts["z"]], axis=1).astype(np.float32)
    u0.tofile(tmp_path / "field.bin")
    xyz.tofile(tmp_path / "points.bin")
    (tmp_path / "params.txt").write_text(f"{n} {m}\n")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "done" in r.stdout
    f32 = lambda v: float(np.float32(v))
    got = np.fromfile(tmp_path / "asm_f32.bin", dtype=np.complex64).reshape(n, n)
    ref = olref.asm_prop(u0, f32(8e-6), f32(8e-6), f32(LAM), f32(0.05), olref.PREC_F32)
    assert olref.rel_l2(got, ref) < 1e-5
    got = np.fromfile(tmp_path / "cgh_f32.bin", dtype=np.complex64).reshape(n, n)
    x, y, z = (xyz[:, k].astype(np.float64) for k in range(3))
    ref = olref.cgh(x, y, z, n, n, f32(8e-6), f32(8e-6), f32(LAM), flags=olref.PREC_F32)
    assert olref.rel_l2(got, ref) < 1e-4
