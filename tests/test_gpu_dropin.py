"""The C++ drop-in headers (include/openol.h) on the GPU: the program
tests/cpp/dropin_test calls ol:: functions with the reference's signatures; every
output is compared with the oracle on the same seeded inputs."""
import os
import subprocess

import numpy as np
import pytest

import olref
from openol_amd import synth

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EXE = os.path.join(ROOT, "tests", "cpp", "dropin_test")
LAM, P = 532e-9, 8e-6


@pytest.mark.gpu
def test_dropin_program_matches_oracle(tmp_path):
    assert os.path.exists(EXE), "build() compiles tests/cpp/dropin_test"
    n, m = 128, 300
    u0 = synth.random_field(n, n, seed=80)
    pts = synth.random_points(m, seed=81, xmax=n * P, ymax=n * P, zmin=0.152, zmax=0.155)
    xyz = np.stack([pts["x"], pts["y"], pts["z"]], axis=1)
    u0.tofile(tmp_path / "field.bin")
    xyz.astype(np.float64).tofile(tmp_path / "points.bin")
    (tmp_path / "params.txt").write_text(f"{n} {m}\n")
    u4k = synth.random_field(2160, 3840, seed=82)
    u4k.tofile(tmp_path / "field4k.bin")
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "done" in r.stdout and "caught:" in r.stdout
    assert "allocation changed" not in r.stdout
    rd = lambda name, dt=np.complex64, k=1: np.fromfile(tmp_path / f"{name}.bin", dtype=dt).reshape(k * n, k * n)
    rel = olref.rel_l2
    # _CghFresnelDiff of the same cloud in pixel units: phase, amplitude-weighted
    ref = olref.cgh_fresnel_diff(xyz[:, 0] / P, xyz[:, 1] / P, xyz[:, 2] / P, n, n, P, LAM)
    got = np.fromfile(tmp_path / "cgh_diff_phase.bin", dtype=np.float32).reshape(n, n)
    dphi = np.abs(np.angle(np.exp(1j * (got - np.angle(ref)))))
    assert np.sum(dphi * np.abs(ref)) / np.sum(np.abs(ref)) < 1e-4
    # tests/Prop_test.cpp:46: FresnelProp(u, 2160, 3840, p, p, lambda, -d), float parameters
    f4k = np.fromfile(tmp_path / "fresnel4k.bin", dtype=np.complex64).reshape(2160, 3840)
    pf, lf, zf = (float(np.float32(v)) for v in (3.74e-6, 532e-9, 0.1))
    assert rel(f4k, olref.fresnel_prop(u4k, pf, pf, lf, zf)) < 1e-5
    assert rel(rd("asm"), olref.asm_prop(u0, P, P, LAM, 0.05)) < 1e-5
    # complex<double> fields: the FP64 passes against the FP64 oracle (the fftw3 path);
    # 1e-9 for ASM, whose 1e5-cycle phase carries ~1e-11 cycle of FP64 rounding
    u0d = u0.astype(np.complex128)
    assert rel(rd("asm_c128", np.complex128), olref.asm_prop(u0d, P, P, LAM, 0.05)) < 1e-9
    assert rel(rd("fresnel_c128", np.complex128), olref.fresnel_prop(u0d, P, P, LAM, 0.2)) < 1e-12
    assert rel(rd("fresnel_bl_c128", np.complex128),
               olref.fresnel_prop(u0d, P, P, LAM, 0.002, olref.FRESNEL_BL_CPU)) < 1e-12
    chosen = olref.prop(u0, P, P, LAM, 0.05)[1]
    ref = olref.asm_prop(u0d, P, P, LAM, 0.05) if chosen == 0 else olref.fresnel_prop(
        u0d, P, P, LAM, 0.05, 0 if chosen == 1 else olref.FRESNEL_BL_CPU)
    assert rel(rd("prop_c128", np.complex128), ref) < 1e-9
    m2 = (2 * n) ** 2
    assert rel(rd("masm_c128", np.complex128), olref.asm_prop(u0d, P, P, LAM, 0.05) / m2) < 1e-9
    assert rel(rd("fresnel"), olref.fresnel_prop(u0, P, P, LAM, 0.2)) < 1e-5
    assert rel(rd("fresnel_bl"), olref.fresnel_prop(u0, P, P, LAM, 0.002, olref.FRESNEL_BL_CPU)) < 1e-5
    assert rel(rd("prop"), olref.prop(u0, P, P, LAM, 0.05)[0]) < 1e-5
    assert rel(rd("shifted"), olref.shifted_fresnel_prop(u0, P, P, LAM, 0.1, 0.5, 1e-4, -2e-4)) < 1e-5
    assert rel(rd("gasm"), olref.asm_prop(u0, P, P, LAM, 0.05)) < 1e-5
    assert rel(rd("arss"), olref.arss_prop(u0, P, P, LAM, 0.15, 0.5)) < 1e-5
    pf, lf = float(np.float32(P)), float(np.float32(LAM))
    f32 = lambda v: float(np.float32(v))
    assert rel(rd("shifted_asm"), olref.shifted_asm(u0, pf, pf, lf, f32(0.004), 0.0, f32(0.5 * n * pf),
                                                   olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)) < 1e-5
    assert rel(rd("split_cpu"), olref.split_prop(u0, pf, pf, lf, f32(0.002), 2, 2, 0,
                                                 olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)) < 1e-5
    assert rel(rd("gsplit"), olref.split_prop(u0, pf, pf, lf, f32(0.002), 2, 2, 0, 0)) < 1e-5
    assert rel(rd("gsplit_fresnel"), olref.split_prop(u0, pf, pf, lf, f32(0.05), 2, 4, 1, 0)) < 1e-5
    assert rel(rd("masm"), olref.asm_prop(u0, P, P, LAM, 0.05) / (4 * n * n)) < 1e-5
    F = np.fft.fft2(u0.astype(np.complex128))
    assert rel(rd("hfft"), F) < 1e-5
    assert rel(rd("hifft"), u0 * (n * n)) < 1e-5
    ang = float(np.float32(np.pi / 180 * 2))
    assert rel(rd("tilted_ref"), olref.tilted(u0, pf, pf, lf, f32(0.004), ang, "bicubic", "asm",
                                              "reference")) < 1e-5
    ang = float(np.float32(np.pi / 180 * -3))
    assert rel(rd("tilted_src"), olref.tilted(u0, pf, pf, lf, f32(0.004), ang, "bilinear", "asm",
                                              "source")) < 1e-5
    assert rel(rd("diffuser"), olref.diffuser_random_phase(u0, olref.PLAY, 3, float(np.float32(2 * np.pi)),
                                                        1, -2)) < 1e-6
    assert rel(rd("lens_mult"), u0 * olref.lens_array(n, n, pf, 16 * pf, f32(0.02), lf, olref.RECORD, 2,
                                                     3)) < 1e-6
    assert rel(rd("lens_circle"), olref.lens_array(n, n, pf, 16 * pf, f32(0.02), lf, olref.PLAY, 0, 0,
                                                   f32(40e-6), True)) < 1e-6
    assert rel(rd("gshifted"), olref.shifted_asm(u0, P, P, LAM, 0.004, -0.25 * n * P, 0.5 * n * P)) < 1e-5
    assert rel(rd("shiftedprop_fresnel"), olref.arss_prop(u0, P, P, LAM, 0.05, 1.0, 1e-4, 0.0)) < 1e-5
    cgh = olref.cgh(pts["x"], pts["y"], pts["z"], n, n, P, P, LAM, dtype=np.complex128)
    assert rel(rd("cgh"), cgh) < 1e-4 and rel(rd("gcgh"), cgh) < 1e-4
    ph = rd("cgh_phase", np.float32)
    assert np.mean(np.abs(np.angle(np.exp(1j * (ph - np.angle(cgh)))))) < 1e-3
    # Cgh(..., normalized=true): pixel units, x read for both axes
    pn = olref.cgh_normalized(pts["x"] / P, pts["y"] / P, pts["z"] / P, n, n, P, LAM)
    assert np.mean(np.abs(np.angle(np.exp(1j * (rd("cgh_norm", np.float32) - pn))))) < 1e-3
    assert rel(rd("wrp"), olref.wrp(pts["x"], pts["y"], pts["z"], n, n, P, LAM, 0.15,
                                    dtype=np.complex128)) < 1e-5
    pz = float([l for l in r.stdout.splitlines() if l.startswith("plane_z")][0].split()[1])
    assert pz == olref.wrp_plane_z(n, n, P, LAM, 0.151, 0.99)
    w = olref.wrp(pts["x"], pts["y"], pts["z"], n, n, P, LAM, pz)
    # WRPStep2 AUTO (include/golcgh.h:494-517): ASM if AsmPropCheck or not FresnelPropCheck
    use_asm = olref.asm_prop_check(n, n, P, P, LAM, pz) or not olref.fresnel_prop_check(n, n, P, P, LAM, pz)
    step2 = olref.asm_prop(w, P, P, LAM, pz) if use_asm else olref.fresnel_prop(w, P, P, LAM, pz)
    assert rel(rd("wrpm"), step2) < 1e-5
    # tables: AsmTransferF (centred, no band limit) and FresnelResponseBandLimit
    M = 2 * n
    du = 1 / (P * M)
    c = (np.arange(M) - M // 2)
    wv = np.sqrt(1 / LAM ** 2 - (c[None, :] * du) ** 2 - (c[:, None] * du) ** 2)
    assert rel(rd("H", k=2), np.exp(2j * np.pi * wv * 0.05)) < 1e-6
    lim = int(np.floor(abs(LAM * 0.002 * 0.5 / P) / P))
    x = c * P
    hh = np.exp(1j * np.pi * (x[None, :] ** 2 + x[:, None] ** 2) / (LAM * 0.002))
    hh *= (np.abs(c)[None, :] < lim) & (np.abs(c)[:, None] < lim)
    assert rel(rd("h", k=2), hh) < 1e-6
    # the same tables in HBM (gAsmTransferF, gFresnelResponse(BandLimit)), and the shifted
    # transfer function: host form with the off-axis window, device form without
    assert rel(rd("gH", k=2), np.exp(2j * np.pi * wv * 0.05)) < 1e-6
    hf = np.exp(1j * np.pi * (x[None, :] ** 2 + x[:, None] ** 2) / (LAM * 0.002))
    assert rel(rd("gh", k=2), hf) < 1e-6
    assert rel(rd("ghbl", k=2), hh) < 1e-6
    hs = olref.shifted_asm_transfer(M, M, du, du, LAM, 0.05, 2e-4, -5e-4, olref.SHIFT_WINDOW)
    assert rel(rd("Hs", k=2), hs) < 1e-6
    assert rel(rd("gHs", k=2), olref.shifted_asm_transfer(M, M, du, du, LAM, 0.05, 2e-4, -5e-4)) < 1e-6
    # the check functions print their limits
    zmax = float([l for l in r.stdout.splitlines() if l.startswith("zmax (from x)")][0].split("=")[1])
    umax = du * n
    assert abs(zmax - np.sqrt(1 / LAM ** 2 - 2 * umax ** 2) / (2 * umax * du)) < 1e-3 * zmax
    assert any(l.startswith("widthlim[px] = ") for l in r.stdout.splitlines())
