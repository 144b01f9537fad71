"""Benchmark: 4096^2 angular-spectrum propagations per second on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--extras]

A step is one band-limited ASM propagation of a seeded 4096 x 4096 complex64
field (lambda 532 nm, 8 um pitch, z = 0.10 m, zero-padded to 8192 x 8192,
cropped back), in place, with the field already resident in HBM.  Prints ONE
JSON line.  `value` is propagations per second.  The line also carries:

* roofline: the dominant kernel of the three-pass propagation, its algorithmic
  HBM bytes per launch over its average launch time (HIP events recorded on the
  stream the kernels run on, inside the timed region), against 8 TB/s;
* cpu_baseline: the oracle's CPU restatement of the reference AsmProp (own
  radix-2 FFT with OpenMP; FFTW is not installed) timed on this host;
* extras (on by default): direct CGH, layered MFMA CGH, WRP + ARSS, the
  multi-plane optimisation step, point-cloud objset / objsort and the 8-bit
  encodings, each with its own throughput.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LAM, PITCH = 532e-9, 8e-6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=4096, help="field side (metric: 4096)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="only the ASM metric")
    ap.add_argument("--only", choices=["ingest", "next", "cgh", "sizes"], default=None,
                    help="run one secondary workload alone and print its figures (profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the plain copy kernel (for PMC byte-counter calibration)")
    return ap.parse_args()


def pass_bytes(n):
    """Algorithmic HBM bytes per launch of each propagation pass (ny = nx = n,
    padded row length 2n, complex64): A reads n^2 c64 and writes n x 2n c64; B
    reads and writes n x 2n c64; C reads n x 2n and writes n^2."""
    return {"prop_rows_pad_fwd": 24 * n * n, "prop_cols_conv": 32 * n * n,
            "prop_rows_crop_inv": 24 * n * n}


TIMING_PERIOD = 8


# Untimed propagations before the warm-up steps: a kernel trace of 200 back-to-back
# propagations at 4096^2 shows the first ~40 (about 20 ms of GPU work) running on a
# rising clock (the column pass falls from 311 to a steady 260 us), so the bench
# first keeps the GPU busy for SETTLE_S seconds of propagations.
SETTLE_S = 0.25


def settle(ol, torch, u, n):
    t0 = time.perf_counter()
    count = 0
    while time.perf_counter() - t0 < SETTLE_S:
        for _ in range(10):
            ol.AsmProp(u, n, n, PITCH, PITCH, LAM, 0.10, bandlimit=True)
        torch.cuda.synchronize()
        count += 10
    return count


def bench_asm(ol, lib, torch, n, steps, warmup, settle_first=False):
    from openol_amd import synth
    u = torch.from_numpy(synth.random_field(n, n, seed=0)).cuda()
    L = lib.load()
    c = lib.ctx()
    settled = settle(ol, torch, u, n) if settle_first else 0
    for _ in range(warmup):
        ol.AsmProp(u, n, n, PITCH, PITCH, LAM, 0.10, bandlimit=True)
    torch.cuda.synchronize()
    L.olx_ctx_reset_timing(c)
    # HIP events bracket every 8th launch (each of the three passes is sampled in
    # turn): timing every launch adds about 20 us per propagation to the steps
    L.olx_ctx_enable_timing(c, TIMING_PERIOD)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ol.AsmProp(u, n, n, PITCH, PITCH, LAM, 0.10, bandlimit=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    L.olx_ctx_enable_timing(c, 0)
    rep = lib.timing_report()
    assert torch.isfinite(torch.view_as_real(u)).all()
    if settle_first:
        rep["_settle_props"] = (settled, 0.0)
    return dt, rep


# Per-launch counters of the three passes at n = 4096 with the default work-group
# shapes, measured with rocprofv3 --pmc: HBM bytes = FETCH_SIZE x 2 + WRITE_SIZE x 1
# (the factors a plain 512 MiB copy kernel reads back), and SQ_INSTS_VALU.
MEASURED_4096 = {
    "prop_rows_pad_fwd": {"traffic": 403.2e6, "valu_instr": 24.94e6},
    "prop_cols_conv": {"traffic": 539.5e6, "valu_instr": 113.12e6},
    "prop_rows_crop_inv": {"traffic": 403.3e6, "valu_instr": 25.59e6},
}
# wave64 VALU instructions per second: 1024 SIMDs x 2.4 GHz / 4 cycles (no packed FP32)
VALU_PEAK = 1024 * 2.4e9 / 4


def roofline_of(rep, n):
    pb = pass_bytes(n)
    kern = {k: v for k, v in rep.items() if k in pb}
    dom = max(kern, key=lambda k: kern[k][1])
    cnt, tot = kern[dom]
    avg_s = tot / cnt / 1e3
    achieved = pb[dom] / avg_s / 1e9
    per_kernel = {k: {"avg_us": round(v[1] / v[0] * 1e3, 2),
                      "GBps": round(pb[k] / (v[1] / v[0] / 1e3) / 1e9, 1)} for k, v in kern.items()}
    # the counters hold for the default work-group shapes only
    meas = MEASURED_4096.get(dom) if (n == 4096 and not os.environ.get("OLX_TILE")
                                     and not os.environ.get("OLX_ORDER")) else None
    out = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
           "traffic": round(meas["traffic"]) if meas else None,
           "algorithmic_bytes_per_launch": pb[dom], "avg_launch_us": round(avg_s * 1e6, 2),
           "per_kernel": per_kernel}
    if meas:
        # the column pass is limited by VALU issue, not by HBM: report that roofline too
        rate = meas["valu_instr"] / avg_s
        out["valu"] = {"instr_per_launch": meas["valu_instr"], "achieved": round(rate / 1e9, 1),
                       "peak": round(VALU_PEAK / 1e9, 1), "unit": "G wave64-instr/s",
                       "frac": round(rate / VALU_PEAK, 4)}
    return out


def gpu_outputs(ol, torch, n):
    """GPU results the CPU leg checks against the oracle: one AsmProp of the metric's
    seeded field, and the first 16 rows of the direct CGH of 10,000 points over 2048^2
    (Fresnel and exact r)."""
    from openol_amd import synth
    u = torch.from_numpy(synth.random_field(n, n, seed=0)).cuda()
    ol.AsmProp(u, n, n, PITCH, PITCH, LAM, 0.10, bandlimit=True)
    N, P = 2048, 10000
    pts = synth.random_points(P, seed=2, xmax=N * PITCH, ymax=N * PITCH, zmin=0.10, zmax=0.20)
    obj = ol.Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    out = {"asm": u.cpu().numpy()}
    for exact in (False, True):
        ol.Cgh(obj, U, N, N, PITCH, PITCH, LAM, exact_r=exact)
        out["cgh_exact" if exact else "cgh_fresnel"] = U[:16].cpu().numpy()
    del U, u
    # WRP of 200,000 points into 4096^2 (sampled pixels) and its ARSS (whole plane)
    N, P = 4096, 200_000
    pts = synth.random_points(P, seed=3, xmax=N * PITCH, ymax=N * PITCH, zmin=0.151, zmax=0.155)
    obj = ol.Object(pts["x"], pts["y"], pts["z"])
    W = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    ol.WRPStep1(obj, W, N, N, PITCH, LAM, 0.15)
    ph, pw = check_pixels(N, 31)
    out["wrp"] = (pts, ph, pw, W.cpu().numpy())
    ol.ARSSFresnelProp(W, W, N, N, PITCH, PITCH, LAM, 0.15, 0.5)
    out["arss"] = W.cpu().numpy()
    # layered CGH of 1,000,000 points in 64 layers over 4096^2 (sampled pixels)
    n, nl, zmin, zmax = 1_000_000, 64, 0.10, 0.20
    pts = synth.random_points(n, seed=4, xmax=N * PITCH, ymax=N * PITCH, zmin=zmin, zmax=zmax)
    obj = ol.Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    ph, pw = check_pixels(N, 32)
    lay = {}
    for prec in ("f16", "bf16x3", "bf16"):
        ol.CghLayered(obj, W, N, N, PITCH, PITCH, LAM, nl, zmin, zmax, precision=prec)
        lay[prec] = W.cpu().numpy()[ph, pw]
    out["layered"] = (pts, ph, pw, lay)
    return out


def check_pixels(n, seed, count=4096):
    """Seeded pixel sample of an n x n plane for the sampled FP64 checks."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, count), rng.integers(0, n, count)


def cpu_baseline(n, gpu=None):
    """The oracle's AsmProp (complex64 field, PREC_T = double) on the host cores:
    1 warm-up, then the median of 5 runs, at the metric size.  With the GPU results
    of gpu_outputs, also their relative L2 error against the oracle: its complex64
    restatement (PREC_T = double) and its FP64 (complex128) form."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import olref
    from openol_amd import synth
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    u = synth.random_field(n, n, seed=0)
    olref.asm_prop(u, PITCH, PITCH, LAM, 0.10, olref.BANDLIMIT)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        ref_asm = olref.asm_prop(u, PITCH, PITCH, LAM, 0.10, olref.BANDLIMIT)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    errors = None
    if gpu is not None:
        ref64 = olref.asm_prop(u.astype(np.complex128), PITCH, PITCH, LAM, 0.10, olref.BANDLIMIT)
        errors = {"asm_metric_vs_c64_restatement": float(f"{olref.rel_l2(gpu['asm'], ref_asm):.3e}"),
                  "asm_metric_vs_fp64": float(f"{olref.rel_l2(gpu['asm'], ref64):.3e}")}
        del ref64
    # the ingest / encoding rows on one core (numpy restatements): objsort of the
    # 1,000,000-point cloud and complex2img (AMP) of a 4096^2 field
    pts = synth.random_points(1_000_000, seed=4, xmax=1.0, ymax=1.0, zmin=-1.0, zmax=1.0)
    t0 = time.perf_counter()
    olref.objsort_order(pts["z"])
    t_sort = time.perf_counter() - t0
    rng = np.random.default_rng(7)
    u4 = (rng.standard_normal((4096, 4096)) + 1j * rng.standard_normal((4096, 4096))).astype(np.complex64)
    t0 = time.perf_counter()
    olref.complex2img(u4, olref.IMG_AMP)
    t_img = time.perf_counter() - t0
    # direct CGH (10,000 points over 2048^2, Fresnel, complex64) on a 16-row band of
    # the hologram, extrapolated linearly in rows
    N, P = 2048, 10000
    pts = synth.random_points(P, seed=2, xmax=N * PITCH, ymax=N * PITCH, zmin=0.10, zmax=0.20)
    rows = 16
    t0 = time.perf_counter()
    band = olref.cgh(pts["x"], pts["y"], pts["z"], rows, N, PITCH, PITCH, LAM, pts["amp"], pts["phase"])
    t_cgh = time.perf_counter() - t0
    if gpu is not None:
        # the device stores amplitude and phase as float32
        amp = pts["amp"].astype(np.float32).astype(np.float64)
        ph = pts["phase"].astype(np.float32).astype(np.float64)
        errors["cgh_10k_2048_fresnel_rows_vs_c64_restatement"] = float(
            f"{olref.rel_l2(gpu['cgh_fresnel'], band):.3e}")
        for name, fl in (("fresnel", 0), ("exact", olref.CGH_EXACT_R)):
            b64 = olref.cgh(pts["x"], pts["y"], pts["z"], rows, N, PITCH, PITCH, LAM, amp, ph,
                            flags=fl, dtype=np.complex128)
            errors[f"cgh_10k_2048_{name}_rows_vs_fp64"] = float(f"{olref.rel_l2(gpu['cgh_' + name], b64):.3e}")
        fmt = lambda v: float(f"{v:.3e}")
        # WRP: FP64 per-pixel sums on the sampled pixels; ARSS of the GPU's WRP plane
        wpts, ph, pw, Wg = gpu["wrp"]
        ref = olref.wrp_pixels(wpts["x"], wpts["y"], wpts["z"], ph, pw, PITCH, LAM, 0.15)
        errors["wrp_200k_4096_sampled_vs_fp64"] = fmt(olref.rel_l2(Wg[ph, pw], ref))
        errors["arss_4096_vs_c64_restatement"] = fmt(olref.rel_l2(
            gpu["arss"], olref.arss_prop(Wg, PITCH, PITCH, LAM, 0.15, 0.5)))
        errors["arss_4096_vs_fp64"] = fmt(olref.rel_l2(
            gpu["arss"], olref.arss_prop(Wg.astype(np.complex128), PITCH, PITCH, LAM, 0.15, 0.5)))
        # layered CGH: FP64 sums of the layer-snapped cloud on the sampled pixels
        lpts, ph, pw, lay = gpu["layered"]
        nl, zmin, zmax = 64, 0.10, 0.20
        lyr = np.clip(np.floor((lpts["z"] - zmin) / (zmax - zmin) * nl), 0, nl - 1)
        zl = zmin + (lyr + 0.5) * (zmax - zmin) / nl
        a32 = lpts["amp"].astype(np.float32).astype(np.float64)
        p32 = lpts["phase"].astype(np.float32).astype(np.float64)
        ref = olref.cgh_pixels(lpts["x"], lpts["y"], zl, ph, pw, PITCH, PITCH, LAM, a32, p32)
        for prec, got in lay.items():
            dphi = np.abs(np.angle(got * np.conj(ref)))
            errors[f"cgh_layered_{prec}_1M_4096_sampled_vs_fp64"] = fmt(olref.rel_l2(got, ref))
            errors[f"cgh_layered_{prec}_phase_err_wmean_rad_vs_fp64"] = fmt(
                np.sum(dphi * np.abs(ref)) / np.sum(np.abs(ref)))
            errors[f"cgh_layered_{prec}_phase_err_max_rad_vs_fp64"] = fmt(dphi.max())
    # one thread: AsmProp at 1024^2 (OpenMP pinned to one thread through libgomp)
    one = None
    try:
        import ctypes
        gomp = ctypes.CDLL("libgomp.so.1")
        gomp.omp_set_num_threads(1)
        u1 = synth.random_field(1024, 1024, seed=0)
        t0 = time.perf_counter()
        olref.asm_prop(u1, PITCH, PITCH, LAM, 0.10, olref.BANDLIMIT)
        one = round(1.0 / (time.perf_counter() - t0), 3)
        gomp.omp_set_num_threads(cores)
    except OSError:
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(1.0 / med, 4), "unit": "props/s", "cores": cores, "kind": "port",
            "sample": f"{n}x{n} AsmProp (band-limited, pad {2 * n}), 1 warm-up + median of 5 "
                      f"= {med:.2f} s/prop; own OpenMP radix-2 FFT (FFTW unavailable - substitute FFT)",
            "cpu_model": model, "nproc": os.cpu_count(),
            "extras": {"asm_1024_props_s_one_thread": one,
                       "cgh_direct_fresnel_Gptpx_s": round(P * rows * N / t_cgh / 1e9, 3),
                       "cgh_sample": f"{rows} rows x {N} px x {P} points (extrapolate linearly)",
                       "objsort_1M_ms": round(t_sort * 1e3, 1),
                       "complex2img_amp_4096_ms": round(t_img * 1e3, 1),
                       "numpy_rows": "objsort and complex2img: numpy, 1 core"},
            "rel_l2_vs_oracle": errors,
            "rel_l2_sample": "AsmProp at the metric size; direct CGH of 10,000 points over 2048^2, "
                             "rows 0-15; WRP of 200,000 points into 4096^2 and layered CGH of "
                             "1,000,000 points / 64 layers over 4096^2 on 4096 seeded pixels (FP64 "
                             "per-pixel sums, layered on the layer-snapped cloud); ARSS s = 0.5 of "
                             "the GPU's WRP plane, whole 4096^2 field"}


def ceilings(lib, torch):
    """Rates this device reaches on kernels that do nothing else: a plain copy
    (1 GiB) and the FMA / sine / MFMA issue loops of olx_peak_probe."""
    import ctypes
    L = lib.load()
    c = lib.ctx()
    out = {}
    nbytes = 1 << 30
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        L.olx_stream_copy(c, b.data_ptr(), a.data_ptr(), nbytes)
    torch.cuda.synchronize()
    L.olx_ctx_reset_timing(c)
    L.olx_ctx_enable_timing(c, 1)
    for _ in range(10):
        L.olx_stream_copy(c, b.data_ptr(), a.data_ptr(), nbytes)
    torch.cuda.synchronize()
    L.olx_ctx_enable_timing(c, 0)
    cnt, tot = lib.timing_report()["stream_copy"]
    out["copy_1GiB_TBps"] = round(2 * nbytes / (tot / cnt / 1e3) / 1e12, 3)
    del a, b
    r = ctypes.c_double()
    for name, kind, scale in (("fma_f32_TFLOPs", 0, 1e12), ("fma_f64_TFLOPs", 1, 1e12),
                              ("sin_f32_Gevals_s", 2, 1e9), ("mfma_f16_TFLOPs", 3, 1e12),
                              ("mfma_bf16_TFLOPs", 4, 1e12)):
        lib.check(L.olx_peak_probe(c, kind, ctypes.byref(r)), "peak_probe")
        out[name] = round(r.value / scale, 1)
    return out


def asm_sizes(ol, lib, torch):
    """ASM props/s at the smaller metric sizes (1024^2 and 2048^2 fields)."""
    out = {}
    for n in (1024, 2048):
        dt, _ = bench_asm(ol, lib, torch, n, 100, 10)
        out[f"asm_{n}_props_s"] = round(100 / dt, 1)
    return out


# VALU lane instructions per (pixel, point) interaction of the direct FP32 kernels,
# from rocprofv3 --pmc SQ_INSTS_VALU on the 10,000-point 2048^2 cloud (wave64
# instructions per launch x 64 / interactions): the separable Fresnel kernel
# (4 + 8 sine / cosine pairs per point for 32 pixels, then 4 FMA per pixel) and the
# exact-r kernel (series and square-root forms).
CGH_VALU_PER_INTERACTION = {"fresnel": 6.254, "exact": 10.77}


def direct_cgh(ol, torch):
    """Direct CGH of 10,000 points over 2048^2, exact r and Fresnel, the
    FP64 kernel, and each FP32 kernel's VALU issue roofline (both are bound by VALU
    issue: sine and cosine are a minority of their instructions)."""
    from openol_amd import synth
    out = {}
    N, n = 2048, 10000
    pts = synth.random_points(n, seed=2, xmax=N * PITCH, ymax=N * PITCH, zmin=0.10, zmax=0.20)
    obj = ol.Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    for exact in (True, False):
        kind = "exact" if exact else "fresnel"
        ol.Cgh(obj, U, N, N, PITCH, PITCH, LAM, exact_r=exact)
        torch.cuda.synchronize()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            ol.Cgh(obj, U, N, N, PITCH, PITCH, LAM, exact_r=exact)
        torch.cuda.synchronize()
        rate = n * N * N * reps / (time.perf_counter() - t0)
        out[f"cgh_direct_{kind}_Gptpx_s"] = round(rate / 1e9, 1)
        out[f"cgh_direct_{kind}_valu_frac"] = round(
            rate * CGH_VALU_PER_INTERACTION[kind] / 64 / VALU_PEAK, 3)
    # the FP64 kernel (exact r), complex128 output
    U64 = torch.empty((N, N), dtype=torch.complex128, device="cuda")
    ol.CghFP64(obj, U64, N, N, PITCH, PITCH, LAM, exact_r=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ol.CghFP64(obj, U64, N, N, PITCH, PITCH, LAM, exact_r=True)
    torch.cuda.synchronize()
    out["cgh_direct_exact_fp64_Gptpx_s"] = round(n * N * N / (time.perf_counter() - t0) / 1e9, 1)
    return out


def extras(ol, lib, torch):
    """Secondary workloads of the metric: direct vs layered CGH, WRP + ARSS,
    the smaller ASM sizes, the multi-plane optimisation and the device ceilings."""
    from openol_amd import synth
    out = {}
    out["ceilings"] = ceilings(lib, torch)
    out.update(asm_sizes(ol, lib, torch))
    out.update(direct_cgh(ol, torch))
    # WRP (200,000 points into 4096^2) + ARSS to the hologram
    N, n = 4096, 200000
    pts = synth.random_points(n, seed=3, xmax=N * PITCH, ymax=N * PITCH, zmin=0.151, zmax=0.155)
    obj = ol.Object(pts["x"], pts["y"], pts["z"])
    W = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    ol.WRPStep1(obj, W, N, N, PITCH, LAM, 0.15)
    ol.ARSSFresnelProp(W, W, N, N, PITCH, PITCH, LAM, 0.15, 0.5)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ol.WRPStep1(obj, W, N, N, PITCH, LAM, 0.15)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps = 20
    for _ in range(reps):
        ol.ARSSFresnelProp(W, W, N, N, PITCH, PITCH, LAM, 0.15, 0.5)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    out["wrp_200k_4096_ms"] = round((t1 - t0) * 1e3, 3)
    del W
    out.update(layered(ol, torch))
    out.update(multiplane(torch))
    out.update(ingest_encode(ol, torch))
    out.update(next_rows(ol, torch))
    out.update(sizes_batch(ol, torch))
    arss_s = (t2 - t1) / reps
    out["arss_4096_ms"] = round(arss_s * 1e3, 3)
    out["arss_4096_GBps_80N2"] = round(80 * N * N / arss_s / 1e9, 1)
    out["wrp_arss_4096_end_to_end_ms"] = round(out["wrp_200k_4096_ms"] + arss_s * 1e3, 3)
    return out


def _event_ms(torch, fn, reps):
    """Average device time of fn() over reps calls (HIP events on torch's current
    stream, which the library context is bound to), after one warm-up call."""
    fn()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def ingest_encode(ol, torch):
    """Point-cloud ingest and 8-bit encodings at full size, each against its HBM
    roofline (algorithmic bytes / time vs 8 TB/s): objset and objsort of the
    1,000,000-point layered cloud, complex2img (AMP: two reads of the field; PHASE:
    one) and phase8bit of a 4096^2 hologram."""
    from openol_amd import synth
    out = {}
    n = 1_000_000
    pts = synth.random_points(n, seed=4, xmax=1.0, ymax=1.0, zmin=-1.0, zmax=1.0)
    obj = ol.Object(pts["x"], pts["y"], pts["z"])
    # bounds (read 32 B / point) then the affine map (read + write 32 B / point)
    ms = _event_ms(torch, lambda: ol.objset(obj, 1e-3, 1e-3, (0.0, 0.0, 0.1), ol.ZMIDDLE), 20)
    out["objset_1M_ms"] = round(ms, 4)
    out["objset_1M_GBps"] = round(96 * n / ms / 1e6, 1)
    ms = _event_ms(torch, lambda: ol.objsort(obj), 10)
    out["objsort_1M_ms"] = round(ms, 3)
    out["objsort_1M_Mpts_s"] = round(n / ms / 1e3, 1)
    del obj
    N = 4096
    rng = np.random.default_rng(7)
    u = torch.from_numpy((rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
                         .astype(np.complex64)).cuda()
    img = torch.empty((N, N), dtype=torch.uint8, device="cuda")
    for mode, nbytes in (("amp", 2 * 8 * N * N + N * N), ("phase", 8 * N * N + N * N)):
        ms = _event_ms(torch, lambda: ol.complex2img(u, img, mode=mode), 20)
        out[f"complex2img_{mode}_4096_ms"] = round(ms, 4)
        out[f"complex2img_{mode}_4096_GBps"] = round(nbytes / ms / 1e6, 1)
        out[f"complex2img_{mode}_4096_frac"] = round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)
    ph = torch.angle(u).contiguous()
    ms = _event_ms(torch, lambda: ol.phase8bit(ph, img), 20)
    out["phase8bit_4096_ms"] = round(ms, 4)
    out["phase8bit_4096_GBps"] = round(5 * N * N / ms / 1e6, 1)
    return out


def next_rows(ol, torch):
    """Fresnel, off-axis, split and tilted propagation and the optical elements at full
    size: device time per call (HIP events) and, for the elementwise passes, GB/s."""
    out = {}
    N = 4096
    rng = np.random.default_rng(8)
    u0 = torch.from_numpy((rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)))
                          .astype(np.complex64)).cuda()
    u = u0.clone()
    # FresnelProp and the CPU-form band-limited variant (cached separable chirp spectra)
    f = u0.clone()
    # 50 calls: a 10-call window right after the memory-bound rows reads up to 15% high
    ms = _event_ms(torch, lambda: ol.FresnelProp(f, N, N, PITCH, PITCH, LAM, 0.3), 50)
    out["fresnel_4096_ms"] = round(ms, 3)
    ms = _event_ms(torch, lambda: ol.FresnelPropBandLimit(f, N, N, PITCH, PITCH, LAM, 0.3,
                                                          variant="cpu"), 50)
    out["fresnel_bl_4096_ms"] = round(ms, 3)
    # a depth sweep: every call a new distance, so both axis spectra are built each time
    zs = iter(0.2 + 1e-4 * k for k in range(1000))
    ms = _event_ms(torch, lambda: ol.FresnelProp(f, N, N, PITCH, PITCH, LAM, next(zs)), 10)
    out["fresnel_4096_new_z_ms"] = round(ms, 3)
    del f
    ms = _event_ms(torch, lambda: ol.gshiftedAsmProp(u, u, N, N, PITCH, PITCH, LAM, 0.01, 1e-3, -2e-3,
                                                     bandlimit=True), 50)
    out["shifted_asm_4096_ms"] = round(ms, 3)
    ms = _event_ms(torch, lambda: ol.gsplitProp(u, N, N, PITCH, PITCH, LAM, 0.01, ol.ASM, 2, 2), 3)
    out["split_asm_4096_2x2_ms"] = round(ms, 3)
    n2 = 2048
    v = u0[:n2, :n2].contiguous()
    for interp, name in ((ol.BILINEAR, "bilinear"), (ol.BICUBIC, "bicubic")):
        ms = _event_ms(torch, lambda: ol.TiltedReferenceProp(v, n2, n2, PITCH, PITCH, LAM, 0.01, 0.05,
                                                              interp), 5)
        out[f"tilted_ref_2048_{name}_ms"] = round(ms, 3)
    lens = ol.Lensarry(N, N, PITCH, 32 * PITCH, 0.05, LAM)
    ms = _event_ms(torch, lambda: lens.mult_complex(u), 20)
    out["lens_mult_4096_ms"] = round(ms, 4)
    out["lens_mult_4096_GBps"] = round(16 * N * N / ms / 1e6, 1)
    d = ol.Diffuser(N, N)
    b = ol.uniform(N * N, seed=1, device=u.device.index)
    ms = _event_ms(torch, lambda: d._apply(u, "phase", b, 6.283), 20)
    out["diffuser_apply_4096_GBps"] = round(20 * N * N / ms / 1e6, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.random_phase(u, ol.PLAY, seed=1)
    torch.cuda.synchronize()
    out["diffuser_random_phase_4096_ms_host_stream"] = round((time.perf_counter() - t0) * 1e3, 1)
    return out


def sizes_batch(ol, torch):
    """The reference's demo sizes (mixed-radix passes), batched small fields, the lens
    fused into the first pass, and the complex128 passes: device time per call (HIP
    events)."""
    from openol_amd import _lib
    out = {}
    rng = np.random.default_rng(9)

    def field(ny, nx, dt=np.complex64):
        return torch.from_numpy((rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx)))
                                .astype(dt)).cuda()
    # tests/Prop_test.cpp:46: FresnelProp(u, 2160, 3840, 3.74e-6, 3.74e-6, 532e-9, d)
    f = field(2160, 3840)
    ms = _event_ms(torch, lambda: ol.FresnelProp(f, 2160, 3840, 3.74e-6, 3.74e-6, LAM, 0.1), 10)
    out["fresnel_2160x3840_ms"] = round(ms, 3)
    out["fresnel_2160x3840_GBps_80N2"] = round(80 * 2160 * 3840 / ms / 1e6, 1)
    g = field(1080, 1920)
    ms = _event_ms(torch, lambda: ol.AsmProp(g, 1080, 1920, PITCH, PITCH, LAM, 0.05, bandlimit=True), 10)
    out["asm_1080x1920_ms"] = round(ms, 3)
    del f, g
    # 16 fields of 1024^2: one batched call against 16 calls
    nb, n = 16, 1024
    b = field(nb * n, n).view(nb, n, n)
    ms = _event_ms(torch, lambda: ol.prop_batch(b, PITCH, PITCH, LAM, 0.1, ol.ASM, _lib.BANDLIMIT), 10)
    out["asm_1024_batch16_props_s"] = round(nb / ms * 1e3, 1)

    def singles():
        for k in range(nb):
            ol.AsmProp(b[k], n, n, PITCH, PITCH, LAM, 0.1, bandlimit=True)
    ms = _event_ms(torch, singles, 10)
    out["asm_1024_single_props_s"] = round(nb / ms * 1e3, 1)
    del b
    # a lens array multiplied in by pass A against a separate multiply, at 4096^2
    N = 4096
    u = field(N, N)
    lens = ol.Lensarry(N, N, PITCH, 32 * PITCH, 0.05, LAM).gen_complex(ol.PLAY)
    uv = u.view(1, N, N)
    ms = _event_ms(torch, lambda: ol.prop_batch(uv, PITCH, PITCH, LAM, 0.1, ol.ASM, _lib.BANDLIMIT,
                                                pre=lens), 10)
    out["asm_4096_fused_lens_ms"] = round(ms, 3)

    def separate():
        ol.mult(u, lens, u)
        ol.AsmProp(u, N, N, PITCH, PITCH, LAM, 0.1, bandlimit=True)
    ms = _event_ms(torch, separate, 10)
    out["asm_4096_lens_then_asm_ms"] = round(ms, 3)
    del u, uv, lens
    # complex128 (the reference's fftw3 path) at configuration 1's size
    d = field(1024, 1024, np.complex128)
    ms = _event_ms(torch, lambda: ol.AsmProp(d, 1024, 1024, PITCH, PITCH, LAM, 0.1, bandlimit=True), 10)
    out["asm_c128_1024_ms"] = round(ms, 3)
    out["asm_c128_1024_GBps_160N2"] = round(160 * 1024 * 1024 / ms / 1e6, 1)
    return out


def layered(ol, torch):
    """Layered-Fresnel CGH on MFMA: 1,000,000 points in 64 layers over z in
    [0.10, 0.20) m, 4096^2 hologram, FP16 and split-BF16 factors; against the
    direct FP32 kernel on the same layer-snapped cloud (rate measured on a 32,768-
    point subset, accuracy on a 256^2 sub-window computed with a shifted cloud)."""
    from openol_amd import synth
    out = {}
    N, n, nl, zmin, zmax = 4096, 1_000_000, 64, 0.10, 0.20
    pts = synth.random_points(n, seed=4, xmax=N * PITCH, ymax=N * PITCH, zmin=zmin, zmax=zmax)
    obj = ol.Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    l = np.clip(np.floor((pts["z"] - zmin) / (zmax - zmin) * nl), 0, nl - 1)
    zl = zmin + (l + 0.5) * (zmax - zmin) / nl
    w0 = h0 = 1920
    win = 256
    sub = ol.Object(pts["x"] - w0 * PITCH, pts["y"] - h0 * PITCH, zl, pts["amp"], pts["phase"])
    D = torch.empty((win, win), dtype=torch.complex64, device="cuda")
    ol.Cgh(sub, D, win, win, PITCH, PITCH, LAM)
    for prec in ("f16", "bf16x3", "bf16"):
        ol.CghLayered(obj, U, N, N, PITCH, PITCH, LAM, nl, zmin, zmax, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        reps = 2
        for _ in range(reps):
            ol.CghLayered(obj, U, N, N, PITCH, PITCH, LAM, nl, zmin, zmax, precision=prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        got = U[h0:h0 + win, w0:w0 + win]
        err = (torch.linalg.norm(got - D) / torch.linalg.norm(D)).item()
        # phase error: amplitude-weighted mean and 99.9th percentile of |arg(got/D)|
        dphi = torch.angle(got * torch.conj(D)).abs()
        amp = D.abs()
        out[f"cgh_layered_{prec}_phase_err_wmean_rad"] = float(f"{(dphi * amp).sum().item() / amp.sum().item():.3e}")
        out[f"cgh_layered_{prec}_phase_err_p999_rad"] = float(f"{torch.quantile(dphi.flatten(), 0.999).item():.3e}")
        out[f"cgh_layered_{prec}_ms"] = round(dt * 1e3, 2)
        out[f"cgh_layered_{prec}_Gptpx_s"] = round(n * N * N / dt / 1e9, 1)
        out[f"cgh_layered_{prec}_relL2_vs_direct_f32"] = float(f"{err:.3e}")
        # 8 flops per complex MAC (x3 products for split bf16)
        fl = 8.0 * n * N * N * (3 if prec == "bf16x3" else 1)
        out[f"cgh_layered_{prec}_TFLOPs"] = round(fl / dt / 1e12, 1)
    m = 32768
    ssub = ol.Object(pts["x"][:m], pts["y"][:m], zl[:m], pts["amp"][:m], pts["phase"][:m])
    ol.Cgh(ssub, U, N, N, PITCH, PITCH, LAM)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ol.Cgh(ssub, U, N, N, PITCH, PITCH, LAM)
    torch.cuda.synchronize()
    rate = m * N * N / (time.perf_counter() - t0)
    out["cgh_direct_fresnel_4096_Gptpx_s"] = round(rate / 1e9, 1)
    out["cgh_layered_f16_speedup_vs_direct"] = round(out["cgh_layered_f16_Gptpx_s"] * 1e9 / rate, 1)
    return out


def multiplane(torch):
    """Config: 2048^2 phase-only hologram, ASM to z = 0.10 / 0.12 / 0.15 m (2x pad),
    seeded 32-Gaussian targets, loss = MSE of mean-normalised intensity against the
    mean-normalised target, 200 Adam steps; loss and gradient from the native
    one-pass multi-plane kernel (per plane: propagate, reduce, dL/df, adjoint
    propagation; no plane field kept)."""
    from openol_amd import optim
    n, zs = 2048, [0.10, 0.12, 0.15]
    targets = optim.gaussian_targets(n, n, zs, seed=5)
    optim.optimize_phase(targets, zs, steps=3, norm="mean")  # warm-up (tables, allocations)
    phi, losses, dt = optim.optimize_phase(targets, zs, steps=200, lr=0.05, norm="mean")
    # the native gradient at the optimised phase against central differences of the
    # loss (evaluated in FP64 from the propagated fields) on 64 seeded pixels
    import openol_amd as ol
    _, g = ol.multiplane_mse(phi, targets, zs, optim.synth.PITCH, optim.synth.PITCH,
                             optim.synth.LAMBDA)
    idx = np.random.default_rng(33).choice(n * n, 64, replace=False)
    fd = optim.central_fd_gradient(phi, targets, zs, idx)
    gk = g.reshape(-1).cpu().numpy()[idx].astype(np.float64)
    fd_err = float(np.linalg.norm(gk - fd) / np.linalg.norm(fd))
    return {"multiplane_2048_steps_per_s": round(200 / dt, 1),
            "multiplane_2048_ms_per_step": round(dt / 200 * 1e3, 3),
            "multiplane_loss_first_last": [round(losses[0], 5), round(losses[-1], 5)],
            "multiplane_2048_grad_vs_central_fd_64px": float(f"{fd_err:.3e}")}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); a propagation "
              "sharded over GPUs (row bands, an all-to-all transpose between the row and "
              "column passes) is not run here.", file=sys.stderr)
        sys.exit(2)
    import torch
    import openol_amd as ol
    from openol_amd import _lib as lib

    if args.only == "ingest":
        print(json.dumps({"ingest_encode": ingest_encode(ol, torch)}), flush=True)
        return
    if args.only == "next":
        print(json.dumps({"next_rows": next_rows(ol, torch)}), flush=True)
        return
    if args.only == "cgh":
        print(json.dumps({"direct_cgh": direct_cgh(ol, torch)}), flush=True)
        return
    if args.only == "sizes":
        print(json.dumps({"sizes_batch": sizes_batch(ol, torch)}), flush=True)
        return
    n = args.n
    dt, rep = bench_asm(ol, lib, torch, n, args.steps, args.warmup, settle_first=True)
    settled = rep.pop("_settle_props")[0]
    ms = dt / args.steps * 1e3
    value = args.steps / dt
    roof = roofline_of(rep, n)
    prop_gbs = 80 * n * n / (dt / args.steps) / 1e9
    line = {
        "metric": "4096^2 ASM props/s (band-limited angular spectrum, 2x pad, complex64)",
        "value": round(value, 2), "unit": "props/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "c64 (f32) with f64 phase",
        "data": "synthetic seeded field (amplitude U[0,1), phase U[0,2pi))",
        "config": {"workload": f"asm_{n}_pad{2 * n}", "n": n, "pad": 2 * n, "lambda": LAM,
                   "pitch": PITCH, "z": 0.10, "bandlimit": True,
                   "clock_settle": f"{settled} untimed propagations ({SETTLE_S} s) before the "
                                   "warm-up steps"},
        "roofline": roof,
        "prop_roofline": {"algorithmic_bytes": 80 * n * n, "achieved_GBps": round(prop_gbs, 1),
                          "frac": round(prop_gbs / HBM_PEAK_GBS, 4)},
    }
    if args.calibrate_copy:
        nbytes = 512 << 20
        a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        L = lib.load()
        c = lib.ctx()
        for _ in range(10):
            L.olx_stream_copy(c, b.data_ptr(), a.data_ptr(), nbytes)
        torch.cuda.synchronize()
        line["copy_calibration_bytes"] = nbytes
    if not args.no_extras:
        try:
            line["extras"] = extras(ol, lib, torch)
        except Exception as e:  # report, do not hide the metric line
            line["extras"] = {"error": repr(e)}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(n, gpu_outputs(ol, torch, n))
        line["gpu_over_cpu"] = round(value / line["cpu_baseline"]["value"], 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
