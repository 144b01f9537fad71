"""GPU parity: direct CGH and WRP vs the oracle (through the C ABI).

CGH tolerance: relative L2 <= 1e-4 against the FP64 oracle (the bar for the CGH
path); WRP: <= 1e-5 against the oracle restatement at PREC_T = double.
"""
import os

import numpy as np
import pytest
import torch

import olref
from openol_amd import synth

pytestmark = pytest.mark.gpu
LAM = 532e-9
P = 8e-6


def ol():
    import openol_amd
    return openol_amd


def rel(a, b):
    return olref.rel_l2(a, b)


def cloud(n, seed, nx, ny, zmin=0.1, zmax=0.2, pitch=P):
    """Seeded cloud; amp and phase rounded to float32, as olx_point stores them."""
    pts = synth.random_points(n, seed=seed, xmax=nx * pitch, ymax=ny * pitch, zmin=zmin, zmax=zmax)
    for k in ("amp", "phase"):
        pts[k] = pts[k].astype(np.float32).astype(np.float64)
    return pts


def obj_of(pts, with_ap=True):
    if with_ap:
        return ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    return ol().Object(pts["x"], pts["y"], pts["z"])


@pytest.mark.parametrize("exact", [False, True])
@pytest.mark.parametrize("shape,n", [((12, 16), 7), ((64, 96), 300), ((100, 70), 257), ((256, 256), 1000),
                                     ((200, 300), 600)])
def test_cgh_matches_fp64_oracle(exact, shape, n):
    ny, nx = shape
    pts = cloud(n, 40 + n, nx, ny)
    U = torch.empty(shape, dtype=torch.complex64, device="cuda")
    ol().Cgh(obj_of(pts), U, ny, nx, P, P, LAM, exact_r=exact)
    ref = olref.cgh(pts["x"], pts["y"], pts["z"], ny, nx, P, P, LAM, pts["amp"], pts["phase"],
                    flags=olref.CGH_EXACT_R if exact else 0, dtype=np.complex128)
    assert rel(U.cpu().numpy(), ref) < 1e-4
    U64 = torch.empty(shape, dtype=torch.complex128, device="cuda")
    ol().CghFP64(obj_of(pts), U64, ny, nx, P, P, LAM, exact_r=exact)
    # k r reaches 3e5 cycles: FP64 rounding of the phase alone is ~1e-10 rad
    assert rel(U64.cpu().numpy(), ref) < 1e-9


def test_cgh_reference_semantics_and_golden():
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
    o = ol().Object(g["cgh_x"], g["cgh_y"], g["cgh_z"])
    U = torch.empty((12, 16), dtype=torch.complex64, device="cuda")
    ol().Cgh(o, U, 12, 16, 6e-6, 8e-6, LAM)  # A = 1, phi = 0: _CghFresnel itself
    assert rel(U.cpu().numpy(), g["cgh_plain"]) < 1e-5
    # the oracle's complex64 restatement (float accumulation) agrees too
    ref32 = olref.cgh(g["cgh_x"], g["cgh_y"], g["cgh_z"], 12, 16, 6e-6, 8e-6, LAM)
    assert rel(U.cpu().numpy(), ref32) < 1e-5


@pytest.mark.parametrize("zmin,zmax", [(2e-3, 5e-3), (2e-3, 0.2)])
def test_cgh_exact_near_field(zmin, zmax):
    """Exact r with points a few millimetres from a 2 mm hologram: r^2 changes by a
    large fraction of ra^2 across a 32 x 32 tile, so the kernel's series for r - ra
    does not apply and the sqrt form runs; the wide z range mixes both forms inside
    every 256-point chunk."""
    ny, nx = 200, 256
    pts = cloud(700, 91, nx, ny, zmin=zmin, zmax=zmax)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().Cgh(obj_of(pts), U, ny, nx, P, P, LAM, exact_r=True)
    ref = olref.cgh(pts["x"], pts["y"], pts["z"], ny, nx, P, P, LAM, pts["amp"], pts["phase"],
                    flags=olref.CGH_EXACT_R, dtype=np.complex128)
    assert rel(U.cpu().numpy(), ref) < 1e-4


def test_cgh_outputs_phase_amp_accumulate():
    ny, nx, n = 48, 80, 200
    pts = cloud(n, 41, nx, ny)
    o = obj_of(pts)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ph = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    am = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    ol().Cgh(o, U, ny, nx, P, P, LAM, phase_out=ph, amp_out=am)
    u = U.cpu().numpy()
    assert np.allclose(ph.cpu().numpy(), np.angle(u), atol=1e-5)
    assert np.allclose(am.cpu().numpy(), np.abs(u), rtol=1e-6)
    # phase-only overload: float32 output tensor (include/olcgh.h:362-365)
    ph2 = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    ol().Cgh(o, ph2, ny, nx, P, P, LAM)
    assert torch.equal(ph2, ph)
    # encode kernel
    p3, a3 = ol().encode(U)
    assert torch.allclose(p3, ph, rtol=0, atol=2e-6) and torch.allclose(a3, am, rtol=1e-6, atol=0)
    # accumulate: two halves sum to the whole
    h1 = ol().Object(pts["x"][:90], pts["y"][:90], pts["z"][:90], pts["amp"][:90], pts["phase"][:90])
    h2 = ol().Object(pts["x"][90:], pts["y"][90:], pts["z"][90:], pts["amp"][90:], pts["phase"][90:])
    V = torch.empty_like(U)
    ol().Cgh(h1, V, ny, nx, P, P, LAM)
    ol().Cgh(h2, V, ny, nx, P, P, LAM, accumulate=True)
    assert rel(V.cpu().numpy(), u) < 1e-5


@pytest.mark.parametrize("exact", [False, True])
def test_cgh_small_hologram_split_over_points(exact):
    """A hologram of fewer tiles than CUs splits its points over work-groups, whose
    partial planes a second kernel sums in slice order: 64 x 96 from 20,000 points
    (9 slices) against the FP64 oracle, with the phase / amplitude outputs and
    accumulation taken on the split path."""
    ny, nx, n = 64, 96, 20000
    pts = cloud(n, 77, nx, ny)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ph = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    am = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    ol().Cgh(obj_of(pts), U, ny, nx, P, P, LAM, exact_r=exact, phase_out=ph, amp_out=am)
    ref = olref.cgh(pts["x"], pts["y"], pts["z"], ny, nx, P, P, LAM, pts["amp"], pts["phase"],
                    flags=olref.CGH_EXACT_R if exact else 0, dtype=np.complex128)
    u = U.cpu().numpy()
    assert rel(u, ref) < 1e-5
    assert np.allclose(ph.cpu().numpy(), np.angle(u), atol=1e-5)
    assert np.allclose(am.cpu().numpy(), np.abs(u), rtol=1e-6)
    # phase only (no complex output) takes the same path
    ph2 = torch.empty_like(ph)
    ol().Cgh(obj_of(pts), ph2, ny, nx, P, P, LAM, exact_r=exact)
    assert torch.equal(ph2, ph)
    # two halves of 10,000 points (each split) accumulate to the whole
    k = n // 2
    half = lambda sl: ol().Object(pts["x"][sl], pts["y"][sl], pts["z"][sl], pts["amp"][sl], pts["phase"][sl])
    V = torch.empty_like(U)
    ol().Cgh(half(slice(0, k)), V, ny, nx, P, P, LAM, exact_r=exact)
    ol().Cgh(half(slice(k, n)), V, ny, nx, P, P, LAM, exact_r=exact, accumulate=True)
    assert rel(V.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("exact,tol", [(False, 6e-6), (True, 8e-6)])
def test_cgh_large_cloud_sums_stay_accurate(exact, tol):
    """200,000 points over 2048^2 (no split), a 64 x 128 window against the FP64 kernel
    on the shifted cloud.  Each 256-point chunk is summed before it is added to the
    pixel's total; accumulating every term into the running sum (|U| ~ 450, two
    roundings per term in the separable kernel) would leave about 1e-5."""
    N, n, h0, w0, wy, wx = 2048, 200000, 1000, 700, 64, 128
    pts = cloud(n, 91, N, N)
    U = torch.empty((N, N), dtype=torch.complex64, device="cuda")
    ol().Cgh(obj_of(pts), U, N, N, P, P, LAM, exact_r=exact)
    sub = ol().Object(pts["x"] - w0 * P, pts["y"] - h0 * P, pts["z"], pts["amp"], pts["phase"])
    D = torch.empty((wy, wx), dtype=torch.complex128, device="cuda")
    ol().CghFP64(sub, D, wy, wx, P, P, LAM, exact_r=exact)
    got = U[h0:h0 + wy, w0:w0 + wx].to(torch.complex128)
    assert (torch.linalg.norm(got - D) / torch.linalg.norm(D)).item() < tol


def test_cgh_gauss_sum_closed_form():
    z = 2 * P * P / LAM
    o = ol().Object([5 * P], [3 * P], [z])
    U = torch.empty((10, 12), dtype=torch.complex64, device="cuda")
    ol().Cgh(o, U, 10, 12, P, P, LAM)
    hh, ww = np.meshgrid(np.arange(10), np.arange(12), indexing="ij")
    expect = 1j ** (((ww - 5) ** 2 + (hh - 3) ** 2) % 4)
    # the FP32 tile offsets carry ~1e-5 rad of rounding at these slopes
    assert np.max(np.abs(U.cpu().numpy() - expect)) < 1e-4


def test_cgh_empty_cloud():
    o = ol().Object(np.zeros(0), np.zeros(0), np.zeros(0))
    U = torch.ones((32, 32), dtype=torch.complex64, device="cuda")
    ol().Cgh(o, U, 32, 32, P, P, LAM)
    assert torch.count_nonzero(U).item() == 0


def test_cgh_10k_points_2048_sampled_fp64():
    """10,000 points, 2048^2 hologram, exact r and Fresnel; FP64 oracle on
    4096 sampled pixels (each pixel is an independent sum)."""
    n, N = 10000, 2048
    pts = cloud(n, 2, N, N, 0.10, 0.20)
    o = obj_of(pts)
    rng = np.random.default_rng(5)
    ph, pw = rng.integers(0, N, 4096), rng.integers(0, N, 4096)
    for exact in (True, False):
        U = torch.empty((N, N), dtype=torch.complex64, device="cuda")
        ol().Cgh(o, U, N, N, P, P, LAM, exact_r=exact)
        got = U[torch.from_numpy(ph).cuda(), torch.from_numpy(pw).cuda()].cpu().numpy()
        ref = olref.cgh_pixels(pts["x"], pts["y"], pts["z"], ph, pw, P, P, LAM, pts["amp"],
                               pts["phase"], flags=olref.CGH_EXACT_R if exact else 0)
        assert rel(got, ref) < 1e-4, exact


@pytest.mark.parametrize("square", [False, True])
def test_wrp_matches_oracle(square):
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
    o = ol().Object(g["wrp_x"], g["wrp_y"], g["wrp_z"])
    U = torch.empty((24, 20), dtype=torch.complex64, device="cuda")
    ol().WRPStep1(o, U, 24, 20, P, LAM, 0.15, square=square)
    assert rel(U.cpu().numpy(), g["wrp_square" if square else "wrp_circle"]) < 1e-5
    # a larger plane with many overlapping windows
    ny, nx, n = 200, 300, 3000
    pts = cloud(n, 42, nx, ny, 0.151, 0.155)
    o = obj_of(pts, with_ap=False)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().WRPStep1(o, U, ny, nx, P, LAM, 0.15, square=square)
    ref = olref.wrp(pts["x"], pts["y"], pts["z"], ny, nx, P, LAM, 0.15, square=square,
                    dtype=np.complex128)
    assert rel(U.cpu().numpy(), ref) < 1e-5
    # deterministic summation order
    U2 = torch.empty_like(U)
    ol().WRPStep1(o, U2, ny, nx, P, LAM, 0.15, square=square)
    assert torch.equal(U, U2)


@pytest.mark.parametrize("square", [False, True])
def test_wrp_dense_tiles_deterministic(square):
    """A cloud dense enough that tile lists pass the 4096 entries the LDS sort holds:
    40,000 points within a 32 x 32-pixel patch plus 2,000 over the whole 160 x 200
    plane, so tiles near the patch take the dense path and the rest the sorted
    lists.  Sampled pixels against the FP64 oracle; two runs bit for bit alike."""
    ny, nx = 160, 200
    a = cloud(40000, 77, 32, 32, 0.151, 0.155)
    b = cloud(2000, 78, nx, ny, 0.151, 0.155)
    pts = {k: np.concatenate([a[k] + (60 * P if k in ("x", "y") else 0), b[k]]) for k in a}
    o = obj_of(pts, with_ap=False)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().WRPStep1(o, U, ny, nx, P, LAM, 0.15, square=square)
    rng = np.random.default_rng(5)
    ph, pw = rng.integers(0, ny, 1500), rng.integers(0, nx, 1500)
    ph[:200], pw[:200] = rng.integers(50, 110, 200), rng.integers(50, 110, 200)  # in the patch
    ref = olref.wrp_pixels(pts["x"], pts["y"], pts["z"], ph, pw, P, LAM, 0.15, square=square)
    got = U.cpu().numpy()[ph, pw]
    assert rel(got, ref) < 1e-5
    for _ in range(2):
        U2 = torch.empty_like(U)
        ol().WRPStep1(o, U2, ny, nx, P, LAM, 0.15, square=square)
        assert torch.equal(U, U2)


@pytest.mark.parametrize("square", [False, True])
def test_wrp_mid_density_tiles(square):
    """Tiles whose lists hold between 1024 and 4096 entries: the accumulate runs the
    tiles of up to 1024 entries in one launch and lists the longer ones for a second
    launch with the larger sort space.  4,000 points in a 64 x 64-pixel patch (the
    windows' bounding boxes, estimated here, put several tiles in that range and none
    above it), plus 1,000 over the 128 x 160 plane.  Sampled pixels against the FP64
    oracle; two runs bit for bit alike."""
    ny, nx = 128, 160
    a = cloud(4000, 79, 64, 64, 0.151, 0.155)
    b = cloud(1000, 80, nx, ny, 0.151, 0.155)
    pts = {k: np.concatenate([a[k] + (40 * P if k in ("x", "y") else 0), b[k]]) for k in a}
    # bounding-box estimate of the tile list lengths (16 x 16 tiles): the same windows
    # as the kernel's tile test, W = |dz| tan(asin(lambda / 2p))
    s = LAM / (2 * P)
    W = np.abs(pts["z"] - 0.15) * s / np.sqrt(1 - s * s)
    counts = []
    for ty in range(ny // 16):
        for tx in range(nx // 16):
            x0, x1, y0, y1 = 16 * tx * P, (16 * tx + 15) * P, 16 * ty * P, (16 * ty + 15) * P
            counts.append(int(np.sum((pts["x"] + W >= x0) & (pts["x"] - W <= x1) &
                                     (pts["y"] + W >= y0) & (pts["y"] - W <= y1))))
    assert sum(1200 < c < 3500 for c in counts) >= 4 and max(counts) < 3500
    o = obj_of(pts, with_ap=False)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().WRPStep1(o, U, ny, nx, P, LAM, 0.15, square=square)
    rng = np.random.default_rng(6)
    ph, pw = rng.integers(0, ny, 1500), rng.integers(0, nx, 1500)
    ph[:300], pw[:300] = rng.integers(40, 104, 300), rng.integers(40, 104, 300)  # in the patch
    ref = olref.wrp_pixels(pts["x"], pts["y"], pts["z"], ph, pw, P, LAM, 0.15, square=square)
    assert rel(U.cpu().numpy()[ph, pw], ref) < 1e-5
    for _ in range(2):
        U2 = torch.empty_like(U)
        ol().WRPStep1(o, U2, ny, nx, P, LAM, 0.15, square=square)
        assert torch.equal(U, U2)


def test_wrp_large_plane_matches_windows():
    """A 33024^2 plane (4.26 M tiles of 16 x 16: the tile-count scan spans 1040 blocks
    and two rounds of its top level) with clusters of points in three far-apart
    windows.  With a power-of-two pitch, translating a cluster by a multiple of 16
    pixels is exact (p w, p (w - w0) and x - w0 p carry no rounding), so each window of
    the large plane must equal, bit for bit, a 64 x 64 plane holding the translated
    cluster; that plane is checked against the oracle."""
    n_side, win = 33024, 64
    p = 2.0 ** -17
    rng = np.random.default_rng(12)
    corners = [(1024, 2048), (16384, 20480), (32960, 32960)]
    xs, ys, zs = [], [], []
    for h0, w0 in corners:
        # windows reach 23 px at dz = 5 mm: keep them inside the 64 x 64 corner
        xs.append((w0 + rng.uniform(24, win - 24, 15)) * p)
        ys.append((h0 + rng.uniform(24, win - 24, 15)) * p)
        zs.append(rng.uniform(0.151, 0.155, 15))
    x, y, z = np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)
    big = torch.empty((n_side, n_side), dtype=torch.complex64, device="cuda")
    ol().WRPStep1(ol().Object(x, y, z), big, n_side, n_side, p, LAM, 0.15)
    for k, (h0, w0) in enumerate(corners):
        sl = slice(15 * k, 15 * (k + 1))
        xk, yk, zk = x[sl] - w0 * p, y[sl] - h0 * p, z[sl]
        small = torch.empty((win, win), dtype=torch.complex64, device="cuda")
        ol().WRPStep1(ol().Object(xk, yk, zk), small, win, win, p, LAM, 0.15)
        assert torch.equal(big[h0:h0 + win, w0:w0 + win], small), (h0, w0)
        ref = olref.wrp(xk, yk, zk, win, win, p, LAM, 0.15, dtype=np.complex128)
        assert rel(small.cpu().numpy(), ref) < 1e-5
    # nothing outside the windows
    for h0, w0 in corners:
        big[h0:h0 + win, w0:w0 + win] = 0
    assert int(torch.count_nonzero(torch.view_as_real(big))) == 0
    del big


def test_wrp_plimit_and_edges():
    # points near and beyond the plane edges, and a p_limit overload (golcgh.h:381-389)
    ny, nx = 64, 48
    x = np.array([-3 * P, 0.0, nx * P + 2 * P, 20.5 * P, 47 * P])
    y = np.array([5 * P, -P, 30 * P, 63 * P, 64 * P + 4 * P])
    z = np.array([0.153, 0.152, 0.155, 0.1505, 0.154])
    o = ol().Object(x, y, z)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().WRPStep1(o, U, ny, nx, P, LAM, 0.15, plimit=2 * P)
    ref = olref.wrp(x, y, z, ny, nx, P, LAM, 0.15, p_limit=2 * P, dtype=np.complex128)
    assert rel(U.cpu().numpy(), ref) < 1e-5


def test_wrp_method_pipeline():
    ny = nx = 256
    pts = cloud(2000, 43, nx, ny, 0.10, 0.104)
    o = obj_of(pts, with_ap=False)
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    zmin = float(pts["z"].min())
    pz = ol().WRPMethod(o, U, ny, nx, P, LAM, zmin, s=0.99)
    assert pz == olref.wrp_plane_z(ny, nx, P, LAM, zmin, 0.99)
    w = olref.wrp(pts["x"], pts["y"], pts["z"], ny, nx, P, LAM, pz)
    ch = 0 if (olref.asm_prop_check(ny, nx, P, P, LAM, pz)
               or not olref.fresnel_prop_check(ny, nx, P, P, LAM, pz)) else 1
    ref = olref.asm_prop(w, P, P, LAM, pz) if ch == 0 else olref.fresnel_prop(w, P, P, LAM, pz)
    assert rel(U.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("shape,n", [((128, 256), 2000), ((97, 61), 300)])
def test_cgh_fresnel_diff_matches_fp64_oracle(shape, n):
    """_CghFresnelDiff (include/olcgh.h:315-349): cloud in pixel units, the p z / lambda
    piston and the row recurrence's exact value; complex field <= 1e-4 against the FP64
    oracle (which carries the recurrence), phase output = atan2 of the field."""
    ny, nx = shape
    p = P
    pts = synth.random_points(n, seed=44, xmax=nx, ymax=ny, zmin=0.10 / p, zmax=0.20 / p)
    for k in ("amp", "phase"):
        pts[k] = pts[k].astype(np.float32).astype(np.float64)
    o = ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ph = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    ol().CghFresnelDiff(o, ny, nx, p, LAM, U=U, phase=ph)
    ref = olref.cgh_fresnel_diff(pts["x"], pts["y"], pts["z"], ny, nx, p, LAM, pts["amp"], pts["phase"])
    assert rel(U.cpu().numpy(), ref) < 1e-4
    u = U.cpu().numpy()
    assert np.abs(np.angle(np.exp(1j * (ph.cpu().numpy() - np.angle(u))))).max() < 1e-5
    # unit amplitude, no per-point phase: the reference's own sum
    o1 = ol().Object(pts["x"], pts["y"], pts["z"])
    _, ph1 = ol().CghFresnelDiff(o1, ny, nx, p, LAM)
    ref1 = olref.cgh_fresnel_diff(pts["x"], pts["y"], pts["z"], ny, nx, p, LAM)
    w = np.abs(ref1)
    dphi = np.abs(np.angle(np.exp(1j * (ph1.cpu().numpy() - np.angle(ref1)))))
    assert np.sum(dphi * w) / np.sum(w) < 1e-4
