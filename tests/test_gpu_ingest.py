"""Point-cloud ingest and 8-bit encodings on the GPU, bit-exact against the oracle:
device bounding box, objset (Object<double> and Object<float> arithmetic), stable
objsort, phase8bit / quant8bit / complex2img / complex2real in float and double,
the edge cases of the reference's loops (NaN seeds, 0 / 0 ranges, wrap modulo
256, out-of-range conversions), and the C++ drop-in program that runs the
reference's tests/Object_test.cpp pipeline."""
import os
import subprocess

import numpy as np
import pytest
import torch

import olref
import openol_amd as ol

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CUBE = os.path.join(GOLDEN, "cube284.3d")
HORN = os.path.join(GOLDEN, "horn02-f000.3d")
ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
EXE = os.path.join(ROOT, "tests", "cpp", "ingest_test")
DEV = torch.device("cuda", 0)


def xyz_of(obj):
    h = obj.host
    return np.stack([h["x"], h["y"], h["z"]], axis=1)


def cloud(xyz):
    return ol.Object(xyz[:, 0], xyz[:, 1], xyz[:, 2], device=0)


# --------------------------------------------------------------------------- clouds
@pytest.mark.gpu
@pytest.mark.parametrize("path", [CUBE, HORN])
def test_objread_bounds(path):
    obj = ol.objread(path, device=0)
    ref = olref.obj_read(path)
    assert np.array_equal(xyz_of(obj), ref)
    mn, mx = ol.objinfo(obj)
    assert list(mn) == ref.min(axis=0).tolist() and list(mx) == ref.max(axis=0).tolist()


@pytest.mark.gpu
@pytest.mark.parametrize("f32", [False, True])
@pytest.mark.parametrize("zmiddle", [False, True])
@pytest.mark.parametrize("two", [False, True])
@pytest.mark.parametrize("path", [CUBE, HORN])
def test_objset_bit_exact(path, f32, zmiddle, two):
    xyz = olref.obj_read(path)
    dt = np.float32 if f32 else np.float64
    src = xyz.astype(dt).astype(np.float64)  # an Object<float> holds float coordinates
    off = (np.float32(3.74e-6 * 128), np.float32(3.74e-6 * 64), np.float32(1.0)) if f32 else (2.5e-4, -1e-4, 0.1)
    W = 256 * np.float32(3.74e-6) * 0.8
    H = 128 * np.float32(3.74e-6) * 0.8
    obj = cloud(src)
    ol.objset(obj, W, H if two else None, tuple(float(o) for o in off),
              ol.ZMIDDLE if zmiddle else ol.ZMIN, f32=f32)
    ref = olref.objset(src, W, H if two else None, off, zmiddle, dt)
    got = xyz_of(obj)
    assert np.array_equal(got, ref.astype(np.float64)), np.abs(got - ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 284, 1023, 1024, 1025, 5197, 70001])
def test_objsort_stable(n):
    rng = np.random.default_rng(n)
    z = rng.integers(-20, 20, size=n).astype(np.float64) * 0.25  # many ties
    z[rng.random(n) < 0.1] = 0.0
    z[rng.random(n) < 0.05] = -0.0
    xyz = np.stack([np.arange(n, dtype=np.float64), rng.random(n), z], axis=1)
    obj = cloud(xyz)
    ol.objsort(obj)
    order = olref.objsort_order(z)
    got = xyz_of(obj)
    assert np.array_equal(got, xyz[order])
    assert np.array_equal(np.signbit(got[:, 2]), np.signbit(xyz[order, 2]))


@pytest.mark.gpu
def test_objsort_fixtures_and_large():
    for path in (CUBE, HORN):
        obj = ol.objread(path, device=0)
        ol.objsort(obj)
        ref = olref.obj_read(path)
        assert np.array_equal(xyz_of(obj), ref[olref.objsort_order(ref[:, 2])])
    # 3 M points (not a power of two): sorted, a permutation, stable on the index column
    n = 3_000_001
    rng = np.random.default_rng(9)
    z = np.round(rng.standard_normal(n) * 100) / 100
    xyz = np.stack([np.arange(n, dtype=np.float64), np.zeros(n), z], axis=1)
    obj = cloud(xyz)
    ol.objsort(obj)
    got = xyz_of(obj)
    assert np.all(np.diff(got[:, 2]) <= 0)
    assert np.array_equal(np.sort(got[:, 0]), np.arange(n, dtype=np.float64))
    same = np.diff(got[:, 2]) == 0
    assert np.all(np.diff(got[:, 0])[same] > 0)
    assert np.array_equal(got[:, 2], z[olref.objsort_order(z)])


@pytest.mark.gpu
def test_cloud_errors():
    obj = ol.Object([0.0], [0.0], [0.0], device=0)
    with pytest.raises(Exception):
        ol.objset(obj, 1e-3, None, (0, 0, 0), 7)
    empty = ol.Object(np.zeros(0), np.zeros(0), np.zeros(0), device=0)
    ol.objsort(empty)  # no-op
    with pytest.raises(Exception):
        ol.objinfo(empty)


# --------------------------------------------------------------------------- encodings
def gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def field(shape, seed, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    u = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(dtype)
    return u


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("offset", [False, True])
def test_phase8bit_bit_exact(dtype, offset):
    rng = np.random.default_rng(1)
    ph = rng.uniform(-np.pi, np.pi, size=(512, 768)).astype(dtype)
    ph.ravel()[:12] = [-np.pi, 0.0, np.pi, np.pi / 2, -0.0, 1e9, -1e12, np.nan, np.inf, -np.inf,
                       np.pi - np.pi / 255, 100.0]
    img = ol.phase8bit(gpu(ph), offset=offset).cpu().numpy()
    assert np.array_equal(img, olref.phase8bit(ph, offset=offset))
    # phases whose code lies at (or one ulp from) a rounding tie k + 1/2: the FP32
    # fast path must hand these to the FP64 path
    k = np.arange(-300, 300) + 0.5
    ties = (k / 255 * 2 * np.pi + (-np.pi if offset else np.pi)).astype(dtype)
    near = np.concatenate([ties, np.nextafter(ties, dtype(np.inf)), np.nextafter(ties, dtype(-np.inf))])
    img = ol.phase8bit(gpu(near), offset=offset).cpu().numpy()
    assert np.array_equal(img, olref.phase8bit(near, offset=offset))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_quant8bit_bit_exact(dtype):
    rng = np.random.default_rng(2)
    v = (rng.standard_normal((1024, 1000)) * 3 + 1).astype(dtype)
    v.ravel()[77] = np.nan
    assert np.array_equal(ol.quant8bit(gpu(v)).cpu().numpy(), olref.quant8bit(v))
    for case in ([5.0] * 7, [np.nan, 1.0, 2.0], [1.0], [2.0, np.inf, 1.0], [-1e30, 1e30, 0.0]):
        a = np.array(case, dtype=dtype)
        assert np.array_equal(ol.quant8bit(gpu(a)).cpu().numpy(), olref.quant8bit(a)), case


MODES = ["phase", "real", "imag", "amp", "real_legacy"]
MODE_ID = {"phase": olref.IMG_PHASE, "real": olref.IMG_REAL, "imag": olref.IMG_IMAG,
           "amp": olref.IMG_AMP, "real_legacy": olref.IMG_REAL_LEGACY}


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("mode", MODES)
def test_complex2img_bit_exact(dtype, mode):
    u = field((768, 1024), 3, dtype)
    u.ravel()[:6] = [-1 + 0j, 1, complex(-1, -0.0), 1j, 0, 3 + 4j]
    got = ol.complex2img(gpu(u), mode=mode).cpu().numpy()
    assert np.array_equal(got, olref.complex2img(u, MODE_ID[mode]))
    # angles at the PHASE code's rounding ties (k + 1/2) / 255 * 2pi - pi, and unit
    # vectors one ulp off them
    th = (np.arange(256) + 0.5) / 255 * 2 * np.pi - np.pi
    rt = np.float32 if dtype == np.complex64 else np.float64
    c, sn = np.cos(th).astype(rt), np.sin(th).astype(rt)
    ties = np.concatenate([c + 1j * sn, np.nextafter(c, rt(2)) + 1j * sn, c + 1j * np.nextafter(sn, rt(2))])
    ties = ties.astype(dtype)
    got = ol.complex2img(gpu(ties), mode=mode).cpu().numpy()
    ref = olref.complex2img(ties, MODE_ID[mode])
    if mode == "phase" and dtype == np.complex128:
        # a double angle within a few ulps of a tie: the device's atan2 and the host
        # libm's (neither correctly rounded) may round the code either way there
        q = 255 * (np.arctan2(ties.imag, ties.real) + np.pi) / (2 * np.pi)
        at_tie = np.abs(q - np.floor(q) - 0.5) < 1e-12
        assert np.array_equal(got[~at_tie], ref[~at_tie])
        assert np.all(np.abs(got[at_tie].astype(int) - ref[at_tie].astype(int)) <= 1)
    else:
        assert np.array_equal(got, ref)
    # all-negative real parts: the legacy mode's maximum stays at 0
    f = field((33, 17), 4, dtype)
    neg = (-np.abs(f.real) + 1j * f.imag).astype(dtype)
    got = ol.complex2img(gpu(neg), mode=mode).cpu().numpy()
    assert np.array_equal(got, olref.complex2img(neg, MODE_ID[mode]))


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_complex2img_edge_cases(dtype):
    for vals in ([np.nan + 0j, 1, 2], [1, np.nan, 2], [2 + 2j] * 5, [complex(np.inf, 0), 1, 0], [0j]):
        u = np.array(vals, dtype=dtype)
        for mode in MODES:
            got = ol.complex2img(gpu(u), mode=mode).cpu().numpy()
            assert np.array_equal(got, olref.complex2img(u, MODE_ID[mode])), (vals, mode)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("mode", ["phase", "real", "imag", "amp"])
def test_complex2real(dtype, mode):
    u = field((512, 512), 5, dtype)
    got = ol.complex2real(gpu(u), mode=mode).cpu().numpy()
    ref = olref.complex2real(u, MODE_ID[mode])
    assert got.dtype == ref.dtype
    if mode == "phase":
        # atan2 in FP64 on both sides, but neither the device library (OCML) nor the
        # host's is correctly rounded: they differ by a few ulps of the double, which
        # survive rounding to float only at a tie.  Tolerance: 1 float ulp for
        # complex64, 4 double ulps for complex128.
        rt = np.float32 if dtype == np.complex64 else np.float64
        k = 1 if dtype == np.complex64 else 4
        ulp = np.spacing(np.abs(ref).astype(rt))
        assert np.all(np.abs(got - ref) <= k * ulp)
        assert np.mean(got == ref) > (0.999 if dtype == np.complex64 else 0.5)
    else:
        assert np.array_equal(got, ref)


@pytest.mark.gpu
def test_encodings_4096_and_bmp(tmp_path):
    """Full hologram size: the PHASE encoding and phase8bit of a 4096^2 field, the
    BMP bytes of the result, and the round trip through the file."""
    u = field((4096, 4096), 6)
    img = ol.complex2img(gpu(u), mode="phase").cpu().numpy()
    assert np.array_equal(img, olref.complex2img(u, olref.IMG_PHASE))
    ph = np.angle(u).astype(np.float32)
    img8 = ol.phase8bit(gpu(ph)).cpu().numpy()
    assert np.array_equal(img8, olref.phase8bit(ph))
    p = str(tmp_path / "h.bmp")
    ol.bmpwrite(p, img8)
    assert open(p, "rb").read() == olref.bmp_bytes(img8)
    assert np.array_equal(ol.bmpread(p, 4096, 4096), img8)
    p2 = str(tmp_path / "s.bmp")
    ol.Save(p2, gpu(u[:100, :130]), mode="amp")
    assert open(p2, "rb").read() == olref.bmp_bytes(olref.complex2img(u[:100, :130], olref.IMG_AMP))


@pytest.mark.gpu
def test_encoding_errors():
    with pytest.raises(Exception):
        ol.phase8bit(torch.zeros(4, dtype=torch.int32, device=DEV))
    with pytest.raises(Exception):
        ol.complex2img(gpu(field((4, 4), 1)), mode="bogus")
    with pytest.raises(Exception):
        ol.Save("/tmp/x.png", gpu(field((4, 4), 1)))


# --------------------------------------------------------------------------- C++ drop-in
@pytest.mark.gpu
def test_ingest_program_matches_oracle(tmp_path):
    """tests/cpp/ingest_test runs tests/Object_test.cpp's pipeline through the drop-in
    headers (PREC_T = float, 256 x 128, p = 3.74 um) on the horn fixture."""
    assert os.path.exists(EXE), "build() compiles tests/cpp/ingest_test"
    r = subprocess.run([EXE, str(tmp_path), HORN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "done" in r.stdout and "caught: Save" in r.stdout
    assert "bmp mismatch" not in r.stdout
    W, H = 256, 128
    rd = lambda name, dt, shape=(-1,): np.fromfile(tmp_path / f"{name}.bin", dtype=dt).reshape(shape)
    xyz = olref.obj_read(HORN)
    assert np.array_equal(rd("obj_read", np.float32, (-1, 3)), xyz.astype(np.float32))
    p = np.float32(3.74e-6)
    off = (np.float32(W // 2 * p), np.float32(H // 2 * p), np.float32(1.0))
    xr = np.float32(np.float64(np.float32(W * p)) * 0.8)
    yr = np.float32(np.float64(np.float32(H * p)) * 0.8)
    ref = olref.objset(xyz.astype(np.float32), xr, yr, off, True, np.float32)
    assert np.array_equal(rd("obj_set", np.float32, (-1, 3)), ref)
    assert np.array_equal(rd("obj_sort", np.float32, (-1, 3)), ref[olref.objsort_order(ref[:, 2])])
    # objinfo prints the fitted cloud's bounds as the reference does
    assert f"zmax: {ref[:, 2].max():g}  zmin: {ref[:, 2].min():g}" in r.stdout
    refd = olref.objset(xyz, 1e-3, None, (1e-4, -2e-4, 0.1), False, np.float64)
    assert np.array_equal(rd("objd_set", np.float64, (-1, 3)), refd)
    assert np.array_equal(rd("objd_gsort", np.float64, (-1, 3)), refd[olref.objsort_order(refd[:, 2])])
    offs = np.array([0.5, -1.25, 3.0], np.float32).astype(np.float64)
    assert np.array_equal(rd("obj_offscale", np.float64, (-1, 3)), xyz * 0.001 + offs)
    ph = rd("phase", np.float32, (H, W))
    img = rd("phase8", np.uint8, (H, W))
    assert np.array_equal(img, olref.phase8bit(ph))
    assert (tmp_path / "olObject.bmp").read_bytes() == olref.bmp_bytes(img)
    assert np.array_equal(rd("quant8", np.uint8, (H, W)), olref.quant8bit(ph))
    assert (tmp_path / "quant.bmp").read_bytes() == olref.bmp_bytes(olref.quant8bit(ph))
    u = rd("u", np.complex64, (H, W))
    assert (tmp_path / "amp.bmp").read_bytes() == olref.bmp_bytes(olref.complex2img(u, olref.IMG_AMP))
    assert np.array_equal(rd("img_legacy_real", np.uint8, (H, W)), olref.complex2img(u, olref.IMG_REAL_LEGACY))
    assert np.array_equal(rd("img_phase", np.uint8, (H, W)), olref.complex2img(u, olref.IMG_PHASE))
    assert np.array_equal(rd("real_amp", np.float32, (H, W)), olref.complex2real(u, olref.IMG_AMP))
    assert (tmp_path / "dev_imag.bmp").read_bytes() == olref.bmp_bytes(olref.complex2img(u, olref.IMG_IMAG))
