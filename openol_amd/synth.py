"""Seeded synthetic inputs for the diffraction / CGH workloads.

A counter-based generator (splitmix64) so that any element of a large seeded
input can be regenerated without storing it:

    u(seed, stream, i) = (splitmix64(splitmix64(seed + G * stream) ^ i) >> 11) * 2**-53

with G = 0x632BE59BD9B4E019.  Streams separate the quantities drawn from one seed
(amplitude, phase, x, y, z, ...).  numpy uint64 arithmetic wraps modulo 2**64,
which is what splitmix64 needs.
"""
import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_G = 0x632BE59BD9B4E019

# stream ids
S_AMP, S_PHASE, S_X, S_Y, S_Z, S_PAMP, S_PPHASE, S_GAUSS = range(1, 9)


def _splitmix64(x):
    x = (x + _GOLD).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(30))) * _M1).astype(np.uint64)
    x = ((x ^ (x >> np.uint64(27))) * _M2).astype(np.uint64)
    return x ^ (x >> np.uint64(31))


def uniform(seed, stream, n, lo=0.0, hi=1.0, offset=0):
    """n doubles uniform in [lo, hi) for (seed, stream), elements offset..offset+n."""
    with np.errstate(over="ignore"):
        key = _splitmix64(np.array([(seed + _G * stream) % (1 << 64)], dtype=np.uint64))[0]
        idx = np.arange(offset, offset + n, dtype=np.uint64)
        r = _splitmix64(idx ^ key)
    u = (r >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return lo + (hi - lo) * u


def random_field(ny, nx, seed, dtype=np.complex64):
    """Complex field: amplitude U[0,1), phase U[0, 2 pi) (the propagation workloads)."""
    a = uniform(seed, S_AMP, ny * nx)
    ph = uniform(seed, S_PHASE, ny * nx, 0.0, 2 * np.pi)
    return (a * np.exp(1j * ph)).astype(dtype).reshape(ny, nx)


def random_points(n, seed, xmax, ymax, zmin, zmax, amin=0.5, amax=1.0):
    """Point cloud: x ~ U[0,xmax), y ~ U[0,ymax), z ~ U[zmin,zmax), A ~ U[amin,amax),
    phi ~ U[0, 2 pi).  Returns a dict of float64 arrays."""
    return dict(
        x=uniform(seed, S_X, n, 0.0, xmax),
        y=uniform(seed, S_Y, n, 0.0, ymax),
        z=uniform(seed, S_Z, n, zmin, zmax),
        amp=uniform(seed, S_PAMP, n, amin, amax),
        phase=uniform(seed, S_PPHASE, n, 0.0, 2 * np.pi),
    )


def gaussian_targets(ny, nx, seed, count=32, smin=20.0, smax=200.0, wmin=0.5, wmax=1.0):
    """Target intensity: sum of `count` seeded Gaussians (centre ~ U(field),
    sigma ~ U[smin,smax) px, weight ~ U[wmin,wmax)), normalised to [0, 1]."""
    cy = uniform(seed, S_GAUSS, count, 0, ny, offset=0)
    cx = uniform(seed, S_GAUSS, count, 0, nx, offset=count)
    sg = uniform(seed, S_GAUSS, count, smin, smax, offset=2 * count)
    wt = uniform(seed, S_GAUSS, count, wmin, wmax, offset=3 * count)
    yy = np.arange(ny, dtype=np.float64)[:, None]
    xx = np.arange(nx, dtype=np.float64)[None, :]
    t = np.zeros((ny, nx))
    for i in range(count):
        t += wt[i] * np.exp(-((yy - cy[i]) ** 2 + (xx - cx[i]) ** 2) / (2 * sg[i] ** 2))
    t -= t.min()
    t /= t.max()
    return t.astype(np.float32)


# Optical constants shared by the workloads
LAMBDA = 532e-9
PITCH = 8e-6
