"""GPU tests of the library context: stream ordering of the shared workspaces,
per-kernel launch-timing samples, the bounded table cache, input validation."""
import numpy as np
import pytest
import torch

import olref
from openol_amd import synth

pytestmark = pytest.mark.gpu
LAM = 532e-9


def ol():
    import openol_amd
    return openol_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).cuda()


def test_two_streams_share_the_workspace_in_order():
    """Propagations issued back to back on two torch streams use the context's one
    intermediate W; the second stream waits for the first, so both results equal
    the single-stream ones bit for bit."""
    n = 2048
    a0, b0 = synth.random_field(n, n, seed=50), synth.random_field(n, n, seed=51)
    ra, rb = dev(a0), dev(b0)
    ol().AsmProp(ra, n, n, 8e-6, 8e-6, LAM, 0.10, bandlimit=True)
    ol().AsmProp(rb, n, n, 8e-6, 8e-6, LAM, 0.15, bandlimit=True)
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):
        a, b = dev(a0), dev(b0)
        torch.cuda.synchronize()
        with torch.cuda.stream(s1):
            ol().AsmProp(a, n, n, 8e-6, 8e-6, LAM, 0.10, bandlimit=True)
        with torch.cuda.stream(s2):
            ol().AsmProp(b, n, n, 8e-6, 8e-6, LAM, 0.15, bandlimit=True)
        torch.cuda.synchronize()
        assert torch.equal(a, ra)
        assert torch.equal(b, rb)


def test_batched_field_is_rejected():
    """A (B, H, W) tensor is not silently propagated as its first plane."""
    u = torch.zeros((2, 64, 64), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        ol().AsmProp(u, 64, 64, 8e-6, 8e-6, LAM, 0.01)
    with pytest.raises(ValueError):
        ol().FresnelProp(u, 64, 64, 8e-6, 8e-6, LAM, 0.01)


def test_autograd_propagates_each_plane():
    """The differentiable wrapper treats leading dimensions as independent fields."""
    from openol_amd import autograd
    u0 = np.stack([synth.random_field(64, 64, seed=s) for s in (52, 53)])
    out = autograd.asm_prop(dev(u0), 8e-6, 8e-6, LAM, 0.01, bandlimit=True).cpu().numpy()
    for k in range(2):
        ref = olref.asm_prop(u0[k], 8e-6, 8e-6, LAM, 0.01, olref.BANDLIMIT)
        assert olref.rel_l2(out[k], ref) < 1e-5


def test_fft2d_timing_samples_both_kernels():
    """Each kernel is sampled on every k-th of its own launches: with two launches
    per 2-D FFT and k = 2, both the row and the column kernel are timed."""
    from openol_amd import _lib
    L, c = _lib.load(), _lib.ctx()
    u = dev(synth.random_field(256, 256, seed=54))
    L.olx_ctx_reset_timing(c)
    assert L.olx_ctx_enable_timing(c, 2) == 0
    for _ in range(4):
        ol().fft(u)
    L.olx_ctx_enable_timing(c, 0)
    rep = _lib.timing_report()
    assert rep["fft_rows"][0] == 2 and rep["fft_cols"][0] == 2, rep


def test_table_cache_evicts_and_clears():
    """A sweep over more distances than the cache holds keeps giving the oracle's
    result (least recently used tables are freed and rebuilt), and
    olx_ctx_clear_tables empties the cache without breaking later calls."""
    from openol_amd import _lib
    n = 64
    u0 = synth.random_field(n, n, seed=55)
    zs = [0.02 + 0.001 * k for k in range(80)]
    for z in zs:
        u = dev(u0)
        ol().FresnelProp(u, n, n, 8e-6, 8e-6, LAM, z)
    ref = olref.fresnel_prop(u0, 8e-6, 8e-6, LAM, zs[-1])
    assert olref.rel_l2(u.cpu().numpy(), ref) < 1e-5
    u = dev(u0)
    ol().FresnelProp(u, n, n, 8e-6, 8e-6, LAM, zs[0])  # evicted earlier: rebuilt
    assert olref.rel_l2(u.cpu().numpy(), olref.fresnel_prop(u0, 8e-6, 8e-6, LAM, zs[0])) < 1e-5
    _lib.check(_lib.load().olx_ctx_clear_tables(_lib.ctx()))
    u = dev(u0)
    ol().FresnelProp(u, n, n, 8e-6, 8e-6, LAM, zs[5])
    assert olref.rel_l2(u.cpu().numpy(), olref.fresnel_prop(u0, 8e-6, 8e-6, LAM, zs[5])) < 1e-5
