"""GPU test of olx_asm_c64_noalloc (the ASM entry point that allocates nothing: its
intermediate lives in a static device array) against olx_asm_c64, through the
C ABI.  Both run the same kernels, so the results must agree bit for bit; two
calls in a row, the second on a smaller field, check that the shared array
keeps no state that matters between calls."""
import ctypes

import pytest
import torch

from openol_amd import _lib, synth

pytestmark = pytest.mark.gpu
LAM = 532e-9


def _run(name, u, n, z, flags):
    rc = _lib.call(name, None, ctypes.c_void_p(u.data_ptr()), n, n, 8e-6, 8e-6, LAM, z, flags)
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return u


@pytest.mark.parametrize("sizes", [(2048, 512), (256, 64)])
def test_noalloc_matches_asm(sizes):
    for k, n in enumerate(sizes):
        u0 = torch.from_numpy(synth.random_field(n, n, seed=60 + k)).cuda()
        for flags, z in ((_lib.BANDLIMIT, 0.10), (_lib.ADJOINT, 0.05)):
            ref = _run("olx_asm_c64", u0.clone(), n, z, flags)
            got = _run("olx_asm_c64_noalloc", u0.clone(), n, z, flags)
            assert torch.isfinite(torch.view_as_real(got)).all()
            assert torch.equal(got, ref), (n, flags)


def test_noalloc_rejects_fields_beyond_its_array():
    u = torch.zeros((4096, 2048), dtype=torch.complex64, device="cuda")
    rc = _lib.call("olx_asm_c64_noalloc", None, ctypes.c_void_p(u.data_ptr()), 4096, 2048, 8e-6,
                   8e-6, LAM, 0.1, 0)
    assert rc == _lib.EINVAL
    assert "static" in _lib.last_error()
