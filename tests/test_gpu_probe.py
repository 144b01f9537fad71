"""GPU check of the on-device ceiling probes: every kind returns a positive rate
below the published dense peak of the part (plus a margin for clock boost)."""
import ctypes

import pytest

from openol_amd import _lib

pytestmark = pytest.mark.gpu

# published dense peaks of the MI355X: FP32 / FP64 vector, FP16 / BF16 MFMA; the
# sine is bounded by the full lane issue rate (16 lanes x 4 SIMDs x 256 CUs x 2.4 GHz)
PEAKS = {0: 157.3e12, 1: 78.6e12, 2: 16 * 4 * 256 * 2.4e9, 3: 2.5e15, 4: 2.5e15}


@pytest.mark.parametrize("kind", sorted(PEAKS))
def test_peak_probe_rates(kind):
    L = _lib.load()
    r = ctypes.c_double()
    _lib.check(L.olx_peak_probe(_lib.ctx(), kind, ctypes.byref(r)), "peak_probe")
    assert 0 < r.value < 1.2 * PEAKS[kind], (kind, r.value)


def test_peak_probe_rejects_unknown_kind():
    L = _lib.load()
    r = ctypes.c_double()
    assert L.olx_peak_probe(_lib.ctx(), 99, ctypes.byref(r)) == _lib.EINVAL
