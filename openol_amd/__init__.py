"""MI355X-native OpenOL diffraction / CGH hot path.

Python host mirror of the reference's operator surface over the C ABI of
libolx.so (include/olx.h): angular-spectrum, Fresnel and ARSS propagation,
direct point-cloud CGH, WRP, layered-Fresnel CGH on MFMA, and an adjoint ASM for
differentiable hologram optimisation.
"""
from . import synth  # noqa: F401  (no GPU needed)

__all__ = ["synth"]


def __getattr__(name):
    # operators import torch lazily so that `import openol_amd` stays cheap
    if name.startswith("_") or name in ("ops", "build", "autograd", "optim"):
        import importlib
        try:
            return importlib.import_module(f"{__name__}.{name}")
        except ModuleNotFoundError:
            raise AttributeError(name) from None
    import importlib
    ops = importlib.import_module(f"{__name__}.ops")
    if hasattr(ops, name):
        return getattr(ops, name)
    raise AttributeError(name)
