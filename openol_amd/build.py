"""Build libolx.so (gfx950) in-tree with hipcc.

    python -m openol_amd.build        # or via __graft_entry__.build()

Every .hip file under openol_amd/csrc is compiled to an object in
openol_amd/_build/ and linked into openol_amd/libolx.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libolx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    # no packed-FP32 instructions (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    # x + 0 -> x: lets the compiler drop the butterflies of statically zero padding
    "-fno-signed-zeros",
]

# Byte-exact ports of the reference's x86-64 arithmetic: one rounding per operation
# (HIP contracts a * b + c into an FMA by default, and __dadd_rn / __dmul_rn are
# plain operators there) and IEEE signed zeros (atan2(-0, x), the +0 / -0 sort key).
FILE_FLAGS = {
    "olx_obj.hip": ["-ffp-contract=off", "-fsigned-zeros"],
    "olx_img.hip": ["-ffp-contract=off", "-fsigned-zeros"],
    # tilted propagation: sample positions must floor() to the reference's cells
    "olx_tilt.hip": ["-ffp-contract=off", "-fsigned-zeros"],
    # optical elements: float phases and products as the reference rounds them
    "olx_optics.hip": ["-ffp-contract=off", "-fsigned-zeros"],
    # WRP: window tests dx*dx + dy*dy < W^2 with each product rounded, as the reference's
    # host arithmetic (and the oracle) evaluate them
    "olx_wrp.hip": ["-ffp-contract=off"],
    # the 256 x 128 GEMM keeps its register-indexed fragment arrays in registers: with
    # one work-group per CU the compiler would otherwise move them into spare LDS
    "olx_layered.hip": ["-mllvm", "-disable-promote-alloca-to-lds"],
}


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, *FILE_FLAGS.get(os.path.basename(src), []), "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{msg}")
    return obj, msg


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    for obj, msg in results:
        if verbose and msg.strip():
            print(msg)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp,
           *[o for o, _ in results]]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
