"""Builds libdgmr_hip.so (the C-ABI library with the gfx950 kernels) in place.

hipcc cross-compiles for gfx950 without a GPU.  Flags:
  -ffp-contract=off          no implicit multiply-add fusion: binning math is a
                             plain IEEE sequence, bit-identical to the host oracle
                             (functions that only compute gradient values opt in
                             with #pragma clang fp contract(fast))
  -munsafe-fp-atomics        hardware float atomics (global_atomic_add_f32)
  -target-feature -packed-fp32-ops   keep packed FP32 instructions out of the code
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["abi_common.hip", "primitives.hip", "splat.hip", "march.hip", "batch.hip", "optim.hip", "raster.hip", "loss.hip", "noalloc.hip"]
OUT = os.path.join(os.path.dirname(HERE), "_lib", "libdgmr_hip.so")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = OUT + ".tmp"
    cmd = ["hipcc", *FLAGS, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 feature flag; that note is expected
    noise = "'-packed-fp32-ops' is not a recognized feature"
    log = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{log}")
    if log.strip() and verbose:
        print(log)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
