"""Builds tests/cpp/_build/libfacade_check.so: the C++ façade check
(facade_check.cpp over include/dgmr.hpp) that tests/test_gpu_facade.py loads,
and libreach_check.so, the host build of the blends' tile-reach test that
tests/test_tile_reach.py checks.  Run by __graft_entry__.build() after the
library itself."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB_DIR = os.path.join(ROOT, "differentiable_gaussian_mixture_renderer_amd", "_lib")
OUT = os.path.join(HERE, "_build", "libfacade_check.so")


REACH_OUT = os.path.join(HERE, "_build", "libreach_check.so")


def build_reach() -> str:
    """The tile-reach test of the blends, built for the host (g++, no
    contraction, as the kernels compile) for a CPU brute-force check."""
    os.makedirs(os.path.dirname(REACH_OUT), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", os.path.join(HERE, "reach_check.cpp"), "-o",
           REACH_OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed:\n{res.stdout}{res.stderr}")
    os.replace(REACH_OUT + ".tmp", REACH_OUT)
    return REACH_OUT


def build() -> str:
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["hipcc", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "facade_check.cpp"),
           "-L", LIB_DIR, "-ldgmr_hip", "-Wl,-rpath,$ORIGIN/../../../differentiable_gaussian_mixture_renderer_amd/_lib", "-o", OUT + ".tmp"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{res.stdout}{res.stderr}")
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build_reach())
    print(build())
