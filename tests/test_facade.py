"""The C++ façade header (include/dgmr.hpp) compiles as host C++17 with g++
against the C ABI header alone (CPU)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_facade_header_compiles(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "dgmr.hpp"\n'
                   "int main() { dgmr::SplatForwardData d; dgmr::VolRasterForwardData r; dgmr::vol_marcher::ForwardData<double> m;\n"
                   "  (void)d; (void)r; (void)m; return 0; }\n")
    res = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I",
                          os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
