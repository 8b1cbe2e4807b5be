"""The CPU oracle and the library's host-only sources (io.cpp, host.cpp)
built with AddressSanitizer and run on small grids near their minimum sizes
(tests/asan/driver.cpp): any out-of-bounds access, use after free or heap
misuse or leak in that host code fails this test.  Host compilers only: the HIP
kernels are not part of it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "multigpu_advectiondiffusion_amd", "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_address_sanitizer(tmp_path):
    san = ["-fsanitize=address", "-fno-omit-frame-pointer", "-g", "-O1"]
    objs = []

    def cc(compiler, src, extra):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run([compiler, *san, *extra, "-c", src, "-o", obj], check=True)
        objs.append(obj)

    cc("gcc", os.path.join(ROOT, "oracle", "oracle.c"), ["-std=c11", "-fopenmp", "-ffp-contract=off"])
    hip_host = ["-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include"]
    for src in (os.path.join(CSRC, "io.cpp"), os.path.join(CSRC, "host.cpp"),
                os.path.join(ROOT, "tests", "asan", "driver.cpp")):
        cc("g++", src, hip_host)
    exe = str(tmp_path / "asan_driver")
    subprocess.run(["g++", *san, "-fopenmp", *objs, f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib",
                    "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([exe, str(tmp_path / "field.bin")], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "asan driver ok" in r.stdout
