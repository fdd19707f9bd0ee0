"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU, no
GPU): the graph loaders and builders (graph.cpp), the label import and the
label-dump loader (index.hip's host side) driven by tests/sanitize_driver.cpp
over valid, random and corrupt inputs."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CS = os.path.join(ROOT, "gpu4hophl_code_amd", "csrc")
ROCM = "/opt/rocm"


@pytest.mark.slow
def test_host_code_under_asan_and_ubsan(tmp_path):
    clang = os.path.join(ROCM, "llvm", "bin", "clang++")
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    if not (os.path.exists(clang) and os.path.exists(hipcc)):
        pytest.skip("ROCm compilers not installed")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CS]
    san = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"]
    objs = []
    # host-only sources with the host compiler
    for src, obj in ((os.path.join(CS, "graph.cpp"), "graph.o"),
                     (os.path.join(ROOT, "tests", "sanitize_driver.cpp"), "drv.o")):
        subprocess.check_call([clang, "-O1", "-g", "-std=c++17", "-fPIC", "-pthread", *san, *inc,
                               "-c", src, "-o", str(tmp_path / obj)])
        objs.append(str(tmp_path / obj))
    # HIP sources: sanitizers on their host side only
    for src in ("index.hip", "parents.hip"):
        obj = str(tmp_path / (src + ".o"))
        subprocess.check_call([hipcc, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
                               *inc, "-Xarch_host", "-fsanitize=address", "-Xarch_host",
                               "-fsanitize=undefined", "-Xarch_host", "-fno-sanitize=function",
                               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                               "-c", os.path.join(CS, src), "-o", obj],
                              stderr=subprocess.DEVNULL)
        objs.append(obj)
    exe = str(tmp_path / "drv")
    subprocess.check_call([clang, *san, *objs, "-o", exe, "-L", os.path.join(ROCM, "lib"),
                           "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib"), "-pthread"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    work = tmp_path / "files"
    work.mkdir()
    out = subprocess.run([exe, str(work)], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stdout + out.stderr[-4000:]
    assert "sanitize pass" in out.stdout
