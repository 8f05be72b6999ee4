"""The host-side C++ (the CPU oracle, the OBJ/MTL loader, the BMP writer and
the scene generators) built with AddressSanitizer and
UndefinedBehaviorSanitizer by g++ and run on small inputs and edge cases
(tests/native/sanitize_driver.cpp). Any memory error, leak or undefined
behaviour aborts the driver; its own consistency checks make it exit
non-zero."""
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
SOURCES = [os.path.join(ROOT, "tests", "native", "sanitize_driver.cpp"), os.path.join(ROOT, "oracle", "rt_oracle.cpp"),
           os.path.join(ROOT, "cuda-rt_amd", "csrc", "rt_io.cpp"), os.path.join(ROOT, "cuda-rt_amd", "csrc", "rt_scenes.cpp")]


@pytest.mark.timeout(600)
def test_host_code_under_asan_and_ubsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "sanitize_driver")
    flags = ["-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-fno-omit-frame-pointer",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-I" + os.path.join(ROOT, "include")]
    r = subprocess.run(["g++", *flags, "-o", exe, *SOURCES], capture_output=True, text=True, timeout=500)
    assert r.returncode == 0, r.stderr[-4000:]
    with gzip.open(os.path.join(GOLDEN, "logcabin.obj.gz"), "rb") as f:
        (tmp_path / "logcabin.obj").write_bytes(f.read())
    shutil.copy(os.path.join(GOLDEN, "logcabin.mtl"), tmp_path / "logcabin.mtl")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="3")
    r = subprocess.run([exe, str(tmp_path / "logcabin.obj"), str(tmp_path) + "/", str(tmp_path)], capture_output=True,
                       text=True, timeout=300, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-6000:]
    assert "all checks passed" in r.stdout
