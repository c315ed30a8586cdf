"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer.

Builds the library's host code with -fsanitize=address,undefined (host side
only: -Xarch_host; the gfx950 device code is built without sanitizers) into a
driver that exercises the C ABI paths that run without a GPU -- FEN read and
write, move generation, the -cpu perft with and without the host TT, TT
snapshot/resume, the workload walk and u128 printing -- and runs it."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_host_code_under_asan_ubsan(tmp_path):
    src = os.path.join(ROOT, "perft_gpu_2026_amd", "csrc")
    flags = ["-O1", "-g", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-Xarch_host", "-fsanitize=address",
             "-Xarch_host", "-fsanitize=undefined", "-Xarch_host", "-fno-omit-frame-pointer",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    objs = []
    procs = []
    for f in ("pg_kernels", "pg_host", "pg_tt", "pg_batch", "pg_serve", "pg_launcher_api", "pg_probe", "pg_magic"):
        o = str(tmp_path / f"{f}.o")
        objs.append(o)
        procs.append(subprocess.Popen([HIPCC] + flags + ["-c", os.path.join(src, f + ".hip"), "-o", o],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
    for p in procs:
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, out.decode()[-2000:]
    drv = str(tmp_path / "driver")
    subprocess.run([HIPCC, "-O1", "-g", "-std=c++17", "-x", "c++", "-fsanitize=address,undefined", "-c",
                    os.path.join(ROOT, "tests", "sanitize", "host_driver.cpp"), "-o", drv + ".o"], check=True,
                   capture_output=True)
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-fsanitize=address,undefined", "-o", drv, drv + ".o"] + objs +
                   ["-lpthread"], check=True, capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([drv], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "asan driver: ok" in r.stdout, r.stdout + r.stderr
