"""Host AddressSanitizer + UndefinedBehaviorSanitizer run of the host C code
the GPU tests rely on: the synthetic-case generator, the FP64 oracle and the
netCDF classic reader/writer (tests/asan_driver.cpp).  GPU code is not built
with a sanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_code_under_asan(tmp_path):
    flags = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    objs = []
    for src, std in ((os.path.join(ROOT, "fv3atm_amd", "csrc", "test_cases.c"), "-std=c11"),
                     (os.path.join(ROOT, "oracle", "fv3_oracle_tp.c"), "-std=c11"),
                     (os.path.join(ROOT, "oracle", "fv3_oracle_mapz.c"), "-std=c11"),
                     (os.path.join(ROOT, "oracle", "fv3_oracle_l2e.c"), "-std=c11"),
                     (os.path.join(ROOT, "oracle", "fv3_oracle_dyn.c"), "-std=c11")):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        subprocess.run(["gcc", std, "-ffp-contract=off", "-fopenmp", *flags, "-c", src, "-o", obj], check=True)
        objs.append(obj)
    exe = str(tmp_path / "asan_driver")
    subprocess.run(["g++", "-std=c++17", *flags, "-fopenmp", os.path.join(ROOT, "tests", "asan_driver.cpp"), *objs,
                    "-o", exe, "-lm"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="2")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240, env=env)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "all checks passed" in r.stdout
