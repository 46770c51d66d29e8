"""Build the native parts in-tree: libfv3atm_amd.so (HIP, gfx950),
libfv3_cases.so (host C) and the Fortran interface layer (amdflang).

Usage: python -m fv3atm_amd.build
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
FC = os.path.join(ROCM, "bin", "amdflang")
ARCH = "gfx950"

# contraction off everywhere except the FP32 fused kernels; packed FP32 math
# off; the FP32 remap takes its inputs to be free of NaNs (no quieting of
# loaded values before each min/max, limiter selects folded into min/max)
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SOURCES = {
    "tp_plain.hip": ["-ffp-contract=off"],
    "mapz_plain.hip": ["-ffp-contract=off"],
    "capi.hip": ["-ffp-contract=off"],
    "fused.hip": ["-ffp-contract=off"],
    "fused32.hip": ["-ffp-contract=fast"],  # compiled once per part (FUSED32_NPART)
    "mapz_fast.hip": ["-ffp-contract=fast", "-fno-honor-nans"],
    "tlad.hip": ["-ffp-contract=off"],
    "noalloc.hip": ["-ffp-contract=off"],
    "halo.hip": ["-ffp-contract=off"],
    "l2e.hip": ["-ffp-contract=off"],
    "io.hip": ["-ffp-contract=off"],
    "dyncore.hip": ["-ffp-contract=off"],
    "gsum.hip": ["-ffp-contract=off"],
}


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops' is not a recognized" not in l
                    and "warning: feature flag" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


# the FP32 fused kernel's launch configurations are spread over this many
# objects (fv3_internal.h FUSED32_NPART) so that they compile in parallel
FUSED32_NPART = 6


def build_hip():
    os.makedirs(OBJ, exist_ok=True)
    jobs = []  # (source, object, extra flags); the longest compiles first
    for s in HIP_SOURCES:
        if not os.path.exists(os.path.join(CSRC, s)):
            continue
        if s == "fused32.hip":
            jobs += [(s, f"fused32_p{p}.o", [f"-DFV3_F32_PART={p}"]) for p in range(FUSED32_NPART)]
        else:
            jobs.append((s, s.replace(".hip", ".o"), []))
    jobs.sort(key=lambda j: j[0] not in ("fused32.hip", "tlad.hip", "fused.hip"))

    def one(job):
        s, o, extra = job
        obj = os.path.join(OBJ, o)
        _run([HIPCC, *HIP_FLAGS, *HIP_SOURCES[s], *extra, "-c", os.path.join(CSRC, s), "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        objs = list(ex.map(one, jobs))
    out = os.path.join(HERE, "libfv3atm_amd.so")
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", out + ".tmp", *objs])
    os.replace(out + ".tmp", out)
    return out


def build_cases():
    out = os.path.join(HERE, "libfv3_cases.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-o", out + ".tmp",
          os.path.join(CSRC, "test_cases.c"), "-lm"])
    os.replace(out + ".tmp", out)
    return out


def build_fortran():
    """The ISO_C_BINDING layer (Fortran interfaces of the path) and its
    harness executable, for the 8-byte build and for the 32-bit build
    (OVERLOAD_R4: 4-byte reals, the FP32 kernels); skipped when amdflang is
    absent."""
    fdir = os.path.join(HERE, "fortran")
    srcs = ["fv_arrays_stub.F90", "fv3atm_amd_cabi.F90", "tp_core_mod.F90", "fv_tracer2d_mod.F90",
            "fv_mapz_mod.F90"]
    if not os.path.exists(FC) or not all(os.path.exists(os.path.join(fdir, s)) for s in srcs):
        return None
    out = []
    for tag, defs in (("", []), ("_r4", ["-DOVERLOAD_R4"])):
        mod = os.path.join(OBJ, "fmod" + tag)
        os.makedirs(mod, exist_ok=True)
        lib = os.path.join(HERE, f"libfv3atm_amd_f{tag}.so")
        _run([FC, "-O2", "-fPIC", "-shared", *defs, "-module-dir", mod, "-o", lib + ".tmp",
              *[os.path.join(fdir, s) for s in srcs], "-L" + HERE, "-lfv3atm_amd",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(lib + ".tmp", lib)
        harness = os.path.join(fdir, "harness.F90")
        if os.path.exists(harness):
            exe = os.path.join(HERE, "fv3_harness" + tag)
            _run([FC, "-O2", *defs, "-I" + mod, "-o", exe + ".tmp", harness, "-L" + HERE,
                  f"-lfv3atm_amd_f{tag}", "-lfv3atm_amd", "-Wl,-rpath,$ORIGIN"])
            os.replace(exe + ".tmp", exe)
        out.append(lib)
    return " ".join(out)


def build_oracle():
    sh = os.path.join(ROOT, "oracle", "build_oracle.sh")
    if os.path.exists(sh):
        _run(["sh", sh])


def build_all(verbose=True):
    steps = [("cases", build_cases), ("oracle", build_oracle), ("hip", build_hip),
             ("fortran", build_fortran)]
    for name, fn in steps:
        r = fn()
        if verbose:
            print(f"[build] {name}: {r or 'ok'}", flush=True)


if __name__ == "__main__":
    build_all()
    sys.exit(0)
