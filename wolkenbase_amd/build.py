"""Builds the native libraries in-tree.

  libwb_gpu.so      HIP kernels + C ABI (include/wb_gpu.h), gfx950 only
  libwb_synth.so    seeded synthetic tiles (C)
  libwb_hostmath.so the device math built for the host, for CPU-side tests

Run as ``python -m wolkenbase_amd.build`` or through __graft_entry__.build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
    "-Wno-unused-result", "-Wno-unused-value",
]
GPU_SOURCES = ["wb_build.hip", "wb_classify.hip", "wb_api.hip", "wb_las.hip", "wb_dump.hip"]
HOST_SOURCES = ["wb_host.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build_gpu(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in GPU_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", "-o", o, os.path.join(CSRC, s)])
    for s in HOST_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append(["g++", "-std=gnu++17", "-O2", "-fPIC", "-ffp-contract=off",
                     "-I", CSRC, "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                     "-c", "-o", o, os.path.join(CSRC, s)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [os.path.join(OBJ, s + ".o") for s in GPU_SOURCES + HOST_SOURCES]
    out = os.path.join(HERE, "libwb_gpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lm"])
    os.replace(tmp, out)
    return out


def build_synth():
    out = os.path.join(HERE, "libwb_synth.so")
    _run(["gcc", "-std=gnu11", "-O2", "-fPIC", "-ffp-contract=off", "-shared", "-o", out,
          os.path.join(CSRC, "wb_synth.c"), "-lm"])
    return out


def build_hostmath():
    out = os.path.join(HERE, "libwb_hostmath.so")
    _run(["g++", "-std=gnu++17", "-O2", "-fPIC", "-ffp-contract=off", "-shared", "-o", out,
          os.path.join(CSRC, "wb_hostcheck.cpp"), os.path.join(CSRC, "wb_host.cpp"), "-lm"])
    return out


def build_all():
    with ThreadPoolExecutor(max_workers=3) as ex:
        fs = [ex.submit(build_gpu), ex.submit(build_synth), ex.submit(build_hostmath)]
        return [f.result() for f in fs]


if __name__ == "__main__":
    for p in build_all():
        print(p)
