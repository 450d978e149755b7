"""Build the gfx950 shared library asend3d_amd/libasend_gpu.so in-tree.

    python -m asend3d_amd.build        (or asend3d_amd.build.build())

also links the job-file driver asend3d_amd/asend_run_job against it.

HIP sources are compiled with hipcc for gfx950 only; the host-only preparation
code with g++.  Objects go to _build/ at the repository root.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "libasend_gpu.so")
CLI = os.path.join(PKG, "asend_run_job")
CLI_SRC = os.path.join(PKG, "cli", "asend_run_job.cpp")
BUILD = os.path.join(ROOT, "_build")

HIP_SOURCES = ["kernels.hip", "precond.hip", "learned.hip", "dynamics.hip", "batched.hip", "shell.hip", "brickim.hip", "modal.hip", "nonlinear.hip", "ldl.hip", "context.hip"]
CXX_SOURCES = ["host_prep.cpp"]
OFFLOAD_ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={OFFLOAD_ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
    "-I", CSRC,
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", CSRC]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        if not os.path.exists(src):
            continue
        obj = os.path.join(BUILD, s + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    for s in CXX_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s + ".o")
        jobs.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    # the job-file driver, linked against the library next to it
    cli_tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), CLI_SRC, "-o", cli_tmp,
          "-L", os.path.dirname(OUT), "-lasend_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(cli_tmp, CLI)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
