"""Build the HIP library ``ecogen_amd/libecogen_amd.so`` for gfx950 in-tree.

Usage: ``python -m ecogen_amd.build`` (also called by ``__graft_entry__.build``).
Every translation unit is compiled by hipcc with -ffp-contract=off (the FP64
path reproduces the reference's uncontracted arithmetic) and packed-FP32
instruction selection disabled.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libecogen_amd.so")

ARCH = os.environ.get("ECG_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

SOURCES = ["ecg_api.hip", "ecg_mlp.hip"] + [f"ecg_step_{m}_{p}.hip" for m in ("euler", "pueq", "ueq") for p in ("f64", "f32")]

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


HOST = os.path.join(HERE, "host")
HOST_SOURCES = ["xml.cpp", "ecogen.cpp", "output.cpp"]
RUNNER = os.path.join(HERE, "ecogen_run")


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(os.path.basename(src))[0] + ".o")
    path = src if os.path.isabs(src) else os.path.join(CSRC, src)
    cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4)))
    srcs = SOURCES + [os.path.join(HOST, h) for h in HOST_SOURCES]
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # Command-line runner for cases in the reference's XML format.
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-O2", "-std=c++17", os.path.join(HOST, "ecogen_run.cpp"), "-o",
           RUNNER + ".tmp", "-L" + HERE, "-lecogen_amd", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"runner build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(RUNNER + ".tmp", RUNNER)
    if verbose:
        print(f"built {LIB} and {RUNNER}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
