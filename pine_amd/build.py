"""Builds the native library pine_amd/libpinehip.so (gfx950) in-tree.

Host translation units are compiled by g++ and the HIP ones by hipcc; all of
them with -ffp-contract=off so that float expressions round step by step as
pine's x86-64 build does.  Packed FP32 instructions are disabled for device
code (-packed-fp32-ops).
"""
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpinehip.so")
CLI = os.path.join(HERE, "pine")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

HOST_SRC = ["bvh_build.cpp", "workload.cpp", "scene_parse.cpp"]
HIP_SRC = ["accel.hip", "render.hip", "guide.hip", "sort.hip"]

CXXFLAGS = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIPFLAGS = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def kernel_resources(src):
    """{kernel symbol: {"VGPRs", "ScratchSize", "Occupancy", ...}} from the
    report the last build of `src` (a HIP_SRC name) wrote, or None."""
    path = os.path.join(OBJ, src + ".resources.txt")
    if not os.path.exists(path):
        return None
    out, name = {}, None
    for line in open(path):
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.search(r"remark:\s+(\w+)(?: \[[^]]*\])?: (\d+)", line)
        if m and name:
            out[name][m.group(1)] = int(m.group(2))
    return out


def build(verbose=False):
    """Compiles every source and links libpinehip.so; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HOST_SRC:
        o = os.path.join(OBJ, s + ".o")
        jobs.append(["g++", *CXXFLAGS, "-c", os.path.join(CSRC, s), "-o", o])
    for s in HIP_SRC:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([HIPCC, *HIPFLAGS, "-Rpass-analysis=kernel-resource-usage", "-c",
                     os.path.join(CSRC, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    for job, out in zip(jobs, outs):
        if job[0] == HIPCC:
            # The compiler's per-kernel register, scratch and occupancy report.
            with open(job[-1][:-2] + ".resources.txt", "w") as f:
                f.write("\n".join(l for l in out.splitlines() if "kernel-resource-usage" in l) + "\n")
        elif verbose and out.strip():
            print(out)
    objs = [os.path.join(OBJ, s + ".o") for s in HOST_SRC + HIP_SRC]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    # The `pine` command line, linked against the library next to it.
    _run(["g++", *CXXFLAGS, "-o", CLI + ".tmp", os.path.join(CSRC, "pine_cli.cpp"), LIB,
          "-Wl,-rpath,$ORIGIN"])
    os.replace(CLI + ".tmp", CLI)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
