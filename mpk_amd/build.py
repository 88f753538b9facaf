"""Build step of libmpkx.so (gfx950) -- run by __graft_entry__.build().

Host sources (*.cpp) are compiled with g++, device sources (*.hip) with
hipcc for gfx950, all as position-independent objects under
mpk_amd/_build/, then linked into mpk_amd/libmpkx.so.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmpkx.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CXXFLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-fvisibility=hidden"]
HIPFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            # packed FP32 instructions are kept out of every device code object
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-fno-gpu-flush-denormals-to-zero", "-Wno-unused-command-line-argument",
            # FP64 kernels must round every product and sum separately (the
            # reference's operation order); FP32 code calls fmaf explicitly
            "-ffp-contract=off",
            "-fvisibility=hidden"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.stderr


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for f in sorted(os.listdir(SRC)):
        src = os.path.join(SRC, f)
        obj = os.path.join(OUT, f + ".o")
        if f.endswith(".cpp"):
            jobs.append(["g++"] + CXXFLAGS + ["-c", src, "-o", obj])
        elif f.endswith(".hip"):
            jobs.append([HIPCC] + HIPFLAGS + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for msg in ex.map(_run, jobs):
            if verbose and msg.strip():
                sys.stderr.write(msg)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
