"""Builds the in-tree shared library ``siconos_amd/lib/libsiconos_amd.so``.

Host C sources (``csrc/host``) are compiled with gcc, HIP sources
(``csrc/hip``) with hipcc for gfx950, and both are linked by hipcc.  Always
rebuilds.  Run as ``python -m siconos_amd.build`` or through
``__graft_entry__.build()``.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
INC = os.path.join(REPO, "include")
OBJ = os.path.join(ROOT, "_build")
LIB_DIR = os.path.join(ROOT, "lib")
LIB = os.path.join(LIB_DIR, "libsiconos_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-fvisibility=hidden", "-ffp-contract=off", "-I", INC]
# -ffp-contract=off: every product is rounded before the sum it feeds, as in
# the reference's C; the packed-fp32 feature is disabled on purpose.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-fvisibility=hidden", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value", "-I", INC]


def _hip_flags(src):
    """The translation unit's flags: *_fp32.hip holds FP32 kernels whose
    products may be fused into multiply-adds; every other unit keeps each
    product rounded (the FP64 sweeps' bit parity with the reference's C)."""
    if src.endswith("_fp32.hip"):
        return ["-ffp-contract=fast" if f == "-ffp-contract=off" else f for f in HIP_FLAGS]
    return HIP_FLAGS


def _sources(sub, ext):
    d = os.path.join(ROOT, "csrc", sub)
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext))


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    steps = []
    objs = []
    for src in _sources("host", ".c"):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        steps.append(["gcc", *C_FLAGS, "-c", src, "-o", obj])
        objs.append(obj)
    for src in _sources("hip", ".hip"):
        obj = os.path.join(OBJ, os.path.basename(src) + ".o")
        steps.append([HIPCC, *_hip_flags(src), "-x", "hip", "-c", src, "-o", obj])
        objs.append(obj)
    with cf.ThreadPoolExecutor(jobs) as ex:
        outs = list(ex.map(_run, steps))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
