"""Build libmfvm.so (hand-written HIP for gfx950) in-tree with hipcc."""

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(PKG), "csrc")
INCLUDE = os.path.join(os.path.dirname(os.path.dirname(PKG)), "include")
LIB = os.path.join(PKG, "lib", "libmfvm.so")
SOURCES = ("capi.hip", "ns.hip", "bc.hip", "heat.hip", "maxwell.hip", "ensemble.hip", "closure.hip", "diag.hip")
ARCH = os.environ.get("MFVM_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # IEEE-faithful arithmetic: no FMA contraction, no fast-math.
    "-ffp-contract=off", "-fno-fast-math",
    # Packed FP32 instructions disabled.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    return res.returncode, out


def build(verbose=False, extra=()):
    """Compile every source to an object (in parallel), then link libmfvm.so."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(os.path.dirname(LIB), "obj")
    os.makedirs(objdir, exist_ok=True)
    comp = [f for f in FLAGS if f != "-shared"]
    jobs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        jobs.append((obj, ["hipcc", *comp, *extra, "-I", INCLUDE, "-c",
                           os.path.join(CSRC, src), "-o", obj]))
    if verbose:
        for _, cmd in jobs:
            print(" ".join(cmd), flush=True)
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "8"))))
    with ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(lambda j: _run(j[1]), jobs))
    for (obj, _), (rc, out) in zip(jobs, results):
        if rc != 0:
            raise RuntimeError(f"hipcc failed ({rc}) for {os.path.basename(obj)}:\n{out}")
        if verbose and out.strip():
            print(out)
    tmp = LIB + ".tmp"
    link = ["hipcc", *FLAGS, *extra, "-Wl,-z,defs", *[o for o, _ in jobs], "-o", tmp]
    if verbose:
        print(" ".join(link), flush=True)
    rc, out = _run(link)
    if rc != 0:
        raise RuntimeError(f"hipcc link failed ({rc}):\n{out}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True, extra=tuple(sys.argv[1:]))
