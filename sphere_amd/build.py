"""Build libsphere_hip.so in-tree with hipcc for gfx950.

Usage: python -m sphere_amd.build   (or __graft_entry__.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(CSRC, "_build")
LIB = os.path.join(HERE, "libsphere_hip.so")

ARCH = "gfx950"
# -ffp-contract=off: no fused multiply-add anywhere, so every FP64 expression
# rounds exactly as written.  Packed FP32 instructions are disabled.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]
SOURCES = ["dem_kernels.hip", "darcy_kernels.hip", "net_kernels.hip", "net.hip", "driver.hip",
           "ns.hip", "fileio.cpp"]


def hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(p) or p == "hipcc":
            return p
    return "hipcc"


def build(verbose=False, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [hipcc()] + HIP_FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if out.returncode != 0:
        raise RuntimeError("link failed:\n" + out.stdout.decode())
    os.replace(tmp, LIB)
    return LIB


def _drain(procs):
    errs = []
    while procs:
        src, p = procs.pop(0)
        out, _ = p.communicate()
        text = out.decode()
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            errs.append("%s:\n%s" % (src, "\n".join(lines)))
        elif lines:
            print("\n".join(lines), file=sys.stderr)
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(errs))


if __name__ == "__main__":
    print(build(verbose=True))
