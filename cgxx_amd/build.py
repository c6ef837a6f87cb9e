"""In-tree build of the native parts (hipcc for gfx950, g++ for host code).

    python -m cgxx_amd.build          # library + cg_hip executable + oracle

Outputs (git-ignored, shipped with the tree):
    cgxx_amd/lib/libcghip.so   the C-ABI library (include/cghip.h)
    cgxx_amd/bin/cg_hip        the CGHIP backend executable
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib", "libcghip.so")
BIN = os.path.join(HERE, "bin", "cg_hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CGHIP_ARCH", "gfx950")

# Contraction off: a*x+y stays a rounded multiply and a rounded add, as in
# the reference's CPU kernels.  Packed FP32 instructions are disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
]

LIB_SOURCES = ["cghip.hip", "kernels.hip", "kernels_mixed.hip", "cghip_mlp.hip", "kernels_mlp.hip",
               "kernels_features.hip", "noalloc.hip"]
HOST_SOURCES = ["CG.cpp", "Matrix.cpp", "Preconditioner.cpp", "CGHIP.cpp", "LearnedPoly.cpp", "main.cpp"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


RESOURCES = os.path.join(HERE, "lib", "kernel_resources.json")
_REMARK_KEYS = {"TotalSGPRs": "sgpr", "VGPRs": "vgpr", "AGPRs": "agpr",
                "ScratchSize [bytes/lane]": "scratch", "Occupancy [waves/SIMD]": "occupancy",
                "LDS Size [bytes/block]": "lds"}


def parse_resource_remarks(text: str) -> dict[str, dict[str, int]]:
    """Per-kernel register, scratch, LDS and occupancy figures from the
    compiler's -Rpass-analysis=kernel-resource-usage remarks."""
    kernels: dict[str, dict[str, int]] = {}
    cur = None
    for line in text.splitlines():
        if "remark:" not in line:
            continue
        body = line.split("remark:", 1)[1].split(" [-Rpass", 1)[0].strip()
        if body.startswith("Function Name:"):
            cur = kernels.setdefault(body.split(":", 1)[1].strip(), {})
            continue
        key, _, val = body.rpartition(":")
        if cur is not None and key.strip() in _REMARK_KEYS:
            try:
                cur[_REMARK_KEYS[key.strip()]] = int(val.strip())
            except ValueError:
                pass
    names = list(kernels)
    try:
        r = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True,
                           check=True)
        demangled = r.stdout.splitlines()
        if len(demangled) == len(names):
            return {d: kernels[m] for m, d in zip(names, demangled)}
    except (OSError, subprocess.CalledProcessError):
        pass
    return kernels


def build_library() -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in LIB_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", "-shared", "-o", tmp, *srcs]
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # drop the remarks and the source lines the compiler echoes under them
    rest = "\n".join(l for l in r.stdout.splitlines()
                     if "packed-fp32-ops" not in l and "remark:" not in l
                     and not re.match(r"^\s*\d*\s*\|", l) and l.strip())
    if rest.strip():
        print(rest, flush=True)
    if r.returncode:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")
    with open(RESOURCES, "w") as f:
        json.dump(parse_resource_remarks(r.stdout), f, indent=1, sort_keys=True)
    os.replace(tmp, LIB)
    return LIB


def build_host() -> str | None:
    srcs = [os.path.join(HOST, s) for s in HOST_SOURCES]
    if not all(os.path.exists(s) for s in srcs):
        return None
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    tmp = BIN + ".tmp"
    _run(["g++", "-std=c++17", "-O3", "-DNDEBUG", "-Wall", "-I" + os.path.join(REPO, "include"),
          "-pthread", "-o", tmp, *srcs, "-L" + os.path.dirname(LIB), "-lcghip",
          "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, BIN)
    return BIN


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


def build_all() -> None:
    build_library()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_library, "host": build_host, "oracle": build_oracle}[w]()
