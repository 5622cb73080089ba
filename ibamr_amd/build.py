"""Builds the HIP library libibamr_amd.so in-tree for gfx950.

    python -m ibamr_amd.build        (or python ibamr_amd/build.py)

Flags:
  --offload-arch=gfx950                    MI355X only
  -ffp-contract=off                        no implicit FMA: the FP64 forms reproduce
                                           the reference's separately rounded operations
                                           (FP32 paths use explicit fmaf where wanted)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed FP32 instructions
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib", "libibamr_amd.so")
SOURCES = ["le_plan.hip", "le_fortran_abi.hip", "ib_forces.hip", "fe_quadrature.hip", "ibtk_facade.cpp"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(SRC, s) for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "include"), "-I", SRC, "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % OUT)
    os.replace(tmp, OUT)
    return OUT


DRIVERS = {"interpolate_01": os.path.join(ROOT, "tests", "cpp", "interpolate_01.cpp")}
DRIVER_DIR = os.path.join(ROOT, "tests", "cpp", "bin")


def build_drivers(verbose=False):
    """C++ test drivers on the ibtk_amd layer (host code only: g++, linked to the library)."""
    os.makedirs(DRIVER_DIR, exist_ok=True)
    outs = []
    for name, src in DRIVERS.items():
        if not os.path.exists(src):
            continue
        out = os.path.join(DRIVER_DIR, name)
        cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", out + ".tmp", src,
               "-L", os.path.dirname(OUT), "-libamr_amd", "-Wl,-rpath,$ORIGIN/../../../ibamr_amd/lib"]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("g++ failed building %s" % out)
        os.replace(out + ".tmp", out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_drivers(verbose=True))
