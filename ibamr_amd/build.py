"""Builds the HIP library libibamr_amd.so in-tree for gfx950.

    python -m ibamr_amd.build        (or python ibamr_amd/build.py)

Every source is compiled to an object in parallel, then the objects are linked
into one shared library.

Flags:
  --offload-arch=gfx950                    MI355X only
  -ffp-contract=off                        no implicit FMA: the FP64 forms reproduce
                                           the reference's separately rounded operations
                                           (FP32 paths use explicit fmaf where wanted)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                           no packed FP32 instructions in device code.
                                           The host pass prints that it does not know
                                           the feature; the device pass honours it.

The device assembly of every unit (kept with -save-temps in a scratch folder) is
scanned for packed FP32 arithmetic (v_pk_add_f32, v_pk_mul_f32, v_pk_fma_f32);
the build fails if any is found, so the flag's effect is checked on every build.
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "lib", "libibamr_amd.so")
SOURCES = ["le_plan.hip", "le_bin.hip", "le_spread.hip", "le_interp.hip", "le_ghost.hip", "le_redist.hip",
           "le_fortran_abi.hip", "ib_forces.hip", "fe_quadrature.hip", "ibtk_facade.cpp"]
DEVICE_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-Wno-unused-command-line-argument"]
PACKED_FP32 = re.compile(r"\bv_pk_(?:add|mul|fma)_f32\b")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


def packed_fp32_lines(asm_text):
    """Lines of device assembly that use packed FP32 arithmetic."""
    return [l.strip() for l in asm_text.splitlines() if PACKED_FP32.search(l) and not l.lstrip().startswith(";")]


def _compile(src, obj, tmpdir, verbose):
    cmd = [hipcc()] + DEVICE_FLAGS + ["-c", "-save-temps", "-I", os.path.join(ROOT, "include"), "-I", SRC,
                                      "-o", obj, src]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=tmpdir)
    if r.returncode != 0:
        return "hipcc failed on %s:\n%s%s" % (src, r.stdout, r.stderr)
    stem = os.path.splitext(os.path.basename(src))[0]
    bad = []
    for f in os.listdir(tmpdir):
        if f.startswith(stem + "-hip-amdgcn") and f.endswith(".s"):
            with open(os.path.join(tmpdir, f)) as fh:
                bad += packed_fp32_lines(fh.read())
    if bad:
        return "packed FP32 instructions in the device code of %s: %s" % (src, "; ".join(bad[:5]))
    return None


def build(verbose=False, jobs=None):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    srcs = [os.path.join(SRC, s) for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    work = tempfile.mkdtemp(prefix="ibamr_amd_build_")
    try:
        tasks = []
        for s in srcs:
            d = os.path.join(work, os.path.basename(s))
            os.makedirs(d)
            tasks.append((s, os.path.join(d, os.path.basename(s) + ".o"), d))
        jobs = jobs or max(1, min(len(tasks), os.cpu_count() or 1, 16))
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            errs = [e for e in ex.map(lambda t: _compile(t[0], t[1], t[2], verbose), tasks) if e]
        if errs:
            sys.stderr.write("\n".join(errs) + "\n")
            raise RuntimeError("building %s failed" % OUT)
        tmp = OUT + ".tmp"
        cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [t[1] for t in tasks]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("linking %s failed" % OUT)
        os.replace(tmp, OUT)
    finally:
        shutil.rmtree(work, ignore_errors=True)
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
