"""Builds libd2dhip.so (gfx950) in-tree with hipcc.

    python 2decomp-fft_amd/build.py            # compile + link
    python 2decomp-fft_amd/build.py --jobs 4
    python 2decomp-fft_amd/build.py --incremental   # only stale objects

Every HIP source is compiled for gfx950 with packed FP32 instructions
disabled, then linked into 2decomp-fft_amd/libd2dhip.so next to this file.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libd2dhip.so")
# (longest first, so that a parallel build starts them first)
SOURCES = ["fast_variants_a.hip", "fast_variants_c.hip", "fast_variants_b.hip",
           "fast_variants_e.hip", "fast_variants_d.hip", "fast_kernels.hip", "mfma_kernels.hip",
           "fft_kernels.hip", "copy_kernels.hip", "halo_kernels.hip", "plan.cpp", "api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LIB = os.path.join(os.path.dirname(os.path.dirname(HIPCC)), "lib")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that warning
    # is expected and dropped here.
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    return src, r.returncode, err, out


def _stale(src):
    """True when src's object is missing or older than src or any header."""
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if not os.path.exists(out):
        return True
    deps = [os.path.join(CSRC, src)]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps += [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))]
    t = os.path.getmtime(out)
    return any(os.path.getmtime(d) > t for d in deps)


def build(jobs=4, verbose=False, incremental=False):
    """Compiles every source (only the stale ones with incremental=True) and
    links the library."""
    os.makedirs(OBJ, exist_ok=True)
    todo = [s for s in SOURCES if not incremental or _stale(s)]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, todo))
    results += [(s, 0, "", os.path.join(OBJ, os.path.splitext(s)[0] + ".o"))
                for s in SOURCES if s not in todo]
    objs = []
    for src, rc, err, out in results:
        if rc != 0:
            raise RuntimeError("hipcc failed on %s:\n%s" % (src, err))
        if verbose and err.strip():
            print(err, file=sys.stderr)
        objs.append(out)
    tmp = LIB + ".tmp"
    # rocTX for the profiler regions, found at run time where it was at link time
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L" + ROCM_LIB, "-lrocprofiler-sdk-roctx", "-Wl,-rpath," + ROCM_LIB]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_EXE = os.path.join(FORTRAN, "fft_hip_tests")


def build_fortran():
    """The Fortran modules (decomp_2d / decomp_2d_fft over the C ABI) and the
    test program that runs the reference's own checks through them.  Skipped
    with a notice when amdflang is absent."""
    flang = os.environ.get("FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(flang):
        print("build: amdflang not found, Fortran shim not built", file=sys.stderr)
        return None
    mod = os.path.join(OBJ, "fmod")
    os.makedirs(mod, exist_ok=True)
    cmd = [flang, "-cpp", "-O2", "-module-dir", mod,
           os.path.join(FORTRAN, "decomp_2d_hip.F90"), os.path.join(FORTRAN, "fft_hip_tests.F90"),
           "-L" + HERE, "-ld2dhip", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib", "-o", FORTRAN_EXE]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=OBJ)
    if r.returncode != 0:
        raise RuntimeError("amdflang failed:\n" + r.stderr)
    return FORTRAN_EXE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true",
                    help="recompile only sources whose object is older than them or a header")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose, args.incremental))
