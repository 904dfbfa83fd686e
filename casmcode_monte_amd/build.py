"""Build recipe of the HIP library (libcasm_monte_mi355x.so) for gfx950.

Every source under csrc/ is compiled with hipcc into one shared library in
casmcode_monte_amd/lib/.  Flags:
  -ffp-contract=off      the EXACT mode's FP64 tails must round every operation
                         the way the reference's host build does (no FMA);
  -target-feature -packed-fp32-ops   no packed FP32 instructions are emitted.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libcasm_monte_mi355x.so")
# measured-peak probes (not part of the Metropolis path; loaded by bench.py)
PROBES_LIB = os.path.join(LIBDIR, "libcasm_monte_probes.so")
SOURCES = ["api.hip", "exact.hip", "canonical.hip", "occupant.hip", "kmc.hip", "nfold.hip",
           "checker.hip",
           "checker_ring.hip", "checker_slab.hip", "fcc.hip", "stats.hip"]
HEADERS = ["internal.hh", "rng.hh", "walk.hh", "fcc.hh"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-value", "-Wno-unused-result",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build(verbose=False):
    """Compile every HIP source for gfx950 and link the shared library."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]

    def compile_one(src):
        obj = os.path.join(objdir, src + ".o")
        out = _run([HIPCC, *FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES) + 1) as ex:
        probes = ex.submit(compile_one, "probes.hip")
        objs = list(ex.map(compile_one, SOURCES))
        probes_obj = probes.result()
    ptmp = PROBES_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", ptmp, probes_obj])
    os.replace(ptmp, PROBES_LIB)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    shutil.rmtree(objdir, ignore_errors=True)
    build_python(verbose)
    return LIB


def build_python(verbose=False):
    """Compile the pybind11 module _libcasm_monte (python/libcasm_monte.cpp): the
    libcasm.monte names over the C++ drop-in, linked to the HIP library."""
    import sysconfig

    import pybind11
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    out = os.path.join(HERE, "_libcasm_monte" + ext)
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
           "-I", os.path.join(ROOT, "include"), "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"],
           os.path.join(HERE, "python", "libcasm_monte.cpp"),
           "-L", LIBDIR, "-lcasm_monte_mi355x", "-Wl,-rpath,$ORIGIN/lib", "-o", tmp]
    o = _run(cmd)
    if verbose and o.strip():
        print(o)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
