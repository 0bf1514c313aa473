"""Builds the native pieces in-tree (run by __graft_entry__.build()):

  libgint_amd/libgint.so        the K-build library (host engine + gfx950 kernels)
  libgint_amd/data/tcg_c0.bin   truncated-Coulomb coefficient table (generated)
  oracle/_build/liboracle.so    CPU oracle (test checker)
  oracle/_ref/libref.so         reference tables, where the reference checkout exists
"""
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
LIB = os.path.join(HERE, "libgint.so")
TCG_BIN = os.path.join(HERE, "data", "tcg_c0.bin")
# SHA-256 of the generated table (207 patches x 22 orders x 105 coefficients)
TCG_SHA256 = None
REF_SRC = "/root/reference/src"

HOST_SRC = ["engine.cpp", "abi.cpp", "plan.cpp", "boys_table.cpp", "c0_file.cpp", "server.cpp"]
DEV_SRC = ["kernels.hip", "kernels_tq.hip", "kernels_eco.hip"]
# kernels_tq_part.hip is built once per part of the templated class kernels
TQ_PARTS = 8
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
HOSTFLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
             "-I" + os.path.join(ROCM, "include"), "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]))
    return r


def build_tcg(verbose=False, force=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(TCG_BIN), exist_ok=True)
    gen = os.path.join(BUILD, "gen_tcg")
    _run(["gcc", "-O2", "-fopenmp", "-o", gen, os.path.join(HERE, "tcg", "gen_tcg.c"), "-lquadmath", "-lm"], verbose)
    if force or not os.path.exists(TCG_BIN):
        r = _run([gen, TCG_BIN + ".tmp", "21"], verbose)
        os.replace(TCG_BIN + ".tmp", TCG_BIN)
        with open(os.path.join(BUILD, "tcg_fit.txt"), "w") as f:
            f.write(r.stdout)
    return TCG_BIN


def tcg_sha256():
    with open(TCG_BIN, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def build_lib(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    steps = []
    for src in HOST_SRC:
        o = os.path.join(BUILD, src + ".o")
        steps.append((["g++"] + HOSTFLAGS + ["-c", os.path.join(CSRC, src), "-o", o], o))
    for src in DEV_SRC:
        o = os.path.join(BUILD, src + ".o")
        steps.append(([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, src), "-o", o], o))
    for part in range(TQ_PARTS):
        o = os.path.join(BUILD, "kernels_tq_part%d.o" % part)
        steps.append(([HIPCC] + HIPFLAGS + ["-DGINT_TQ_PART=%d" % part, "-c", os.path.join(CSRC, "kernels_tq_part.hip"),
                       "-o", o], o))
    # the compiles are independent: run them side by side (at most 16 jobs)
    jobs = jobs or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda st: _run(st[0], verbose), steps))
    objs = [o for _, o in steps]
    tmp = LIB + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp] + objs +
         ["-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")], verbose)
    os.replace(tmp, LIB)
    return LIB


DRV = os.path.join(HERE, "libgintdrv.so")


def build_driver(verbose=False):
    """the caller-side driver (synthetic systems, screening, call streams)"""
    os.makedirs(BUILD, exist_ok=True)
    tmp = DRV + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", "-o", tmp, os.path.join(CSRC, "driver.cpp")], verbose)
    os.replace(tmp, DRV)
    return DRV


def build_oracle(verbose=False):
    out = os.path.join(ROOT, "oracle", "_build")
    os.makedirs(out, exist_ok=True)
    so = os.path.join(out, "liboracle.so")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-std=c99", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
          "-o", so, os.path.join(ROOT, "oracle", "oracle.c"), "-lm"], verbose)
    return so


def build_ref(verbose=False):
    """the reference's self-contained plain-C tables, compiled where they lie,
    and its Fortran binding module linked with a caller against libgint.so"""
    if not os.path.isdir(REF_SRC):
        return None
    out = os.path.join(ROOT, "oracle", "_ref")
    os.makedirs(out, exist_ok=True)
    so = os.path.join(out, "libref.so")
    _run(["g++", "-O1", "-fPIC", "-shared", "-w", "-I" + REF_SRC, "-o", so,
          os.path.join(ROOT, "oracle", "ref_shim.cpp")], verbose)
    fc = os.path.join(ROCM, "llvm", "bin", "amdflang")
    mod = os.path.join(REF_SRC, "interface_libgint.F90")
    if os.path.exists(fc) and os.path.exists(mod) and os.path.exists(LIB):
        modir = os.path.join(BUILD, "fmod")
        os.makedirs(modir, exist_ok=True)
        _run([fc, "-O1", "-J", modir, "-o", os.path.join(out, "drive_libgint"), mod,
              os.path.join(ROOT, "tests", "fortran", "drive_libgint.F90"), LIB,
              "-Wl,-rpath,$ORIGIN/../../libgint_amd", "-Wl,-rpath," + os.path.join(ROCM, "lib")], verbose)
    return so


def build_all(verbose=False):
    build_tcg(verbose)
    build_oracle(verbose)
    build_driver(verbose)
    build_lib(verbose)
    build_ref(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
