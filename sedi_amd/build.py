"""Builds libsedi_amd.so (HIP for gfx950) and the sedi-estimator driver, in-tree.

    python -m sedi_amd.build          # or: python sedi_amd/build.py

Objects go to sedi_amd/_build/, the library to sedi_amd/libsedi_amd.so and the
CLI driver to sedi_amd/bin/sedi-estimator.  Every compile line carries
-ffp-contract=off: the kernels round each product and difference separately,
as the reference's expressions do (Stencil.hpp:43-50, SliceOps.hpp:38-40).
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libsedi_amd.so")
BIN = os.path.join(PKG, "bin", "sedi-estimator")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("SEDI_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# Packed-FP32 instructions are switched off for every device compile.  The
# march unrolls a tile's planes fully (#pragma unroll); the largest pass body
# (batched pipelined CG's x step, 32 planes) exceeds LLVM's default size limit
# for that, which left its ring of loads in scratch memory.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-mllvm", "-pragma-unroll-threshold=65536"]

LIB_SOURCES = ["runtime.cpp", "api.cpp", "mt_jump.cpp", "h5io.cpp", "tuning.cpp", "plain.hip", "rng.hip", "batched.hip",
               "batched_pcg.hip", "batched_bicg.hip", "estimator.hip", "serve.hip"]
DRIVER_SOURCE = os.path.join(CSRC, "driver", "sedi_estimator.cpp")


def _compile(src, extra=(), build_dir=BUILD):
    path = os.path.join(CSRC, src)
    obj = os.path.join(build_dir, src + ".o")
    # every source is HIP (the .cpp files hold host code that includes the
    # device headers), so all of them get the gfx950 device flags
    cmd = [HIPCC, "-x", "hip"] + DEVICE + COMMON + list(extra) + ["-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), res.stdout, res.stderr))
    # hipcc's host pass warns about the device-only target feature; keep the rest
    warn = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l
                     and "target-feature" not in l).strip()
    return obj, warn


def build(verbose=True, extra=(), lib=LIB, build_dir=BUILD, driver=True):
    """Compiles the library (and the CLI driver unless driver=False); `extra`
    adds compiler flags, `lib`/`build_dir` redirect the outputs (variants)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    jobs = min(len(LIB_SOURCES), max(1, (os.cpu_count() or 4)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, extra, build_dir), LIB_SOURCES))
    objs = []
    for obj, warn in results:
        objs.append(obj)
        if warn and verbose:
            print(warn, file=sys.stderr)
    tmp = lib + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-ldl"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    if not driver:
        return lib
    # CLI driver: plain C++ host code over the C ABI
    cmd = [HIPCC, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), DRIVER_SOURCE,
           "-o", BIN + ".tmp", "-L" + PKG, "-lsedi_amd", "-Wl,-rpath,$ORIGIN/.."]
    subprocess.run(cmd, check=True)
    os.replace(BIN + ".tmp", BIN)
    if verbose:
        print("built", LIB, "and", BIN)
    return LIB


VARIANTS = {
    # the cross-workgroup reduction published with a release fence and read
    # with an acquire fence (the C++ memory model's own guarantee) instead of
    # agent-scope atomic stores and a wait on their completion
    "fence": ["-DSEDI_GRID_SUM_FENCE"],
    # the mixed-precision pipelined CG and BiCGStab computing their stencil
    # and vector updates in FP32 registers (the default computes them in FP64
    # on the FP32-stored vectors); bench.py --mode arith measures both
    "f32compute": ["-DSEDI_MIXED_COMPUTE_F64=0"],
}


def build_variant(name, verbose=False):
    """Builds sedi_amd/variants/libsedi_amd_<name>.so with the flags of
    VARIANTS[name] (no CLI driver); tests load it through SEDI_AMD_LIB."""
    out_dir = os.path.join(PKG, "variants")
    os.makedirs(out_dir, exist_ok=True)
    return build(verbose=verbose, extra=VARIANTS[name], lib=os.path.join(out_dir, "libsedi_amd_%s.so" % name),
                 build_dir=os.path.join(BUILD, "variant_" + name), driver=False)


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build()
