"""Build the gfx950 hydro library in-tree: ramses2_amd/lib/libramses2_hip.so.

hipcc cross-compiles for gfx950 without a GPU.  Packed-FP32 instructions are
disabled (-packed-fp32-ops) so that float code is emitted as scalar VALU ops.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIBNAME = "libramses2_hip.so"
SOURCES = ["rh_abi.hip", "rh_godunov.hip", "rh_ext.hip", "rh_serve.hip", "rh_amr.hip", "rh_flag.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RH_OFFLOAD_ARCH", "gfx950")

FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function"]


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FSRC = os.path.join(HERE, "fortran")


def lib_path():
    return os.path.join(LIBDIR, LIBNAME)


def fortran_driver_path():
    return os.path.join(LIBDIR, "sod_driver")


def build_fortran():
    """Compile the Fortran binding and the Sod driver against the library
    (flang from ROCm); skipped when flang is not installed."""
    if not os.path.exists(FLANG):
        return None
    mod = os.path.join(LIBDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    out = fortran_driver_path()
    cmd = [FLANG, "-O2", "-J", mod, "-o", out, os.path.join(FSRC, "hip_interface.f90"),
           os.path.join(FSRC, "sod_driver.f90"), "-L" + LIBDIR, "-lramses2_hip", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("flang failed building the Fortran driver")
    return out


def build(verbose=False, out=None, defines=()):
    """Compile every HIP source into the shared library (always recompiles)."""
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    out = out or lib_path()
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-o", tmp] + srcs
    res = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % LIBNAME)
    if verbose:
        sys.stderr.write(res.stderr)
    os.replace(tmp, out)
    if out == lib_path():
        build_fortran()
    return out


if __name__ == "__main__":
    print(build(verbose=True))
