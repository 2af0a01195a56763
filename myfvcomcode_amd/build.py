"""In-tree build of the native libraries.

    libfvgrid.so   host C++ (g++): cases, grid metrics, cold start
    libfvgpu.so    HIP for gfx950 (hipcc): device state and the time step

Both land in ``myfvcomcode_amd/lib/``.  The FP64 device code is compiled
with ``-ffp-contract=off`` so that every ``a*b+c`` rounds twice, as in the
reference's scalar Fortran; kernels of the mixed-precision form opt back in
to contraction with a pragma.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_grid(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfvgrid.so")
    _run([CXX, "-O2", "-fPIC", "-shared", "-std=c++17", "-I" + INCLUDE,
          os.path.join(CSRC, "fvgrid.cpp"), "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_gpu(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfvgpu.so")
    srcs = [os.path.join(CSRC, f) for f in ("fvgpu.hip",)]
    cmd = [HIPCC, "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
           "-ffp-contract=off", "-fno-fast-math",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I" + INCLUDE] + srcs + ["-o", out + ".tmp"]
    _run(cmd, verbose)
    os.replace(out + ".tmp", out)
    return out


FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_fortran(verbose=False):
    """The Fortran drivers (fvcom_gpu: the seeded cases; fvcom_mesh: a
    host's own mesh through FVGPU_SETUP_MESH) over the ISO_C_BINDING
    module."""
    fdir = os.path.join(HERE, "fortran")
    mdir = os.path.join(LIB, "_fmod")
    os.makedirs(mdir, exist_ok=True)
    out = os.path.join(LIB, "fvcom_gpu")
    _run([FC, "-O2", "-J", mdir, "-c", os.path.join(fdir, "fvgpu_mod.f90"),
          "-o", os.path.join(mdir, "fvgpu_mod.o")], verbose)
    _run([FC, "-O2", "-J", mdir, os.path.join(fdir, "fvcom_driver.f90"), os.path.join(mdir, "fvgpu_mod.o"),
          "-L" + LIB, "-lfvgpu", "-lfvgrid", "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"], verbose)
    os.replace(out + ".tmp", out)
    out2 = os.path.join(LIB, "fvcom_mesh")
    _run([FC, "-O2", "-J", mdir, os.path.join(fdir, "fvcom_mesh_driver.f90"), os.path.join(mdir, "fvgpu_mod.o"),
          "-L" + LIB, "-lfvgpu", "-lfvgrid", "-Wl,-rpath,$ORIGIN", "-o", out2 + ".tmp"], verbose)
    os.replace(out2 + ".tmp", out2)
    return out


def build_all(verbose=False):
    build_grid(verbose)
    build_gpu(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
