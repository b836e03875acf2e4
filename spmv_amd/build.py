"""Build recipe of libspmv_amd.so (hipcc, gfx950) -- invoked by
__graft_entry__.build() and usable standalone:  python -m spmv_amd.build
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libspmv_amd.so")
CXXLIB = os.path.join(HERE, "libspmv.so")
BIN = os.path.join(HERE, "bin")
MPI_HOME = os.environ.get("MPI_HOME", "/opt/conda")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPMV_AMD_ARCH", "gfx950")

SOURCES = ["capi.cpp", "petsc.cpp", "matrix.hip", "transpose.hip", "l2gmap.hip", "cg.hip", "batch.hip"]

# -ffp-contract=off: products and sums are rounded separately (the FP64
# kernels are compared bit for bit with the reference's Eigen arithmetic).
# -packed-fp32-ops disabled: packed FP32 instructions are not used.
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *FLAGS, "-x", "c++", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build_cpp(verbose=False):
    """The C++ API (libspmv.so: spmv::L2GMap, spmv::Matrix, spmv::cg over the
    C ABI) and the demo drivers (bin/matvecmain, bin/cgmain, bin/restrictmain, and
    bin/createmain, a driver of Matrix::create_matrix for the tests).  Needs an MPI
    installation (the API keeps the reference's MPI_Comm arguments)."""
    mpi_inc = os.path.join(MPI_HOME, "include")
    mpi_lib = os.path.join(MPI_HOME, "lib")
    if not os.path.exists(os.path.join(mpi_inc, "mpi.h")):
        print(f"no MPI under {MPI_HOME}: C++ API not built", file=sys.stderr)
        return None
    inc = [f"-I{os.path.join(ROOT, 'include')}", f"-I{mpi_inc}"]
    # a directory holding only the MPI library and the Fortran runtime it
    # needs: an rpath to the MPI prefix's lib directory would also pick up the
    # older C++ runtime it ships
    mpi_dir = os.path.join(HERE, "_mpi")
    os.makedirs(mpi_dir, exist_ok=True)
    for name in os.listdir(mpi_lib):
        if name.startswith(("libmpi.so", "libgfortran.so", "libquadmath.so")):
            link = os.path.join(mpi_dir, name)
            if os.path.lexists(link):
                os.remove(link)
            os.symlink(os.path.join(mpi_lib, name), link)
    tmp = CXXLIB + ".tmp"
    subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-x", "hip",
                           f"--offload-arch={ARCH}", *inc,
                           os.path.join(HERE, "cpp", "spmv_api.cpp"), "-o", tmp,
                           f"-L{HERE}", "-lspmv_amd", f"-L{mpi_dir}", "-lmpi",
                           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/_mpi"])
    os.replace(tmp, CXXLIB)
    os.makedirs(BIN, exist_ok=True)
    for demo in ("matvecmain", "cgmain", "restrictmain", "createmain"):
        out = os.path.join(BIN, demo)
        subprocess.check_call([HIPCC, "-O2", "-std=c++17", "-x", "c++", *inc,
                               os.path.join(HERE, "demos", demo + ".cpp"), "-o", out + ".tmp",
                               f"-L{HERE}", "-lspmv", "-lspmv_amd", f"-L{mpi_dir}", "-lmpi",
                               "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,$ORIGIN/../_mpi"])
        os.replace(out + ".tmp", out)
    if verbose:
        print("built", CXXLIB, "and demos in", BIN, file=sys.stderr)
    return CXXLIB


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB, file=sys.stderr)
    build_cpp(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
