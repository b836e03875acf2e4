"""Writes tests/golden/l2gmap/*.json: the L2GMap plan and one update /
reverse_update per rank, from real MPI runs (mpirun -n 1..4) of
oracle/l2gmap_mpi.c, a restatement of spmv/L2GMap.cpp:15-170 on MPICH.

Run in a container with MPICH (default /opt/conda):
    python tests/golden/make_l2gmap_golden.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
MPI_HOME = os.environ.get("MPI_HOME", "/opt/conda")
CASES = ["demo1d", "grid2d", "slab27", "scatter"]


def main():
    out = os.path.join(HERE, "l2gmap")
    os.makedirs(out, exist_ok=True)
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "l2gmap_mpi",
                           f"MPI_HOME={MPI_HOME}"])
    exe = os.path.join(ROOT, "oracle", "l2gmap_mpi")
    for case in CASES:
        for n in (1, 2, 3, 4):
            subprocess.check_call([os.path.join(MPI_HOME, "bin", "mpirun"), "-n", str(n),
                                   exe, case, out])
    print("wrote", len(os.listdir(out)), "files to", out, file=sys.stderr)


if __name__ == "__main__":
    main()
