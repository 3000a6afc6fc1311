"""Build recipe of the CPU oracle (test infrastructure).

    python oracle/build.py          -> oracle/liborc.so
    python oracle/build.py --ref    -> also oracle/_ref/ (reference competitor
                                       binaries compiled from /root/reference
                                       sources, when that tree is present)

The oracle is compiled with -ffp-contract=off and no -march, like the
reference's own build (Code/nomad/CMakeLists.txt:9,40), so its float and
double arithmetic is exactly the operation sequence in the source.
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path("/root/reference")
MPI_INC = Path("/opt/conda/include")
MPI_LIB = Path("/opt/conda/lib/libmpi.so.12")


def run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"failed: {' '.join(map(str, cmd))}\n{p.stdout}")


def build_oracle() -> Path:
    out = HERE / "liborc.so"
    run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-pthread", "-Wall",
         str(HERE / "nomad_oracle.cpp"), "-o", str(out)])
    return out


def build_ref() -> list[Path]:
    """Compile the reference's DSGD and CCD++ competitors from their own
    sources (Code/competitor), with MPI from the image, into oracle/_ref.
    The NOMAD binary itself needs TBB, Boost and glog headers that the image
    lacks, so it is not built."""
    src = REF / "Code" / "competitor"
    if not src.exists() or not MPI_LIB.exists() or not (MPI_INC / "mpi.h").exists():
        return []
    dst = HERE / "_ref"
    dst.mkdir(exist_ok=True)
    outs = []
    recipes = {
        "mpi-sgd-omp": ["sgd-mpi-omp.cpp", "util.cpp"],
        "mpi-ccd-r1-omp": ["ccdr1-mpi-omp.cpp", "util.cpp", "smat.cpp"],
    }
    for name, files in recipes.items():
        out = dst / name
        run(["g++", "-std=c++11", "-O2", "-fopenmp", "-w", f"-I{MPI_INC}", *[str(src / f) for f in files],
             str(MPI_LIB), "-o", str(out)])
        outs.append(out)
    return outs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", action="store_true")
    a = ap.parse_args()
    build_oracle()
    if a.ref:
        outs = build_ref()
        print("\n".join(map(str, outs)) or "reference sources absent; _ref not built", file=sys.stderr)


if __name__ == "__main__":
    main()
