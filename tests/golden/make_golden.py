"""Writes the golden fixtures of tests/golden from the reference's own programs.

Run in a container that has /root/reference and the competitor binaries built
by `python oracle/build.py --ref` (oracle/_ref/).  Nothing at test time runs
the reference: the tests read the files this script writes.

    python tests/golden/make_golden.py

Fixtures:
  libmf_smalldata.txt  copy of the reference's data fixture
                       Code/libmf-1.0/smalldata ("user item rating" lines)
  ref_problem.npz      a 300 x 150 synthetic problem (train/test CSC), the
                       input of the two traces below
  ref_traces.json      printed by the reference binaries on that problem:
    dsgd_obj0    Code/competitor/sgd-mpi-omp (rank 8, lambda 0.05, one
                 iteration with learning rate 0 and the bold-driver strategy,
                 one thread): its printed objective of the initial factors
                 W ~ mt19937_64(12345 + 1), H ~ mt19937_64(12345 + 139),
                 U(0, 1/sqrt(k)) (sgd-mpi-omp.cpp:122-164, 663-705)
    ccd          Code/competitor/ccdr1-mpi-omp (-k 8 -n 1 -l 0.1 -t 3 -T 3
                 -e 1e-3 -p 1 -q 1): per (outer iteration, rank) the printed
                 obj (%.6g) and test rmse (%.10g); W ~ mt19937_64(12345 + 1)
                 (one thread), H = 0 (ccdr1-mpi-omp.cpp:207-239, 414-420)
"""
from __future__ import annotations

import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

REF = Path("/root/reference")
MPIEXEC = "/opt/conda/bin/mpiexec"


def run_ref(args, cwd):
    env = dict(os.environ, LD_LIBRARY_PATH="/usr/lib/x86_64-linux-gnu:/opt/conda/lib", OMP_NUM_THREADS="1")
    p = subprocess.run([MPIEXEC, "-n", "1", *map(str, args)], cwd=cwd, env=env, capture_output=True, text=True,
                       timeout=600)
    if p.returncode != 0:
        raise RuntimeError(f"{args[0]} failed ({p.returncode}):\n{p.stdout}\n{p.stderr}")
    return p.stdout


def main() -> int:
    from nomad_amd import lib

    shutil.copyfile(REF / "Code" / "libmf-1.0" / "smalldata", HERE / "libmf_smalldata.txt")
    os.chmod(HERE / "libmf_smalldata.txt", 0o644)

    tr, te = lib.synth(lib.preset(1, 0.15))
    np.savez(HERE / "ref_problem.npz", m=tr.m, n=tr.n, tr_colptr=tr.colptr, tr_rowidx=tr.rowidx, tr_val=tr.val,
             te_colptr=te.colptr, te_rowidx=te.rowidx, te_val=te.val)
    refbin = ROOT / "oracle" / "_ref"
    with tempfile.TemporaryDirectory() as d:
        lib.write_dat(Path(d) / "train.dat", tr)
        lib.write_dat(Path(d) / "test.dat", te)
        out = run_ref([refbin / "mpi-sgd-omp", 8, 0.05, 1, 0.0, 0.5, 1, 0, 1, d], d)
        m = re.search(r"iter 1 .* obj (\S+)", out)
        dsgd_obj0 = float(m.group(1))
        out = run_ref([refbin / "mpi-ccd-r1-omp", "-k", 8, "-n", 1, "-l", 0.1, "-t", 3, "-T", 3, "-e", 1e-3, "-p", 1,
                       "-q", 1, d], d)
        ccd = []
        for line in out.splitlines():
            mm = re.match(r"iter (\d+) rank (\d+) .* obj (\S+) loss (\S+) reg (\S+) .* rmse (\S+)", line)
            if mm:
                ccd.append(dict(iter=int(mm.group(1)), rank=int(mm.group(2)), obj=float(mm.group(3)),
                                loss=float(mm.group(4)), reg=float(mm.group(5)), rmse=float(mm.group(6))))
    (HERE / "ref_traces.json").write_text(json.dumps(dict(
        problem="ref_problem.npz", k=8,
        dsgd=dict(args="rank 8 lambda 0.05 maxiter 1 lrate 0 decay 0.5 strategy 1 stratum 0 threads 1",
                  W_seed=12345 + 1, H_seed=12345 + 139, obj0=dsgd_obj0, reg=0.05),
        ccd=dict(args="-k 8 -n 1 -l 0.1 -t 3 -T 3 -e 1e-3 -p 1 -q 1", W_seed=12345 + 1, lam=0.1, iters=3,
                 inner=3, eps=1e-3, trace=ccd)), indent=1))
    print(f"dsgd obj0 {dsgd_obj0}; ccd {len(ccd)} trace points", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
