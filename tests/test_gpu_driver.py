"""The drop-in driver (nomad_float / nomad_double) end to end on the GPU:
reference options, train.dat/test.dat in, the reference's Statistics block
and --output dump out (Code/nomad/nb_main.cpp:264-300, 350-366, 683-705)."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nomad_amd import lib
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "nomad_amd" / "bin"

pytestmark = pytest.mark.gpu


def parse_dump(path, m, n, k):
    W = np.zeros((m, k))
    H = np.zeros((n, k))
    seen_r = seen_c = 0
    for line in Path(path).read_text().splitlines():
        f = line.split(",")
        vals = np.array([float(x) for x in f[2:]])
        assert len(vals) == k
        if f[0] == "row":
            W[int(f[1])] = vals
            seen_r += 1
        else:
            assert f[0] == "column"
            H[int(f[1])] = vals
            seen_c += 1
    assert (seen_r, seen_c) == (m, n)
    return W, H


@pytest.mark.parametrize("exe,storage", [("nomad_float", O.FP32), ("nomad_double", O.FP64)])
def test_driver_end_to_end(tmp_path, small, exe, storage):
    tr, te = small
    lib.write_dat(tmp_path / "train.dat", tr)
    lib.write_dat(tmp_path / "test.dat", te)
    out = tmp_path / "factors"
    print(f"[driver] {exe} m={tr.m} n={tr.n}", flush=True)
    r = subprocess.run([str(BIN / exe), "-i", str(tmp_path), "--timeout", "0.2", "0.4", "--pause", "0",
                        "--dim", "16", "--lrate", "0.001", "--drate", "0.1", "--reg", "0.05", "--seed", "12345",
                        "-o", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    blocks = r.stdout.count("Statistics:")
    assert blocks == 2
    test_rmse = [float(x) for x in re.findall(r"current test RMSE: (\d+\.\d{10})\n", r.stdout)]
    assert len(test_rmse) == 2
    det = re.findall(r"detail: train: s=([\d.]+), c=(\d+); test: s=([\d.]+), c=(\d+); u=(\d+), f=(\d+), s=0",
                     r.stdout)
    assert len(det) == 2 and int(det[-1][1]) == tr.nnz and int(det[-1][3]) == te.nnz
    assert int(det[-1][4]) > int(det[0][4]) > 0
    # the dump reproduces the printed RMSE (values are printed with 6 digits)
    W, H = parse_dump(str(out) + "0", tr.m, tr.n, 16)
    rmse = np.sqrt(O.sse(te, W, H) / te.nnz)
    assert abs(rmse - test_rmse[-1]) < 1e-4 * test_rmse[-1]
