"""Checkpoint and resume of device state (the reference's checkpoint files,
Code/nomad/nb_checkpoint.cpp, plus the step-size state it does not save):
an interrupted and resumed stratified run equals the uninterrupted run bit
for bit; an async run resumes with its visit counters; the drop-in driver
writes and recovers checkpoints."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "nomad_amd" / "bin"

pytestmark = pytest.mark.gpu

HYP = dict(lrate=0.002, drate=0.05, reg=0.05)


def _ctx(lib, tr, te, k, **kw):
    c = lib.Context(tr.m, tr.n, k, seed=12345, **{**HYP, **kw})
    c.upload(tr, te)
    return c


@pytest.mark.parametrize("storage,rounding", [("fp32", "nearest"), ("bf16", "stochastic"), ("fp64", "nearest")])
def test_strat_interrupted_and_resumed_equals_uninterrupted(gpu, small, tmp_path, storage, rounding):
    tr, te = small
    k = 16
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 12345)
    kw = dict(storage=storage, rounding=rounding, mode="strat", nworkers=8)
    with _ctx(gpu, tr, te, k, **kw) as a:
        a.set_factors(W0, H0)
        a.train(epochs=4)
        Wu, Hu = a.get_factors()
    with _ctx(gpu, tr, te, k, **kw) as b:
        b.set_factors(W0, H0)
        b.train(epochs=2)
        W, H = b.get_factors()
        gpu.write_checkpoint(tmp_path, 1, 0, 4, W, H)
        progress = b.progress
    assert progress == 2
    with _ctx(gpu, tr, te, k, **kw) as c:
        W, H = gpu.read_checkpoint(tmp_path, 1, 0, 4, tr.m, tr.n, k)
        c.set_factors(W, H)
        c.progress = progress
        c.train(epochs=2)
        Wr, Hr = c.get_factors()
    assert np.array_equal(Wr, Wu) and np.array_equal(Hr, Hu)


def test_async_resume_keeps_visit_counters(gpu, small, tmp_path):
    tr, te = small
    k = 16
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 12345)
    kw = dict(storage="fp32", mode="async", nworkers=512)
    with _ctx(gpu, tr, te, k, **kw) as a:
        a.set_factors(W0, H0)
        a.train(seconds=0.002)
        a.train(seconds=0.002)
        W, H = a.get_factors()
        rs = a.route_state()
        prog = a.progress
        ev_a = a.eval()
    gpu.write_checkpoint(tmp_path, 3, 1, 4, W, H)
    with _ctx(gpu, tr, te, k, **kw) as b:
        W2, H2 = gpu.read_checkpoint(tmp_path, 3, 1, 4, tr.m, tr.n, k)
        b.set_factors(W2, H2)
        b.progress = prog
        b.set_visit_counts(rs["count"])
        rb = b.route_state()
        assert np.array_equal(rb["count"], rs["count"]) and (rb["tok_cycle"] == prog).all()
        ev_b = b.eval()
        assert ev_b.test_sse == ev_a.test_sse  # the same factors
        st = b.train(seconds=0.002)
        assert st.updates > 0
        rc = b.route_state()
        assert (rc["count"] >= rs["count"]).all() and rc["count"].sum() > rs["count"].sum()
        with pytest.raises(gpu.NomadError):
            b.set_visit_counts(rs["count"][:-1])  # wrong length


def test_driver_writes_and_recovers_checkpoints(gpu, tiny, tmp_path):
    tr, te = tiny
    gpu.write_dat(tmp_path / "train.dat", tr)
    gpu.write_dat(tmp_path / "test.dat", te)
    cp = tmp_path / "cp"
    cp.mkdir()
    base = ["--nthreads", "3", "--dim", "8", "--lrate", "0.002", "--pause", "0", "-i", str(tmp_path),
            "--cptype", "sync", "--cppath", str(cp), "--jobid", "7"]
    r = subprocess.run([str(BIN / "nomad_float"), *base, "--timeout", "0.2", "0.4"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    for e in (0, 1):
        for p in range(3):
            assert (cp / f"cp-t7-e{e}-p{p}.state").exists() and (cp / f"cp-t7-e{e}-p{p}.msg").exists()
        assert (cp / f"cp-t7-e{e}.nomad").exists()
    W, H = gpu.read_checkpoint(cp, 7, 1, 3, tr.m, tr.n, 8)
    assert np.isfinite(W).all() and np.abs(H).max() > 0
    r2 = subprocess.run([str(BIN / "nomad_float"), *base, "--timeout", "0.2", "--recover", "1", "--cptype", "none"],
                        capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    assert "recovered checkpoint 1" in r2.stdout
    first = [float(x) for x in re.findall(r"current test RMSE: ([0-9.]+)", r.stdout)]
    resumed = [float(x) for x in re.findall(r"current test RMSE: ([0-9.]+)", r2.stdout)]
    assert len(first) == 2 and len(resumed) == 1
    assert resumed[0] <= first[0] * 1.01  # training continues from the later factors
