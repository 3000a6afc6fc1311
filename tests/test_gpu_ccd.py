"""CCD++ on the GPU (nomad_ccd) against the reference CCD++ program's printed
trace (tests/golden/ref_traces.json, written by tests/golden/make_golden.py
from Code/competitor/ccdr1-mpi-omp) and against the oracle's restatement."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from nomad_amd import lib
from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"

pytestmark = pytest.mark.gpu


def ref_problem():
    z = np.load(GOLDEN / "ref_problem.npz")
    m, n = int(z["m"]), int(z["n"])
    tr = lib.Csc(m, n, z["tr_colptr"], z["tr_rowidx"], z["tr_val"])
    te = lib.Csc(m, n, z["te_colptr"], z["te_rowidx"], z["te_val"])
    return tr, te, json.loads((GOLDEN / "ref_traces.json").read_text())


def test_ccd_matches_reference_trace():
    tr, te, g = ref_problem()
    k, c = g["k"], g["ccd"]
    W0, _ = O.init_factors(tr.m, tr.n, k, 1, 12345)  # mt19937_64(12345 + 1): one thread
    H0 = np.zeros((tr.n, k))
    print(f"[ccd] reference problem {tr.m}x{tr.n} k={k}", flush=True)
    W, H, obj, rmse, st = lib.ccd(tr, W0, H0, c["lam"], c["iters"], c["inner"], c["eps"], test=te)
    printed = set()
    for p in c["trace"]:
        q = (p["iter"] - 1) * k + p["rank"] - 1
        printed.add(q)
        assert obj[q] == pytest.approx(p["obj"], rel=6e-6), p  # printed with %.6g
        assert rmse[q] == pytest.approx(p["rmse"], rel=1e-9), p  # printed with %.10g
    assert all(np.isnan(obj[q]) for q in range(len(obj)) if q not in printed)
    assert st.ranks_done == len(c["trace"])


@pytest.mark.parametrize("k,lam,eps,iters", [(16, 0.05, 1e-3, 2), (5, 0.5, 1e-3, 2), (16, 0.05, 0.3, 3),
                                            (1, 0.05, 1e-3, 3)])
def test_ccd_matches_oracle(small, k, lam, eps, iters):
    """Values and decisions against the oracle; with eps = 0.3 ranks stop at
    their first inner iteration, so outer iterations end early and the next
    one starts at rank 0 (the residual passes of consecutive ranks are fused,
    and the fused pass picks the next rank on the device); with k = 1 every
    fused pass hands over to rank 0 of the next outer iteration."""
    tr, te = small
    rng = np.random.default_rng(k)
    W0 = rng.random((tr.m, k)) / np.sqrt(k)
    H0 = np.zeros((tr.n, k))
    print(f"[ccd] {tr.m}x{tr.n} k={k} lambda={lam} eps={eps} iters={iters}", flush=True)
    Wg, Hg, og, rg, st = lib.ccd(tr, W0, H0, lam, iters, 3, eps, test=te)
    Wo, Ho, oo, ro = O.ccd(tr, W0, H0, lam, iters, 3, eps, test=te)
    if eps > 0.1:  # the early-stop path is exercised: an outer iteration ends early and the next one runs
        skipped = np.isnan(oo).reshape(iters, k)
        assert skipped[0].any() and not skipped[1:].all(axis=1).any(), skipped
    # same decisions (skipped ranks) and the same values up to FP64 summation order
    np.testing.assert_array_equal(np.isnan(og), np.isnan(oo))
    ok = ~np.isnan(oo)
    np.testing.assert_allclose(og[ok], oo[ok], rtol=1e-9)
    np.testing.assert_allclose(rg[ok], ro[ok], rtol=1e-9)
    np.testing.assert_allclose(Wg, Wo, rtol=1e-7, atol=1e-10)
    np.testing.assert_allclose(Hg, Ho, rtol=1e-7, atol=1e-10)
    assert st.alg_bytes > 0 and st.inner_done >= st.ranks_done


def test_ccd_bad_arguments():
    tr, _, _ = ref_problem()
    with pytest.raises(lib.NomadError):
        lib.ccd(tr, np.zeros((tr.m, 4)), np.zeros((tr.n, 4)), 0.1, 1, 0)
    with pytest.raises(lib.NomadError):
        lib.ccd(tr, np.zeros((tr.m, 4)), np.zeros((tr.n, 4)), -1.0, 1, 1)
