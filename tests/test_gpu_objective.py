"""Objective and gradient on the GPU (nomad_obj_*), through the C ABI and the
torch.autograd binding: against the oracle (FP64), against the reference
DSGD program's printed objective (tests/golden), and by finite differences
(torch.autograd.gradcheck on a small problem; central differences on 1,000
entries at the 100,000 x 20,000, 10 M-rating, k = 64 shape)."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from nomad_amd import lib
from oracle import oracle as O

torch = pytest.importorskip("torch")
from nomad_amd.torch_ops import factor_objective  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden"
pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64, ldf=None):
    a = np.asarray(a)
    if ldf is not None and a.shape[1] < ldf:
        a = np.concatenate([a, np.zeros((a.shape[0], ldf - a.shape[1]))], axis=1)
    return torch.tensor(a, dtype=dtype, device="cuda")


def test_objective_matches_oracle_fp64(small):
    tr, _ = small
    k = 15  # odd: one zero padding column
    rng = np.random.default_rng(3)
    W = rng.random((tr.m, k)) / 4
    H = rng.random((tr.n, k)) / 4
    print(f"[obj] {tr.m}x{tr.n} nnz={tr.nnz} k={k}", flush=True)
    obj = lib.Objective(tr, k, "fp64")
    Wt = dev(W, ldf=obj.ldf).requires_grad_()
    Ht = dev(H, ldf=obj.ldf).requires_grad_()
    L = factor_objective(Wt, Ht, obj, 0.05)
    L.backward()
    Lo, gWo, gHo = O.loss_grad(tr, W, H, 0.05)
    assert float(L) == pytest.approx(Lo, rel=1e-12)
    np.testing.assert_allclose(Wt.grad.cpu().numpy()[:, :k], gWo, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(Ht.grad.cpu().numpy()[:, :k], gHo, rtol=1e-10, atol=1e-10)
    assert float(Wt.grad[:, k:].abs().max()) == 0.0


def test_objective_fp32_close_to_fp64(small):
    tr, _ = small
    k = 16
    rng = np.random.default_rng(4)
    W = rng.random((tr.m, k)) / 4
    H = rng.random((tr.n, k)) / 4
    obj = lib.Objective(tr, k, "fp32")
    Wt = dev(W, torch.float32).requires_grad_()
    Ht = dev(H, torch.float32).requires_grad_()
    L = factor_objective(Wt, Ht, obj, 0.05)
    L.backward()
    Lo, gWo, gHo = O.loss_grad(tr, W, H, 0.05)
    assert float(L) == pytest.approx(Lo, rel=1e-5)
    np.testing.assert_allclose(Wt.grad.cpu().numpy(), gWo, rtol=1e-3, atol=1e-3)


def test_objective_matches_reference_dsgd_print():
    z = np.load(GOLDEN / "ref_problem.npz")
    g = json.loads((GOLDEN / "ref_traces.json").read_text())
    tr = lib.Csc(int(z["m"]), int(z["n"]), z["tr_colptr"], z["tr_rowidx"], z["tr_val"])
    k = g["k"]
    W, H = O.init_factors(tr.m, tr.n, k, 1, 12345)  # the program's initial factors
    obj = lib.Objective(tr, k, "fp64")
    L = factor_objective(dev(W), dev(H), obj, g["dsgd"]["reg"])
    assert float(L) == pytest.approx(g["dsgd"]["obj0"], rel=5e-6)


def test_gradcheck_small():
    tr, _ = lib.synth(lib.preset(1, 0.05))
    k = 4
    rng = np.random.default_rng(5)
    obj = lib.Objective(tr, k, "fp64")
    W = dev(rng.random((tr.m, k)) / 2).requires_grad_()
    H = dev(rng.random((tr.n, k)) / 2).requires_grad_()
    assert torch.autograd.gradcheck(lambda a, b: factor_objective(a, b, obj, 0.1), (W, H), eps=1e-6, atol=1e-5,
                                    rtol=1e-4)


def test_central_differences_config4_shape():
    spec = lib.preset(4, 1.0)
    tr, _ = lib.synth(spec)
    k = 64
    print(f"[obj] config-4 shape {tr.m}x{tr.n} nnz={tr.nnz} k={k}", flush=True)
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    obj = lib.Objective(tr, k, "fp64")
    W = dev(W0).requires_grad_()
    H = dev(H0).requires_grad_()
    L = factor_objective(W, H, obj, 0.05)
    L.backward()
    gW, gH = W.grad.detach(), H.grad.detach()
    rng = np.random.default_rng(6)
    eps = 1e-6
    Wd, Hd = W.detach().clone(), H.detach().clone()
    worst = 0.0
    for q in range(1000):
        on_w = q % 2 == 0
        T, G = (Wd, gW) if on_w else (Hd, gH)
        i = int(rng.integers(T.shape[0]))
        d = int(rng.integers(k))
        x = float(T[i, d])
        T[i, d] = x + eps
        lp = factor_objective(Wd, Hd, obj, 0.05)
        T[i, d] = x - eps
        lm = factor_objective(Wd, Hd, obj, 0.05)
        T[i, d] = x
        fd = (float(lp) - float(lm)) / (2 * eps)
        g = float(G[i, d])
        err = abs(fd - g) / (1e-3 + 1e-4 * abs(g))
        worst = max(worst, err)
    assert worst <= 1.0, worst
