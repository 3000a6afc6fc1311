"""Objective and gradient on the GPU (nomad_obj_*), through the C ABI and the
torch.autograd binding: against the oracle (FP64), against the reference
DSGD program's printed objective (tests/golden), and by finite differences
(torch.autograd.gradcheck on a small problem; central differences on 1,000
entries at the 100,000 x 20,000, 10 M-rating, k = 64 shape).  Tensors live in
host memory: torch's wheel brings its own HIP runtime, so the harness hands
them to the library's kernels through host buffers."""
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


def dev(a, dtype=torch.float64):
    # host tensors: the harness hands them to the library through host memory
    return torch.tensor(np.asarray(a), dtype=dtype)


@pytest.mark.parametrize("k", [15, 201])
def test_objective_matches_oracle_fp64(small, k):
    # 15: odd, padded to 16 on the device; 201: two element pairs per lane
    tr, _ = small
    rng = np.random.default_rng(3)
    W = rng.random((tr.m, k)) / 4
    H = rng.random((tr.n, k)) / 4
    print(f"[obj] {tr.m}x{tr.n} nnz={tr.nnz} k={k}", flush=True)
    obj = lib.Objective(tr, k, "fp64")
    Wt = dev(W).requires_grad_()
    Ht = dev(H).requires_grad_()
    L = factor_objective(Wt, Ht, obj, 0.05)
    L.backward()
    Lo, gWo, gHo = O.loss_grad(tr, W, H, 0.05)
    assert float(L) == pytest.approx(Lo, rel=1e-12)
    np.testing.assert_allclose(Wt.grad.numpy(), gWo, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(Ht.grad.numpy(), gHo, rtol=1e-10, atol=1e-10)


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
    np.testing.assert_allclose(Wt.grad.numpy(), gWo, rtol=1e-3, atol=1e-3)


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
    """GPU gradient vs FP64 central differences (eps 1e-6) of the objective on
    1,000 random entries at the 100,000 x 20,000, 10 M-rating, k = 64 shape.
    Perturbing w_i (h_j) changes only the terms of row i (column j); those
    are evaluated in FP64 numpy, independently of the GPU code."""
    spec = lib.preset(4, 1.0)
    tr, _ = lib.synth(spec)
    k, reg = 64, 0.05
    print(f"[obj] config-4 shape {tr.m}x{tr.n} nnz={tr.nnz} k={k}", flush=True)
    W, H = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    obj = lib.Objective(tr, k, "fp64")
    _, gW, gH = obj.eval_host(W, H, reg)
    cols = np.repeat(np.arange(tr.n), np.diff(tr.colptr))
    order = np.argsort(tr.rowidx, kind="stable")
    rowptr = np.concatenate([[0], np.cumsum(np.bincount(tr.rowidx, minlength=tr.m))])
    csr_cols, csr_vals = cols[order], tr.val[order]

    def part_row(i, w):
        js = csr_cols[rowptr[i]:rowptr[i + 1]]
        e = H[js] @ w - csr_vals[rowptr[i]:rowptr[i + 1]]
        return e @ e + reg * len(js) * (w @ w)

    def part_col(j, h):
        a, b = tr.colptr[j], tr.colptr[j + 1]
        e = W[tr.rowidx[a:b]] @ h - tr.val[a:b]
        return e @ e + reg * (b - a) * (h @ h)

    rng = np.random.default_rng(6)
    eps = 1e-6
    worst = 0.0
    for q in range(1000):
        on_w = q % 2 == 0
        T, G, part = (W, gW, part_row) if on_w else (H, gH, part_col)
        i = int(rng.integers(T.shape[0]))
        d = int(rng.integers(k))
        x = T[i].copy()
        xp, xm = x.copy(), x.copy()
        xp[d] += eps
        xm[d] -= eps
        fd = (part(i, xp) - part(i, xm)) / (2 * eps)
        worst = max(worst, abs(fd - G[i, d]) / (1e-6 + 1e-6 * abs(G[i, d])))
    assert worst <= 1.0, worst


def test_objective_fp32_config4_shape():
    """The FP32 leg at the full 100,000 x 20,000, 10 M-rating, k = 64 shape:
    FP32 arithmetic with FP64 sums stays within 1e-5 (objective) and 1e-4
    (gradient, relative Frobenius norm) of the FP64 evaluation of the same
    factors, and the FP64 value equals the oracle's."""
    tr, _ = lib.synth(lib.preset(4, 1.0))
    k, reg = 64, 0.05
    print(f"[obj fp32] config-4 shape {tr.m}x{tr.n} nnz={tr.nnz} k={k}", flush=True)
    W, H = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    W = W.astype(np.float32).astype(np.float64)  # values exact in both precisions
    H = H.astype(np.float32).astype(np.float64)
    L64, gW64, gH64 = lib.Objective(tr, k, "fp64").eval_host(W, H, reg)
    L32, gW32, gH32 = lib.Objective(tr, k, "fp32").eval_host(W, H, reg)
    Lo, _, _ = O.loss_grad(tr, W, H, reg, want_grad=False)
    assert L64 == pytest.approx(Lo, rel=1e-12)
    assert L32 == pytest.approx(L64, rel=1e-5)
    num = np.sqrt(np.sum((gW32 - gW64) ** 2) + np.sum((gH32 - gH64) ** 2))
    den = np.sqrt(np.sum(gW64 ** 2) + np.sum(gH64 ** 2))
    print(f"[obj fp32] L rel {abs(L32 - L64) / L64:.3e}, gradient rel {num / den:.3e}", flush=True)
    assert num <= 1e-4 * den
