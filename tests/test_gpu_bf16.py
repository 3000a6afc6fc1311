"""BF16 storage on the GPU: stochastic rounding keeps sub-ulp updates
(unbiased), round-to-nearest drops them, and the BF16 model's test RMSE
against an FP64 oracle run of the same schedule (the stated gap)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rmse(ctx):
    ev = ctx.eval()
    return np.sqrt(ev.train_sse / ev.train_count), np.sqrt(ev.test_sse / ev.test_count)


def test_sr_tracks_fp32_where_nearest_stalls(gpu, small):
    """Steps mostly below half a bf16 ulp: round-to-nearest drops them and
    returns the factor to its old value, stochastic rounding moves it by the
    step on average, so its training progress matches FP32's.  (The oracle in
    the GPU's arithmetic gives the same gains: fp32 0.394, sr 0.394, rne
    0.226 at these settings.)"""
    tr, te = small
    k = 16
    hyp = dict(lrate=2e-4, drate=0.0, reg=0.05)
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 12345)
    gain = {}
    for name, storage, rounding in (("fp32", "fp32", "nearest"), ("sr", "bf16", "stochastic"),
                                    ("rne", "bf16", "nearest")):
        ctx = gpu.Context(tr.m, tr.n, k, storage=storage, rounding=rounding, mode="strat", nworkers=8, seed=7, **hyp)
        ctx.upload(tr, te)
        ctx.set_factors(W0, H0)
        r0, _ = _rmse(ctx)
        ctx.train(epochs=10)
        r1, _ = _rmse(ctx)
        ctx.close()
        gain[name] = r0 - r1
    print(f"[bf16] train RMSE gain over 10 epochs: {gain}", flush=True)
    assert gain["fp32"] > 0.01
    assert abs(gain["sr"] - gain["fp32"]) <= 0.05 * gain["fp32"]
    assert gain["rne"] <= 0.75 * gain["fp32"]


def test_bf16_gap_vs_fp64_oracle(gpu, small):
    """Same deterministic schedule (strat, B = 8, 10 epochs): GPU BF16 with
    stochastic rounding vs the FP64 oracle in the reference's arithmetic."""
    tr, te = small
    k = 16
    hyp = dict(lrate=0.005, drate=0.1, reg=0.05)
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 12345)
    ctx = gpu.Context(tr.m, tr.n, k, storage="bf16", rounding="stochastic", mode="strat", nworkers=8, seed=7, **hyp)
    ctx.upload(tr, te)
    ctx.set_factors(W0, H0)
    ctx.train(epochs=10)
    _, rb = _rmse(ctx)
    ctx.close()
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=8, epochs=10, **hyp)
    ro = np.sqrt(O.sse(te, Wo, Ho) / te.nnz)
    gap = (rb - ro) / ro
    print(f"[bf16] test RMSE bf16-sr {rb:.6f} fp64 oracle {ro:.6f} gap {gap * 100:+.3f} %", flush=True)
    assert abs(gap) < 0.01
