"""Factor rows of more than 128 elements (--dim up to 256): every kernel then
holds element pairs l and l + 64 of a row in lane l.

Bars, as for the narrow rows (test_gpu_sgd.py): strat mode bit-identical to
the oracle's GPU-order replay for every storage, the batched launch
bit-identical to the single runs, exact update counts and test RMSE within
0.5 % of the CPU async reference in async mode, and the SSE kernel equal to
the oracle's GPU order to 1e-11 (and to the reference order to 1e-5).
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

HYP = dict(lrate=0.001, drate=0.1, reg=0.05)
SR_SEED = lambda seed: (seed ^ (seed >> 32)) & 0xFFFFFFFF ^ 0xA5A5A5A5  # noqa: E731  (capi.hip sr_seed)


def _ctx(lib, tr, te, k, **kw):
    ctx = lib.Context(tr.m, tr.n, k, **{**HYP, **kw})
    ctx.upload(tr, te)
    return ctx


@pytest.mark.parametrize("k", [200, 255, 256])
def test_wide_strat_fp32_bitexact(gpu, tiny, k):
    tr, te = tiny
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 12345)
    print(f"wide strat fp32 k={k}", flush=True)
    with _ctx(gpu, tr, te, k, storage="fp32", mode="strat", nworkers=4) as ctx:
        ctx.set_factors(W0, H0)
        Wo, Ho = W0, H0
        for ep in range(2):
            ctx.train(epochs=1)
            Wg, Hg = ctx.get_factors()
            Wo, Ho = O.sgd_strat(tr, Wo, Ho, storage=O.FP32, arith=O.ARITH_GPU, B=4, epoch0=ep, epochs=1, **HYP)
            assert np.array_equal(Wg, Wo), f"epoch {ep}: W differs, max {np.abs(Wg - Wo).max()}"
            assert np.array_equal(Hg, Ho), f"epoch {ep}: H differs, max {np.abs(Hg - Ho).max()}"


def test_wide_strat_fp64_bitexact(gpu, tiny):
    tr, te = tiny
    k = 180
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 2, 7)
    with _ctx(gpu, tr, te, k, storage="fp64", mode="strat", nworkers=5) as ctx:
        ctx.set_factors(W0, H0)
        ctx.train(epochs=2)
        Wg, Hg = ctx.get_factors()
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_GPU, B=5, epochs=2, **HYP)
    assert np.array_equal(Wg, Wo) and np.array_equal(Hg, Ho)


@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
def test_wide_strat_bf16_bitexact(gpu, tiny, rounding):
    tr, te = tiny
    k = 230
    seed = 12345
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, seed)
    with _ctx(gpu, tr, te, k, storage="bf16", mode="strat", nworkers=4, rounding=rounding, seed=seed) as ctx:
        ctx.set_factors(W0, H0)
        ctx.train(epochs=2)
        Wg, Hg = ctx.get_factors()
    r = O.ROUND_STOCHASTIC if rounding == "stochastic" else O.ROUND_NEAREST
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=O.BF16, arith=O.ARITH_GPU, rounding=r, sr_seed=SR_SEED(seed), B=4,
                         epochs=2, **HYP)
    assert np.array_equal(Wg, Wo) and np.array_equal(Hg, Ho)


@pytest.mark.parametrize("storage", ["fp32", "bf16"])
def test_wide_strat_balanced_rows_bitexact(gpu, tiny, storage):
    """Row blocks balanced by ratings (NOMAD_FLAG_BALANCE) with two element
    pairs per lane: single and batched runs equal the oracle's replay with
    the same row bounds."""
    tr, te = tiny
    k, B, seed = 210, 5, 99
    rb = gpu.row_blocks(tr.rowidx, tr.m, B)
    assert not np.array_equal(rb, np.minimum(np.arange(B + 1) * -(-tr.m // B), tr.m))
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, seed)
    rnd = "stochastic" if storage == "bf16" else "nearest"
    a = _ctx(gpu, tr, te, k, storage=storage, mode="strat", nworkers=B, balance=True, rounding=rnd, seed=seed)
    b = _ctx(gpu, tr, te, k, storage=storage, mode="strat", nworkers=B, balance=True, rounding=rnd, seed=seed)
    for c in (a, b):
        c.set_factors(W0, H0)
    a.train(epochs=2)
    gpu.train_batch([b], 2)
    ost = O.BF16 if storage == "bf16" else O.FP32
    orn = O.ROUND_STOCHASTIC if storage == "bf16" else O.ROUND_NEAREST
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=ost, arith=O.ARITH_GPU, rounding=orn, sr_seed=SR_SEED(seed), B=B,
                         epochs=2, row_bounds=rb, **HYP)
    for c, name in ((a, "single"), (b, "batched")):
        Wg, Hg = c.get_factors()
        assert np.array_equal(Wg, Wo) and np.array_equal(Hg, Ho), (name, np.abs(Wg - Wo).max())
        c.close()


def test_wide_batch_equals_single(gpu, tiny):
    """Two wide problems (k 150 and 256) in one batched launch, each equal bit
    for bit to its own run."""
    tr, te = tiny
    runs = []
    for k, B in ((150, 3), (256, 6)):
        W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, k)
        runs.append((k, B, W0, H0))
    batch = []
    for k, B, W0, H0 in runs:
        c = _ctx(gpu, tr, te, k, storage="bf16", rounding="stochastic", mode="strat", nworkers=B, seed=k)
        c.set_factors(W0, H0)
        batch.append(c)
    gpu.train_batch(batch, 2)
    for c, (k, B, W0, H0) in zip(batch, runs):
        with _ctx(gpu, tr, te, k, storage="bf16", rounding="stochastic", mode="strat", nworkers=B, seed=k) as s:
            s.set_factors(W0, H0)
            s.train(epochs=2)
            for x, y in zip(c.get_factors(), s.get_factors()):
                assert np.array_equal(x, y), k
        c.close()


def test_wide_batch_rejects_mixed_layouts(gpu, tiny):
    tr, te = tiny
    a = _ctx(gpu, tr, te, 100, storage="fp32", mode="strat", nworkers=2)
    b = _ctx(gpu, tr, te, 200, storage="fp32", mode="strat", nworkers=2)
    with pytest.raises(gpu.NomadError):
        gpu.train_batch([a, b], 1)
    a.close()
    b.close()


@pytest.mark.parametrize("storage,nworkers", [("fp32", 1024), ("bf16", 1024), ("fp64", 1024), ("bf16", 2048),
                                              ("fp64", 2048), ("fp32", 4096)])
def test_wide_async_counts_and_rmse(gpu, small, storage, nworkers):
    """k = 200 in async mode (1024 workers: 4 waves per workgroup, 32 LDS
    slots; 2048: 8 waves, 64 slots (32 for fp64); 4096: 16 waves): exactly
    nnz updates per epoch and, after 6 epochs, the test RMSE within 0.5 % of
    the CPU async reference."""
    tr, te = small
    k = 200
    epochs = 6
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 12345)
    rnd = "stochastic" if storage == "bf16" else "nearest"
    print(f"wide async {storage} {nworkers}", flush=True)
    with _ctx(gpu, tr, te, k, storage=storage, rounding=rnd, mode="async", nworkers=nworkers) as ctx:
        ctx.set_factors(W0, H0)
        for ep in range(epochs):
            st = ctx.train(epochs=1)
            assert st.updates == tr.nnz and st.epochs_done == ep + 1, (ep, st.updates)
        ev = ctx.eval()
    gpu_rmse = np.sqrt(ev.test_sse / ev.test_count)
    ost = O.FP64 if storage == "fp64" else O.FP32
    Wc, Hc, _ = O.sgd_async(tr, W0, H0, storage=ost, T=4, seed=12345, epochs=epochs, **HYP)
    cpu_rmse = np.sqrt(O.sse(te, Wc, Hc, storage=ost) / te.nnz)
    assert abs(gpu_rmse - cpu_rmse) <= 0.005 * cpu_rmse, (gpu_rmse, cpu_rmse)


def test_wide_async_timed_segments(gpu, small):
    """Timed segments with k = 256: updates every segment and the counts
    equal to the visits, as for narrow rows."""
    tr, te = small
    k = 256
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 4, 3)
    with _ctx(gpu, tr, te, k, storage="bf16", rounding="nearest", mode="async", nworkers=512) as ctx:
        ctx.set_factors(W0, H0)
        total = 0
        for _ in range(3):
            st = ctx.train(seconds=0.02)
            assert st.updates > 0
            total += st.updates
        rs = ctx.route_state()
        assert int(rs["count"].astype(np.int64) @ rs["nratings"].astype(np.int64)) == total
        ev = ctx.eval()
        assert np.isfinite(ev.test_sse) and ev.test_sse > 0


def test_wide_eval_matches_oracle(gpu, small):
    tr, te = small
    k = 210
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 3, 5)
    for storage, ost in (("fp32", O.FP32), ("fp64", O.FP64), ("bf16", O.BF16)):
        with _ctx(gpu, tr, te, k, storage=storage, mode="async", nworkers=16) as ctx:
            ctx.set_factors(W0, H0)
            ev = ctx.eval()
            W, H = ctx.get_factors()
        for a, got in ((tr, ev.train_sse), (te, ev.test_sse)):
            ref = O.sse(a, W, H, storage=ost, arith=O.ARITH_GPU)
            assert abs(got - ref) <= 1e-11 * ref, (storage, got, ref)
            ref_order = O.sse(a, W, H, storage=ost, arith=O.ARITH_REF)
            assert abs(got - ref_order) <= 1e-5 * ref_order


def test_wide_driver_dim_200(tmp_path, small):
    """The drop-in driver takes --dim up to 256 (the reference's --dim has no
    upper bound; above 256 the library refuses)."""
    import re
    import subprocess
    from pathlib import Path

    tr, te = small
    gpu_bin = Path(__file__).resolve().parents[1] / "nomad_amd" / "bin" / "nomad_float"
    gpu_lib = __import__("nomad_amd.lib", fromlist=["lib"])
    gpu_lib.write_dat(tmp_path / "train.dat", tr)
    gpu_lib.write_dat(tmp_path / "test.dat", te)
    print("[driver] nomad_float --dim 200", flush=True)
    r = subprocess.run([str(gpu_bin), "-i", str(tmp_path), "--timeout", "0.2", "--pause", "0", "--dim", "200",
                        "--lrate", "0.001", "--drate", "0.1", "--reg", "0.05"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rmse = [float(x) for x in re.findall(r"current test RMSE: (\d+\.\d{10})\n", r.stdout)]
    assert len(rmse) == 1 and np.isfinite(rmse[0])
    r = subprocess.run([str(gpu_bin), "-i", str(tmp_path), "--timeout", "0.1", "--pause", "0", "--dim", "257"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
