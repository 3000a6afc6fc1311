"""Batched stratified sweep (nomad_train_batch) and the autotuner on the GPU.

The bar for the batch: every problem's factors after a batched launch are
bit-identical to the same problem trained alone (nomad_train, strat mode),
for problems of different sizes, ranks, regularisations and strata counts,
in every storage type, for any persistent grid size, and across calls.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

STRATA = [3, 8, 16, 5, 1, 32, 12, 7]


def _problems(gpu, count, scale, first=0):
    out = []
    for p in range(first, first + count):
        spec, k, reg = gpu.preset_batch(p, scale)
        tr, te = gpu.synth(spec)
        out.append((tr, te, k, reg))
    return out


def _ctx(gpu, tr, te, k, reg, B, storage, seed, balance=False):
    rounding = "stochastic" if storage == "bf16" else "nearest"
    c = gpu.Context(tr.m, tr.n, k, storage=storage, rounding=rounding, mode="strat", nworkers=B, seed=seed,
                    lrate=0.002, drate=0.05, reg=reg, balance=balance)
    c.upload(tr, te)
    W0, H0 = gpu.init_factors(tr.m, tr.n, k, 2, 1000 + seed)
    c.set_factors(W0, H0)
    return c


@pytest.mark.parametrize("storage", ["fp32", "bf16", "fp64"])
def test_batch_bitwise_equals_single_runs(gpu, storage):
    probs = _problems(gpu, 6, 0.01)
    single, batch = [], []
    for i, (tr, te, k, reg) in enumerate(probs):
        B = STRATA[i % len(STRATA)]
        a = _ctx(gpu, tr, te, k, reg, B, storage, seed=i)
        a.train(epochs=2)
        a.train(epochs=1)
        single.append(a)
        batch.append(_ctx(gpu, tr, te, k, reg, B, storage, seed=i))
    st = gpu.train_batch(batch, 2)
    st2 = gpu.train_batch(batch, 1)
    for i, (a, b) in enumerate(zip(single, batch)):
        Wa, Ha = a.get_factors()
        Wb, Hb = b.get_factors()
        tr = probs[i][0]
        print(f"[batch {storage}] problem {i}: m={tr.m} n={tr.n} nnz={tr.nnz} k={probs[i][2]} B={a.nworkers}",
              flush=True)
        assert np.array_equal(Wa, Wb), f"problem {i}: W differs"
        assert np.array_equal(Ha, Hb), f"problem {i}: H differs"
        assert st[i].updates == 2 * tr.nnz and st2[i].updates == tr.nnz
        assert st2[i].epochs_done == 3
        assert np.isfinite(Wb).all() and np.abs(Wb).max() > 0
        a.close()
        b.close()


@pytest.mark.parametrize("waves", [1, 5, 64])
def test_batch_any_grid_size(gpu, waves):
    """Few waves (down to one workgroup) must give the same bits: the launch
    needs no co-residency."""
    probs = _problems(gpu, 3, 0.01, first=10)
    ref, got = [], []
    for i, (tr, te, k, reg) in enumerate(probs):
        ref.append(_ctx(gpu, tr, te, k, reg, STRATA[i], "fp32", seed=i))
        got.append(_ctx(gpu, tr, te, k, reg, STRATA[i], "fp32", seed=i))
    gpu.train_batch(ref, 2)
    gpu.train_batch(got, 2, waves=waves)
    for a, b in zip(ref, got):
        Wa, Ha = a.get_factors()
        Wb, Hb = b.get_factors()
        assert np.array_equal(Wa, Wb) and np.array_equal(Ha, Hb)
        a.close()
        b.close()


def test_batch_sweep_32_problems(gpu):
    """The sweep shape (32 uneven problems, k in {16..128}, reg in {0.01..0.5},
    BF16 with stochastic rounding) at 1/100 of its sizes: batched vs one run
    per problem, bitwise, and the batched launch's time beside the loop's."""
    probs = _problems(gpu, 32, 0.01)
    single, batch = [], []
    t_single = 0.0
    for i, (tr, te, k, reg) in enumerate(probs):
        B = 16
        a = _ctx(gpu, tr, te, k, reg, B, "bf16", seed=i)
        t_single += a.train(epochs=2).train_seconds
        single.append(a)
        batch.append(_ctx(gpu, tr, te, k, reg, B, "bf16", seed=i))
    t0 = time.perf_counter()
    st = gpu.train_batch(batch, 2)
    wall = time.perf_counter() - t0
    print(f"[batch sweep] 32 problems, {sum(p[0].nnz for p in probs)} ratings: batched {st[0].train_seconds * 1e3:.2f} ms"
          f" (host wall {wall * 1e3:.1f} ms), one launch sequence per problem {t_single * 1e3:.2f} ms", flush=True)
    for i, (a, b) in enumerate(zip(single, batch)):
        Wa, Ha = a.get_factors()
        Wb, Hb = b.get_factors()
        assert np.array_equal(Wa, Wb) and np.array_equal(Ha, Hb), f"problem {i} differs"
        ev = b.eval()
        assert np.isfinite(ev.test_sse)
        a.close()
        b.close()


def test_batch_rejects_mixed_contexts(gpu, tiny):
    tr, te = tiny
    a = gpu.Context(tr.m, tr.n, 8, mode="strat", nworkers=4)
    b = gpu.Context(tr.m, tr.n, 8, mode="async", nworkers=64)
    c = gpu.Context(tr.m, tr.n, 8, storage="fp64", mode="strat", nworkers=4)
    for x in (a, b, c):
        x.upload(tr, te)
    for bad in ([a, b], [a, c], [a, a]):
        with pytest.raises(gpu.NomadError) as e:
            gpu.train_batch(bad, 1)
        assert e.value.code == -1
    d = gpu.Context(tr.m, tr.n, 8, mode="strat", nworkers=4)  # no data uploaded
    with pytest.raises(gpu.NomadError) as e:
        gpu.train_batch([a, d], 1)
    assert e.value.code == -3
    for x in (a, b, c, d):
        x.close()


@pytest.mark.parametrize("mode", ["async", "strat"])
def test_autotune_picks_a_valid_fastest_config(gpu, small, mode):
    tr, te = small
    best, trials = gpu.autotune(tr, 16, mode=mode, storage="fp32", lrate=0.001, drate=0.1, reg=0.05,
                                trial_seconds=0.02)
    ok = [t for t in trials if t.status == 0]
    print(f"[autotune {mode}] " + ", ".join(f"({t.nworkers},{t.flags}):{t.updates_per_s / 1e9:.3f}G" for t in trials),
          flush=True)
    assert ok and best.status == 0
    assert best.updates_per_s == max(t.updates_per_s for t in ok)
    assert any(t.nworkers == best.nworkers and t.flags == best.flags for t in ok)
    # the chosen configuration trains
    c = gpu.Context(tr.m, tr.n, 16, mode=mode, nworkers=best.nworkers, balance=bool(best.flags & gpu.FLAG_BALANCE),
                    lrate=0.001, drate=0.1, reg=0.05)
    assert c.nworkers == best.nworkers
    c.upload(tr, te)
    W0, H0 = gpu.init_factors(tr.m, tr.n, 16, 4, 12345)
    c.set_factors(W0, H0)
    st = c.train(epochs=2)
    assert st.updates >= tr.nnz
    ev = c.eval()
    assert np.isfinite(ev.test_sse)
    c.close()


def test_autotune_batch_picks_valid_strata_counts(gpu):
    """Every chosen count is a candidate (a common count in 64..1024, or 2x
    or 4x it for a slow problem) capped at the problem's rows; the chosen
    counts and row balances train."""
    probs = _problems(gpu, 4, 0.01, first=20)
    strata, balanced, rate = gpu.autotune_batch([(tr, k, reg) for tr, te, k, reg in probs], tune_balance=True)
    print(f"[autotune batch] strata {strata}, balanced {balanced}, {rate / 1e9:.3f} G updates/s", flush=True)
    assert rate > 0
    uncapped = [B for (tr, te, k, reg), B in zip(probs, strata) if B != tr.m]
    common = min(uncapped) if uncapped else 0
    for (tr, te, k, reg), B in zip(probs, strata):
        assert 1 <= B <= tr.m
        assert B == tr.m or (B in (64, 128, 256, 512, 1024, 2048, 4096) and B in (common, 2 * common, 4 * common))
    cs = [_ctx(gpu, tr, te, k, reg, B, "bf16", seed=i, balance=bl)
          for i, ((tr, te, k, reg), B, bl) in enumerate(zip(probs, strata, balanced))]
    st = gpu.train_batch(cs, 1)
    assert all(s.updates == p[0].nnz for s, p in zip(st, probs))
    for c in cs:
        c.close()


def _sr_seed(seed):
    return ((seed ^ (seed >> 32)) & 0xFFFFFFFF) ^ 0xA5A5A5A5  # the library's stochastic-rounding key


@pytest.mark.parametrize("storage", ["bf16", "fp32"])
def test_batch_k64_k128_B256_vs_oracle(gpu, storage):
    """Problems of the sweep with k = 64 and k = 128 at 1/100 of their sizes,
    256 strata each (capped at the rows), trained in one batched launch, are
    bit-identical to the oracle's replay of the same stratified schedule in
    the kernels' arithmetic (fp32, and bf16 with stochastic rounding)."""
    found = {}
    p = 0
    while len(found) < 2 and p < 256:
        spec, k, reg = gpu.preset_batch(p, 0.01)
        if k in (64, 128) and k not in found:
            found[k] = (p, spec, reg)
        p += 1
    assert set(found) == {64, 128}
    ctxs, probs = [], []
    for i, (k, (p, spec, reg)) in enumerate(sorted(found.items())):
        tr, te = gpu.synth(spec)
        B = min(256, tr.m)
        c = _ctx(gpu, tr, te, k, reg, B, storage, seed=i)
        ctxs.append(c)
        probs.append((tr, k, reg, B, i))
    gpu.train_batch(ctxs, 2)
    for c, (tr, k, reg, B, seed) in zip(ctxs, probs):
        W0, H0 = gpu.init_factors(tr.m, tr.n, k, 2, 1000 + seed)
        ost = O.BF16 if storage == "bf16" else O.FP32
        rnd = O.ROUND_STOCHASTIC if storage == "bf16" else O.ROUND_NEAREST
        Wo, Ho = O.sgd_strat(tr, W0, H0, storage=ost, arith=O.ARITH_GPU, rounding=rnd, sr_seed=_sr_seed(seed), B=B,
                             epochs=2, lrate=0.002, drate=0.05, reg=reg)
        Wg, Hg = c.get_factors()
        print(f"[batch vs oracle {storage}] m={tr.m} n={tr.n} nnz={tr.nnz} k={k} B={B}", flush=True)
        assert np.array_equal(Wg, Wo) and np.array_equal(Hg, Ho), (k, np.abs(Wg - Wo).max())
        c.close()


@pytest.mark.parametrize("storage", ["fp32", "bf16", "fp64"])
def test_balanced_strata_vs_oracle(gpu, storage):
    """NOMAD_FLAG_BALANCE in strat mode: the row blocks are the balanced split
    of nomad_row_blocks.  A single run and a batched run (beside an
    unbalanced problem) are bit-identical to the oracle's replay with those
    row bounds, on problems whose balanced split differs from ceil(m / B)."""
    probs = _problems(gpu, 3, 0.01, first=21)
    rounding = "stochastic" if storage == "bf16" else "nearest"
    spec = [(16, True), (7, False), (32, True)]

    def make(i):
        tr, te, k, reg = probs[i]
        B, bal = spec[i]
        c = gpu.Context(tr.m, tr.n, k, storage=storage, rounding=rounding, mode="strat", nworkers=B, seed=i,
                        lrate=0.002, drate=0.05, reg=reg, balance=bal)
        c.upload(tr, te)
        W0, H0 = gpu.init_factors(tr.m, tr.n, k, 2, 1000 + i)
        c.set_factors(W0, H0)
        return c

    batch = [make(i) for i in range(3)]
    gpu.train_batch(batch, 2)
    for i, (tr, te, k, reg) in enumerate(probs):
        B, bal = spec[i]
        rb = gpu.row_blocks(tr.rowidx, tr.m, B) if bal else None
        if bal:
            eq = np.minimum(np.arange(B + 1) * -(-tr.m // B), tr.m)
            assert not np.array_equal(rb, eq), "the balanced split must differ from the default one"
        single = make(i)
        single.train(epochs=2)
        W0, H0 = gpu.init_factors(tr.m, tr.n, k, 2, 1000 + i)
        ost = {"fp32": O.FP32, "bf16": O.BF16, "fp64": O.FP64}[storage]
        rnd = O.ROUND_STOCHASTIC if storage == "bf16" else O.ROUND_NEAREST
        Wo, Ho = O.sgd_strat(tr, W0, H0, storage=ost, arith=O.ARITH_GPU, rounding=rnd, sr_seed=_sr_seed(i), B=B,
                             epochs=2, lrate=0.002, drate=0.05, reg=reg, row_bounds=rb)
        print(f"[balanced {storage}] problem {i}: m={tr.m} n={tr.n} nnz={tr.nnz} k={k} B={B} balanced={bal}",
              flush=True)
        for c, name in ((single, "single"), (batch[i], "batched")):
            Wg, Hg = c.get_factors()
            assert np.array_equal(Wg, Wo) and np.array_equal(Hg, Ho), (name, i, np.abs(Wg - Wo).max())
        single.close()
        batch[i].close()


def test_batch_sweep_full_size_autotuned(gpu):
    """The full batched sweep: 32 problems at full size (about 161 M ratings),
    the batch autotuner's common strata count, one batched epoch bit-identical
    to the same epoch of every problem run on its own."""
    probs = _problems(gpu, 32, 1.0)
    total = sum(p[0].nnz for p in probs)
    print(f"[sweep full] 32 problems, {total} training ratings", flush=True)
    strata, balanced, rate = gpu.autotune_batch([(tr, k, reg) for tr, te, k, reg in probs], balance=True,
                                                tune_balance=True)
    print(f"[sweep full] autotuned strata {sorted(set(strata))}, {sum(balanced)} balanced, {rate / 1e9:.3f} G "
          f"updates/s", flush=True)
    batch = [_ctx(gpu, tr, te, k, reg, B, "bf16", seed=i, balance=bl)
             for i, ((tr, te, k, reg), B, bl) in enumerate(zip(probs, strata, balanced))]
    st = gpu.train_batch(batch, 1)
    print(f"[sweep full] batched epoch {st[0].train_seconds * 1e3:.1f} ms", flush=True)
    t_seq = 0.0
    for i, ((tr, te, k, reg), B, bl) in enumerate(zip(probs, strata, balanced)):
        a = _ctx(gpu, tr, te, k, reg, B, "bf16", seed=i, balance=bl)
        t_seq += a.train(epochs=1).train_seconds
        Wa, Ha = a.get_factors()
        Wb, Hb = batch[i].get_factors()
        assert st[i].updates == tr.nnz
        assert np.array_equal(Wa, Wb) and np.array_equal(Ha, Hb), f"problem {i} differs"
        a.close()
    print(f"[sweep full] one problem after another {t_seq * 1e3:.1f} ms", flush=True)
    for c in batch:
        c.close()
