"""CPU checks that pin the oracle against things other than itself: the C++
standard's MT19937-64 definition, closed forms, hand-computed updates, numpy
FP64 replays, finite differences and invariants."""
import json
from pathlib import Path

import numpy as np
import pytest

from nomad_amd import lib
from oracle import oracle as O

HYP = dict(lrate=0.001, drate=0.1, reg=0.05)


# --- an independent MT19937-64 (Matsumoto & Nishimura), pinned by the standard
class MT64:
    def __init__(self, seed):
        self.mt = [0] * 312
        self.mt[0] = seed & 0xFFFFFFFFFFFFFFFF
        for i in range(1, 312):
            x = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (x ^ (x >> 62)) + i) & 0xFFFFFFFFFFFFFFFF
        self.i = 312

    def __call__(self):
        if self.i >= 312:
            mt = self.mt
            for i in range(312):
                x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
                mt[i] = mt[(i + 156) % 312] ^ (x >> 1) ^ (0xB5026F5AA96619E9 if x & 1 else 0)
            self.i = 0
        x = self.mt[self.i]
        self.i += 1
        x ^= (x >> 29) & 0x5555555555555555
        x ^= (x << 17) & 0x71D67FFFEDA60000
        x ^= (x << 37) & 0xFFF7EEE000000000
        x ^= x >> 43
        return x & 0xFFFFFFFFFFFFFFFF


def test_mt64_standard_value():
    g = MT64(5489)
    for _ in range(9999):
        g()
    assert g() == 9981545732273789042  # [rand.predef] mt19937_64, 10000th output


def _uniform(g, b):
    # libstdc++ generate_canonical<double, 53> with a 64-bit engine: one draw,
    # double(x) / 2^64, then a + (b - a) * u with a = 0
    u = float(g()) / 2.0 ** 64
    if u >= 1.0:
        u = np.nextafter(1.0, 0.0)
    return u * b + 0.0


@pytest.mark.parametrize("T", [1, 3])
def test_init_factors_streams(T):
    m, n, k, seed = 10, 7, 5, 12345
    W, H = O.init_factors(m, n, k, T, seed)
    b = 1.0 / np.sqrt(k)
    per = m // T + (m % T > 0)
    for t in range(T):
        g = MT64(seed + t + 1)
        rows = range(t * per, min((t + 1) * per, m))
        exp = [_uniform(g, b) for _ in range(len(rows) * k)]
        assert np.array_equal(W[rows.start:rows.stop].ravel(), np.array(exp))
    g = MT64(seed + 139)
    assert np.array_equal(H.ravel(), np.array([_uniform(g, b) for _ in range(n * k)]))
    # the product's initialiser draws the same streams
    W2, H2 = lib.init_factors(m, n, k, T, seed)
    assert np.array_equal(W, W2) and np.array_equal(H, H2)


def test_init_perms_are_permutations():
    P = O.init_perms(50, 6, 12345)
    assert (np.sort(P, axis=1) == np.arange(6)).all()
    assert np.array_equal(P, O.init_perms(50, 6, 12345))
    assert not np.array_equal(P, O.init_perms(50, 6, 12346))


M64 = (1 << 64) - 1


def _uid(g, lo, hi):
    """libstdc++'s uniform_int_distribution{lo, hi} on a 64-bit engine:
    Lemire's multiply-and-shift draw over 128 bits, redrawn while the low
    half falls below (2^64 - range) mod range."""
    rng = hi - lo + 1
    prod = g() * rng
    if (prod & M64) < rng:
        thr = ((1 << 64) - rng) % rng
        while (prod & M64) < thr:
            prod = g() * rng
    return (prod >> 64) + lo


def _shuffle(v, g):
    """libstdc++'s std::shuffle for a 64-bit engine and a short range: an
    even-length range first swaps element 1 with a draw in {0, 1}; then two
    elements per draw, positions x / (r + 1) and x % (r + 1) of one draw
    x in [0, r (r + 1)) with r = i + 1."""
    n = len(v)
    if n == 0:
        return
    assert M64 // n >= n
    i = 1
    if n % 2 == 0:
        r = _uid(g, 0, 1)
        v[i], v[r] = v[r], v[i]
        i += 1
    while i != n:
        r = i + 1
        x = _uid(g, 0, r * (r + 1) - 1)
        p1, p2 = x // (r + 1), x % (r + 1)
        v[i], v[p1] = v[p1], v[i]
        i += 1
        v[i], v[p2] = v[p2], v[i]
        i += 1


@pytest.mark.parametrize("T,seed", [(1, 5), (2, 12345), (3, 7), (6, 12345), (7, 99), (64, 1), (256, 12345)])
def test_init_perms_match_libstdcxx_shuffle(T, seed):
    """orc_init_perms (set_perm with mt19937_64(seed + 137), pool.cpp:32-38,
    nb_main.cpp:79-87) equals the restated shuffle draw for draw; a wrong
    seed offset, a missing first swap or a one-swap-per-draw shuffle fails."""
    n = 40
    P = O.init_perms(n, T, seed)
    g = MT64(seed + 137)
    for j in range(n):
        v = list(range(T))
        _shuffle(v, g)
        assert P[j].tolist() == v, (T, seed, j)


def test_step_closed_form():
    for c in range(5):
        exp = 0.001 * 1.5 / (1.0 + 0.1 * (c + 1) ** 1.5)
        assert O.step(O.FP64, 0.001, 0.1, c) == pytest.approx(exp, rel=1e-15)
        assert O.step(O.FP32, 0.001, 0.1, c) == np.float32(
            np.float64(np.float32(0.001)) * 1.5 / (1.0 + np.float64(np.float32(0.1)) * (c + 1) ** 1.5))


def _one_rating(k=2):
    # a 1 x 1 matrix with a single rating r = 3
    return lib.Csc(1, 1, np.array([0, 1], np.int64), np.array([0], np.int32), np.array([3.0]))


def test_one_rating_hand_computed():
    a = _one_rating()
    w = np.array([[0.5, -0.25]])
    h = np.array([[0.2, 0.4]])
    eta = 0.001 * 1.5 / (1.0 + 0.1 * 1.0)  # count 0
    lam = 0.05
    err = -3.0 + 0.2 * 0.5 + 0.4 * -0.25   # = -3.0
    w_exp = w - eta * (err * h + lam * w)
    h_exp = h - eta * (err * w + lam * h)
    W1, H1 = O.sgd_strat(a, w, h, storage=O.FP64, arith=O.ARITH_REF, B=1, epochs=1, **HYP)
    assert np.allclose(W1, w_exp, rtol=0, atol=1e-16) and np.allclose(H1, h_exp, rtol=0, atol=1e-16)
    # both factors move from the *old* values (a dropped "tmp" would fail this)
    assert not np.allclose(H1, h - eta * (err * W1 + lam * h))
    W2, H2 = O.sgd_strat(a, w, h, storage=O.FP64, arith=O.ARITH_GPU, B=1, epochs=1, **HYP)
    assert np.allclose(W2, W1, rtol=1e-15) and np.allclose(H2, H1, rtol=1e-15)


def _numpy_serial(a, W, H, epochs, lrate, drate, reg):
    W = W.copy()
    H = H.copy()
    for e in range(epochs):
        eta = lrate * 1.5 / (1.0 + drate * (e + 1) ** 1.5)
        for j in range(a.n):
            for o in range(a.colptr[j], a.colptr[j + 1]):
                i = a.rowidx[o]
                err = W[i] @ H[j] - a.val[o]
                wi = W[i].copy()
                W[i] -= eta * (err * H[j] + reg * wi)
                H[j] -= eta * (err * wi + reg * H[j])
    return W, H


def test_serial_replay_matches_numpy_fp64():
    tr, _ = lib.synth(lib.preset(1, 0.05))
    k = 6
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 1, 3)
    Wn, Hn = _numpy_serial(tr, W0, H0, 2, **HYP)
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=1, epochs=2, **HYP)
    assert np.allclose(Wo, Wn, rtol=1e-12, atol=1e-14) and np.allclose(Ho, Hn, rtol=1e-12, atol=1e-14)


def test_strat_tiles_commute():
    """Tiles of one sub-epoch touch disjoint rows and columns: B = 4 equals a
    replay whose tiles run in another order -- checked by B=4 vs the numpy
    replay of the same tile sequence."""
    tr, _ = lib.synth(lib.preset(1, 0.05))
    k, B = 4, 4
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 2, 9)
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=B, epochs=1, **HYP)
    W, H = W0.copy(), H0.copy()
    rpb = -(-tr.m // B)
    cpb = -(-tr.n // B)
    eta = HYP["lrate"] * 1.5 / (1.0 + HYP["drate"])
    for s in range(B):
        for b in reversed(range(B)):
            cb = (b + s) % B
            for j in range(cb * cpb, min(tr.n, (cb + 1) * cpb)):
                for o in range(tr.colptr[j], tr.colptr[j + 1]):
                    i = tr.rowidx[o]
                    if not (b * rpb <= i < (b + 1) * rpb):
                        continue
                    err = W[i] @ H[j] - tr.val[o]
                    wi = W[i].copy()
                    W[i] -= eta * (err * H[j] + HYP["reg"] * wi)
                    H[j] -= eta * (err * wi + HYP["reg"] * H[j])
    assert np.allclose(Wo, W, rtol=1e-12, atol=1e-14) and np.allclose(Ho, H, rtol=1e-12, atol=1e-14)


def test_strat_row_bounds():
    """Row blocks given as bounds (uneven, one of them empty): the oracle
    equals a numpy replay of the same tile sequence with those blocks; the
    ceil(m / B) bounds give the default split bit for bit; malformed bounds
    are rejected."""
    tr, _ = lib.synth(lib.preset(1, 0.05))
    k, B = 4, 4
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 2, 11)
    rb = np.array([0, 3, 3, tr.m // 2, tr.m], np.int32)
    Wo, Ho = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=B, epochs=1, row_bounds=rb, **HYP)
    W, H = W0.copy(), H0.copy()
    cpb = -(-tr.n // B)
    eta = HYP["lrate"] * 1.5 / (1.0 + HYP["drate"])
    for s in range(B):
        for b in range(B):
            cb = (b + s) % B
            for j in range(cb * cpb, min(tr.n, (cb + 1) * cpb)):
                for o in range(tr.colptr[j], tr.colptr[j + 1]):
                    i = tr.rowidx[o]
                    if not (rb[b] <= i < rb[b + 1]):
                        continue
                    err = W[i] @ H[j] - tr.val[o]
                    wi = W[i].copy()
                    W[i] -= eta * (err * H[j] + HYP["reg"] * wi)
                    H[j] -= eta * (err * wi + HYP["reg"] * H[j])
    assert np.allclose(Wo, W, rtol=1e-12, atol=1e-14) and np.allclose(Ho, H, rtol=1e-12, atol=1e-14)
    # a different split is a different schedule
    Wd, _ = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=B, epochs=1, **HYP)
    assert not np.array_equal(Wd, Wo)
    rpb = -(-tr.m // B)
    eq = np.minimum(np.arange(B + 1) * rpb, tr.m).astype(np.int32)
    for st in (O.FP32, O.BF16):
        a = O.sgd_strat(tr, W0, H0, storage=st, arith=O.ARITH_GPU, B=B, epochs=2, row_bounds=eq, **HYP)
        d = O.sgd_strat(tr, W0, H0, storage=st, arith=O.ARITH_GPU, B=B, epochs=2, **HYP)
        assert np.array_equal(a[0], d[0]) and np.array_equal(a[1], d[1])
    for bad in ([1, 3, 3, 5, tr.m], [0, 3, 2, 5, tr.m], [0, 3, 3, 5, tr.m - 1]):
        with pytest.raises(RuntimeError):
            O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=B, epochs=1,
                        row_bounds=np.array(bad, np.int32), **HYP)


def test_gpu_order_close_to_reference_order():
    tr, _ = lib.synth(lib.preset(1, 0.2))
    k = 16
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    Wr, Hr = O.sgd_strat(tr, W0, H0, storage=O.FP32, arith=O.ARITH_REF, B=4, epochs=2, **HYP)
    Wg, Hg = O.sgd_strat(tr, W0, H0, storage=O.FP32, arith=O.ARITH_GPU, B=4, epochs=2, **HYP)
    assert np.abs(Wr - Wg).max() <= 1e-5 * np.abs(Wr).max()
    assert not np.array_equal(Wr, Wg)  # the two orders are really different


@pytest.mark.parametrize("k", [200, 256])
def test_gpu_order_wide_rows(k):
    """Rows of more than 128 elements: the GPU order takes pairs l and l + 64
    in lane l.  FP64 agrees with the reference order to rounding (a dropped,
    doubled or misplaced second pair would not), and FP32 stays within 1e-5
    of it while really differing from it."""
    tr, _ = lib.synth(lib.preset(1, 0.15))
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    Wr, Hr = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_REF, B=3, epochs=2, **HYP)
    Wg, Hg = O.sgd_strat(tr, W0, H0, storage=O.FP64, arith=O.ARITH_GPU, B=3, epochs=2, **HYP)
    assert np.abs(Wr - Wg).max() <= 1e-12 * np.abs(Wr).max()
    assert np.abs(Hr - Hg).max() <= 1e-12 * np.abs(Hr).max()
    Wr, _ = O.sgd_strat(tr, W0, H0, storage=O.FP32, arith=O.ARITH_REF, B=3, epochs=2, **HYP)
    Wg, _ = O.sgd_strat(tr, W0, H0, storage=O.FP32, arith=O.ARITH_GPU, B=3, epochs=2, **HYP)
    assert np.abs(Wr - Wg).max() <= 1e-5 * np.abs(Wr).max()
    assert not np.array_equal(Wr, Wg)


def test_sse_gpu_order_wide_rows():
    """The SSE in the GPU order for k = 200 equals numpy's FP64 dot products
    to rounding."""
    tr, te = lib.synth(lib.preset(1, 0.15))
    k = 200
    W, H = lib.init_factors(tr.m, tr.n, k, 2, 1)
    r, c, v = te.to_triplets()
    want = float(np.sum((np.einsum("ij,ij->i", W[r], H[c]) - v) ** 2))
    got = O.sse(te, W, H, storage=O.FP64, arith=O.ARITH_GPU)
    assert abs(got - want) <= 1e-12 * want


def test_async_one_thread_is_serial():
    tr, _ = lib.synth(lib.preset(1, 0.1))
    W0, H0 = lib.init_factors(tr.m, tr.n, 8, 1, 4)
    Wa, Ha, st = O.sgd_async(tr, W0, H0, storage=O.FP32, T=1, seed=4, epochs=3, **HYP)
    Ws, Hs = O.sgd_strat(tr, W0, H0, storage=O.FP32, arith=O.ARITH_REF, B=1, epochs=3, **HYP)
    assert st.updates == 3 * tr.nnz and st.min_cycles == 3
    assert np.array_equal(Wa, Ws) and np.array_equal(Ha, Hs)


def _bf16_rne(x):
    """Round float32 values to the nearest bf16 (torch's conversion, ties to even)."""
    import torch

    return torch.from_numpy(np.asarray(x, np.float32).copy()).to(torch.bfloat16).to(torch.float32).numpy()


def test_async_bf16_one_thread_matches_float32_replay():
    """bf16 storage with round to nearest in the async restatement at T = 1
    equals a float32 replay of the serial order in which the initial factors,
    each row after its update and each column after its visit are rounded by
    torch's bf16 conversion; steps large enough that rounding changes results."""
    tr, _ = lib.synth(lib.preset(1, 0.05))
    k = 6
    hyp = dict(lrate=0.05, drate=0.1, reg=0.05)
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 1, 3)
    W = _bf16_rne(W0.astype(np.float32))
    H = _bf16_rne(H0.astype(np.float32))
    f = np.float32
    reg = f(hyp["reg"])
    for e in range(2):
        step = f(O.step(O.BF16, hyp["lrate"], hyp["drate"], e))
        for j in range(tr.n):
            h = H[j].copy()
            for o in range(tr.colptr[j], tr.colptr[j + 1]):
                i = tr.rowidx[o]
                w = W[i].copy()
                err = -f(tr.val[o])
                for d in range(k):
                    err = f(err + f(h[d] * w[d]))
                for d in range(k):
                    t = w[d]
                    w[d] = f(w[d] - f(step * f(f(err * h[d]) + f(reg * t))))
                    h[d] = f(h[d] - f(step * f(f(err * t) + f(reg * h[d]))))
                W[i] = _bf16_rne(w)
            H[j] = _bf16_rne(h)
    Wa, Ha, st = O.sgd_async(tr, W0, H0, storage=O.BF16, T=1, seed=4, epochs=2, **hyp)
    assert st.updates == 2 * tr.nnz
    assert np.array_equal(Wa, W.astype(np.float64)) and np.array_equal(Ha, H.astype(np.float64))
    assert not np.any(Wa.astype(np.float32).view(np.uint32) & 0xFFFF)
    Wf, _, _ = O.sgd_async(tr, W0, H0, storage=O.FP32, T=1, seed=4, epochs=2, **hyp)
    assert not np.array_equal(Wf, Wa)  # the rounding is really applied


@pytest.mark.parametrize("storage", [O.FP32, O.FP64])
def test_async_segments_one_thread_is_serial(storage):
    """Timeout segments at T = 1 (nb_main.cpp:109-328): segments cut by update
    budgets mid-epoch and mid-column leave the column rotation and the visit
    counts running, so after three epochs' worth of updates the result equals
    three uninterrupted epochs of the serial order bit for bit, and thread 0
    counts three visits of every column.  Restarting each segment at column
    0, or a count reset by set_perm, would change the steps and fail it."""
    tr, _ = lib.synth(lib.preset(1, 0.1))
    W0, H0 = lib.init_factors(tr.m, tr.n, 8, 1, 4)
    hyp = dict(lrate=0.01, drate=0.1, reg=0.05)
    run = O.AsyncRun(tr, W0, H0, storage=storage, T=1, seed=4, **hyp)
    total = 3 * tr.nnz
    budgets = [tr.nnz // 3 + 7, tr.nnz, tr.nnz // 2 + 1]
    done = 0
    for b in budgets:
        st = run.segment(max_updates=b)
        assert b <= st.updates <= b + int(np.diff(tr.colptr).max())
        done += st.updates
    st = run.segment(max_updates=total - done)
    done += st.updates
    assert done == total
    Wa, Ha = run.factors()
    Ws, Hs = O.sgd_strat(tr, W0, H0, storage=storage, arith=O.ARITH_REF, B=1, epochs=3, **hyp)
    assert np.array_equal(Wa, Ws) and np.array_equal(Ha, Hs)
    c = run.counts()
    nonempty = np.diff(tr.colptr) > 0
    last = np.nonzero(nonempty)[0].max()
    # columns after the last non-empty one are visited (empty) once the next
    # segment starts; every column up to it has exactly three visits
    assert (c[0, : last + 1] == 3).all()
    run.close()


def test_async_segments_threads_counts_match_updates():
    """At T > 1 each thread's visit count of a column, times the column's
    ratings in that thread's rows, summed over threads and columns, equals
    the updates of all segments; and a column's count at any thread lies
    between its completed route cycles and that plus the segments run."""
    tr, _ = lib.synth(lib.preset(1, 0.1))
    T = 4
    W0, H0 = lib.init_factors(tr.m, tr.n, 8, T, 4)
    run = O.AsyncRun(tr, W0, H0, storage=O.FP32, T=T, seed=4, **HYP)
    ups = 0
    for b in (tr.nnz // 2, tr.nnz, tr.nnz // 3):
        ups += run.segment(max_updates=b).updates
    ups += run.segment(epochs=1).updates
    c = run.counts()
    per = -(-tr.m // T)
    part = np.zeros((T, tr.n), np.int64)
    r, col, _ = tr.to_triplets()
    np.add.at(part, (r // per, col), 1)
    assert int((c.astype(np.int64) * part).sum()) == ups
    assert c.min() >= 1 and c.max() - c.min() <= 4 + 2
    run.close()


def test_async_threads_update_count():
    tr, _ = lib.synth(lib.preset(1, 0.1))
    W0, H0 = lib.init_factors(tr.m, tr.n, 8, 4, 4)
    _, _, st = O.sgd_async(tr, W0, H0, storage=O.FP64, T=4, seed=4, epochs=2, **HYP)
    assert st.updates == 2 * tr.nnz and st.min_cycles == 2


def test_sse_matches_numpy():
    tr, te = lib.synth(lib.preset(1, 0.1))
    W, H = lib.init_factors(tr.m, tr.n, 12, 2, 1)
    r, c, v = te.to_triplets()
    exp = float(np.sum((np.einsum("ij,ij->i", W[r], H[c]) - v) ** 2))
    assert O.sse(te, W, H, storage=O.FP64) == pytest.approx(exp, rel=1e-12)
    assert O.sse(te, W, H, storage=O.FP32) == pytest.approx(exp, rel=1e-5)


def test_loss_grad_numpy_and_finite_differences():
    tr, _ = lib.synth(lib.preset(4, 0.004))
    rng = np.random.default_rng(0)
    k, reg = 5, 0.1
    W = rng.normal(size=(tr.m, k)) * 0.3
    H = rng.normal(size=(tr.n, k)) * 0.3
    L, gW, gH = O.loss_grad(tr, W, H, reg)
    r, c, v = tr.to_triplets()
    e = np.einsum("ij,ij->i", W[r], H[c]) - v
    degr = np.bincount(r, minlength=tr.m)
    degc = np.bincount(c, minlength=tr.n)
    Lnp = np.sum(e ** 2) + reg * (np.sum(degr * np.sum(W ** 2, 1)) + np.sum(degc * np.sum(H ** 2, 1)))
    assert L == pytest.approx(Lnp, rel=1e-12)
    eps = 1e-6
    for _ in range(20):
        which = rng.integers(2)
        M, G = (W, gW) if which == 0 else (H, gH)
        i, d = rng.integers(M.shape[0]), rng.integers(k)
        old = M[i, d]
        M[i, d] = old + eps
        Lp = O.loss_grad(tr, W, H, reg, want_grad=False)[0]
        M[i, d] = old - eps
        Lm = O.loss_grad(tr, W, H, reg, want_grad=False)[0]
        M[i, d] = old
        fd = (Lp - Lm) / (2 * eps)
        assert fd == pytest.approx(G[i, d], rel=1e-5, abs=1e-6)


def test_ccd_decreases_objective():
    tr, _ = lib.synth(lib.preset(1, 0.1))
    k = 4
    rng = np.random.default_rng(1)
    W = rng.random((tr.m, k))
    H = np.zeros((tr.n, k))
    _, _, obj = O.ccd(tr, W, H, 0.1, 3, 3)
    o = obj[~np.isnan(obj)]
    assert (np.diff(o) <= 1e-9 * o[:-1]).all()


# ---- pinned against the reference competitor programs (tests/golden) --------
GOLDEN = Path(__file__).resolve().parent / "golden"


def ref_problem():
    z = np.load(GOLDEN / "ref_problem.npz")
    m, n = int(z["m"]), int(z["n"])
    tr = lib.Csc(m, n, z["tr_colptr"], z["tr_rowidx"], z["tr_val"])
    te = lib.Csc(m, n, z["te_colptr"], z["te_rowidx"], z["te_val"])
    return tr, te, json.loads((GOLDEN / "ref_traces.json").read_text())


def test_objective_matches_reference_dsgd_print():
    # sgd-mpi-omp with learning rate 0 prints calobj_mpi of its initial
    # factors: W from mt19937_64(12345 + 1), H from mt19937_64(12345 + 139)
    tr, _, g = ref_problem()
    k = g["k"]
    W, H = O.init_factors(tr.m, tr.n, k, 1, 12345)
    L, _, _ = O.loss_grad(tr, W, H, g["dsgd"]["reg"], want_grad=False)
    assert L == pytest.approx(g["dsgd"]["obj0"], rel=5e-6)


def test_ccd_matches_reference_ccdr1_trace():
    tr, te, g = ref_problem()
    k, c = g["k"], g["ccd"]
    W, _ = O.init_factors(tr.m, tr.n, k, 1, 12345)  # one thread: mt19937_64(12345 + 1)
    H = np.zeros((tr.n, k))
    _, _, obj, rmse = O.ccd(tr, W, H, c["lam"], c["iters"], c["inner"], c["eps"], test=te)
    for p in c["trace"]:
        q = (p["iter"] - 1) * k + p["rank"] - 1
        assert obj[q] == pytest.approx(p["obj"], rel=6e-6), p
        assert rmse[q] == pytest.approx(p["rmse"], rel=1e-9), p
    # ranks the reference did not print were skipped by its early stop
    printed = {(p["iter"] - 1) * k + p["rank"] - 1 for p in c["trace"]}
    assert all(np.isnan(obj[q]) for q in range(len(obj)) if q not in printed)


def test_stochastic_rounding_unbiased():
    """The kernels' stochastic rounding (bit-identical to this restatement,
    tests/test_gpu_sgd.py) rounds up with probability equal to the fraction
    of the bf16 ulp below which the value lies: over 40,000 hashed draws the
    mean of the rounded values matches the value within 4 sigma, for both
    halves of the random word; nearest-even rounding is biased at the same
    points."""
    rng = np.random.default_rng(11)
    for frac in (0.1, 0.3, 0.5, 0.77):
        lo = np.float32(1.25)  # bf16-representable; its bf16 ulp is 2**-7
        ulp = 2.0 ** -7
        x = float(np.float32(lo + frac * ulp))
        frac = (x - lo) / ulp
        n = 20000
        ups = np.zeros(2)
        for t in range(n):
            a, b, lane = (int(v) for v in rng.integers(0, 2**32, 3, dtype=np.uint64))
            r0, r1 = O.bf16_store_pair(x, x, O.ROUND_STOCHASTIC, 0xA5A5A5A5, a, b, lane % 64)
            assert r0 in (lo, lo + ulp) and r1 in (lo, lo + ulp)
            ups += [(r0 > lo), (r1 > lo)]
        p_hat = ups / n
        sigma = np.sqrt(frac * (1 - frac) / n)
        assert np.all(np.abs(p_hat - frac) < 4 * sigma), (frac, p_hat)
        rne, _ = O.bf16_store_pair(x, x, O.ROUND_NEAREST, 0, 0, 0, 0)
        assert rne == (lo + ulp if frac > 0.5 else lo)
