"""nomad_row_blocks: the row blocks of NOMAD_FLAG_BALANCE (host code, no
device).  Checked against an exhaustive min-max partition on small inputs:
the blocks are contiguous, non-empty, cover every row, and the largest
block holds no more ratings than the best contiguous split into P blocks."""
import numpy as np
import pytest

from nomad_amd import lib


def best_max_load(deg, P):
    """Smallest possible largest block load over contiguous splits of the
    rows into exactly P non-empty blocks (dynamic programming)."""
    m = len(deg)
    pref = np.concatenate([[0], np.cumsum(deg)])
    INF = float("inf")
    f = [[INF] * (m + 1) for _ in range(P + 1)]  # f[b][r]: rows 0..r-1 in b blocks
    f[0][0] = 0
    for b in range(1, P + 1):
        for r in range(b, m + 1):
            f[b][r] = min(max(f[b - 1][s], pref[r] - pref[s]) for s in range(b - 1, r))
    return f[P][m]


@pytest.mark.parametrize("seed", range(12))
def test_row_blocks_min_max_contiguous(seed):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(3, 28))
    P = int(rng.integers(1, m + 1))
    # heavy-tailed degrees, some rows empty
    deg = (rng.pareto(1.2, m) * 4).astype(np.int64) * (rng.random(m) < 0.85)
    rowidx = np.repeat(np.arange(m, dtype=np.int32), deg)
    rng.shuffle(rowidx)
    b = lib.row_blocks(rowidx, m, P)
    assert b[0] == 0 and b[-1] == m and len(b) == P + 1
    assert np.all(np.diff(b) >= 1), b
    pref = np.concatenate([[0], np.cumsum(deg)])
    load = pref[b[1:]] - pref[b[:-1]]
    assert load.max() == best_max_load(deg, P), (deg, P, b)


def test_row_blocks_one_row_each_and_errors():
    rowidx = np.array([0, 0, 2, 1, 2, 2], np.int32)
    assert list(lib.row_blocks(rowidx, 3, 3)) == [0, 1, 2, 3]
    assert list(lib.row_blocks(np.zeros(0, np.int32), 4, 2)) in ([0, 1, 4], [0, 2, 4], [0, 3, 4])
    for m, P in ((3, 4), (3, 0), (0, 1)):
        with pytest.raises(lib.NomadError):
            lib.row_blocks(rowidx[rowidx < max(m, 1)], m, P)
    with pytest.raises(lib.NomadError):
        lib.row_blocks(np.array([5], np.int32), 3, 2)  # row index out of range


def best_max_ratio(deg, cap):
    """Smallest possible max over blocks of load / capacity (contiguous,
    exactly len(cap) non-empty blocks, block b with capacity cap[b])."""
    m, P = len(deg), len(cap)
    pref = np.concatenate([[0], np.cumsum(deg)])
    INF = float("inf")
    f = [[INF] * (m + 1) for _ in range(P + 1)]
    f[0][0] = 0.0
    for b in range(1, P + 1):
        for r in range(b, m + 1):
            f[b][r] = min(max(f[b - 1][s], (pref[r] - pref[s]) / cap[b - 1]) for s in range(b - 1, r))
    return f[P][m]


@pytest.mark.parametrize("seed", range(8))
def test_row_blocks_with_capacities(seed):
    """With capacities (the async context lowers the last wave's of each
    workgroup), on light rows where no block needs splitting, the largest
    load / capacity is within the integer bound of the best split."""
    rng = np.random.default_rng(100 + seed)
    P = int(rng.integers(2, 9))
    m = int(rng.integers(4 * P, 6 * P))
    deg = rng.integers(1, 7, m)
    cap = np.where(np.arange(P) % 4 == 3, 0.7, 1.0) * rng.uniform(0.8, 1.0, P)
    rowidx = np.repeat(np.arange(m, dtype=np.int32), deg)
    b = lib.row_blocks(rowidx, m, P, cap)
    assert b[0] == 0 and b[-1] == m and np.all(np.diff(b) >= 1)
    pref = np.concatenate([[0], np.cumsum(deg)])
    ratio = ((pref[b[1:]] - pref[b[:-1]]) / cap).max()
    best = best_max_ratio(deg, cap)
    assert ratio <= np.ceil(best - 1e-9) + 1e-9, (ratio, best)
    with pytest.raises(lib.NomadError):
        lib.row_blocks(rowidx, m, P, np.full(P, 1.5))
