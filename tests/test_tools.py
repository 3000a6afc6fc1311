"""Host-side data pieces: the .dat format (Code/nomad/nb_data.cpp:19-176,
Scripts/convert.py layout), its reader/writer, the partitioned reading of the
reference, the drop-in driver's option handling.  CPU only."""
from __future__ import annotations

import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

from nomad_amd import lib
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = Path(__file__).resolve().parent / "golden"


def tiny():
    # 3 x 4, 5 ratings, given as triplets (row, col, value)
    trip = [(0, 1, 4.0), (0, 3, 1.5), (1, 0, 5.0), (2, 1, 3.0), (2, 2, -2.25)]
    rows = np.array([t[0] for t in trip]); cols = np.array([t[1] for t in trip])
    vals = np.array([t[2] for t in trip])
    order = np.lexsort((rows, cols))
    colptr = np.zeros(5, np.int64)
    np.add.at(colptr, cols + 1, 1)
    return lib.Csc(3, 4, np.cumsum(colptr), rows[order].astype(np.int32), vals[order]), trip


def expected_bytes(trip, m, n, long_format):
    # the format written by hand from its definition: header, per-row counts,
    # column indices and values in row-major order
    trip = sorted(trip)
    head = struct.pack("<iiiq", 1015, m, n, len(trip)) if long_format else struct.pack("<iiii", 1211216, m, n,
                                                                                       len(trip))
    counts = [sum(1 for t in trip if t[0] == i) for i in range(m)]
    return (head + struct.pack(f"<{m}i", *counts) + struct.pack(f"<{len(trip)}i", *[t[1] for t in trip])
            + struct.pack(f"<{len(trip)}d", *[t[2] for t in trip]))


@pytest.mark.parametrize("long_format", [False, True])
def test_dat_bytes_match_format(tmp_path, long_format):
    a, trip = tiny()
    p = tmp_path / "t.dat"
    lib.write_dat(p, a, long_format=long_format)
    assert p.read_bytes() == expected_bytes(trip, 3, 4, long_format)


@pytest.mark.parametrize("long_format", [False, True])
def test_dat_read_hand_written_file(tmp_path, long_format):
    a, trip = tiny()
    p = tmp_path / "h.dat"
    p.write_bytes(expected_bytes(trip, 3, 4, long_format))
    b = lib.read_dat(p)
    assert (b.m, b.n) == (3, 4)
    np.testing.assert_array_equal(b.colptr, a.colptr)
    np.testing.assert_array_equal(b.rowidx, a.rowidx)
    np.testing.assert_array_equal(b.val, a.val)


def smalldata():
    """libmf's smalldata fixture (Code/libmf-1.0/smalldata): 'user item rating'
    lines, ids from 0."""
    t = np.loadtxt(GOLDEN / "libmf_smalldata.txt", dtype=np.float64)
    rows, cols, vals = t[:, 0].astype(np.int64), t[:, 1].astype(np.int64), t[:, 2]
    m, n = int(rows.max()) + 1, int(cols.max()) + 1
    order = np.lexsort((rows, cols))
    colptr = np.zeros(n + 1, np.int64)
    np.add.at(colptr, cols + 1, 1)
    return lib.Csc(m, n, np.cumsum(colptr), rows[order].astype(np.int32), vals[order])


@pytest.mark.parametrize("nparts", [1, 3, 7])
def test_smalldata_round_trip_and_reference_partition(tmp_path, nparts):
    a = smalldata()
    assert a.nnz == 10000
    p = tmp_path / "small.dat"
    lib.write_dat(p, a)
    b = lib.read_dat(p)
    np.testing.assert_array_equal(b.colptr, a.colptr)
    np.testing.assert_array_equal(b.rowidx, a.rowidx)
    np.testing.assert_array_equal(b.val, a.val)
    # the reference's partitioned read: part q holds rows [q*ceil(m/P), ...)
    # as a CSC with local row ids
    rpb = -(-a.m // nparts)
    total = 0
    for q in range(nparts):
        m, n, tot, min_row, col, rows, vals = O.dat_part(p, q, nparts)
        assert (m, n, tot, min_row) == (a.m, a.n, a.nnz, q * rpb)
        sel = (a.rowidx >= q * rpb) & (a.rowidx < min((q + 1) * rpb, a.m))
        cols = np.repeat(np.arange(a.n), np.diff(a.colptr))[sel]
        np.testing.assert_array_equal(np.diff(col), np.bincount(cols, minlength=a.n))
        np.testing.assert_array_equal(rows, a.rowidx[sel] - q * rpb)
        np.testing.assert_array_equal(vals, a.val[sel])
        total += len(rows)
    assert total == a.nnz


def test_preset_sizes_exact():
    tr, te = lib.synth(lib.preset(1, 1.0))
    assert (tr.m, tr.n, tr.nnz + te.nnz) == (2000, 1000, 100000)
    # train and test are disjoint, rows ascend inside each column
    for a in (tr, te):
        for j in range(0, a.n, 97):
            r = a.rowidx[a.colptr[j]:a.colptr[j + 1]]
            assert np.all(np.diff(r) > 0)
    key_tr = set(zip(np.repeat(np.arange(tr.n), np.diff(tr.colptr)).tolist(), tr.rowidx.tolist()))
    key_te = set(zip(np.repeat(np.arange(te.n), np.diff(te.colptr)).tolist(), te.rowidx.tolist()))
    assert not key_tr & key_te


DRIVER = ROOT / "nomad_amd" / "bin" / "nomad_float"


def run_driver(*args):
    return subprocess.run([str(DRIVER), *args], capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not DRIVER.exists(), reason="driver not built")
def test_driver_rejects_bad_options(tmp_path):
    r = run_driver("--bogus", "1")
    assert r.returncode != 0 and "unrecognised option" in r.stderr
    r = run_driver("--timeout", "1")
    assert r.returncode != 0 and "input path has to be specified" in r.stderr
    r = run_driver("-i", str(tmp_path), "--timeout", "2", "1")
    assert r.returncode != 0 and "increasing" in r.stderr
    r = run_driver("-i", str(tmp_path), "--storage", "fp16")
    assert r.returncode != 0
    r = run_driver("-i", str(tmp_path), "--timeout", "1")  # no train.dat there
    assert r.returncode != 0 and "train.dat" in r.stderr
    # README spellings: --path for the data, --ptoken accepted (non-negative)
    r = run_driver("--path", str(tmp_path), "--ptoken", "1024", "--timeout", "1")
    assert r.returncode != 0 and "train.dat" in r.stderr
    r = run_driver("--path", str(tmp_path), "--ptoken", "-1")  # rejected before any file is read
    assert r.returncode != 0 and "cannot open" not in r.stderr


@pytest.mark.parametrize("nparts", [1, 4, 13])
def test_global_csc_helper_merges_reference_parts(tmp_path, nparts):
    """include/nomad_csc.h merges the reference's per-thread parts (local row
    ids, Code/nomad/nomad_body.h:34-46) into the global CSC the library takes:
    equal to the whole file read at once, for 1, 4 and 13 parts."""
    a = smalldata()
    p = tmp_path / "small.dat"
    lib.write_dat(p, a)
    exe = tmp_path / "csc_parts"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", f"-I{ROOT / 'oracle'}",
                    str(ROOT / "tests" / "csc_parts_main.cpp"), str(ROOT / "oracle" / "liborc.so"),
                    f"-Wl,-rpath,{ROOT / 'oracle'}", "-o", str(exe)], check=True)
    out = tmp_path / "g.bin"
    subprocess.run([str(exe), str(p), str(nparts), str(out)], check=True)
    raw = out.read_bytes()
    n, nnz = a.n, a.nnz
    colptr = np.frombuffer(raw, np.int64, n + 1)
    rowidx = np.frombuffer(raw, np.int32, nnz, 8 * (n + 1))
    val = np.frombuffer(raw, np.float64, nnz, 8 * (n + 1) + 4 * nnz)
    b = lib.read_dat(p)
    np.testing.assert_array_equal(colptr, b.colptr)
    np.testing.assert_array_equal(rowidx, b.rowidx)
    np.testing.assert_array_equal(val, b.val)


# ---- the reference's data tools: Scripts/convert.py, synthgen-conf, synthgen ----
class MT32:
    """std::mt19937 (Matsumoto & Nishimura, 32-bit), seeded like the standard
    engine's constructor."""

    def __init__(self, seed):
        self.mt = [seed & 0xFFFFFFFF]
        for i in range(1, 624):
            x = self.mt[-1]
            self.mt.append((1812433253 * (x ^ (x >> 30)) + i) & 0xFFFFFFFF)
        self.i = 624

    def __call__(self):
        if self.i >= 624:
            mt = self.mt
            for k in range(624):
                y = (mt[k] & 0x80000000) | (mt[(k + 1) % 624] & 0x7FFFFFFF)
                mt[k] = mt[(k + 397) % 624] ^ (y >> 1) ^ (0x9908B0DF if y & 1 else 0)
            self.i = 0
        y = self.mt[self.i]
        self.i += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9D2C5680
        y ^= (y << 15) & 0xEFC60000
        y ^= y >> 18
        return y & 0xFFFFFFFF


def test_mt32_standard_value():
    g = MT32(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995  # [rand.predef] mt19937, 10000th output


def _canonical(g):
    # libstdc++ generate_canonical<double, 53> on a 32-bit engine: two draws
    s = float(g()) + float(g()) * 4294967296.0
    u = s / 18446744073709551616.0
    return u if u < 1.0 else np.nextafter(1.0, 0.0)


class Normal:
    """libstdc++ normal_distribution: Marsaglia's polar method, the second
    value of each pair kept for the next call."""

    def __init__(self, mean, sd):
        self.mean, self.sd, self.saved = mean, sd, None

    def __call__(self, g):
        import math

        if self.saved is not None:
            r, self.saved = self.saved, None
        else:
            while True:
                x = 2.0 * _canonical(g) - 1.0
                y = 2.0 * _canonical(g) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2 * math.log(r2) / r2)
            self.saved = x * mult
            r = y * mult
        return r * self.sd + self.mean


class Discrete:
    """libstdc++ discrete_distribution: weights over their sum, cumulative
    sums (last set to 1), the first index whose cumulative sum reaches one
    canonical draw."""

    def __init__(self, w):
        import itertools

        s = 0.0
        for x in w:
            s += float(x)
        self.cp = list(itertools.accumulate(float(x) / s for x in w))
        self.cp[-1] = 1.0

    def __call__(self, g):
        import bisect

        return bisect.bisect_left(self.cp, _canonical(g))


def _restated_conf(nrow, ncol, seed):
    import math

    g = MT32(seed)
    nd = Normal(20.0, 5.0)
    rows = []
    for i in range(ncol + 1):
        v = int(nd(g))
        while v < 0 or v >= ncol:
            v = int(nd(g))
        rows.append((i, v))
    cols = []
    ex = 1.0 / (1 - 2.2)
    for i in range(nrow + 1):
        v = int(1.0 * math.pow(1.0 - _canonical(g), ex))
        while v > nrow:
            v = int(1.0 * math.pow(1.0 - _canonical(g), ex))
        cols.append((i, v))
    return rows, cols


def _restated_gen(rows, cols, nrow, ncol, dim, seed):
    g = MT32(seed)
    col_deg_dist = Discrete([f for _, f in cols])
    row_deg_dist = Discrete([f for _, f in rows])
    nd = Normal(0.0, 1.0)
    weights = [cols[col_deg_dist(g)][0] for _ in range(ncol)]
    col_dist = Discrete(weights)
    H = [nd(g) for _ in range(dim * ncol)]
    out = {"train": ([], [], []), "test": ([], [], [])}
    for _ in range(nrow):
        u = [nd(g) for _ in range(dim)]
        deg = min(rows[row_deg_dist(g)][0] * 5 // 4, ncol)
        ntr = int(deg * 0.8)
        tr, te = set(), set()
        for j in range(deg):
            while True:
                c = col_dist(g)
                s = tr if j < ntr else te
                if c not in s:
                    s.add(c)
                    break
        for name, s in (("train", tr), ("test", te)):
            sizes, ind, val = out[name]
            for c in sorted(s):
                t = 0.0
                for q in range(dim):
                    t += u[q] * H[dim * c + q]
                t += nd(g) * 0.1
                ind.append(c)
                val.append(t)
            sizes.append(len(s))
    files = {}
    for name, (sizes, ind, val) in out.items():
        files[name] = (struct.pack("<iiiq", 1015, nrow, ncol, sum(sizes)) + struct.pack(f"<{len(sizes)}I", *sizes)
                       + struct.pack(f"<{len(ind)}i", *ind) + struct.pack(f"<{len(val)}d", *val))
    return files


@pytest.mark.parametrize("nrow,ncol,seed", [(30, 12, 12345), (7, 40, 3)])
def test_synthgen_conf_matches_restatement(tmp_path, nrow, ncol, seed):
    """nomad_synthgen_conf's degree files equal an independent restatement of
    synthgen-conf.cpp:74-122 on libstdc++'s mt19937, polar normal and
    canonical draws, value for value."""
    rf, cf = tmp_path / "row.txt", tmp_path / "col.txt"
    lib.synthgen_conf(nrow, ncol, seed, rf, cf)
    rows, cols = _restated_conf(nrow, ncol, seed)
    assert rf.read_text() == "".join(f"{i},{v}\n" for i, v in rows)
    assert cf.read_text() == "".join(f"{i},{v}\n" for i, v in cols)


@pytest.mark.parametrize("nrow,ncol,dim,seed", [(40, 12, 3, 12345), (25, 30, 2, 9)])
def test_synthgen_matches_restatement(tmp_path, nrow, ncol, dim, seed):
    """nomad_synthgen's train.dat / test.dat (gen.sh's concatenation of
    header, row sizes, column indexes and values) equal, byte for byte, an
    independent restatement of synthgen.cpp:182-311 on libstdc++'s
    discrete and normal distributions."""
    rf, cf = tmp_path / "row.txt", tmp_path / "col.txt"
    lib.synthgen_conf(nrow, ncol, seed, rf, cf)
    ntr, nte = lib.synthgen(rf, cf, tmp_path, nrow, ncol, dim, seed)
    rows, cols = _restated_conf(nrow, ncol, seed)
    want = _restated_gen(rows, cols, nrow, ncol, dim, seed)
    assert (tmp_path / "train.dat").read_bytes() == want["train"]
    assert (tmp_path / "test.dat").read_bytes() == want["test"]
    tr = lib.read_dat(tmp_path / "train.dat")
    assert (tr.m, tr.n, tr.nnz) == (nrow, ncol, ntr)


def test_synthgen_moments(tmp_path):
    """At 2,000 x 400 with dim 8: row-degree frequencies ~ int(N(20, 5))
    (mean 19.5 after truncation, sd 5), column-degree frequencies follow the
    2.2 power law (P(>= 2) = 2^-1.2, P(>= 10) = 10^-1.2), each row's split is
    80/20, and the values have variance dim + 0.01 (unit-normal factors plus
    0.1 N(0, 1) noise)."""
    nrow, ncol, dim = 2000, 400, 8
    rf, cf = tmp_path / "row.txt", tmp_path / "col.txt"
    lib.synthgen_conf(nrow, ncol, 12345, rf, cf)
    rfreq = np.array([int(l.split(",")[1]) for l in rf.read_text().splitlines()])
    cfreq = np.array([int(l.split(",")[1]) for l in cf.read_text().splitlines()])
    assert abs(rfreq.mean() - 19.5) < 0.6 and abs(rfreq.std() - 5.0) < 0.5
    for x in (2, 10):
        p = (cfreq >= x).mean()
        assert abs(p - x ** -1.2) < 4 * np.sqrt(x ** -1.2 * (1 - x ** -1.2) / cfreq.size), (x, p)
    lib.synthgen(rf, cf, tmp_path, nrow, ncol, dim, 12345)
    raw = (tmp_path / "train.dat").read_bytes()
    nnz = struct.unpack_from("<q", raw, 12)[0]
    sizes = np.frombuffer(raw, np.uint32, nrow, 20)
    raw_te = (tmp_path / "test.dat").read_bytes()
    sizes_te = np.frombuffer(raw_te, np.uint32, nrow, 20)
    deg = sizes + sizes_te
    assert (sizes == (deg * 0.8).astype(np.int64)).all()
    vals = np.frombuffer(raw, np.float64, nnz, 20 + 4 * nrow + 4 * nnz)
    assert abs(vals.var() / (dim + 0.01) - 1) < 0.1 and abs(vals.mean()) < 0.1


def test_convert_tsv_matches_convert_py_procedure(tmp_path):
    """nomad_convert: ids in order of first appearance, shuffled by Python's
    own random.seed(12345) + the Python 2 shuffle (j = int(random() * (i+1))),
    the files written as scipy's csr_matrix would (duplicates summed, columns
    ascending), plus the id lists."""
    import random

    import scipy.sparse as sp

    rng = np.random.default_rng(0)
    users = [f"u{v}" for v in rng.integers(1000, 2000, 60)]
    items = [f"i{v}" for v in rng.integers(10, 90, 60)]
    lines_tr = [f"{users[q]}\t{items[q]}\t{q % 5 + 1}\t{q}\n" for q in range(50)]
    lines_tr.append(f"{users[3]}\t{items[3]}\t2.5\t0\n")  # a duplicate (user, item): summed
    lines_te = [f"{users[q]}\t{items[q]}\t{(q % 3) + 0.5}\n" for q in range(45, 60)]
    (tmp_path / "tr.tsv").write_text("".join(lines_tr))
    (tmp_path / "te.tsv").write_text("".join(lines_te))
    out = tmp_path / "out"
    out.mkdir()
    nu, ni = lib.convert_tsv(tmp_path / "tr.tsv", tmp_path / "te.tsv", out)
    uo, io = [], []
    for ln in lines_tr + lines_te:
        u, i = ln.split("\t")[:2]
        if u not in uo:
            uo.append(u)
        if i not in io:
            io.append(i)
    r = random.Random(12345)
    for v in (uo, io):
        for i in reversed(range(1, len(v))):
            j = int(r.random() * (i + 1))
            v[i], v[j] = v[j], v[i]
    assert (nu, ni) == (len(uo), len(io))
    assert (out / "user_ids.txt").read_text() == "".join(f"{u}\n" for u in uo)
    assert (out / "item_ids.txt").read_text() == "".join(f"{i}\n" for i in io)
    ui = {u: q for q, u in enumerate(uo)}
    ii = {i: q for q, i in enumerate(io)}
    for name, lines in (("train.dat", lines_tr), ("test.dat", lines_te)):
        t = [ln.split("\t") for ln in lines]
        m = sp.csr_matrix(([float(x[2]) for x in t], ([ui[x[0]] for x in t], [ii[x[1]] for x in t])),
                          shape=(len(uo), len(io)))
        sizes = m.indptr[1:] - m.indptr[:-1]
        want = (struct.pack("=iiii", 1211216, len(uo), len(io), m.getnnz()) + struct.pack(f"={len(sizes)}i", *sizes)
                + struct.pack(f"={len(m.indices)}i", *m.indices) + struct.pack(f"={len(m.data)}d", *m.data))
        assert (out / name).read_bytes() == want, name


def test_tool_executables(tmp_path):
    """The command-line front ends run under their reference names."""
    b = ROOT / "nomad_amd" / "bin"
    r = subprocess.run([str(b / "nomad_gen"), str(tmp_path / "g"), "60", "20", "4", "5"], capture_output=True,
                       text=True)
    assert r.returncode == 0 and "success" in r.stdout, r.stderr
    a = lib.read_dat(tmp_path / "g" / "train.dat")
    assert (a.m, a.n) == (60, 20) and a.nnz > 0
    r = subprocess.run([str(b / "nomad_synthgen_conf"), "--nrow", "10", "--ncol", "5", "--rowfile",
                        str(tmp_path / "r.txt"), "--colfile", str(tmp_path / "c.txt")], capture_output=True, text=True)
    assert r.returncode == 0 and (tmp_path / "r.txt").exists(), r.stderr
    r = subprocess.run([str(b / "nomad_convert")], capture_output=True, text=True)
    assert r.returncode == 1 and "usage" in r.stderr


# ---- checkpoint files (Code/nomad/nb_checkpoint.cpp layout) --------------------
def test_checkpoint_files_layout_and_round_trip(tmp_path):
    """nomad_tools_write_checkpoint writes the reference's files: per part
    p of ceil(m / P) rows a .state file (int32 rows, int32 dim, the rows as
    doubles, archive_local) and a .msg file of packed column records (int32
    column, int64 flag, dim doubles, ColumnData::serialize) with the columns
    j = p (mod P); an empty .smsg for part 0.  Reading them back restores W
    and H exactly."""
    m, n, k, P = 10, 7, 4, 3
    rng = np.random.default_rng(2)
    W, H = rng.normal(size=(m, k)), rng.normal(size=(n, k))
    lib.write_checkpoint(tmp_path, 5, 2, P, W, H)
    per = -(-m // P)
    for p in range(P):
        base = tmp_path / f"cp-t5-e2-p{p}"
        rows = W[p * per:min((p + 1) * per, m)]
        assert (base.with_suffix(".state")).read_bytes() == struct.pack("<ii", rows.shape[0], k) + rows.tobytes()
        want = b"".join(struct.pack("<iq", j, 0) + H[j].tobytes() for j in range(p, n, P))
        assert (base.with_suffix(".msg")).read_bytes() == want
    assert (tmp_path / "cp-t5-e2-p0.smsg").read_bytes() == b""
    W2, H2 = lib.read_checkpoint(tmp_path, 5, 2, P, m, n, k)
    assert np.array_equal(W2, W) and np.array_equal(H2, H)


def test_checkpoint_reader_rejects_bad_files(tmp_path):
    m, n, k = 6, 5, 3
    W, H = np.ones((m, k)), np.ones((n, k))
    lib.write_checkpoint(tmp_path, 1, 0, 2, W, H)
    with pytest.raises(lib.NomadError, match="rows of dimension"):
        lib.read_checkpoint(tmp_path, 1, 0, 2, m, n, k + 1)
    msg = tmp_path / "cp-t1-e0-p1.msg"
    raw = msg.read_bytes()
    msg.write_bytes(raw[:-5])  # a truncated record
    with pytest.raises(lib.NomadError, match="truncated"):
        lib.read_checkpoint(tmp_path, 1, 0, 2, m, n, k)
    rec = 12 + 8 * k
    msg.write_bytes(raw[:len(raw) - rec])  # a column missing
    with pytest.raises(lib.NomadError, match="no record of column"):
        lib.read_checkpoint(tmp_path, 1, 0, 2, m, n, k)


@pytest.mark.skipif(not DRIVER.exists(), reason="driver not built")
def test_driver_checkpoint_options(tmp_path):
    r = run_driver("-i", str(tmp_path), "--cptype", "sync")
    assert r.returncode != 0 and "checkpoint path has to be specified" in r.stderr
    r = run_driver("-i", str(tmp_path), "--cptype", "often", "--cppath", str(tmp_path))
    assert r.returncode != 0 and "checkpoint type" in r.stderr
