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
