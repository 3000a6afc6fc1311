"""ctypes binding of the CPU oracle (liborc.so) -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  See nomad_oracle.cpp for what pins each function.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
FP32, BF16, FP64 = 0, 1, 2
ARITH_REF, ARITH_GPU = 0, 1
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)


class AsyncStats(C.Structure):
    _fields_ = [("updates", C.c_int64), ("failures", C.c_int64), ("seconds", C.c_double),
                ("min_cycles", C.c_int64)]


_SYMS = {
    "orc_dat_part_info": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _i32p, _i32p, _i64p, _i32p, _i32p, _i64p]),
    "orc_dat_part_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]),
    "orc_init_factors": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, _f64p, _f64p]),
    "orc_init_perms": (C.c_int, [C.c_int, C.c_int, C.c_uint64, _i32p]),
    "orc_step": (C.c_double, [C.c_int, C.c_double, C.c_double, C.c_int]),
    "orc_sgd_strat": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                  _i64p, _i32p, _f64p, _f64p, _f64p, C.c_int, _i32p, C.c_int, C.c_int,
                                  C.c_double, C.c_double, C.c_double]),
    "orc_sgd_async": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p,
                                C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int64,
                                C.POINTER(AsyncStats)]),
    "orc_async_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p,
                                      _f64p, C.c_uint64, C.c_double, C.c_double, C.c_double]),
    "orc_async_segment": (C.c_int, [C.c_void_p, C.c_double, C.c_int64, C.c_int64, C.POINTER(AsyncStats)]),
    "orc_async_get": (C.c_int, [C.c_void_p, _f64p, _f64p]),
    "orc_async_counts": (C.c_int, [C.c_void_p, _i32p]),
    "orc_async_destroy": (None, [C.c_void_p]),
    "orc_bf16_store_pair": (None, [C.c_float, C.c_float, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                   C.POINTER(C.c_float)]),
    "orc_sse": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p]),
    "orc_loss_grad": (C.c_double, [C.c_int, C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                                   _f64p, _f64p]),
    "orc_ccd": (C.c_int, [C.c_int, C.c_int, C.c_int, _i64p, _i32p, _f64p, _f64p, _f64p, C.c_double, C.c_int,
                          C.c_int, C.c_double, _f64p, _i64p, _i32p, _f64p, _f64p]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = HERE / "liborc.so"
        if not path.exists():
            raise ImportError(f"{path} not built; run `python oracle/build.py`")
        _lib = C.CDLL(str(path))
        for s, (r, a) in _SYMS.items():
            f = getattr(_lib, s)
            f.restype, f.argtypes = r, a
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _csc(a):
    return (np.ascontiguousarray(a.colptr, np.int64), np.ascontiguousarray(a.rowidx, np.int32),
            np.ascontiguousarray(a.val, np.float64))


def dat_part(path, part, nparts):
    """Part `part` of `nparts` of a .dat file as the reference reads it
    (Code/nomad/nb_data.cpp:19-176): (nrows, ncols, total_nnz, min_row,
    col_offset[ncols+1], local row ids, values)."""
    m, n, mn, lr = (C.c_int() for _ in range(4))
    tot, ln = C.c_int64(), C.c_int64()
    rc = lib().orc_dat_part_info(str(path).encode(), part, nparts, C.byref(m), C.byref(n), C.byref(tot),
                                 C.byref(mn), C.byref(lr), C.byref(ln))
    if rc:
        raise RuntimeError(f"orc_dat_part_info failed ({rc})")
    col = np.zeros(n.value + 1, np.int32)
    rows = np.zeros(ln.value, np.int32)
    vals = np.zeros(ln.value, np.float64)
    rc = lib().orc_dat_part_read(str(path).encode(), part, nparts, _p(col, C.c_int32), _p(rows, C.c_int32),
                                 _p(vals, C.c_double))
    if rc:
        raise RuntimeError(f"orc_dat_part_read failed ({rc})")
    return m.value, n.value, tot.value, mn.value, col, rows, vals


def bf16_store_pair(x0, x1, rounding, seed, a, b, lane):
    out = (C.c_float * 2)()
    lib().orc_bf16_store_pair(x0, x1, rounding, seed, a, b, lane, out)
    return out[0], out[1]


def init_factors(m, n, k, nthreads, seed):
    W = np.zeros((m, k))
    H = np.zeros((n, k))
    lib().orc_init_factors(m, n, k, nthreads, seed, _p(W, C.c_double), _p(H, C.c_double))
    return W, H


def init_perms(n, nthreads, seed):
    P = np.zeros((n, nthreads), np.int32)
    lib().orc_init_perms(n, nthreads, seed, _p(P, C.c_int32))
    return P


def step(storage, lrate, drate, count):
    return lib().orc_step(storage, lrate, drate, count)


def sgd_strat(a, W, H, *, storage=FP32, arith=ARITH_REF, rounding=ROUND_NEAREST, sr_seed=0, B=1, epoch0=0,
              epochs=1, lrate=0.001, drate=0.1, reg=0.05, row_bounds=None):
    """row_bounds: B + 1 first rows of the row blocks (None: ceil(m / B) rows each)."""
    colptr, rowidx, val = _csc(a)
    W = np.array(W, np.float64, copy=True, order="C")
    H = np.array(H, np.float64, copy=True, order="C")
    k = W.shape[1]
    rb = None if row_bounds is None else np.ascontiguousarray(row_bounds, np.int32)
    if rb is not None and rb.shape != (B + 1,):
        raise ValueError("row_bounds needs B + 1 entries")
    r = lib().orc_sgd_strat(storage, arith, rounding, sr_seed, a.m, a.n, k, _p(colptr, C.c_int64),
                            _p(rowidx, C.c_int32), _p(val, C.c_double), _p(W, C.c_double), _p(H, C.c_double),
                            B, None if rb is None else _p(rb, C.c_int32), epoch0, epochs, lrate, drate, reg)
    if r < 0:
        raise RuntimeError("orc_sgd_strat failed")
    return W, H


def sgd_async(a, W, H, *, storage=FP32, T=4, seed=12345, lrate=0.001, drate=0.1, reg=0.05, seconds=0.0,
              epochs=0):
    colptr, rowidx, val = _csc(a)
    W = np.array(W, np.float64, copy=True, order="C")
    H = np.array(H, np.float64, copy=True, order="C")
    st = AsyncStats()
    rc = lib().orc_sgd_async(storage, T, a.m, a.n, W.shape[1], _p(colptr, C.c_int64), _p(rowidx, C.c_int32),
                             _p(val, C.c_double), _p(W, C.c_double), _p(H, C.c_double), seed, lrate, drate, reg,
                             seconds, epochs, C.byref(st))
    if rc != 0:
        raise RuntimeError("orc_sgd_async failed")
    return W, H, st


class AsyncRun:
    """The async restatement kept across timeout segments (orc_async_*):
    segment() runs one segment (seconds, an update budget or epochs) and
    re-permutes every column's route afterwards."""

    def __init__(self, a, W, H, *, storage=FP32, T=4, seed=12345, lrate=0.001, drate=0.1, reg=0.05):
        colptr, rowidx, val = _csc(a)
        W = np.ascontiguousarray(W, np.float64)
        H = np.ascontiguousarray(H, np.float64)
        self.m, self.n, self.k, self.T = a.m, a.n, W.shape[1], T
        self._h = lib().orc_async_create(storage, T, a.m, a.n, self.k, _p(colptr, C.c_int64), _p(rowidx, C.c_int32),
                                         _p(val, C.c_double), _p(W, C.c_double), _p(H, C.c_double), seed, lrate,
                                         drate, reg)
        if not self._h:
            raise RuntimeError("orc_async_create failed")

    def segment(self, seconds=0.0, max_updates=0, epochs=0):
        st = AsyncStats()
        if lib().orc_async_segment(self._h, seconds, max_updates, epochs, C.byref(st)) != 0:
            raise RuntimeError("orc_async_segment failed")
        return st

    def factors(self):
        W = np.zeros((self.m, self.k))
        H = np.zeros((self.n, self.k))
        lib().orc_async_get(self._h, _p(W, C.c_double), _p(H, C.c_double))
        return W, H

    def counts(self):
        c = np.zeros((self.T, self.n), np.int32)
        lib().orc_async_counts(self._h, _p(c, C.c_int32))
        return c

    def close(self):
        if self._h:
            lib().orc_async_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def sse(a, W, H, *, storage=FP64, arith=ARITH_REF):
    colptr, rowidx, val = _csc(a)
    W = np.ascontiguousarray(W, np.float64)
    H = np.ascontiguousarray(H, np.float64)
    return lib().orc_sse(storage, arith, a.m, a.n, W.shape[1], _p(colptr, C.c_int64), _p(rowidx, C.c_int32),
                         _p(val, C.c_double), _p(W, C.c_double), _p(H, C.c_double))


def loss_grad(a, W, H, reg, want_grad=True):
    colptr, rowidx, val = _csc(a)
    W = np.ascontiguousarray(W, np.float64)
    H = np.ascontiguousarray(H, np.float64)
    gW = np.zeros_like(W) if want_grad else None
    gH = np.zeros_like(H) if want_grad else None
    L = lib().orc_loss_grad(a.m, a.n, W.shape[1], _p(colptr, C.c_int64), _p(rowidx, C.c_int32),
                            _p(val, C.c_double), _p(W, C.c_double), _p(H, C.c_double), reg,
                            None if gW is None else _p(gW, C.c_double), None if gH is None else _p(gH, C.c_double))
    return L, gW, gH


def ccd(a, W, H, lam, iters, inner, eps=1e-3, test=None):
    """CCD++ restatement; returns W, H, obj[iters*k] and, with a test CSC,
    rmse[iters*k] (NaN where the early stop skipped a rank)."""
    colptr, rowidx, val = _csc(a)
    W = np.array(W, np.float64, copy=True, order="C")
    H = np.array(H, np.float64, copy=True, order="C")
    k = W.shape[1]
    obj = np.zeros(iters * k)
    rmse = np.zeros(iters * k)
    te = _csc(test) if test is not None else None
    rc = lib().orc_ccd(a.m, a.n, k, _p(colptr, C.c_int64), _p(rowidx, C.c_int32), _p(val, C.c_double),
                       _p(W, C.c_double), _p(H, C.c_double), lam, iters, inner, eps, _p(obj, C.c_double),
                       None if te is None else _p(te[0], C.c_int64), None if te is None else _p(te[1], C.c_int32),
                       None if te is None else _p(te[2], C.c_double), None if te is None else _p(rmse, C.c_double))
    if rc != 0:
        raise RuntimeError("orc_ccd failed")
    if test is None:
        return W, H, obj
    return W, H, obj, rmse
