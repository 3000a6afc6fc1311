"""ctypes bindings of libnomad_hip.so (include/nomad_hip.h) and
libnomad_tools.so (include/nomad_tools.h).

The Python layer is plumbing: every computation runs in the native libraries.
If libnomad_hip.so is missing the import of `hip()` raises; there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIBDIR = Path(__file__).resolve().parent / "lib"

STORE_FP32, STORE_BF16, STORE_FP64 = 0, 1, 2
MODE_ASYNC, MODE_STRAT = 0, 1
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
FLAG_BALANCE = 1
FLAG_TIMING = 2
STORAGE_NAMES = {"fp32": STORE_FP32, "bf16": STORE_BF16, "fp64": STORE_FP64}


class NomadError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nomad error {code}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32),
        ("storage", C.c_int32), ("mode", C.c_int32), ("rounding", C.c_int32),
        ("nworkers", C.c_int32), ("ptoken", C.c_int32), ("device", C.c_int32), ("flags", C.c_int32),
        ("lrate", C.c_double), ("drate", C.c_double), ("reg", C.c_double),
        ("seed", C.c_uint64),
    ]


class TrainStats(C.Structure):
    _fields_ = [
        ("updates", C.c_int64), ("visits", C.c_int64), ("failed_pops", C.c_int64),
        ("epochs_done", C.c_int64), ("train_seconds", C.c_double), ("alg_bytes", C.c_double),
    ]


class EvalStats(C.Structure):
    _fields_ = [
        ("train_sse", C.c_double), ("train_count", C.c_int64),
        ("test_sse", C.c_double), ("test_count", C.c_int64), ("seconds", C.c_double),
    ]


class CcdParams(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32), ("iters", C.c_int32), ("inner", C.c_int32),
        ("device", C.c_int32), ("lam", C.c_double), ("eps", C.c_double),
    ]


class CcdStats(C.Structure):
    _fields_ = [
        ("ranks_done", C.c_int64), ("inner_done", C.c_int64), ("seconds", C.c_double), ("alg_bytes", C.c_double),
        ("solve_seconds", C.c_double),
    ]


class ObjParams(C.Structure):
    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("k", C.c_int32), ("storage", C.c_int32), ("device", C.c_int32)]


class TuneTrial(C.Structure):
    _fields_ = [("nworkers", C.c_int32), ("flags", C.c_int32), ("status", C.c_int32), ("pad", C.c_int32),
                ("updates_per_s", C.c_double)]


class BatchProblem(C.Structure):
    _fields_ = [("params", Params), ("colptr", C.POINTER(C.c_int64)), ("rowidx", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double)), ("nnz", C.c_int64)]


class SynthSpec(C.Structure):
    _fields_ = [
        ("m", C.c_int32), ("n", C.c_int32), ("nnz_train", C.c_int64), ("nnz_test", C.c_int64),
        ("rank", C.c_int32), ("degree", C.c_int32), ("rating", C.c_int32), ("threads", C.c_int32),
        ("factor_var", C.c_double), ("noise_std", C.c_double), ("row_exp", C.c_double),
        ("col_exp", C.c_double), ("center", C.c_double), ("scale", C.c_double),
        ("lo", C.c_double), ("hi", C.c_double), ("seed", C.c_uint64),
    ]


_P = C.c_void_p
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)

# symbol -> (restype, argtypes); the header declarations, one for one
HIP_SYMBOLS = {
    "nomad_create": (C.c_int, [C.POINTER(Params), C.POINTER(_P)]),
    "nomad_destroy": (C.c_int, [_P]),
    "nomad_last_error": (C.c_char_p, []),
    "nomad_abi_version": (C.c_int, []),
    "nomad_upload_train": (C.c_int, [_P, _i64p, _i32p, _f64p, C.c_int64]),
    "nomad_upload_test": (C.c_int, [_P, _i64p, _i32p, _f64p, C.c_int64]),
    "nomad_set_factors": (C.c_int, [_P, _f64p, _f64p]),
    "nomad_get_factors": (C.c_int, [_P, _f64p, _f64p]),
    "nomad_train": (C.c_int, [_P, C.c_double, C.c_int64, C.POINTER(TrainStats)]),
    "nomad_eval": (C.c_int, [_P, C.POINTER(EvalStats)]),
    "nomad_eval_noalloc": (C.c_int, [_P, C.POINTER(EvalStats)]),
    "nomad_get_nworkers": (C.c_int, [_P, _i32p]),
    "nomad_row_blocks": (C.c_int, [_i32p, C.c_int64, C.c_int32, C.c_int32, _f64p, _i32p]),
    "nomad_get_counters": (C.c_int, [_P, _i64p, C.c_int32]),
    "nomad_get_route_state": (C.c_int, [_P, _i64p, _i32p, _i32p, _i32p, _i32p, _i32p, _i32p, _i64p]),
    "nomad_set_visit_counts": (C.c_int, [_P, _i32p, C.c_int64]),
    "nomad_get_progress": (C.c_int, [_P, _i64p]),
    "nomad_set_progress": (C.c_int, [_P, C.c_int64]),
    "nomad_train_batch": (C.c_int, [C.POINTER(_P), C.c_int32, C.c_int64, C.c_int32, C.POINTER(TrainStats)]),
    "nomad_autotune": (C.c_int, [C.POINTER(Params), _i64p, _i32p, _f64p, C.c_int64, C.c_double, C.POINTER(TuneTrial),
                                 C.POINTER(TuneTrial), C.c_int32, _i32p]),
    "nomad_autotune_batch": (C.c_int, [C.POINTER(BatchProblem), C.c_int32, _i32p, _i32p, C.POINTER(C.c_double)]),
    "nomad_obj_create": (C.c_int, [C.POINTER(ObjParams), _i64p, _i32p, _f64p, C.c_int64, C.POINTER(_P)]),
    "nomad_obj_eval": (C.c_int, [_P, _P, _P, C.c_double, _P, _P, C.POINTER(C.c_double), _P]),
    "nomad_obj_destroy": (C.c_int, [_P]),
    "nomad_obj_eval_host": (C.c_int, [_P, _f64p, _f64p, C.c_double, _f64p, _f64p, C.POINTER(C.c_double)]),
    "nomad_ccd": (C.c_int, [C.POINTER(CcdParams), _i64p, _i32p, _f64p, C.c_int64, _i64p, _i32p, _f64p, C.c_int64,
                            _f64p, _f64p, _f64p, _f64p, C.POINTER(CcdStats)]),
}

TOOLS_SYMBOLS = {
    "nomad_tools_last_error": (C.c_char_p, []),
    "nomad_tools_read_dat": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "nomad_tools_read_dat_header": (C.c_int, [C.c_char_p, _i32p, _i32p, _i64p]),
    "nomad_tools_write_dat": (C.c_int, [C.c_char_p, _P, C.c_int32]),
    "nomad_tools_matrix_from_csc": (C.c_int, [C.c_int32, C.c_int32, _i64p, _i32p, _f64p, C.POINTER(_P)]),
    "nomad_tools_matrix_info": (C.c_int, [_P, _i32p, _i32p, _i64p]),
    "nomad_tools_matrix_colptr": (_i64p, [_P]),
    "nomad_tools_matrix_rowidx": (_i32p, [_P]),
    "nomad_tools_matrix_val": (_f64p, [_P]),
    "nomad_tools_matrix_free": (None, [_P]),
    "nomad_tools_synth": (C.c_int, [C.POINTER(SynthSpec), C.POINTER(_P), C.POINTER(_P)]),
    "nomad_tools_preset": (C.c_int, [C.c_int32, C.c_double, C.POINTER(SynthSpec)]),
    "nomad_tools_preset_batch": (C.c_int, [C.c_int32, C.c_double, C.POINTER(SynthSpec), _i32p, _f64p]),
    "nomad_tools_init_factors": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                           _f64p, _f64p]),
    "nomad_tools_convert_tsv": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64, _i64p, _i64p]),
    "nomad_tools_synthgen_conf": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_char_p, C.c_char_p, C.c_char_p]),
    "nomad_tools_synthgen": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, _i64p, _i64p]),
    "nomad_tools_write_checkpoint": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                               C.c_int32, _f64p, _f64p]),
    "nomad_tools_read_checkpoint": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                              C.c_int32, _f64p, _f64p]),
}

_libs: dict[str, C.CDLL] = {}


def _load(name: str, symbols: dict) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = LIBDIR / name
    if not path.exists():
        raise ImportError(f"{path} is not built; run `python -m nomad_amd.build` "
                          "(or __graft_entry__.build())")
    lib = C.CDLL(str(path), mode=C.RTLD_GLOBAL)
    for sym, (res, args) in symbols.items():
        fn = getattr(lib, sym)
        fn.restype = res
        fn.argtypes = args
    _libs[name] = lib
    return lib


def hip() -> C.CDLL:
    return _load("libnomad_hip.so", HIP_SYMBOLS)


def tools() -> C.CDLL:
    return _load("libnomad_tools.so", TOOLS_SYMBOLS)


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# ---------------------------------------------------------------------------
@dataclass
class Csc:
    """Global CSC of a rating matrix, rows ascending inside each column."""
    m: int
    n: int
    colptr: np.ndarray  # int64[n+1]
    rowidx: np.ndarray  # int32[nnz]
    val: np.ndarray     # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.colptr[-1])

    @staticmethod
    def from_triplets(m: int, n: int, rows, cols, vals) -> "Csc":
        rows = np.asarray(rows, np.int64)
        cols = np.asarray(cols, np.int64)
        vals = np.asarray(vals, np.float64)
        order = np.lexsort((rows, cols))
        colptr = np.zeros(n + 1, np.int64)
        np.add.at(colptr, cols + 1, 1)
        return Csc(m, n, np.cumsum(colptr), rows[order].astype(np.int32), vals[order].copy())

    def to_triplets(self):
        cols = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.colptr))
        return self.rowidx.astype(np.int64), cols, self.val


def _tools_check(rc: int) -> None:
    if rc != 0:
        raise NomadError(rc, tools().nomad_tools_last_error().decode())


def _matrix_to_csc(h) -> Csc:
    L = tools()
    m, n, nnz = C.c_int32(), C.c_int32(), C.c_int64()
    _tools_check(L.nomad_tools_matrix_info(h, C.byref(m), C.byref(n), C.byref(nnz)))
    colptr = np.ctypeslib.as_array(L.nomad_tools_matrix_colptr(h), (n.value + 1,)).copy()
    if nnz.value:
        rowidx = np.ctypeslib.as_array(L.nomad_tools_matrix_rowidx(h), (nnz.value,)).copy()
        val = np.ctypeslib.as_array(L.nomad_tools_matrix_val(h), (nnz.value,)).copy()
    else:
        rowidx = np.zeros(0, np.int32)
        val = np.zeros(0, np.float64)
    return Csc(m.value, n.value, colptr, rowidx, val)


def read_dat(path: str | os.PathLike) -> Csc:
    h = C.c_void_p()
    _tools_check(tools().nomad_tools_read_dat(str(path).encode(), C.byref(h)))
    try:
        return _matrix_to_csc(h)
    finally:
        tools().nomad_tools_matrix_free(h)


def write_dat(path: str | os.PathLike, a: Csc, long_format: bool = False) -> None:
    L = tools()
    h = C.c_void_p()
    colptr = np.ascontiguousarray(a.colptr, np.int64)
    rowidx = np.ascontiguousarray(a.rowidx, np.int32)
    val = np.ascontiguousarray(a.val, np.float64)
    _tools_check(L.nomad_tools_matrix_from_csc(a.m, a.n, _ptr(colptr, C.c_int64), _ptr(rowidx, C.c_int32),
                                               _ptr(val, C.c_double), C.byref(h)))
    try:
        _tools_check(L.nomad_tools_write_dat(str(path).encode(), h, int(long_format)))
    finally:
        L.nomad_tools_matrix_free(h)


def convert_tsv(train_path, test_path, out_dir, seed: int = 12345) -> tuple[int, int]:
    """Scripts/convert.py's conversion (nomad_tools_convert_tsv): returns (users, items)."""
    nu, ni = C.c_int64(), C.c_int64()
    _tools_check(tools().nomad_tools_convert_tsv(str(train_path).encode(), str(test_path).encode(),
                                                 str(out_dir).encode(), seed, C.byref(nu), C.byref(ni)))
    return nu.value, ni.value


def synthgen_conf(nrow: int, ncol: int, seed: int, row_path, col_path, sep: str = ",") -> None:
    _tools_check(tools().nomad_tools_synthgen_conf(nrow, ncol, seed, str(row_path).encode(), str(col_path).encode(),
                                                   sep.encode()))


def synthgen(row_path, col_path, out_dir, nrow: int, ncol: int, dim: int, seed: int) -> tuple[int, int]:
    a, b = C.c_int64(), C.c_int64()
    _tools_check(tools().nomad_tools_synthgen(str(row_path).encode(), str(col_path).encode(), str(out_dir).encode(),
                                              nrow, ncol, dim, seed, C.byref(a), C.byref(b)))
    return a.value, b.value


def write_checkpoint(dir_path, job_id: int, epoch: int, nparts: int, W, H) -> None:
    """W and H in the reference's checkpoint files (nomad_tools_write_checkpoint)."""
    W = np.ascontiguousarray(W, np.float64)
    H = np.ascontiguousarray(H, np.float64)
    _tools_check(tools().nomad_tools_write_checkpoint(str(dir_path).encode(), job_id, epoch, nparts, W.shape[0],
                                                      H.shape[0], W.shape[1], _ptr(W, C.c_double),
                                                      _ptr(H, C.c_double)))


def read_checkpoint(dir_path, job_id: int, epoch: int, nparts: int, m: int, n: int, k: int):
    W = np.zeros((m, k))
    H = np.zeros((n, k))
    _tools_check(tools().nomad_tools_read_checkpoint(str(dir_path).encode(), job_id, epoch, nparts, m, n, k,
                                                     _ptr(W, C.c_double), _ptr(H, C.c_double)))
    return W, H


def preset(config: int, scale: float = 1.0) -> SynthSpec:
    s = SynthSpec()
    _tools_check(tools().nomad_tools_preset(config, scale, C.byref(s)))
    return s


def preset_batch(p: int, scale: float = 1.0):
    s = SynthSpec()
    k = C.c_int32()
    reg = C.c_double()
    _tools_check(tools().nomad_tools_preset_batch(p, scale, C.byref(s), C.byref(k), C.byref(reg)))
    return s, k.value, reg.value


def synth(spec: SynthSpec) -> tuple[Csc, Csc]:
    L = tools()
    a, b = C.c_void_p(), C.c_void_p()
    _tools_check(L.nomad_tools_synth(C.byref(spec), C.byref(a), C.byref(b)))
    try:
        return _matrix_to_csc(a), _matrix_to_csc(b)
    finally:
        L.nomad_tools_matrix_free(a)
        L.nomad_tools_matrix_free(b)


def row_blocks(rowidx, m: int, P: int, cap=None) -> np.ndarray:
    """First rows of the P balanced row blocks (NOMAD_FLAG_BALANCE), plus m;
    cap: optional per-block capacities in (0, 1].  Host only."""
    r = np.ascontiguousarray(rowidx, np.int32)
    out = np.zeros(P + 1, np.int32)
    cp = None if cap is None else np.ascontiguousarray(cap, np.float64)
    rc = hip().nomad_row_blocks(_ptr(r, C.c_int32), r.size, m, P, None if cp is None else _ptr(cp, C.c_double),
                                _ptr(out, C.c_int32))
    if rc != 0:
        raise NomadError(rc, hip().nomad_last_error().decode())
    return out


def init_factors(m: int, n: int, k: int, nthreads: int, seed: int):
    W = np.zeros((m, k), np.float64)
    H = np.zeros((n, k), np.float64)
    _tools_check(tools().nomad_tools_init_factors(m, n, k, nthreads, seed, _ptr(W, C.c_double),
                                                  _ptr(H, C.c_double)))
    return W, H


def ccd(train: Csc, W, H, lam: float, iters: int, inner: int, eps: float = 1e-3, test: Csc | None = None,
        device: int = 0):
    """CCD++ rank-one solver on the GPU (nomad_ccd): returns W, H, obj[iters*k],
    rmse[iters*k] (NaN where skipped; rmse all NaN without a test set), stats."""
    h = hip()
    W = np.array(W, np.float64, copy=True, order="C")
    H = np.array(H, np.float64, copy=True, order="C")
    k = W.shape[1]
    p = CcdParams(train.m, train.n, k, iters, inner, device, lam, eps)
    obj = np.zeros(iters * k)
    rmse = np.zeros(iters * k)
    st = CcdStats()
    cp, ri, va = (np.ascontiguousarray(train.colptr, np.int64), np.ascontiguousarray(train.rowidx, np.int32),
                  np.ascontiguousarray(train.val, np.float64))
    if test is not None:
        tcp, tri, tva = (np.ascontiguousarray(test.colptr, np.int64), np.ascontiguousarray(test.rowidx, np.int32),
                         np.ascontiguousarray(test.val, np.float64))
        targs = (_ptr(tcp, C.c_int64), _ptr(tri, C.c_int32), _ptr(tva, C.c_double), int(tcp[-1]))
    else:
        targs = (None, None, None, 0)
    rc = h.nomad_ccd(C.byref(p), _ptr(cp, C.c_int64), _ptr(ri, C.c_int32), _ptr(va, C.c_double), int(cp[-1]),
                     *targs, _ptr(W, C.c_double), _ptr(H, C.c_double), _ptr(obj, C.c_double),
                     _ptr(rmse, C.c_double), C.byref(st))
    if rc != 0:
        raise NomadError(rc, h.nomad_last_error().decode())
    return W, H, obj, rmse, st


class Objective:
    """The factorisation objective and its gradient on the GPU
    (nomad_obj_create / nomad_obj_eval[_host]).  eval_host() takes and
    returns host arrays; eval() takes device addresses of m x ldf / n x ldf
    arrays allocated through this library's own HIP runtime (torch wheels
    bundle a different one: hand torch tensors over with eval_host)."""

    def __init__(self, train: Csc, k: int, storage: str = "fp64", device: int = 0):
        self._lib = hip()
        self.m, self.n, self.k = train.m, train.n, k
        self.ldf = (k + 1) & ~1
        self.storage = storage
        p = ObjParams(train.m, train.n, k, STORAGE_NAMES[storage], device)
        cp, ri, va = (np.ascontiguousarray(train.colptr, np.int64), np.ascontiguousarray(train.rowidx, np.int32),
                      np.ascontiguousarray(train.val, np.float64))
        self._h = C.c_void_p()
        rc = self._lib.nomad_obj_create(C.byref(p), _ptr(cp, C.c_int64), _ptr(ri, C.c_int32), _ptr(va, C.c_double),
                                        int(cp[-1]), C.byref(self._h))
        if rc != 0:
            raise NomadError(rc, self._lib.nomad_last_error().decode())

    def eval(self, W_ptr: int, H_ptr: int, reg: float, gW_ptr: int, gH_ptr: int, stream: int = 0) -> float:
        loss = C.c_double()
        rc = self._lib.nomad_obj_eval(self._h, C.c_void_p(W_ptr), C.c_void_p(H_ptr), reg, C.c_void_p(gW_ptr),
                                      C.c_void_p(gH_ptr), C.byref(loss), C.c_void_p(stream or None))
        if rc != 0:
            raise NomadError(rc, self._lib.nomad_last_error().decode())
        return loss.value

    def eval_host(self, W, H, reg: float):
        """Loss and gradients for host arrays W (m x k), H (n x k)."""
        W = np.ascontiguousarray(W, np.float64)
        H = np.ascontiguousarray(H, np.float64)
        gW = np.zeros_like(W)
        gH = np.zeros_like(H)
        loss = C.c_double()
        rc = self._lib.nomad_obj_eval_host(self._h, _ptr(W, C.c_double), _ptr(H, C.c_double), reg,
                                           _ptr(gW, C.c_double), _ptr(gH, C.c_double), C.byref(loss))
        if rc != 0:
            raise NomadError(rc, self._lib.nomad_last_error().decode())
        return loss.value, gW, gH

    def close(self) -> None:
        if self._h:
            self._lib.nomad_obj_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
class Context:
    """One factorisation on one GPU, driven through the C ABI."""

    def __init__(self, m: int, n: int, k: int, *, storage: str | int = "fp32", mode: str | int = "async",
                 rounding: str | int = "nearest", nworkers: int = 0, ptoken: int = 0, device: int = 0,
                 lrate: float = 0.001, drate: float = 0.1, reg: float = 1.0, seed: int = 12345,
                 balance: bool = False, timing: bool = False):
        self._lib = hip()
        p = Params()
        p.m, p.n, p.k = m, n, k
        p.storage = STORAGE_NAMES[storage] if isinstance(storage, str) else storage
        p.mode = {"async": MODE_ASYNC, "strat": MODE_STRAT}[mode] if isinstance(mode, str) else mode
        p.rounding = ({"nearest": ROUND_NEAREST, "stochastic": ROUND_STOCHASTIC}[rounding]
                      if isinstance(rounding, str) else rounding)
        p.nworkers, p.ptoken, p.device = nworkers, ptoken, device
        p.flags = (FLAG_BALANCE if balance else 0) | (FLAG_TIMING if timing else 0)
        p.lrate, p.drate, p.reg, p.seed = lrate, drate, reg, seed
        self.params = p
        self._h = C.c_void_p()
        self._check(self._lib.nomad_create(C.byref(p), C.byref(self._h)))
        self.m, self.n, self.k = m, n, k

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise NomadError(rc, self._lib.nomad_last_error().decode())

    def close(self) -> None:
        if self._h:
            self._lib.nomad_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nworkers(self) -> int:
        v = C.c_int32()
        self._check(self._lib.nomad_get_nworkers(self._h, C.byref(v)))
        return v.value

    COUNTER_NAMES = ["updates", "visits", "idle_polls", "claims", "claimed", "exits", "resident_sum", "blocked",
                     "ticks_visit", "ticks_exit", "ticks_claim", "ticks_chain", "ticks_transit", "ticks_inbox",
                     "depth_sum", "work_items"]

    def counters(self) -> dict:
        """Device counters of the last async train call (totals)."""
        out = np.zeros(16, np.int64)
        self._check(self._lib.nomad_get_counters(self._h, _ptr(out, C.c_int64), 16))
        return dict(zip(self.COUNTER_NAMES, out.tolist()))

    def unit_counters(self, units: int) -> np.ndarray:
        """The same counters per worker (wave): array (units, 16)."""
        out = np.zeros(16 * (units + 1), np.int64)
        self._check(self._lib.nomad_get_counters(self._h, _ptr(out, C.c_int64), out.size))
        return out[16:].reshape(units, 16)

    def route_state(self) -> dict:
        """Route state of an async context (nomad_get_route_state): per
        (column, workgroup) pair in column-then-workgroup order its workgroup,
        ratings, visit count and position in the current route; per column
        the position of its next hop and its completed route cycles."""
        V = C.c_int64()
        self._check(self._lib.nomad_get_route_state(self._h, None, None, None, None, None, None, None, C.byref(V)))
        V = V.value
        out = dict(vptr=np.zeros(self.n + 1, np.int64), unit=np.zeros(V, np.int32),
                   nratings=np.zeros(V, np.int32), count=np.zeros(V, np.int32), route_pos=np.zeros(V, np.int32),
                   tok_pos=np.zeros(self.n, np.int32), tok_cycle=np.zeros(self.n, np.int32))
        self._check(self._lib.nomad_get_route_state(
            self._h, _ptr(out["vptr"], C.c_int64), _ptr(out["unit"], C.c_int32), _ptr(out["nratings"], C.c_int32),
            _ptr(out["count"], C.c_int32), _ptr(out["route_pos"], C.c_int32), _ptr(out["tok_pos"], C.c_int32),
            _ptr(out["tok_cycle"], C.c_int32), None))
        return out

    @property
    def progress(self) -> int:
        v = C.c_int64()
        self._check(self._lib.nomad_get_progress(self._h, C.byref(v)))
        return v.value

    @progress.setter
    def progress(self, epochs: int) -> None:
        self._check(self._lib.nomad_set_progress(self._h, epochs))

    def set_visit_counts(self, count) -> None:
        c = np.ascontiguousarray(count, np.int32)
        self._check(self._lib.nomad_set_visit_counts(self._h, _ptr(c, C.c_int32), c.size))

    def upload(self, train: Csc, test: Csc | None = None) -> None:
        for a, fn in ((train, self._lib.nomad_upload_train), (test, self._lib.nomad_upload_test)):
            if a is None:
                continue
            colptr = np.ascontiguousarray(a.colptr, np.int64)
            rowidx = np.ascontiguousarray(a.rowidx, np.int32)
            val = np.ascontiguousarray(a.val, np.float64)
            self._check(fn(self._h, _ptr(colptr, C.c_int64), _ptr(rowidx, C.c_int32), _ptr(val, C.c_double),
                           a.nnz))

    def set_factors(self, W: np.ndarray | None, H: np.ndarray | None) -> None:
        Wc = None if W is None else np.ascontiguousarray(W, np.float64)
        Hc = None if H is None else np.ascontiguousarray(H, np.float64)
        # the library reads m*k (n*k) doubles: a wrongly shaped array is refused here
        if Wc is not None and Wc.shape != (self.m, self.k):
            raise ValueError(f"W has shape {Wc.shape}, expected {(self.m, self.k)}")
        if Hc is not None and Hc.shape != (self.n, self.k):
            raise ValueError(f"H has shape {Hc.shape}, expected {(self.n, self.k)}")
        self._check(self._lib.nomad_set_factors(self._h, None if Wc is None else _ptr(Wc, C.c_double),
                                                None if Hc is None else _ptr(Hc, C.c_double)))

    def get_factors(self):
        W = np.zeros((self.m, self.k), np.float64)
        H = np.zeros((self.n, self.k), np.float64)
        self._check(self._lib.nomad_get_factors(self._h, _ptr(W, C.c_double), _ptr(H, C.c_double)))
        return W, H

    def train(self, seconds: float = 0.0, epochs: int = 0) -> TrainStats:
        st = TrainStats()
        self._check(self._lib.nomad_train(self._h, seconds, epochs, C.byref(st)))
        return st

    def eval(self) -> EvalStats:
        st = EvalStats()
        self._check(self._lib.nomad_eval(self._h, C.byref(st)))
        return st

    def eval_noalloc(self) -> EvalStats:
        """nomad_eval_noalloc: the same sums with no device allocation (the
        library's fixed scratch array holds the partials)."""
        st = EvalStats()
        self._check(self._lib.nomad_eval_noalloc(self._h, C.byref(st)))
        return st


def train_batch(ctxs: list[Context], epochs: int, waves: int = 0) -> list[TrainStats]:
    """Runs `epochs` stratified epochs of every context in one persistent
    launch (nomad_train_batch); each context's result equals its own
    Context.train(epochs=...) bit for bit."""
    lib = hip()
    arr = (_P * len(ctxs))(*[c._h for c in ctxs])
    stats = (TrainStats * len(ctxs))()
    rc = lib.nomad_train_batch(arr, len(ctxs), epochs, waves, stats)
    if rc != 0:
        raise NomadError(rc, lib.nomad_last_error().decode())
    return list(stats)


def autotune(train: Csc, k: int, *, mode: str = "async", storage: str = "fp32", rounding: str = "nearest",
             lrate: float = 0.001, drate: float = 0.1, reg: float = 1.0, seed: int = 12345, device: int = 0,
             trial_seconds: float = 0.05):
    """nomad_autotune: returns (best TuneTrial, list of all TuneTrials)."""
    lib = hip()
    p = Params()
    p.m, p.n, p.k = train.m, train.n, k
    p.storage = STORAGE_NAMES[storage]
    p.mode = {"async": MODE_ASYNC, "strat": MODE_STRAT}[mode]
    p.rounding = {"nearest": ROUND_NEAREST, "stochastic": ROUND_STOCHASTIC}[rounding]
    p.device = device
    p.lrate, p.drate, p.reg, p.seed = lrate, drate, reg, seed
    colptr = np.ascontiguousarray(train.colptr, np.int64)
    rowidx = np.ascontiguousarray(train.rowidx, np.int32)
    val = np.ascontiguousarray(train.val, np.float64)
    best = TuneTrial()
    trials = (TuneTrial * 32)()
    nt = C.c_int32()
    rc = lib.nomad_autotune(C.byref(p), _ptr(colptr, C.c_int64), _ptr(rowidx, C.c_int32), _ptr(val, C.c_double),
                            train.nnz, trial_seconds, C.byref(best), trials, 32, C.byref(nt))
    if rc != 0:
        raise NomadError(rc, lib.nomad_last_error().decode())
    return best, list(trials[: min(nt.value, 32)])


def autotune_batch(problems, *, storage: str = "bf16", rounding: str = "stochastic", lrate: float = 0.005,
                   drate: float = 0.05, seed: int = 12345, device: int = 0, balance: bool = False,
                   tune_balance: bool = False):
    """nomad_autotune_batch for [(train Csc, k, reg), ...]: returns (strata per
    problem, balanced row blocks per problem (bools), the batch's updates/s
    with them).  balance: the starting row balance (NOMAD_FLAG_BALANCE) of
    every problem; tune_balance: let the tuner toggle it per problem."""
    lib = hip()
    arr = (BatchProblem * len(problems))()
    keep = []
    for i, (tr, k, reg) in enumerate(problems):
        p = arr[i].params
        p.m, p.n, p.k = tr.m, tr.n, k
        p.storage = STORAGE_NAMES[storage]
        p.mode = MODE_STRAT
        p.rounding = {"nearest": ROUND_NEAREST, "stochastic": ROUND_STOCHASTIC}[rounding]
        p.device = device
        p.lrate, p.drate, p.reg, p.seed = lrate, drate, reg, seed + i
        p.flags = FLAG_BALANCE if balance else 0
        cp = np.ascontiguousarray(tr.colptr, np.int64)
        ri = np.ascontiguousarray(tr.rowidx, np.int32)
        va = np.ascontiguousarray(tr.val, np.float64)
        keep += [cp, ri, va]
        arr[i].colptr, arr[i].rowidx, arr[i].val, arr[i].nnz = (_ptr(cp, C.c_int64), _ptr(ri, C.c_int32),
                                                                _ptr(va, C.c_double), tr.nnz)
    out = np.zeros(len(problems), np.int32)
    flags = np.full(len(problems), FLAG_BALANCE if balance else 0, np.int32)
    rate = C.c_double()
    rc = lib.nomad_autotune_batch(arr, len(problems), _ptr(out, C.c_int32),
                                  _ptr(flags, C.c_int32) if tune_balance else None, C.byref(rate))
    if rc != 0:
        raise NomadError(rc, lib.nomad_last_error().decode())
    return out.tolist(), [bool(f & FLAG_BALANCE) for f in flags.tolist()], rate.value
