"""Benchmark: NOMAD SGD updates/s and time-to-target test RMSE on one MI355X.

Workload (default): Netflix-shaped synthetic ratings -- 480,189 users x 17,770
items, 99,072,112 train + 1,408,395 test ratings, Zipf row/column degrees,
integer ratings 1-5 from a rank-50 truth -- k = 100, reg 0.05, lrate 0.012,
drate 0.05, factors initialised with the reference's streams.

One step = one training segment of the asynchronous token kernel: the
columns circulate for a fixed device time (--segment seconds) with inputs
resident in HBM, then training stops and the test RMSE is evaluated.  This is
the reference's timeout schedule (Code/nomad/nb_main.cpp:109-134: train until
the next timeout, stop, evaluate), and updates/s is its u= count over training
time, evaluation excluded (nb_main.cpp:265-266, 284).  The FP32-stored path
runs first; its final test RMSE x 1.005 is the target.  Two BF16-stored paths
(FP32 arithmetic) then run the same schedule, each on a fresh context: "bf16"
rounds every stored value to the nearest bf16, "bf16_sr" rounds
stochastically; for each the training time at which it first reaches the
target is reported.  `value` is the "bf16" path's updates/s.  Warmup segments
run on a separate context that is then discarded.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--segment S] [--scale F]

Prints one JSON line.

Other workloads (not the default line): `--workload yahoo` (BF16 vs FP32 vs
FP64 test RMSE after equal epochs at the Yahoo! Music shape, with the
oracle's FP64 async run on the host), `--workload ccd` (CCD++ rank-one
sweeps at the Netflix shape), `--workload objective` (objective and gradient
at 100,000 x 20,000 with 10 M cells, k = 64), and `--workload sweep`: the batched
hyper-parameter sweep -- 32 independent problems of uneven sizes (users ~
logU(10k, 200k), items ~ logU(2k, 50k), ratings ~ logU(0.5M, 20M), k in
{16, 32, 64, 128}, reg in {0.01, 0.05, 0.1, 0.5}), BF16 storage with
stochastic rounding -- with each problem's strata count and row balance
picked by the library's autotuner, then all problems trained in one
persistent launch per
step (one step = one epoch of every problem), beside the same epochs run one
problem after another.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_ctx(lib, tr, te, W0, H0, *, storage, rounding, k, hyp, nworkers):
    ctx = lib.Context(tr.m, tr.n, k, storage=storage, rounding=rounding, mode="async", nworkers=nworkers,
                      seed=12345, balance=True, **hyp)
    ctx.upload(tr, te)
    ctx.set_factors(W0, H0)
    return ctx


def interp_hit(rmse0, times, curve, target):
    """Device time at which the test RMSE first reaches target, linearly
    interpolated between the two evaluations around the crossing (the curve
    starts at rmse0 at time 0); None if it is never reached."""
    if target is None:
        return None
    if rmse0 <= target:
        return 0.0
    t_prev, r_prev = 0.0, rmse0
    for t, r in zip(times, curve):
        if r <= target:
            return round(t_prev + (r_prev - target) / (r_prev - r) * (t - t_prev), 4)
        t_prev, r_prev = t, r
    return None


def grad_norm(obj, ctx, reg):
    """Stationarity of the current iterate: the objective L = sum (w.h - r)^2 +
    reg (sum deg(i) |w_i|^2 + sum deg(j) |h_j|^2) and the Frobenius norm of its
    full gradient, evaluated in FP64 by the library's objective kernels on the
    context's factors (outside the timed region)."""
    W, H = ctx.get_factors()
    L, gW, gH = obj.eval_host(W, H, reg)
    return float(L), float(np.sqrt(np.sum(gW * gW) + np.sum(gH * gH)))


def run_path(lib, tr, te, W0, H0, *, storage, rounding, k, hyp, nworkers, warmup, steps, segment, target=None,
             log_name=None, obj=None):
    name = log_name or storage
    if warmup > 0:
        ctx = make_ctx(lib, tr, te, W0, H0, storage=storage, rounding=rounding, k=k, hyp=hyp, nworkers=nworkers)
        for _ in range(warmup):
            ctx.train(seconds=segment)
        ctx.close()
    ctx = make_ctx(lib, tr, te, W0, H0, storage=storage, rounding=rounding, k=k, hyp=hyp, nworkers=nworkers)
    curve, times, grads, losses = [], [], [], []
    upd = 0
    dev_s = wall_s = alg = 0.0
    visits = hops = 0
    t_hit = None
    ev0 = ctx.eval()
    rmse0 = float(np.sqrt(ev0.test_sse / max(ev0.test_count, 1)))  # before training
    if obj is not None:
        L, g = grad_norm(obj, ctx, hyp["reg"])
        losses.append(L)
        grads.append(g)
    for step in range(steps):
        t0 = time.perf_counter()
        st = ctx.train(seconds=segment)
        wall_s += time.perf_counter() - t0
        dev_s += st.train_seconds
        upd += st.updates
        alg += st.alg_bytes
        visits += st.visits
        hops += ctx.counters()["claimed"]
        ev = ctx.eval()
        rmse = float(np.sqrt(ev.test_sse / max(ev.test_count, 1)))
        curve.append(round(rmse, 6))
        times.append(round(dev_s, 4))
        if target is not None and t_hit is None and rmse <= target:
            t_hit = dev_s
        gtxt = ""
        if obj is not None:
            L, g = grad_norm(obj, ctx, hyp["reg"])
            losses.append(L)
            grads.append(g)
            gtxt = f" L={L:.6g} |grad L|={g:.6g}"
        log(f"{name}: segment {step + 1} updates={st.updates} dev={st.train_seconds * 1e3:.1f} ms "
            f"epochs(min over columns)={st.epochs_done} test RMSE={rmse:.6f}{gtxt}")
    nw = ctx.nworkers
    rs = ctx.route_state()
    ctx.close()
    return dict(updates=upd, dev_s=dev_s, wall_s=wall_s, alg_bytes=alg, visits=visits, hops=hops, curve=curve,
                times=times, rmse0=rmse0, t_hit=t_hit, nworkers=nw, grads=grads, losses=losses,
                visit_count_range=[int(rs["count"].min()), int(rs["count"].max())])


def profiled_traffic(storage: str) -> float | None:
    """HBM (L2-to-fabric) bytes per update of the async kernel measured by the
    two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE, each scaled by a
    plain copy of known size) summarised in profiles/; None if absent."""
    return profile_figure("sgd_async_kernel_netflix_k100_pmc.md", f"traffic_bytes_per_update[{storage}]")


def profile_figure(name: str, key: str) -> float | None:
    """A `key: value` line of profiles/<name> (copy-calibrated counter
    figures written from a rocprofv3 run of the same workload); None if the
    file or the line is absent."""
    import re

    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    m = re.search(rf"^{re.escape(key)}: (\S+)$", f.read_text(), flags=re.M)
    return float(m.group(1)) if m else None


def copy_calibrate(nbytes: int) -> None:
    """One hipMemcpy device-to-device of nbytes, which the HIP runtime runs as
    a blit kernel reading and writing nbytes: under rocprofv3 --pmc its
    FETCH_SIZE / WRITE_SIZE give the counters' scale (KiB per byte moved)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
        raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    log(f"copy calibration: {nbytes} bytes device to device")


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(tr, te, W0, H0, k, hyp, seconds, segment, target):
    """The oracle's restatement of the reference nomad_float async loop
    (nb_main.cpp:415-578, per-thread visit counts, set_perm after every
    timeout) through the same timeout schedule as the GPU paths, with longer
    segments: train a segment, stop, evaluate the test RMSE (excluded from
    the training time, nb_main.cpp:265-266), until `seconds` of training or
    the target RMSE is reached."""
    from oracle import oracle as O

    # the host cores this job may use: the box gives one GPU's job 16
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16)))
    run = O.AsyncRun(tr, W0, H0, storage=O.FP32, T=threads, seed=12345, **hyp)
    train_s, upd = 0.0, 0
    curve, times = [], []
    t_hit = None
    while train_s < seconds - 1e-9:
        st = run.segment(seconds=min(segment, seconds - train_s))
        train_s += st.seconds
        upd += st.updates
        W, H = run.factors()
        rmse = float(np.sqrt(O.sse(te, W, H, storage=O.FP32, arith=O.ARITH_REF) / te.nnz))
        curve.append(round(rmse, 6))
        times.append(round(train_s, 3))
        log(f"cpu: {train_s:.1f} s {upd} updates test RMSE {rmse:.6f}")
        if target is not None and rmse <= target:
            t_hit = train_s
            break
    run.close()
    return dict(value=upd / train_s, unit="updates/s", cores=threads, kind="port", cpu_model=cpu_model(),
                sample=f"CPU restatement of the reference nomad_float async loop (nb_main.cpp:415-578), "
                       f"{threads} threads, {segment:g} s timeout segments on the full training set until "
                       f"{seconds:g} s of training or the target RMSE, {upd} updates",
                train_seconds=times, test_rmse_curve=curve, time_to_target_s=t_hit,
                rmse0=None)


def sweep(a, lib) -> dict:
    """Batched sweep: autotuned strata per problem, one launch for all."""
    hyp = dict(lrate=0.005, drate=0.05)
    probs = []
    t0 = time.perf_counter()
    for p in range(a.problems):
        spec, k, reg = lib.preset_batch(p, a.scale)
        tr, te = lib.synth(spec)
        probs.append((tr, te, k, reg))
    total = sum(x[0].nnz for x in probs)
    log(f"sweep: generated {len(probs)} problems, {total} training ratings in {time.perf_counter() - t0:.1f} s")
    strata, tune_s = [], time.perf_counter()
    fixed = [int(x) for x in a.strata.split(",")] if a.strata else []
    bal = [x == "1" for x in a.balance.split(",")] if a.balance else []
    if fixed:
        strata = [fixed[i % len(fixed)] for i in range(len(probs))]
        balanced = [bal[i % len(bal)] for i in range(len(probs))] if bal else [False] * len(probs)
    else:
        # the batch is tuned as a whole: the problems share the device
        strata, balanced, rate = lib.autotune_batch([(tr, k, reg) for tr, te, k, reg in probs], storage="bf16",
                                                    rounding="stochastic", balance=True, tune_balance=True, **hyp)
        log(f"sweep: autotuned strata {sorted(set(strata))}, {sum(balanced)} of {len(probs)} problems with "
            f"balanced row blocks ({rate / 1e9:.3f} G updates/s in the trial)")
    tune_s = time.perf_counter() - tune_s

    def contexts():
        cs = []
        for i, (tr, te, k, reg) in enumerate(probs):
            c = lib.Context(tr.m, tr.n, k, storage="bf16", rounding="stochastic", mode="strat", nworkers=strata[i],
                            seed=12345 + i, reg=reg, balance=balanced[i], **hyp)
            c.upload(tr, te)
            W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 12345 + i)
            c.set_factors(W0, H0)
            cs.append(c)
        return cs

    batch = contexts()
    for _ in range(a.warmup):
        lib.train_batch(batch, 1)
    dev = alg = 0.0
    upd = 0
    for step in range(a.steps):
        st = lib.train_batch(batch, 1)
        dev += st[0].train_seconds
        upd += sum(x.updates for x in st)
        alg += sum(x.alg_bytes for x in st)
        log(f"sweep: batched epoch {step + 1}: {st[0].train_seconds * 1e3:.2f} ms")
    seq = contexts()
    for c in seq:
        for _ in range(a.warmup):
            c.train(epochs=1)
    seq_dev = 0.0
    for step in range(a.steps):
        seq_dev += sum(c.train(epochs=1).train_seconds for c in seq)
    same = all(np.array_equal(x, y) for cb, cs in zip(batch, seq) for x, y in zip(cb.get_factors(), cs.get_factors()))
    rmse = [float(np.sqrt(c.eval().test_sse / max(pr[1].nnz, 1))) for c, pr in zip(batch, probs)]
    for c in batch + seq:
        c.close()
    achieved = alg / dev / 1e9
    tb = profile_figure("sgd_batch_kernel_sweep32.md", "traffic_bytes_per_update[sweep]")
    traffic = tb * upd / a.steps if tb else None
    return {
        "metric": "batched sweep updates/s, 32 uneven problems, BF16 storage, 1 GPU",
        "value": upd / dev,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dev / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched strat sweep, {len(probs)} problems, one epoch of each per launch",
            "problems": [dict(m=p[0].m, n=p[0].n, nnz=p[0].nnz, k=p[2], reg=p[3], B=b, balanced=bl)
                         for p, b, bl in zip(probs, strata, balanced)],
            "storage": "bf16", "rounding": "stochastic", **hyp, "scale": a.scale,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "sgd_batch_kernel<bf16>",
            "frac_measured_traffic": traffic / (dev / a.steps) / 1e9 / HBM_PEAK_GBS if traffic else None,
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per update (copy-calibrated), "
                              "profiles/sgd_batch_kernel_sweep32.md, x updates per launch",
            "alg_bytes_per_launch": alg / a.steps, "avg_launch_ms": dev / a.steps * 1e3,
        },
        "sequential": {"updates_per_s": upd / seq_dev, "ms_per_step": seq_dev / a.steps * 1e3,
                       "bitwise_equal_to_batched": bool(same)},
        "speedup_vs_sequential": seq_dev / dev,
        "autotune_seconds": tune_s,
        "strata_source": "fixed" if fixed else "autotuner",
        "test_rmse_after": rmse,
    }


def ccd_workload(a, lib) -> dict:
    """CCD++ rank-one sweeps (nomad_ccd) on the Netflix shape: one outer
    iteration over all k ranks, a.ccd_inner inner iterations per rank."""
    k = a.ccd_rank
    tr, te = lib.synth(lib.preset(2, a.scale))
    W0, _ = lib.init_factors(tr.m, tr.n, k, 1, 12345)
    H0 = np.zeros((tr.n, k))
    _, _, obj, rmse, st = lib.ccd(tr, W0, H0, 0.05, 1, a.ccd_inner, eps=1e-3, test=te)
    log(f"ccd: {st.ranks_done} ranks, {st.inner_done} inner iterations in {st.solve_seconds:.3f} s "
        f"({st.seconds:.3f} s with host set-up), final obj {obj[-1]:.6g} test RMSE {rmse[-1]:.6f}")
    gbs = st.alg_bytes / st.solve_seconds / 1e9
    pr = profile_figure("ccd_pass_netflix_k20.md", "traffic_bytes_per_rank[ccd]")
    pi = profile_figure("ccd_pass_netflix_k20.md", "traffic_bytes_per_inner[ccd]")
    traffic = st.ranks_done * pr + st.inner_done * pi if (pr and pi and a.scale == 1.0) else None
    return {"metric": "CCD++ rank-one sweep, Netflix-shape, FP64, 1 GPU", "value": st.inner_done / st.solve_seconds,
            "unit": "inner iterations/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic", "seconds": st.solve_seconds, "seconds_with_setup": st.seconds,
            "config": {"workload": f"ccd++ k={k}, 1 outer iteration, {a.ccd_inner} inner, lambda 0.05",
                       "m": tr.m, "n": tr.n, "nnz_train": tr.nnz},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                         "frac_measured_traffic": traffic / st.solve_seconds / 1e9 / HBM_PEAK_GBS if traffic else None,
                         "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per rank and per inner iteration "
                                           "(copy-calibrated), profiles/ccd_pass_netflix_k20.md",
                         "note": "wall time of the rank / inner-iteration loop (host set-up excluded)"},
            "final_test_rmse": float(rmse[-1])}


def objective_workload(a, lib) -> dict:
    """Objective + full gradient (nomad_obj_eval) at the gradient-check shape
    (100,000 x 20,000, 10 M cells, k = 64), device arrays, a.steps evaluations."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    tr, _ = lib.synth(lib.preset(4, a.scale))
    k = 64
    out = {}
    for storage, esz in (("fp64", 8), ("fp32", 4)):
        ob = lib.Objective(tr, k, storage=storage)
        ldf = (k + 1) & ~1
        rng = np.random.default_rng(1)
        Wh = (rng.random((tr.m, ldf)) / 8).astype(np.float64 if esz == 8 else np.float32)
        Hh = (rng.random((tr.n, ldf)) / 8).astype(Wh.dtype)
        ptrs = []
        for arr in (Wh, Hh, Wh, Hh):
            d = C.c_void_p()
            if hip.hipMalloc(C.byref(d), C.c_size_t(arr.nbytes)) != 0:
                raise RuntimeError("hipMalloc failed")
            hip.hipMemcpy(d, arr.ctypes.data_as(C.c_void_p), C.c_size_t(arr.nbytes), 1)
            ptrs.append(d.value)
        ob.eval(ptrs[0], ptrs[1], 0.05, ptrs[2], ptrs[3])  # warm-up
        t0 = time.perf_counter()
        for _ in range(a.steps):
            L = ob.eval(ptrs[0], ptrs[1], 0.05, ptrs[2], ptrs[3])
        dt = (time.perf_counter() - t0) / a.steps
        for d in ptrs:
            hip.hipFree(C.c_void_p(d))
        ob.close()
        # per rating and pass (columns, then rows): index + value (in the
        # arithmetic type) and the other side's row gathered; per line element:
        # its own factor read once and its gradient written once
        alg = 2.0 * tr.nnz * (4 + esz + k * esz) + 2.0 * (tr.m + tr.n) * k * esz
        out[storage] = {"ms_per_eval": dt * 1e3, "loss": L, "alg_GBps": alg / dt / 1e9}
        log(f"objective {storage}: {dt * 1e3:.3f} ms per evaluation (loss {L:.6g})")
    gbs = out["fp64"]["alg_GBps"]
    tb = profile_figure("obj_pass_100k_20k_10m_k64.md", "traffic_bytes_per_eval[fp64]")
    traffic = tb if (tb and a.scale == 1.0) else None
    return {"metric": "objective + gradient evaluations/s, 100k x 20k x 10M, k=64, FP64, 1 GPU",
            "value": 1e3 / out["fp64"]["ms_per_eval"], "unit": "evaluations/s", "n_gpus": 1,
            "higher_is_better": True, "dtype": "f64", "data": "synthetic", "steps": a.steps,
            "config": {"workload": "objective and full gradient, device arrays", "m": tr.m, "n": tr.n,
                       "nnz": tr.nnz, "k": k},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                         "frac_measured_traffic": (traffic / (out["fp64"]["ms_per_eval"] * 1e-3) / 1e9 /
                                                   HBM_PEAK_GBS) if traffic else None,
                         "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per evaluation (copy-calibrated), "
                                           "profiles/obj_pass_100k_20k_10m_k64.md",
                         "note": "host wall time per evaluation (launches + one loss read back); the gathered "
                                 "rows (W: 51 MB, H: 10 MB in FP64) are served largely by the L2 and the "
                                 "memory-side cache"},
            "paths": out}


def yahoo_workload(a, lib) -> dict:
    """Yahoo! Music shape (1,000,990 x 624,961, 252.8 M training ratings,
    integer ratings 0-100, k = 100, reg 1.0): the async kernel with BF16
    (stochastic rounding), FP32 and FP64 storage for a.epochs epochs each,
    beside the oracle's FP64 async restatement on the host for the same
    epochs; test RMSE of each and its gap to the FP64 oracle."""
    from oracle import oracle as O

    k = 100
    hyp = dict(lrate=0.001, drate=0.05, reg=1.0)
    t0 = time.perf_counter()
    tr, te = lib.synth(lib.preset(3, a.scale))
    log(f"yahoo: generated {tr.m} x {tr.n}, train {tr.nnz}, test {te.nnz} in {time.perf_counter() - t0:.1f} s")
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    paths = {}
    for name, storage, rnd in (("bf16", "bf16", "stochastic"), ("fp32", "fp32", "nearest"),
                               ("fp64", "fp64", "nearest")):
        ctx = lib.Context(tr.m, tr.n, k, storage=storage, rounding=rnd, mode="async", nworkers=a.nworkers,
                          seed=12345, balance=True, **hyp)
        ctx.upload(tr, te)
        ctx.set_factors(W0, H0)
        st = ctx.train(epochs=a.epochs)
        ev = ctx.eval()
        rmse = float(np.sqrt(ev.test_sse / ev.test_count))
        paths[name] = {"test_rmse": rmse, "updates_per_s": st.updates / st.train_seconds,
                       "train_seconds": st.train_seconds, "epochs": int(st.epochs_done)}
        log(f"yahoo: gpu {name}: {a.epochs} epochs in {st.train_seconds:.2f} s "
            f"({st.updates / st.train_seconds / 1e9:.3f} G updates/s), test RMSE {rmse:.5f}")
        # timed segments (the benchmark's mode): columns circulate for a
        # fixed device time, so no launch waits for the heaviest column
        upd = dev = 0.0
        for _ in range(a.yahoo_segments):
            s2 = ctx.train(seconds=a.segment)
            upd += s2.updates
            dev += s2.train_seconds
        paths[name]["timed_updates_per_s"] = upd / dev
        paths[name]["timed_segments"] = a.yahoo_segments
        log(f"yahoo: gpu {name}: {a.yahoo_segments} timed {a.segment} s segments: {upd / dev / 1e9:.3f} G updates/s")
        ctx.close()
    # cpu_baseline leg: the oracle's FP64 async restatement on the host, the
    # same epochs; its test RMSE is the FP64 reference of the gaps
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    Wc, Hc, st = O.sgd_async(tr, W0, H0, storage=O.FP64, T=threads, seed=12345, epochs=a.epochs, **hyp)
    cpu_s = time.perf_counter() - t0
    ro = float(np.sqrt(O.sse(te, Wc, Hc, storage=O.FP64) / te.nnz))
    log(f"yahoo: cpu fp64 oracle ({threads} threads): {a.epochs} epochs in {cpu_s:.1f} s, test RMSE {ro:.5f}")
    for v in paths.values():
        v["gap_vs_fp64_oracle"] = (v["test_rmse"] - ro) / ro
    head = paths["bf16"]
    return {"metric": "BF16 vs FP32 vs FP64 test RMSE after equal epochs, Yahoo-shape k=100, 1 GPU",
            "value": head["updates_per_s"], "unit": "updates/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "bf16 storage, f32 arithmetic", "data": "synthetic", "epochs": a.epochs,
            "config": {"workload": "yahoo-shape synthetic, async NOMAD token kernel, epoch mode",
                       "m": tr.m, "n": tr.n, "nnz_train": tr.nnz, "nnz_test": te.nnz, "k": k, **hyp,
                       "workers": a.nworkers},
            "paths": paths,
            "cpu_baseline": {"value": st.updates / cpu_s, "unit": "updates/s", "cores": threads, "kind": "port",
                             "sample": f"oracle async restatement (orc_sgd_async), FP64 storage and arithmetic, "
                                       f"{a.epochs} epochs of the full training set",
                             "test_rmse": ro}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed training segments per path")
    ap.add_argument("--warmup", type=int, default=1, help="untimed segments on a throw-away context")
    ap.add_argument("--segment", type=float, default=0.1, help="device seconds per segment")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--nworkers", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=24.0, help="CPU baseline: training seconds at most")
    ap.add_argument("--cpu-segment", type=float, default=3.0, help="CPU baseline: seconds per timeout segment")
    ap.add_argument("--no-grad", action="store_true", help="skip the per-segment gradient norms")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--only", choices=["fp32", "bf16", "bf16_sr"], default=None)
    ap.add_argument("--workload", choices=["netflix", "sweep", "ccd", "objective", "yahoo"], default="netflix")
    ap.add_argument("--epochs", type=int, default=3, help="yahoo: epochs per path")
    ap.add_argument("--yahoo-segments", type=int, default=5, help="yahoo: timed segments per path after the epochs")
    ap.add_argument("--ccd-rank", type=int, default=100)
    ap.add_argument("--ccd-inner", type=int, default=3)
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES through the HIP runtime (a plain copy "
                         "kernel of known size, the reference for the HBM traffic counters)")
    ap.add_argument("--problems", type=int, default=32, help="sweep: number of problems")
    ap.add_argument("--strata", default="", help="sweep: comma-separated strata per problem (cycled) instead of "
                                                   "the autotuner's choice")
    ap.add_argument("--balance", default="", help="sweep, with --strata: comma-separated 0/1 per problem (cycled), "
                                                 "1 = row blocks balanced by ratings (default all 0)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); multi-GPU token exchange is a design, "
              "not part of this build", file=sys.stderr)
        return 2

    from nomad_amd import lib

    lib.hip()
    if a.copy_calibrate > 0:
        copy_calibrate(a.copy_calibrate)
    if a.workload != "netflix":
        fn = {"sweep": sweep, "ccd": ccd_workload, "objective": objective_workload,
              "yahoo": yahoo_workload}[a.workload]
        print(json.dumps(fn(a, lib)), flush=True)
        return 0
    k = 100
    hyp = dict(lrate=0.012, drate=0.05, reg=0.05)
    t0 = time.perf_counter()
    spec = lib.preset(2, a.scale)
    tr, te = lib.synth(spec)
    log(f"generated {tr.m} x {tr.n}, train {tr.nnz}, test {te.nnz} in {time.perf_counter() - t0:.1f} s")
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 12345)
    obj = None if a.no_grad else lib.Objective(tr, k, storage="fp64")
    common = dict(k=k, hyp=hyp, nworkers=a.nworkers, warmup=a.warmup, steps=a.steps, segment=a.segment, obj=obj)

    res = {}
    target = None
    if a.only in (None, "fp32"):
        res["fp32"] = run_path(lib, tr, te, W0, H0, storage="fp32", rounding="nearest", **common)
        target = res["fp32"]["curve"][-1] * 1.005
        res["fp32"]["t_hit"] = next((t for t, r in zip(res["fp32"]["times"], res["fp32"]["curve"]) if r <= target),
                                    None)
    if a.only in (None, "bf16"):
        res["bf16"] = run_path(lib, tr, te, W0, H0, storage="bf16", rounding="nearest", target=target, **common)
    if a.only in (None, "bf16_sr"):
        res["bf16_sr"] = run_path(lib, tr, te, W0, H0, storage="bf16", rounding="stochastic", target=target,
                                  log_name="bf16_sr", **common)

    for r in res.values():
        r["t_hit_interp"] = interp_hit(r["rmse0"], r["times"], r["curve"], target)
    if obj is not None:
        obj.close()

    head_name = next(n for n in ("bf16", "bf16_sr", "fp32") if n in res)
    head = res[head_name]
    value = head["updates"] / head["dev_s"]
    achieved = head["alg_bytes"] / head["dev_s"] / 1e9
    tpu = profiled_traffic(head_name)
    traffic = tpu * head["updates"] / a.steps if tpu is not None else None
    # the h term counted once per (column, workgroup) hop instead of once per
    # (column, wave) visit: inside a workgroup h moves through LDS
    esz = 4 if head_name == "fp32" else 2
    alg_hop = head["alg_bytes"] - (head["visits"] - head["hops"]) * 2.0 * k * esz
    achieved_hop = alg_hop / head["dev_s"] / 1e9
    out = {
        "metric": "updates/s + time-to-target test RMSE, Netflix-shape k=100, BF16 vs FP32, 1 GPU",
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["dev_s"] / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"netflix-shape synthetic, async NOMAD token kernel, {a.segment} s training segments",
            "m": tr.m, "n": tr.n, "nnz_train": tr.nnz, "nnz_test": te.nnz, "k": k, **hyp,
            "storage": "fp32" if head_name == "fp32" else "bf16",
            "rounding": "stochastic" if head_name == "bf16_sr" else "nearest",
            "workers": head["nworkers"], "row_blocks": "balanced by ratings", "scale": a.scale,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "frac_h_per_hop": achieved_hop / HBM_PEAK_GBS,
            "frac_measured_traffic": (traffic / (head["dev_s"] / a.steps) / 1e9 / HBM_PEAK_GBS
                                      if traffic is not None else None),
            "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE per update (copy-calibrated), "
                              "profiles/sgd_async_kernel_netflix_k100_pmc.md, x updates per launch",
            "kernel": "sgd_async_kernel<" + {"fp32": "fp32", "bf16": "bf16, nearest",
                                             "bf16_sr": "bf16, stochastic"}[head_name] + ">",
            "alg_bytes_per_launch": head["alg_bytes"] / a.steps,
            "avg_launch_ms": head["dev_s"] / a.steps * 1e3,
        },
        "paths": {
            name: {
                "updates_per_s": r["updates"] / r["dev_s"],
                "segment_ms": r["dev_s"] / a.steps * 1e3,
                "wall_segment_ms": r["wall_s"] / a.steps * 1e3,
                "alg_GBps": r["alg_bytes"] / r["dev_s"] / 1e9,
                "visits_per_segment": r["visits"] / a.steps,
                "train_seconds": r["times"],
                "test_rmse_curve": r["curve"],
                "time_to_target_s": r["t_hit"],
                "time_to_target_interp_s": r["t_hit_interp"],
                "hops_per_segment": r["hops"] / a.steps,
                "objective_curve": [round(x, 6) for x in r["losses"]],
                "grad_norm_curve": [round(x, 6) for x in r["grads"]],
                "visit_count_range": r["visit_count_range"],
            } for name, r in res.items()
        },
        "target_test_rmse": target,
    }
    if not a.skip_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(tr, te, W0, H0, k, hyp, a.cpu_seconds, a.cpu_segment, target)
        out["cpu_baseline"]["rmse0"] = head["rmse0"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
