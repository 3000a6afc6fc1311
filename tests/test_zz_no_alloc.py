"""nomad_eval_noalloc (the entry point that allocates no device memory): its
sums equal nomad_eval's bit for bit, on two calls in a row, the second on a
smaller problem, so partials left in the shared scratch array by the first
call cannot leak into the second."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ctx(lib, tr, te, k, storage):
    ctx = lib.Context(tr.m, tr.n, k, storage=storage, mode="async", nworkers=256, lrate=0.012, drate=0.05,
                      reg=0.05, seed=7)
    ctx.upload(tr, te)
    W0, H0 = lib.init_factors(tr.m, tr.n, k, 4, 7)
    ctx.set_factors(W0, H0)
    return ctx


@pytest.mark.parametrize("storage", ["fp32", "bf16", "fp64"])
def test_eval_noalloc_equals_eval(gpu, storage):
    big = gpu.synth(gpu.preset(2, 0.1))    # Netflix shape at 1/10: ~9.9 M training ratings
    small = gpu.synth(gpu.preset(1, 0.3))  # ~27 K training ratings
    for name, (tr, te), k in (("big", big, 64), ("small", small, 16)):
        print(f"[noalloc] {storage} {name} nnz={tr.nnz}", flush=True)
        with _ctx(gpu, tr, te, k, storage) as ctx:
            ctx.train(epochs=1)
            a = ctx.eval_noalloc()
            b = ctx.eval()
            c = ctx.eval_noalloc()
        assert a.train_count == tr.nnz and a.test_count == te.nnz
        for x in (a, c):
            assert x.train_sse == b.train_sse and x.test_sse == b.test_sse, (name, x.train_sse, b.train_sse)
        assert np.isfinite(b.test_sse) and b.test_sse > 0
