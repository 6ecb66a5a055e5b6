"""fz_solver_batched: the solves of several same-shaped contexts in one launch
per stage give each context's fz_solver result bit for bit; contexts of
another shape are refused."""
import ctypes as C

import numpy as np
import pytest

import fizzy_amd as F
from test_gpu_kernels import Pair, rand

pytestmark = pytest.mark.gpu


def init(pr):
    F.call("fz_initsolver", pr.ctx.h, None, None, None, None, None, None, None, None, None)


@pytest.mark.parametrize("n,cbc,gr", [([64, 64, 32], "PPPPNN", 0.6), ([32, 24, 16], "NNPPDD", 0.0),
                                      ([2, 64, 48], "PPNNNN", 0.0)], ids=["pow2", "generic", "two-d"])
def test_batched_solves_equal_single_solves(n, cbc, gr, rng):
    k = 5
    prs = [Pair(n, [1.0, 0.9, 1.7], cbc, gr=gr) for _ in range(k)]
    try:
        for pr in prs:
            init(pr)
        fs = [rand(rng, n) for _ in range(k)]
        singles, batch = [], []
        for pr, f in zip(prs, fs):
            d = pr.h().upload(f)
            F.call("fz_solver", pr.ctx.h, d.ptr)
            singles.append(d.download())
            batch.append(pr.h().upload(f))
        ctxs = (C.c_void_p * k)(*[pr.ctx.h.value for pr in prs])
        ptrs = (C.c_void_p * k)(*[b.ptr.value for b in batch])
        F.call("fz_solver_batched", ctxs, ptrs, k)
        for q in range(k):
            got = batch[q].download()
            assert np.array_equal(got[1:-1, 1:-1, 1:-1], singles[q][1:-1, 1:-1, 1:-1]), q
    finally:
        for pr in prs:
            pr.close()


def test_batched_refuses_another_shape():
    a, b = Pair([16, 16, 8], [1, 1, 1], "PPPPNN"), Pair([16, 16, 10], [1, 1, 1], "PPPPNN")
    try:
        init(a)
        init(b)
        da, db = a.h(), b.h()
        ctxs = (C.c_void_p * 2)(a.ctx.h.value, b.ctx.h.value)
        ptrs = (C.c_void_p * 2)(da.ptr.value, db.ptr.value)
        rc = F.lib().fz_solver_batched(ctxs, ptrs, 2)
        assert rc != 0 and b"differs" in F.lib().fz_last_error()
    finally:
        a.close()
        b.close()
