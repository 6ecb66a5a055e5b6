"""fz_chkdiv_noalloc, the chkdiv variant that allocates nothing (its partial
sums live in a static device array shared by every call on the GPU), against
fz_chkdiv through the C ABI: two calls in a row, the second on a smaller grid,
so the second reads a scratch array still holding the first call's partials.
Both run the same kernels with the same block count, so the results are
compared bit for bit; divmax is also compared with the oracle's chkdiv
(src/chkdiv.f90:15-49)."""
import ctypes as C

import numpy as np
import pytest

import fizzy_amd as F
import oracle as O
from fizzy_amd.device import DArray

pytestmark = pytest.mark.gpu

# the first grid needs the maximum of 1024 partial sums, the second 9
GRIDS = [([96, 80, 64], [1.0, 0.8, 1.3], 0.0), ([20, 12, 9], [1.0, 1.0, 1.0], 1.1)]


def _rand(rng, n):
    return np.asfortranarray(rng.uniform(-1.0, 1.0, (n[0] + 2, n[1] + 2, n[2] + 2)))


def test_noalloc_chkdiv_equals_chkdiv_on_two_calls(rng):
    ctxs = []
    try:
        for n, l, gr in GRIDS:
            ctx = F.Context(n, l, "PPPPNN", gr=gr)
            ctxs.append(ctx)
            u, v, w = _rand(rng, n), _rand(rng, n), _rand(rng, n)
            du, dv, dw = ctx.halo(u), ctx.halo(v), ctx.halo(w)
            out = DArray(ctx, (2,))
            ctx._arrays.append(out)
            out.upload(np.full(2, np.nan))
            F.call("fz_chkdiv_noalloc", ctx.h, du.ptr, dv.ptr, dw.ptr, out.ptr)
            ctx.sync()
            got = out.download()
            divtot, divmax = C.c_double(), C.c_double()
            F.call("fz_chkdiv", ctx.h, du.ptr, dv.ptr, dw.ptr, C.byref(divtot), C.byref(divmax))
            assert got[0] == divtot.value and got[1] == divmax.value, (n, got, divtot.value, divmax.value)
            _, ref_max = O.chkdiv(n, ctx.dli, ctx.dzfi, u, v, w)
            assert got[1] == ref_max, (n, got[1], ref_max)
    finally:
        for ctx in ctxs:
            ctx.close()


def test_noalloc_chkdiv_rejects_null_arguments():
    ctx = F.Context([8, 8, 8], [1.0, 1.0, 1.0], "PPPPNN")
    try:
        rc = F.lib().fz_chkdiv_noalloc(ctx.h, None, None, None, None)
        assert rc == -1
        assert b"null argument" in F.lib().fz_last_error()
    finally:
        ctx.close()
