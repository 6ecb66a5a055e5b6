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


# ---------------------------------------------------------------------------
# batched steps of whole cases (fz_case_batch_*)
from fizzy_amd import cases as CS  # noqa: E402
from fizzy_amd import serve  # noqa: E402

STEP_FIELDS = ["u", "v", "w", "p", "po", "psi", "phi", "kappa", "normx", "normy", "normz", "psiflx_x", "psiflx_y",
               "psiflx_z", "dpsidt_old", "dudt_old", "dvdt_old", "dwdt_old"]


def mix_of(base, ratios):
    """Cases of one shape with their own densities, viscosities and surface tension."""
    out = []
    for q, r in enumerate(ratios):
        out.append(dict(base, rho12=[r, 1.0], mu12=[10.0 / (1 + q), 0.5 + 0.1 * q], sigma=24.5 / (1 + 0.5 * q),
                        icheck=4))
    return out


def run_alone(spec, nsteps):
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(nsteps)
        return {k: dc.get(k) for k in STEP_FIELDS}, dc.scalars()
    finally:
        dc.close()


BATCH_SHAPES = {
    "bubble3d": CS.rising_bubble_3d(ng=(32, 32, 64)),
    "bubble2d": serve.bubble_case(64, two_d=True),
    "mixed": dict(serve.bubble_case(32, 48), precision="mixed", refine=-1),
}


@pytest.mark.parametrize("shape", list(BATCH_SHAPES))
def test_batched_steps_equal_each_case_alone(shape):
    """Four cases of one shape (different density ratios, viscosities and
    surface tension), each first stepped alone (one or two steps, so their
    buffer rotations are out of phase), then 11 steps together (checks every
    4 steps inside): every field and scalar equals the case run alone for the
    same number of steps, bit for bit."""
    specs = mix_of(BATCH_SHAPES[shape], [1000.0, 20.0, 300.0, 2.0])
    pre = [1, 2, 1, 2]
    ref = [run_alone(s, p + 11) for s, p in zip(specs, pre)]
    dcs = [F.DeviceCase(s) for s in specs]
    try:
        for dc, p in zip(dcs, pre):
            dc.init()
            dc.step(p)
        b = F.CaseBatch(dcs)
        try:
            assert b.step(11) == []
        finally:
            b.close()
        for q, (dc, (fields, sc)) in enumerate(zip(dcs, ref)):
            for k in STEP_FIELDS:
                assert np.array_equal(dc.get(k), fields[k]), (q, k)
            assert np.array_equal(dc.scalars(), sc), q
        dcs[0].step(3)  # a case steps alone again after its batch
        f, sc = run_alone(specs[0], pre[0] + 14)
        assert np.array_equal(dcs[0].get("u"), f["u"]) and np.array_equal(dcs[0].scalars(), sc)
    finally:
        for dc in dcs:
            dc.close()


def test_batch_refuses_what_it_does_not_cover():
    base = serve.bubble_case(16, 32)
    a = F.DeviceCase(base)
    others = [F.DeviceCase(serve.bubble_case(16, 24)),  # another grid
              F.DeviceCase(dict(base, precision="mixed", refine=1)),  # FP32 projection
              F.DeviceCase(CS.vof_droplets(ng=(16, 16, 16), ndrop=1))]  # THINC/QQ
    fresh = F.DeviceCase(base)  # first step not taken
    try:
        for dc in [a] + others:
            dc.init()
            dc.step(1)
        fresh.init()
        for o, why in zip(others + [fresh], ["another grid", "FP32 projection", "THINC/QQ", "first step"]):
            with pytest.raises(F.FizzyError, match=why):
                F.CaseBatch([a, o])
        with pytest.raises(F.FizzyError, match="twice"):
            F.CaseBatch([a, a])
    finally:
        for dc in [a, fresh] + others:
            dc.close()
