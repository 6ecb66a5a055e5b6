"""Mixed-precision Poisson solve (fz_solver_mixed): FP32 transforms and line
solves with FP64 pivots, wrapped in FP64-residual iterative refinement,
against the FP64 solver (itself bitwise equal to the oracle) on the same
seeded right-hand sides.

Stated bounds: after refinement the FP64 residual max|b - A x|/max|b|
(checked independently with fz_chk_poisson, src/debug.f90:81-114) is at most
1e-12 or 4x the plain FP64 solve's own residual, whichever is larger, and
the solution agrees with the FP64 one to 1e-10 relative (mean removed where
the operator is singular). A single FP32 solve (no refinement) is accurate
to about 1e-6 relative, which the test also states."""
import ctypes as C

import numpy as np
import pytest

import fizzy_amd as F
import oracle as O
from test_gpu_kernels import Pair, rand

pytestmark = pytest.mark.gpu

CASES = [([16, 12, 10], "PPPPNN", 0.0), ([64, 64, 32], "PPPPNN", 0.6), ([30, 18, 7], "NNPPNN", 0.0),
         ([12, 10, 9], "DDNNDD", 0.6), ([64, 2, 128], "NNPPNN", 0.0), ([20, 14, 8], "NDDNPP", 0.0),
         ([24, 16, 16], "PPPPPP", 0.0), ([36, 10, 12], "DNNDDN", 0.0), ([128, 128, 64], "PPPPNN", 0.6),
         ([16, 64, 8], "PPNDNN", 0.0)]


def solve_pair(n, cbc, gr, rng, tol, maxit):
    pr = Pair(n, [1.0, 0.9, 1.7], cbc, gr=gr)
    try:
        F.call("fz_initsolver", pr.ctx.h, None, None, None, None, None, None, None, None, None)
        f = rand(rng, n)
        singular = "D" not in cbc
        if singular:
            # a compatible right-hand side: the operator's left null vector is
            # the cell volume (dzf), which the reference's grid makes
            # nonuniform at 1e-7 even for a uniform spacing
            wz = pr.dzf[1:-1]
            fi = f[1:-1, 1:-1, 1:-1]
            fi -= (fi * wz).sum() / (wz.sum() * n[0] * n[1])
        b = pr.h().upload(f)
        x64 = pr.h().upload(f)
        F.call("fz_solver", pr.ctx.h, x64.ptr)
        xm = pr.h().upload(f)
        it, rr = C.c_int(), C.c_double()
        F.call("fz_solver_mixed", pr.ctx.h, xm.ptr, tol, maxit, C.byref(it), C.byref(rr))
        zero = (C.c_double * 6)()
        res = {}
        for name, x in (("fp64", x64), ("mixed", xm)):
            F.call("fz_boundp", pr.ctx.h, cbc.encode(), zero, x.ptr)
            dm = C.c_double()
            F.call("fz_chk_poisson", pr.ctx.h, b.ptr, x.ptr, C.byref(dm))
            res[name] = dm.value / np.abs(f[1:-1, 1:-1, 1:-1]).max()
        g64 = x64.download()[1:-1, 1:-1, 1:-1]
        gm = xm.download()[1:-1, 1:-1, 1:-1]
        c64 = np.abs(g64).max()  # with the null-mode constant the reference's pivots leave
        if singular:
            g64, gm = g64 - g64.mean(), gm - gm.mean()
        err = np.abs(gm - g64).max() / np.abs(g64).max()
        # the FP64 solution itself is exact only to ~eps times its constant
        fp64_acc = 64 * 2.2e-16 * c64 / np.abs(g64).max()
        return it.value, rr.value, res, err, fp64_acc
    finally:
        pr.close()


@pytest.mark.parametrize("n,cbc,gr", CASES, ids=[f"{'x'.join(map(str, a))}-{b}-{g}" for a, b, g in CASES])
def test_mixed_solver_refines_to_fp64(n, cbc, gr, rng):
    it, rr, res, err, fp64_acc = solve_pair(n, cbc, gr, rng, 1e-13, 8)
    print(n, cbc, "iterations", it, "relres", rr, "chk_poisson", res, "vs fp64", err, "fp64 accuracy", fp64_acc)
    assert it >= 1
    assert res["mixed"] <= max(1e-12, 4 * res["fp64"])
    assert rr <= 2 * res["mixed"] + 1e-13 and res["mixed"] <= 2 * rr + 1e-13  # two roundings of one operator
    assert err <= max(1e-10, fp64_acc)


@pytest.mark.parametrize("cbc", ["PPPPNN", "PPPPPP"])
@pytest.mark.parametrize("rows", ["8", "16", "-8", "-16", "rc8", "rc16"])
def test_thomas_variants_agree_bitwise(rows, cbc, monkeypatch):
    """Every FP32 Thomas variant (rows per chunk, one or two columns per
    lane, rc: no pivot array; a context reads FZ_THOMAS_ROWS32 and
    FZ_THOMAS_RC when it is created), for Neumann and periodic z, gives the
    default variant's solution bit for bit."""
    out = []
    for env in (None, rows):
        if env is None:
            monkeypatch.delenv("FZ_THOMAS_ROWS32", raising=False)
            monkeypatch.delenv("FZ_THOMAS_RC", raising=False)
        else:
            monkeypatch.setenv("FZ_THOMAS_ROWS32", env.lstrip("rc"))
            if env.startswith("rc"):
                monkeypatch.setenv("FZ_THOMAS_RC", "1")
        it, rr, res, err, acc = solve_pair([64, 64, 32], cbc, 0.6 if cbc == "PPPPNN" else 0.0,
                                           np.random.default_rng(7), 1e-13, 2)
        out.append((it, rr, err))
    assert out[0] == out[1]


@pytest.mark.parametrize("n,cbc,gr", CASES[:4], ids=[f"{'x'.join(map(str, a))}-{b}" for a, b, g in CASES[:4]])
def test_single_fp32_solve_is_fp32_accurate(n, cbc, gr, rng):
    it, rr, res, err, fp64_acc = solve_pair(n, cbc, gr, rng, 1e-13, 0)
    print(n, cbc, "relres", rr, "vs fp64", err, "fp64 accuracy", fp64_acc)
    assert it == 0
    assert 1e-9 < err <= 1e-5
    assert rr <= 1e-5


# ---------------------------------------------------------------------------
# the mixed-precision case step against the FP64 one
# ---------------------------------------------------------------------------
from fizzy_amd import cases  # noqa: E402

CASE_FIELDS = ["u", "v", "w", "p", "psi", "kappa"]


def run_case(spec, nsteps, graphs=False):
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        if graphs:
            F.call("fz_case_set_graphs", dc._h, 1)
        vol0 = dc.get("psi")[1:-1, 1:-1, 1:-1].sum()
        dc.step(nsteps)
        out = {k: dc.get(k)[1:-1, 1:-1, 1:-1] for k in CASE_FIELDS}
        out["p"] = out["p"] - out["p"].mean()
        return out, dc.scalars(), vol0
    finally:
        dc.close()


def field_errors(got, ref):
    """Per-field error of a mixed run against the FP64 run: velocities
    relative to the largest FP64 velocity component, p (mean removed) and psi
    relative to their maxima, curvature relative to its maximum inside the
    interface band 0.05 < psi < 0.95 (outside it both runs hold the
    meaningless curvature of a saturated phase field)."""
    vs = max(np.abs(ref[k]).max() for k in ("u", "v", "w"))
    e = {k: float(np.abs(got[k] - ref[k]).max() / vs) for k in ("u", "v", "w")}
    for k in ("p", "psi"):
        e[k] = float(np.abs(got[k] - ref[k]).max() / np.abs(ref[k]).max())
    band = (ref["psi"] > 0.05) & (ref["psi"] < 0.95)
    e["kappa"] = float(np.abs(got["kappa"] - ref["kappa"])[band].max() / np.abs(ref["kappa"][band]).max())
    return e


MCASES = [("bubble3d-32x32x64", cases.rising_bubble_3d(ng=(32, 32, 64), rho12=(1000.0, 1.0), mu12=(10.0, 0.1))),
          ("hysing2d", cases.hysing_2d()),
          ("vof-droplets-32", cases.vof_droplets(ng=(32, 32, 32)))]


@pytest.mark.parametrize("name,spec", MCASES, ids=[m[0] for m in MCASES])
def test_mixed_case_against_fp64(name, spec):
    """40 steps (checks every 10) from the same initial state. Stated bounds
    with one refinement: chkdiv divmax <= 1e-10; velocity <= 1e-2, pressure
    <= 1e-3, psi <= 1e-5 (THINC/QQ: 1e-3) and interface curvature <= 1e-2 of
    their scales. Without refinement the FP32 solve leaves divmax at O(1e-6).
    With the plain step's FP64 projection (refine -1) divmax stays within
    10x of the FP64 step's and the fields within the same bounds."""
    nsteps = 40
    ref, sref, _ = run_case(spec, nsteps)
    out = {}
    for nref in (0, 1, 2, -1):
        got, sc, vol0 = run_case(dict(spec, precision="mixed", refine=nref), nsteps)
        errs = field_errors(got, ref)
        vol = float(got["psi"].sum())
        out[nref] = (errs, float(sc[7]), abs(vol - vol0) / vol0)
        print(name, "refinements", nref, "errors vs fp64", errs, "divmax", sc[7], "fp64 divmax", sref[7],
              "phase volume drift", out[nref][2])
    errs, divmax, drift = out[1]
    assert divmax <= 1e-10
    assert out[2][1] <= 1e-10
    assert out[0][1] > 1e-9  # the refinement is what brings the divergence down
    # THINC's sharp profile (an FP32 exp per face) moves psi more than ACDI's
    bounds = {"u": 1e-2, "v": 1e-2, "w": 1e-2, "p": 1e-3, "psi": 1e-3 if spec.get("vof") else 1e-5, "kappa": 1e-2}
    for k in CASE_FIELDS:
        assert errs[k] <= bounds[k], (k, errs[k])
    assert drift <= 1e-5
    errs64, divmax64, drift64 = out[-1]
    assert divmax64 <= max(10 * float(sref[7]), 1e-13)
    for k in CASE_FIELDS:
        assert errs64[k] <= bounds[k], (k, errs64[k])
    assert drift64 <= 1e-5


GCASES = [("fp64", cases.rising_bubble_3d(ng=(32, 32, 64))),
          ("mixed", dict(cases.rising_bubble_3d(ng=(32, 32, 64)), precision="mixed", refine=1)),
          ("mixed-fp64-projection", dict(cases.rising_bubble_3d(ng=(32, 32, 64)), precision="mixed", refine=-1)),
          ("hysing2d", cases.hysing_2d()),
          # THINC/QQ with the Taylor-Green forcing (dt read from device memory)
          ("vof-forced", cases.vof_droplets(ng=(32, 32, 32), ndrop=3, r=0.6)),
          ("vof-forced-mixed", dict(cases.vof_droplets(ng=(32, 32, 32), ndrop=3, r=0.6), precision="mixed",
                                    refine=1))]


@pytest.mark.parametrize("name,spec", GCASES, ids=[g[0] for g in GCASES])
def test_step_graphs_give_the_same_fields(name, spec):
    """Steps replayed from captured graphs (two parities, replayed many times
    with the eager dt/divergence checks in between) equal eager steps bit for
    bit."""
    a, sa, _ = run_case(spec, 35)
    b, sb, _ = run_case(spec, 35, graphs=True)
    for k in CASE_FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(sa, sb)
