"""Routine-level parity: every hot-path entry point of libfizzy (called through
the C ABI) against the CPU oracle on the same seeded inputs.

Tolerances: the library is built without FMA contraction and keeps the
reference's operation order, so routines made of +,-,*,/,sqrt agree to the
last bit (asserted as rel <= 1e-15); routines with tanh/log/exp/sin/cos differ
only by the ulp-level differences of the two math libraries (rel <= 1e-13).
The Poisson solve uses the same transform algorithm on both sides and is
compared bit for bit, null mode included."""
import ctypes as C

import numpy as np
import pytest

import fizzy_amd as F
import oracle as O
from conftest import rel_err
from fizzy_amd.device import DArray, dvec

pytestmark = pytest.mark.gpu
L = O.lib()
I, D, P, S = O.ivec, O.dvec, O.ptr, O.cstr


class Pair:
    """A GPU context and the matching oracle grid."""

    def __init__(self, n, l, cbcpre, gr=0.0):
        self.n = list(n)
        self.ctx = F.Context(n, l, cbcpre, gr=gr)
        self.dl, self.dli = self.ctx.dl, self.ctx.dli
        self.dzc, self.dzf, self.zc, self.zf = self.ctx.dzc, self.ctx.dzf, self.ctx.zc, self.ctx.zf
        self.dzci, self.dzfi = 1.0 / self.dzc, 1.0 / self.dzf
        self.cbcpre = "".join(cbcpre)

    def h(self):
        return self.ctx.halo()

    def i(self):
        return self.ctx.interior()

    def close(self):
        self.ctx.close()


def rand(rng, n, lo=-1.0, hi=1.0):
    return np.asfortranarray(rng.uniform(lo, hi, (n[0] + 2, n[1] + 2, n[2] + 2)))


SHAPES = [([16, 12, 10], [1.0, 0.8, 1.3], 0.0), ([33, 7, 9], [1.0, 1.0, 1.0], 1.1), ([64, 64, 32], [1, 1, 0.5], 0.0)]


@pytest.fixture(params=SHAPES, ids=lambda s: "x".join(map(str, s[0])))
def pair(request):
    n, l, gr = request.param
    p = Pair(n, l, "PPPPNN", gr)
    yield p
    p.close()


def test_fillps_correc_updatep(pair, rng):
    n = pair.n
    u, v, w, pp = rand(rng, n), rand(rng, n), rand(rng, n), rand(rng, n)
    du, dv, dw, dp, dpp = pair.h().upload(u), pair.h().upload(v), pair.h().upload(w), pair.h(), pair.h().upload(pp)
    p = rand(rng, n)
    dp.upload(p)
    F.call("fz_fillps", pair.ctx.h, 1.0 / 3e-3, 0.7, du.ptr, dv.ptr, dw.ptr, dp.ptr)
    L.or_fillps(I(n), D(pair.dli), P(pair.dzfi), C.c_double(1.0 / 3e-3), C.c_double(0.7), P(u), P(v), P(w), P(p))
    assert rel_err(dp.download(), p) <= 1e-15
    F.call("fz_correc", pair.ctx.h, 0.7, dvec([1, 2]), 3e-3, dp.ptr, None, du.ptr, dv.ptr, dw.ptr)
    L.or_correc(I(n), D(pair.dli), P(pair.dzci), C.c_double(0.7), C.c_double(3e-3), P(p), P(u), P(v), P(w))
    for d, h in ((du, u), (dv, v), (dw, w)):
        assert rel_err(d.download(), h) <= 1e-15
    F.call("fz_updatep", pair.ctx.h, dpp.ptr, dp.ptr)
    L.or_updatep(I(n), P(pp), P(p))
    assert rel_err(dp.download(), p) <= 1e-15


def test_acdi_transport_phi_gamma(pair, rng):
    n = pair.n
    psi = rand(rng, n, 0.0, 1.0)
    u, v, w = rand(rng, n), rand(rng, n), rand(rng, n)
    nx, ny, nz = rand(rng, n), rand(rng, n), rand(rng, n)
    seps, gam = 0.51 * pair.dl[0], 0.8
    phi = O.halo(n)
    L.or_acdi_cmpt_phi(I(n), C.c_double(seps), P(psi), P(phi))
    dpsi_ = pair.h().upload(psi)
    dphi = pair.h()
    F.call("fz_acdi_cmpt_phi", pair.ctx.h, seps, dpsi_.ptr, dphi.ptr)
    assert rel_err(dphi.download(), phi) <= 1e-13
    dphi.upload(phi)  # identical phi into the transport
    ins = [pair.h().upload(a) for a in (u, v, w, nx, ny, nz)]
    ddt, fx, fy, fz = pair.i(), pair.h(), pair.h(), pair.h()
    F.call("fz_acdi_transport_pf", pair.ctx.h, gam, seps, *[a.ptr for a in ins], dphi.ptr, dpsi_.ptr, ddt.ptr, fx.ptr,
           fy.ptr, fz.ptr)
    odt, ofx, ofy, ofz = O.interior(n), O.halo(n), O.halo(n), O.halo(n)
    L.or_acdi_transport_pf(I(n), D(pair.dli), P(pair.dzci), P(pair.dzfi), C.c_double(gam), C.c_double(seps), P(u),
                           P(v), P(w), P(nx), P(ny), P(nz), P(phi), P(psi), P(odt), P(ofx), P(ofy), P(ofz))
    c = (slice(1, -1),) * 3
    assert rel_err(ddt.download(), odt) <= 1e-13
    for d, h in ((fx, ofx), (fy, ofy), (fz, ofz)):
        assert rel_err(d.download()[c], h[c]) <= 1e-13
    g = C.c_double()
    F.call("fz_acdi_set_gamma", pair.ctx.h, 1.0, ins[0].ptr, ins[1].ptr, ins[2].ptr, C.byref(g))
    assert g.value == L.or_acdi_set_gamma(I(n), C.c_double(1.0), P(u), P(v), P(w))


def test_youngs_and_clip(pair, rng):
    n = pair.n
    phi = rand(rng, n)
    dphi = pair.h().upload(phi)
    outs = [pair.h() for _ in range(4)]
    F.call("fz_cmpt_norm_curv_youngs", pair.ctx.h, dphi.ptr, *[o.ptr for o in outs])
    ref = [O.halo(n) for _ in range(4)]
    L.or_cmpt_norm_curv_youngs(I(n), D(pair.dli), P(pair.dzci), P(pair.dzfi), P(phi), *[P(r) for r in ref])
    c = (slice(1, -1),) * 3
    for o, r in zip(outs, ref):  # +,-,*,/ and sqrt only: correctly rounded on both sides
        assert np.array_equal(o.download()[c], r[c]), rel_err(o.download()[c], r[c])
    x = rand(rng, n, -0.5, 1.5)
    dx = pair.h().upload(x)
    F.call("fz_clip_field", pair.ctx.h, dvec([0.0, 1.0]), dx.ptr)
    L.or_clip_field(I(n), C.c_double(0.0), C.c_double(1.0), P(x))
    assert np.array_equal(dx.download(), x)


def test_youngs_outside_the_fast_range(pair, rng):
    """The FP64 Youngs kernel normalises with unscaled square roots and one
    shared reciprocal per corner; cells whose gradients leave that sequence's
    range (components below 2^-300, sums above 2^600, zeros, subnormals,
    infinities) take the plain sqrt and '/'. A phase field mixing such values
    with ordinary ones must still give the oracle's normals and curvature bit
    for bit (NaN where the oracle has NaN)."""
    n = pair.n
    phi = rand(rng, n)
    pick = rng.random(phi.shape)
    phi[pick < 0.08] = 0.0
    phi[(pick >= 0.08) & (pick < 0.12)] *= 2.0 ** -320
    phi[(pick >= 0.12) & (pick < 0.14)] = 5e-324 * rng.integers(1, 1000, phi.shape)[(pick >= 0.12) & (pick < 0.14)]
    phi[(pick >= 0.14) & (pick < 0.17)] *= 2.0 ** 310
    phi[(pick >= 0.17) & (pick < 0.175)] = np.inf
    phi = np.asfortranarray(phi)
    dphi = pair.h().upload(phi)
    outs = [pair.h() for _ in range(4)]
    F.call("fz_cmpt_norm_curv_youngs", pair.ctx.h, dphi.ptr, *[o.ptr for o in outs])
    ref = [O.halo(n) for _ in range(4)]
    with np.errstate(all="ignore"):
        L.or_cmpt_norm_curv_youngs(I(n), D(pair.dli), P(pair.dzci), P(pair.dzfi), P(phi), *[P(r) for r in ref])
    c = (slice(1, -1),) * 3
    for o, r in zip(outs, ref):
        got = o.download()[c]
        assert np.array_equal(got, r[c], equal_nan=True)
        keep = ~np.isnan(got)  # NaN sign bits differ between the host's and the device's arithmetic
        assert np.array_equal(np.signbit(got[keep]), np.signbit(r[c][keep]))


@pytest.mark.parametrize("seps_cells, raised", [(0.51, False), (2.0 ** 24, True)], ids=["flag-clear", "flag-raised"])
def test_youngs_under_the_range_flag(pair, rng, seps_cells, raised):
    """The FP64 step's phase geometry (fz_phase_geometry_eval): phi from psi
    raising the range flag, then the check-free Youngs kernel when the flag
    stays clear (quartered grid factors, no division fix-ups) or its checked
    twin when it is raised (here by phi above 2^20 from a wide interface),
    against the oracle's acdi_cmpt_phi and cmpt_norm_curv_youngs bit for bit.
    psi holds cells of exactly 0, 1 and 0.5 and a slab of zeros, so that many
    vertex gradients and some phi are exactly zero."""
    n = pair.n
    psi = rand(rng, n, 0.0, 1.0)
    pick = rng.random(psi.shape)
    psi[pick < 0.15] = 0.0
    psi[(pick >= 0.15) & (pick < 0.3)] = 1.0
    psi[(pick >= 0.3) & (pick < 0.35)] = 0.5
    psi[:, :, : n[2] // 3] = 0.0
    psi = np.asfortranarray(psi)
    seps = seps_cells * pair.dl[0]
    phi = O.halo(n)
    L.or_acdi_cmpt_phi(I(n), C.c_double(seps), P(psi), P(phi))
    ref = [O.halo(n) for _ in range(4)]
    L.or_cmpt_norm_curv_youngs(I(n), D(pair.dli), P(pair.dzci), P(pair.dzfi), P(phi), *[P(r) for r in ref])
    dpsi, dphi = pair.h().upload(psi), pair.h()
    outs = [pair.h() for _ in range(4)]
    word = DArray(pair.ctx, (1,))
    try:
        F.call("fz_phase_geometry_eval", pair.ctx.h, seps, dpsi.ptr, dphi.ptr, *[o.ptr for o in outs], word.ptr)
        flag = int(word.download().view(np.uint32)[0])
    finally:
        word.free()
    assert (flag != 0) == raised
    assert np.array_equal(dphi.download(), phi)
    c = (slice(1, -1),) * 3
    for o, r in zip(outs, ref):
        assert np.array_equal(o.download()[c], r[c]), rel_err(o.download()[c], r[c])


def test_vof_thinc(pair, rng):
    n = pair.n
    psi = rand(rng, n, 0.0, 1.0)
    beta = 1.0
    phi = O.halo(n)
    L.or_vof_thinc_cmpt_phi(I(n), C.c_double(beta), P(psi), P(phi))
    dpsi_, dphi = pair.h().upload(psi), pair.h()
    F.call("fz_vof_thinc_cmpt_phi", pair.ctx.h, beta, dpsi_.ptr, dphi.ptr)
    assert rel_err(dphi.download(), phi) <= 1e-13
    dphi.upload(phi)
    u, v, w = rand(rng, n), rand(rng, n), rand(rng, n)
    nx, ny, nz = rand(rng, n), rand(rng, n), rand(rng, n)
    ins = [pair.h().upload(a) for a in (u, v, w, nx, ny, nz)]
    ddt, fx, fy, fz = pair.i(), pair.h(), pair.h(), pair.h()
    F.call("fz_vof_thinc_transport_psi", pair.ctx.h, beta, *[a.ptr for a in ins], dphi.ptr, dpsi_.ptr, ddt.ptr, fx.ptr,
           fy.ptr, fz.ptr)
    odt, ofx, ofy, ofz = O.interior(n), O.halo(n), O.halo(n), O.halo(n)
    L.or_vof_thinc_transport_psi(I(n), D(pair.dli), P(pair.dzfi), C.c_double(beta), P(u), P(v), P(w), P(nx), P(ny),
                                 P(nz), P(phi), P(odt), P(ofx), P(ofy), P(ofz))
    assert rel_err(ddt.download(), odt) <= 1e-13
    s = (slice(0, n[0] + 1), slice(0, n[1] + 1), slice(0, n[2] + 1))
    for d, h in ((fx, ofx), (fy, ofy), (fz, ofz)):
        assert rel_err(d.download()[s], h[s]) <= 1e-13


def test_momentum_and_ab2(pair, rng):
    n = pair.n
    u, v, w = rand(rng, n), rand(rng, n), rand(rng, n)
    psi, psio = rand(rng, n, 0, 1), rand(rng, n, 0, 1)
    fx, fy, fz = rand(rng, n), rand(rng, n), rand(rng, n)
    rho12, mu12 = [1000.0, 100.0], [10.0, 1.0]
    dins = {k: pair.h().upload(a) for k, a in dict(u=u, v=v, w=w, psi=psi, psio=psio, fx=fx, fy=fy, fz=fz).items()}
    dR = [pair.i() for _ in range(3)]
    F.call("fz_mom_xyz_ad", pair.ctx.h, dvec(rho12), dvec(mu12), dins["fx"].ptr, dins["fy"].ptr, dins["fz"].ptr,
           dins["u"].ptr, dins["v"].ptr, dins["w"].ptr, dins["psi"].ptr, *[r.ptr for r in dR])
    oR = [O.interior(n) for _ in range(3)]
    L.or_mom_xyz_ad(I(n), D(pair.dli), P(pair.dzci), P(pair.dzfi), D(rho12), D(mu12), P(fx), P(fy), P(fz), P(u), P(v),
                    P(w), P(psi), *[P(r) for r in oR])
    for d, h in zip(dR, oR):
        assert rel_err(d.download(), h) <= 1e-15
    p, pn, po, kap = rand(rng, n), rand(rng, n), rand(rng, n), rand(rng, n)
    dp = [pair.h().upload(a) for a in (p, kap, pn, po)]
    dR2 = [pair.i() for _ in range(3)]
    bf, ga, rav = [0.1, 0.0, -0.2], [0.0, 0.0, -0.98], [0.0, 0.0, 3.0]
    F.call("fz_mom_xyz_oth", pair.ctx.h, 1.3, dvec(rho12), dvec([1, 1]), dvec(bf), dvec(ga), 24.5, 100.0, dvec(rav),
           dp[0].ptr, dins["psi"].ptr, dp[1].ptr, dp[2].ptr, dp[3].ptr, *[r.ptr for r in dR2])
    oR2 = [O.interior(n) for _ in range(3)]
    L.or_mom_xyz_oth(I(n), D(pair.dli), P(pair.dzci), P(pair.dzfi), C.c_double(1.3), D(rho12), D(bf), D(ga),
                     C.c_double(24.5), C.c_double(100.0), D(rav), P(p), P(psi), P(kap), P(pn), P(po), *[P(r) for r in oR2])
    for d, h in zip(dR2, oR2):
        assert rel_err(d.download(), h) <= 1e-15
    # rk: AB2 + rescale, then the oth increment, with an explicit history object
    st = C.c_void_p()
    F.call("fz_rk_state_create", pair.ctx.h, C.byref(st))
    try:
        dt, dt_r = 2e-3, 1.3  # same dt_r as the mom_xyz_oth call above
        rkpar = [1.0 + 0.5 * dt_r, -0.5 * dt_r]
        F.call("fz_rk", pair.ctx.h, st, dvec(rkpar), dt, dt_r, dvec(bf), dvec(ga), 24.5, dvec(rav), dvec(rho12),
               dvec(mu12), dvec([1, 1]), 100.0, dins["psi"].ptr, dp[1].ptr, dp[0].ptr, dp[2].ptr, dp[3].ptr,
               dins["psio"].ptr, dins["fx"].ptr, dins["fy"].ptr, dins["fz"].ptr, dins["u"].ptr, dins["v"].ptr,
               dins["w"].ptr)
        f1, f2 = rkpar[0] * dt, rkpar[1] * dt
        L.or_rk_mom_ab2(I(n), C.c_double(f1), C.c_double(f2), D(rho12), P(psio), P(psi), *[P(r) for r in oR],
                        *[P(r) for r in oR], P(u), P(v), P(w))
        L.or_rk_mom_oth(I(n), C.c_double(f1 + f2), D(rho12), P(psi), *[P(r) for r in oR2], P(u), P(v), P(w))
        for k, h in (("u", u), ("v", v), ("w", w)):
            assert rel_err(dins[k].download(), h) <= 1e-15
    finally:
        F.lib().fz_rk_state_destroy(st)


BCS = ["PPPPNN", "NNPPNN", "PPDDPP", "DNNDPP", "NNNNNN", "PPPPPP"]


@pytest.mark.parametrize("cbc", BCS)
def test_boundary_conditions(cbc, rng):
    n = [9, 7, 6]
    pr = Pair(n, [1.0, 1.0, 1.0], cbc, gr=0.7)
    try:
        x = rand(rng, n)
        codes = [cbc, cbc.replace("N", "D"), cbc.replace("D", "N")]
        for code in codes:
            bc = list(rng.uniform(-1, 1, 6))
            dxa = pr.h().upload(x)
            F.call("fz_boundp", pr.ctx.h, code.encode(), dvec(bc), dxa.ptr)
            y = x.copy(order="F")
            L.or_boundp(I(n), S(cbc), S(code), D(bc), D(pr.dl), P(pr.dzc), P(y))
            assert np.array_equal(dxa.download(), y), code
        cbcvel = []
        for comp in range(3):
            for d in range(3):
                pairc = cbc[2 * d:2 * d + 2]
                cbcvel += list(pairc) if pairc == "PP" else [("D", "N")[(comp + d) % 2]] * 2
        bcv = list(rng.uniform(-1, 1, 18))
        for is_correc in (0, 1):
            us = [rand(rng, n) for _ in range(3)]
            dus = [pr.h().upload(a) for a in us]
            F.call("fz_bounduvw", pr.ctx.h, "".join(cbcvel).encode(), dvec(bcv), is_correc, *[d.ptr for d in dus])
            L.or_bounduvw(I(n), S(cbc), S(cbcvel), D(bcv), is_correc, D(pr.dl), P(pr.dzc), P(pr.dzf),
                          *[P(a) for a in us])
            for d, a in zip(dus, us):
                assert np.array_equal(d.download(), a)
    finally:
        pr.close()


KINDS = {"R2HC": 0, "HC2R": 1, "REDFT01": 4, "REDFT10": 5, "REDFT11": 6, "RODFT01": 8, "RODFT10": 9, "RODFT11": 10}


@pytest.mark.parametrize("kind", list(KINDS))
def test_r2r_lines_bitwise(kind, rng):
    """fz_r2r (the fft routine, src/fft.f90:176-187) against the oracle's r2r, line by line."""
    ctx = F.Context([4, 4, 4], [1, 1, 1], "PPPPPP")
    try:
        for n in [1, 2, 3, 5, 6, 7, 8, 10, 12, 13, 16, 18, 32, 36, 64, 96, 128, 256, 512, 1024, 2048]:
            if KINDS[kind] in (6, 10) and 2 * n > 2048:
                continue
            howmany = 7
            x = rng.standard_normal((howmany, n))
            d = F.DArray(ctx, (n * howmany,))
            d.upload(np.ascontiguousarray(x).reshape(-1))
            F.call("fz_r2r", ctx.h, KINDS[kind], n, howmany, d.ptr)
            got = d.download().reshape(howmany, n)
            ref = O.r2r(KINDS[kind], x)
            d.free()
            assert np.array_equal(got, ref), (kind, n, rel_err(got, ref))
    finally:
        ctx.close()


POISSON = [([16, 12, 10], "PPPPNN"), ([64, 64, 32], "PPPPNN"), ([30, 18, 7], "NNPPNN"), ([12, 10, 9], "DDNNDD"),
           ([2, 64, 64], "PPNNNN"), ([64, 2, 128], "NNPPNN"), ([20, 14, 8], "NDDNPP"), ([24, 16, 16], "PPPPPP"),
           ([36, 10, 12], "DNNDDN"), ([128, 128, 64], "PPPPNN"), ([256, 256, 8], "PPPPNN"),
           ([512, 40, 4], "PPNNNN"), ([24, 512, 5], "NNPPDD"), ([16, 64, 8], "PPNDNN"), ([1024, 6, 3], "DNPPNN")]


@pytest.mark.parametrize("n,cbc", POISSON, ids=[f"{'x'.join(map(str, a))}-{b}" for a, b in POISSON])
def test_poisson_solve_bitwise(n, cbc, rng):
    poisson_solve_bitwise(n, cbc, rng)


NOT_PERIODIC_Z = [(n, cbc) for n, cbc in POISSON if cbc[4:] != "PP"]


@pytest.mark.parametrize("rows", ["8", "16"])
@pytest.mark.parametrize("n,cbc", NOT_PERIODIC_Z, ids=[f"{'x'.join(map(str, a))}-{b}" for a, b in NOT_PERIODIC_Z])
def test_poisson_solve_bitwise_without_pivot_array(n, cbc, rows, rng, monkeypatch):
    """The Thomas kernel that re-runs the elimination per chunk instead of
    storing pivots (a context reads FZ_THOMAS_RC and FZ_THOMAS_ROWS64 when
    it is created) matches the oracle bit for bit as the default one does."""
    monkeypatch.setenv("FZ_THOMAS_RC", "1")
    monkeypatch.setenv("FZ_THOMAS_ROWS64", rows)
    poisson_solve_bitwise(n, cbc, rng)


@pytest.mark.parametrize("n,cbc", NOT_PERIODIC_Z, ids=[f"{'x'.join(map(str, a))}-{b}" for a, b in NOT_PERIODIC_Z])
def test_poisson_solve_bitwise_with_pivots_made_at_set_up(n, cbc, rng, monkeypatch):
    """The Thomas kernel reading pivots made once by fz_initsolver (a context
    reads FZ_THOMAS_PRE when it is created) matches the oracle bit for bit."""
    monkeypatch.setenv("FZ_THOMAS_PRE", "1")
    poisson_solve_bitwise(n, cbc, rng)


def poisson_solve_bitwise(n, cbc, rng):
    pr = Pair(n, [1.0, 0.9, 1.7], cbc, gr=0.6 if cbc[4:] != "PP" else 0.0)
    try:
        a, b, c = np.zeros(n[2]), np.zeros(n[2]), np.zeros(n[2])
        lam = np.zeros((n[0], n[1]), order="F")
        nf = C.c_double()
        dptr = C.POINTER(C.c_double)
        F.call("fz_initsolver", pr.ctx.h, None, lam.ctypes.data_as(dptr), a.ctypes.data_as(dptr),
               b.ctypes.data_as(dptr), c.ctypes.data_as(dptr), C.byref(nf), None, None, None)
        sol = O.Solver(n, pr.dli, pr.dzci, pr.dzfi, list(cbc))
        assert np.array_equal(lam, sol.lambdaxy) and np.array_equal(a, sol.a) and np.array_equal(b, sol.b)
        assert nf.value == sol.normfft
        f = rand(rng, n)
        if "D" not in cbc:
            f[1:-1, 1:-1, 1:-1] -= f[1:-1, 1:-1, 1:-1].mean()
        dpx = pr.h().upload(f)
        F.call("fz_solver", pr.ctx.h, dpx.ptr)
        ref = f.copy(order="F")
        sol.solve(ref)
        got = dpx.download()
        assert np.array_equal(got, ref), rel_err(got, ref)
    finally:
        pr.close()


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s[0])))
def test_checks_and_bulk_mean(shape, rng):
    """chkdiv (src/chkdiv.f90:15-49), chkdt (src/chkdt.f90:15-94) and
    bulk_mean_12_stag (src/utils.f90:68-136) through the C ABI against the
    oracle. Maxima are order-independent and compared bit for bit; the sums
    differ only in summation order (divtot: the round-off of a sum of N
    terms of size divmax; the bulk mean: a sum of positive terms)."""
    n, l, gr = shape
    pr = Pair(n, l, "PPPPNN", gr)
    try:
        u, v, w = rand(rng, n), rand(rng, n), rand(rng, n)
        du, dv, dw = pr.h().upload(u), pr.h().upload(v), pr.h().upload(w)
        dt_, dm_ = C.c_double(), C.c_double()
        F.call("fz_chkdiv", pr.ctx.h, du.ptr, dv.ptr, dw.ptr, C.byref(dt_), C.byref(dm_))
        divtot, divmax = O.chkdiv(n, pr.dli, pr.dzfi, u, v, w)
        assert dm_.value == divmax
        ncell = n[0] * n[1] * n[2]
        assert abs(dt_.value - divtot) <= 1e-13 * ncell * divmax
        for prm in [(1, 1, 0, (0.3, 0.02), (50.0, 1.0), 0.07, (0.0, 0.0, -9.8), 0.7, 0.01),
                    (1, 0, 0, (0.0, 0.0), (1.0, 1.0), 0.0, (0.0, 0.0, 0.0), 0.0, 0.0),
                    (0, 1, 1, (0.1, 0.1), (2.0, 1.0), 0.5, (1.0, 0.0, 0.0), 0.0, 0.0)]:
            solve_ns, track, vof, mu12, rho12, sigma, gacc, gam, seps = prm
            got = C.c_double()
            F.call("fz_chkdt", pr.ctx.h, solve_ns, track, vof, dvec(mu12), dvec(rho12), sigma, dvec(gacc), du.ptr,
                   dv.ptr, dw.ptr, gam, seps, C.byref(got))
            ref = L.or_chkdt(I(n), D(pr.dl), P(pr.dzci), P(pr.dzfi), solve_ns, track, vof, D(mu12), D(rho12),
                             C.c_double(sigma), D(gacc), P(u), P(v), P(w), C.c_double(gam), C.c_double(seps))
            assert got.value == ref, prm
        psi = rand(rng, n, 0.0, 1.0)
        dpsi = pr.h().upload(psi)
        vol = l[0] * l[1] * l[2]
        for idir in (1, 2, 3):
            gvr = np.ascontiguousarray(pr.dl[0] * pr.dl[1] * (pr.dzc if idir < 3 else pr.dzf) / vol)
            got = C.c_double()
            F.call("fz_bulk_mean_12_stag", pr.ctx.h, idir, dpsi.ptr, dvec([40.0, 1.5]), C.byref(got))
            ref = L.or_bulk_mean_12_stag(I(n), idir, P(gvr), P(psi), D([40.0, 1.5]))
            assert abs(got.value - ref) <= 1e-14 * abs(ref), (idir, got.value, ref)
    finally:
        pr.close()


RHSB = [([12, 10, 9], "DDNNDD"), ([20, 14, 8], "NDDNPP"), ([16, 12, 10], "PPPPND"), ([10, 8, 12], "DNNDDN")]


@pytest.mark.parametrize("n,cbc", RHSB, ids=[f"{'x'.join(map(str, a))}-{b}" for a, b in RHSB])
def test_updt_rhs_b_with_boundary_values(n, cbc, rng):
    """updt_rhs_b (src/bound.f90:341-396) with non-zero Dirichlet and Neumann
    pressure values: the boundary terms made by fz_initsolver (bc_rhs,
    src/initsolver.f90:184-228) added to p, then the solve, bit for bit
    against the oracle."""
    pr = Pair(n, [1.0, 0.9, 1.7], cbc, gr=0.6 if cbc[4:] != "PP" else 0.0)
    try:
        bc = list(rng.uniform(-1, 1, 6))
        a, b, c = np.zeros(n[2]), np.zeros(n[2]), np.zeros(n[2])
        lam = np.zeros((n[0], n[1]), order="F")
        rbx, rby, rbz = (np.zeros(s, order="F") for s in ((n[1], n[2], 2), (n[0], n[2], 2), (n[0], n[1], 2)))
        nf = C.c_double()
        dptr = C.POINTER(C.c_double)
        F.call("fz_initsolver", pr.ctx.h, dvec(bc), lam.ctypes.data_as(dptr), a.ctypes.data_as(dptr),
               b.ctypes.data_as(dptr), c.ctypes.data_as(dptr), C.byref(nf), rbx.ctypes.data_as(dptr),
               rby.ctypes.data_as(dptr), rbz.ctypes.data_as(dptr))
        sol = O.Solver(n, pr.dli, pr.dzci, pr.dzfi, list(cbc), bc=bc)
        for got, ref in ((rbx, sol.rhsbx), (rby, sol.rhsby), (rbz, sol.rhsbz)):
            assert np.array_equal(got, ref)
        assert any(np.abs(r).max() > 0 for r in (rbx, rby, rbz))
        f = rand(rng, n)
        dpx = pr.h().upload(f)
        F.call("fz_updt_rhs_b", pr.ctx.h, dpx.ptr)
        ref = f.copy(order="F")
        L.or_updt_rhs_b(I(n), S(cbc), P(sol.rhsbx), P(sol.rhsby), P(sol.rhsbz), P(ref))
        assert np.array_equal(dpx.download(), ref)
        F.call("fz_solver", pr.ctx.h, dpx.ptr)
        sol.solve(ref)
        assert np.array_equal(dpx.download(), ref)
    finally:
        pr.close()


def test_lscale_forcing(pair, rng):
    """lscale_forcing (src/forcing.f90:48-63) from the context's per-coordinate
    sine/cosine tables: Taylor-Green against the oracle bit for bit, on two
    calls in a row (the first makes the tables); the ABC form against the same
    expressions in numpy (rel <= 1e-13: numpy's sin/cos)."""
    n, dt, alpha = pair.n, 2e-3, 0.5
    u, v, w = rand(rng, n), rand(rng, n), rand(rng, n)
    du, dv, dw = pair.h().upload(u), pair.h().upload(v), pair.h().upload(w)
    l = [float(x) for x in pair.ctx.l]
    for _ in range(2):
        F.call("fz_lscale_forcing", pair.ctx.h, 2, alpha, dt, du.ptr, dv.ptr, dw.ptr)
        L.or_lscale_forcing_tg(I(n), C.c_double(alpha), C.c_double(dt), D(l), D(pair.dl), P(pair.zc), P(pair.zf),
                               P(u), P(v))
        assert np.array_equal(du.download(), u) and np.array_equal(dv.download(), v)
        assert np.array_equal(dw.download(), w)
    F.call("fz_lscale_forcing", pair.ctx.h, 1, alpha, dt, du.ptr, dv.ptr, dw.ptr)
    k0, f0 = 2.0, alpha * (2.0 * np.pi / l[2]) ** 2
    i = np.arange(1, n[0] + 1)[:, None, None]
    j = np.arange(1, n[1] + 1)[None, :, None]
    k = np.arange(1, n[2] + 1)[None, None, :]
    xc = (i.astype(np.float32) - np.float32(0.5)).astype(np.float64) * pair.dl[0] / l[0]
    yc = (j.astype(np.float32) - np.float32(0.5)).astype(np.float64) * pair.dl[1] / l[1]
    zc = np.asarray(pair.zc)[1:n[2] + 1][None, None, :] / l[2]
    s = (slice(1, -1),) * 3
    ref = [u[s] + dt * (f0 * (np.sin(k0 * zc) + np.cos(k0 * yc))),
           v[s] + dt * (f0 * (np.sin(k0 * xc) + np.cos(k0 * zc))),
           w[s] + dt * (f0 * (np.sin(k0 * yc) + np.cos(k0 * xc)))]
    for d, r in zip((du, dv, dw), ref):
        assert rel_err(d.download()[s], np.broadcast_to(r, d.download()[s].shape)) <= 1e-13
