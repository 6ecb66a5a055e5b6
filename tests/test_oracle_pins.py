"""Pins for the oracle routines that the stencil tests leave open: the
two-phase terms of mom_xyz_ad (src/mom.f90:545-850), THINC/QQ conservation
(src/vof_thinc_qq.f90:39-133), acdi_set_gamma (src/acdi.f90:32-62), the AB2
updates of rk/rk_2fl (src/rk.f90:86-130, :240-247), updatep
(src/updatep.f90:21-23), fillps/correc on linear fields (src/fillps.f90:39-50,
src/correc.f90:34-66), the Taylor-Green forcing (src/forcing.f90:48-63),
bulk_mean_12_stag (src/utils.f90:68-136), chkdiv (src/chkdiv.f90:15-49) and
the clustered grids of initgrid (src/initgrid.f90:14-78). Each check is a
property the mathematics fixes (conservation of a flux form, a closed form
on a field where the answer is known, a symmetry), chosen so that a wrong
index, sign or dropped term in the restatement fails it."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()
I, D, P, S = O.ivec, O.dvec, O.ptr, O.cstr


def grid(n, l, gr=0.0, gtype=1):
    dzc, dzf, zc, zf = O.initgrid(gtype, n[2], gr, l[2])
    dl = [l[d] / float(np.float32(n[d])) for d in range(3)]
    return dl, [1 / x for x in dl], dzc, dzf, zc, zf, 1 / dzc, 1 / dzf


def periodic(a):
    a[0] = a[-2]
    a[-1] = a[1]
    a[:, 0] = a[:, -2]
    a[:, -1] = a[:, 1]
    a[:, :, 0] = a[:, :, -2]
    a[:, :, -1] = a[:, :, 1]
    return a


def rand_periodic(rng, n, lo=-1.0, hi=1.0):
    return periodic(np.asfortranarray(rng.uniform(lo, hi, (n[0] + 2, n[1] + 2, n[2] + 2))))


def test_mom_ad_two_phase_is_a_conservative_flux_form(rng):
    """With variable density flux and viscosity, every term of mom_xyz_ad is a
    difference of face fluxes, so on a triply periodic uniform grid the domain
    sum of each component vanishes for arbitrary two-phase fields. A face
    quantity taken at the wrong cell (an index slip in the mass flux or in a
    viscosity average) breaks the telescoping."""
    n = [12, 10, 8]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.0, 0.9, 0.8])
    u, v, w = (rand_periodic(rng, n) for _ in range(3))
    psi = rand_periodic(rng, n, 0.0, 1.0)
    fx, fy, fz = (rand_periodic(rng, n) for _ in range(3))
    du, dv, dw = O.interior(n), O.interior(n), O.interior(n)
    L.or_mom_xyz_ad(I(n), D(dli), P(dzci), P(dzfi), D([1000.0, 1.0]), D([10.0, 0.1]), P(fx), P(fy), P(fz), P(u),
                    P(v), P(w), P(psi), P(du), P(dv), P(dw))
    for r in (du, dv, dw):
        scale = np.abs(r).sum()
        assert scale > 1.0
        assert abs(r.sum()) < 1e-13 * scale


def test_mom_ad_uniform_phase_reduces_to_the_mixture(rng):
    """psi uniform (=c) with the consistent phase flux c*u: the two-phase right-hand
    side equals the single-phase one with rho = rho2 + c*(rho1-rho2) and
    mu = mu2 + c*(mu1-mu2) (the mixture rules of src/mom.f90:608-630, :784-842)."""
    n = [8, 6, 8]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.0, 1.0, 1.0], gr=0.0)
    c = 0.375
    rho12, mu12 = (7.0, 3.0), (0.5, 0.25)
    u, v, w = (np.asfortranarray(rng.standard_normal((n[0] + 2, n[1] + 2, n[2] + 2))) for _ in range(3))
    psi = O.halo(n, c)
    out2 = [O.interior(n) for _ in range(3)]
    L.or_mom_xyz_ad(I(n), D(dli), P(dzci), P(dzfi), D(rho12), D(mu12), P(np.asfortranarray(c * u)),
                    P(np.asfortranarray(c * v)), P(np.asfortranarray(c * w)), P(u), P(v), P(w), P(psi),
                    *[P(a) for a in out2])
    rho = rho12[1] + c * (rho12[0] - rho12[1])
    mu = mu12[1] + c * (mu12[0] - mu12[1])
    out1 = [O.interior(n) for _ in range(3)]
    z = O.halo(n)
    L.or_mom_xyz_ad(I(n), D(dli), P(dzci), P(dzfi), D([rho, rho]), D([mu, mu]), P(z), P(z), P(z), P(u), P(v), P(w),
                    P(psi), *[P(a) for a in out1])
    for a, b in zip(out2, out1):
        assert np.abs(a - b).max() < 1e-12 * np.abs(b).max()


def test_thinc_transport_conserves_the_phase(rng):
    """THINC/QQ's right-hand side is minus the divergence of its face fluxes:
    on a periodic grid the domain sum vanishes, and it equals the flux
    differences cell by cell."""
    n = [10, 8, 8]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.0, 1.0, 1.0])
    u, v, w = (rand_periodic(rng, n) for _ in range(3))
    nx, ny, nz = (rand_periodic(rng, n) for _ in range(3))
    psi = rand_periodic(rng, n, 0.0, 1.0)
    phi = O.halo(n)
    beta = 1.0
    L.or_vof_thinc_cmpt_phi(I(n), C.c_double(beta), P(psi), P(phi))
    dpsi = O.interior(n)
    fx, fy, fz = O.halo(n), O.halo(n), O.halo(n)
    L.or_vof_thinc_transport_psi(I(n), D(dli), P(dzfi), C.c_double(beta), P(u), P(v), P(w), P(nx), P(ny), P(nz),
                                 P(phi), P(dpsi), P(fx), P(fy), P(fz))
    ref = -((fx[1:-1, 1:-1, 1:-1] - fx[:-2, 1:-1, 1:-1]) * dli[0] + (fy[1:-1, 1:-1, 1:-1] - fy[1:-1, :-2, 1:-1]) * dli[1]
            + (fz[1:-1, 1:-1, 1:-1] - fz[1:-1, 1:-1, :-2]) * dzfi[None, None, 1:-1])
    assert np.abs(dpsi - ref).max() < 1e-12 * np.abs(ref).max()
    assert abs(dpsi.sum()) < 1e-12 * np.abs(dpsi).sum()


def test_acdi_gamma_is_the_largest_cell_centred_speed():
    n = [6, 5, 4]
    u, v, w = O.halo(n, 0.3), O.halo(n, -0.4), O.halo(n, 1.2)
    g = L.or_acdi_set_gamma(I(n), C.c_double(2.0), P(u), P(v), P(w))
    assert g == pytest.approx(2.0 * 1.3, rel=1e-15)
    # one faster face: the cell-centred average of the two faces of cell i sees half of it
    u[3, 2, 2] = 5.0
    g = L.or_acdi_set_gamma(I(n), C.c_double(1.0), P(u), P(v), P(w))
    uc = 0.5 * (5.0 + 0.3)
    assert g == pytest.approx(np.sqrt(uc * uc + 0.16 + 1.44), rel=1e-15)


def test_ab2_updates_closed_form(rng):
    n = [5, 4, 3]
    f1, f2 = 0.3, -0.1
    psi = O.halo(n)
    psi[1:-1, 1:-1, 1:-1] = rng.uniform(0, 1, n)
    r, ro = np.asfortranarray(rng.standard_normal(n)), np.asfortranarray(rng.standard_normal(n))
    expect = psi[1:-1, 1:-1, 1:-1] + f1 * r + f2 * ro
    L.or_rk_psi_ab2(I(n), C.c_double(f1), C.c_double(f2), P(r), P(ro), P(psi))
    assert np.allclose(psi[1:-1, 1:-1, 1:-1], expect, rtol=1e-15, atol=1e-15)
    # momentum: with psi = psio = c the density rescale is 1 and the increment is divided by rho(c)
    c, rho12 = 0.25, (9.0, 1.0)
    rho = rho12[1] + (rho12[0] - rho12[1]) * c
    ph = O.halo(n, c)
    u0 = [np.asfortranarray(rng.standard_normal((n[0] + 2, n[1] + 2, n[2] + 2))) for _ in range(3)]
    uu = [a.copy(order="F") for a in u0]
    R = [np.asfortranarray(rng.standard_normal(n)) for _ in range(3)]
    Ro = [np.asfortranarray(rng.standard_normal(n)) for _ in range(3)]
    L.or_rk_mom_ab2(I(n), C.c_double(f1), C.c_double(f2), D(rho12), P(ph), P(ph), *[P(a) for a in R],
                    *[P(a) for a in Ro], *[P(a) for a in uu])
    for a0, a, r1, r0 in zip(u0, uu, R, Ro):
        assert np.allclose(a[1:-1, 1:-1, 1:-1], a0[1:-1, 1:-1, 1:-1] + (f1 * r1 + f2 * r0) / rho, rtol=1e-14,
                           atol=1e-14)
    before = [a.copy(order="F") for a in uu]
    L.or_rk_mom_oth(I(n), C.c_double(f1 + f2), D(rho12), P(ph), *[P(a) for a in R], *[P(a) for a in uu])
    for a0, a, r1 in zip(before, uu, R):
        assert np.allclose(a[1:-1, 1:-1, 1:-1], a0[1:-1, 1:-1, 1:-1] + (f1 + f2) * r1 / rho, rtol=1e-14, atol=1e-14)


def test_updatep_adds_the_correction(rng):
    n = [4, 3, 5]
    p = np.asfortranarray(rng.standard_normal((6, 5, 7)))
    pp = np.asfortranarray(rng.standard_normal((6, 5, 7)))
    expect = p + pp
    L.or_updatep(I(n), P(pp), P(p))
    assert np.array_equal(p, expect)


def test_fillps_and_correc_on_linear_fields():
    """u = a x, v = b y, w = c z (face positions): the divergence is a+b+c in every
    cell; p = g . x gives the constant correction -dt/rho0 * g."""
    n = [6, 5, 7]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.2, 1.0, 0.9], gr=1.3)
    xf = np.arange(n[0] + 2) * dl[0]
    yf = np.arange(n[1] + 2) * dl[1]
    a, b, c = 0.7, -1.1, 2.5
    u = np.asfortranarray(np.broadcast_to(a * xf[:, None, None], (n[0] + 2, n[1] + 2, n[2] + 2)).copy())
    v = np.asfortranarray(np.broadcast_to(b * yf[None, :, None], u.shape).copy())
    w = np.asfortranarray(np.broadcast_to(c * zf[None, None, :], u.shape).copy())
    p = O.halo(n)
    dt, rho0 = 0.01, 2.0
    L.or_fillps(I(n), D(dli), P(dzfi), C.c_double(1 / dt), C.c_double(rho0), P(u), P(v), P(w), P(p))
    assert np.allclose(p[1:-1, 1:-1, 1:-1], rho0 * (a + b + c) / dt, rtol=1e-10)
    xc = (np.arange(n[0] + 2) - 0.5) * dl[0]
    yc = (np.arange(n[1] + 2) - 0.5) * dl[1]
    g = (0.3, -0.2, 0.5)
    p = np.asfortranarray(g[0] * xc[:, None, None] + g[1] * yc[None, :, None] + g[2] * zc[None, None, :])
    u0, v0, w0 = u.copy(order="F"), v.copy(order="F"), w.copy(order="F")
    L.or_correc(I(n), D(dli), P(dzci), C.c_double(rho0), C.c_double(dt), P(p), P(u), P(v), P(w))
    s = (slice(1, -1),) * 3
    for new, old, gg in zip((u, v, w), (u0, v0, w0), g):
        assert np.allclose(new[s] - old[s], -dt / rho0 * gg, rtol=1e-9, atol=1e-14)


def test_taylor_green_forcing_is_solenoidal(rng):
    """The Taylor-Green body force has the same wavenumber in x and y, so on a
    grid with equal normalised spacings its staggered discrete divergence
    vanishes."""
    n = [16, 16, 8]
    l = [2.0, 2.0, 1.0]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, l)
    u, v = O.halo(n), O.halo(n)
    L.or_lscale_forcing_tg(I(n), C.c_double(0.5), C.c_double(0.1), D(l), D(dl), P(zc), P(zf), P(u), P(v))
    # cells 2..n (their i-1 and j-1 faces are forced too)
    div = (u[2:-1, 2:-1, 1:-1] - u[1:-2, 2:-1, 1:-1]) * dli[0] + (v[2:-1, 2:-1, 1:-1] - v[2:-1, 1:-2, 1:-1]) * dli[1]
    assert np.abs(u).max() > 0.01
    assert np.abs(div).max() < 1e-12 * np.abs(u).max() * dli[0]


def test_bulk_mean_mixture_and_face_average(rng):
    n = [6, 5, 4]
    l = [1.0, 1.0, 1.0]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, l)
    gvr = np.asfortranarray(dl[0] * dl[1] * dzf / (l[0] * l[1] * l[2]))
    psi = O.halo(n, 0.3)
    for idir in (1, 2, 3):
        m = L.or_bulk_mean_12_stag(I(n), idir, P(gvr), P(psi), D([5.0, 1.0]))
        assert m == pytest.approx(1.0 + 4.0 * 0.3, rel=1e-14)
    # psi varying along x only: the x-staggered mean averages neighbouring cells
    psi = O.halo(n)
    x = rng.uniform(0, 1, n[0] + 2)
    psi[...] = x[:, None, None]
    m = L.or_bulk_mean_12_stag(I(n), 1, P(gvr), P(psi), D([5.0, 1.0]))
    face = 0.5 * (x[2:] + x[1:-1])
    assert m == pytest.approx(np.mean(1.0 + 4.0 * face), rel=1e-13)


def test_chkdiv_closed_form():
    n = [5, 4, 6]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.0, 1.0, 1.0], gr=0.9)
    u = O.halo(n)
    u[...] = (np.arange(n[0] + 2) * dl[0])[:, None, None]
    z = O.halo(n)
    divtot, divmax = O.chkdiv(n, dli, dzfi, u, z, z)
    assert divmax == pytest.approx(1.0, rel=1e-12)
    assert divtot == pytest.approx(n[0] * n[1] * n[2], rel=1e-12)


@pytest.mark.parametrize("gtype", [1, 2, 3, 4])
def test_initgrid_clustering_properties(gtype):
    """Every clustering law maps [0,1] onto [0,lz] monotonically, the uniform
    grid is its limit for vanishing stretching, and the two-end and middle
    laws (1 and 4) are mirror-symmetric; 2 and 3 are mirror images."""
    nz, lz = 32, 2.0
    dzc, dzf, zc, zf = O.initgrid(gtype, nz, 1.5, lz)
    assert zf[nz] == pytest.approx(lz, rel=1e-12) and zf[0] == 0.0
    assert np.all(dzf[1:-1] > 0) and dzf[1:-1].sum() == pytest.approx(lz, rel=1e-12)
    assert np.allclose(zc[1:-1], 0.5 * (zf[1:-1] + zf[:-2]), rtol=1e-12)
    dz0 = O.initgrid(gtype, nz, 0.0, lz)[1]
    assert np.allclose(dz0[1:-1], lz / nz, rtol=1e-6)
    if gtype in (1, 4):
        assert np.allclose(dzf[1:-1], dzf[1:-1][::-1], rtol=1e-5)
        assert not np.allclose(dzf[1:-1], lz / nz, rtol=1e-3)
    if gtype == 2:
        d3 = O.initgrid(3, nz, 1.5, lz)[1]
        assert np.allclose(dzf[1:-1], d3[1:-1][::-1], rtol=1e-5)
