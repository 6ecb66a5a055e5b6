"""Pins the oracle's stencil routines by closed forms and invariants the
mathematics fixes (the restated formulas are not retyped here):
mom_xyz_ad (src/mom.f90:505-851), mom_xyz_oth (:853-1039), acdi
(src/acdi.f90), THINC (src/vof_thinc_qq.f90), Youngs (src/two_fluid.f90),
fillps/correc (src/fillps.f90, src/correc.f90), boundary conditions
(src/bound.f90), chkdt (src/chkdt.f90) and init2fl (src/two_fluid.f90)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

L = O.lib()
I, D, P, S = O.ivec, O.dvec, O.ptr, O.cstr


def grid(n, l, gr=0.0):
    dzc, dzf, zc, zf = O.initgrid(1, n[2], gr, l[2])
    dl = [l[d] / float(np.float32(n[d])) for d in range(3)]
    return dl, [1 / x for x in dl], dzc, dzf, zc, zf, 1 / dzc, 1 / dzf


def periodic_fill(a):
    a[0] = a[-2]
    a[-1] = a[1]
    a[:, 0] = a[:, -2]
    a[:, -1] = a[:, 1]
    a[:, :, 0] = a[:, :, -2]
    a[:, :, -1] = a[:, :, 1]
    return a


def test_mom_ad_uniform_flow_has_zero_rhs():
    n = [8, 6, 5]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1])
    u, v, w, psi = O.halo(n, 0.3), O.halo(n, -0.7), O.halo(n, 1.1), O.halo(n, 0.4)
    fx, fy, fz = O.halo(n, 0.4 * 0.3), O.halo(n, 0.4 * -0.7), O.halo(n, 0.4 * 1.1)
    du, dv, dw = O.interior(n), O.interior(n), O.interior(n)
    L.or_mom_xyz_ad(I(n), D(dli), P(dzci), P(dzfi), D([3.0, 1.0]), D([0.2, 0.1]), P(fx), P(fy), P(fz), P(u), P(v),
                    P(w), P(psi), P(du), P(dv), P(dw))
    assert max(np.abs(du).max(), np.abs(dv).max(), np.abs(dw).max()) < 1e-12


def test_mom_ad_single_phase_is_advection_plus_laplacian_plus_graddiv(rng):
    """rho, mu constant: R = -div(rho u u) + mu (lap u + grad div u), written here
    in array form on the MAC grid independently of the per-point restatement."""
    n = [10, 8, 8]  # n3 a power of two: the single-precision face coordinate k/n3 is exact
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.0, 0.8, 0.6])
    rho, mu = 2.5, 0.03
    u, v, w = [np.asfortranarray(rng.standard_normal((n[0] + 2, n[1] + 2, n[2] + 2))) for _ in range(3)]
    psi = O.halo(n, 0.5)
    fx, fy, fz = O.halo(n), O.halo(n), O.halo(n)
    du, dv, dw = O.interior(n), O.interior(n), O.interior(n)
    L.or_mom_xyz_ad(I(n), D(dli), P(dzci), P(dzfi), D([rho, rho]), D([mu, mu]), P(fx), P(fy), P(fz), P(u), P(v),
                    P(w), P(psi), P(du), P(dv), P(dw))
    dx, dy, dz = dl[0], dl[1], dl[2]
    c = slice(1, -1)

    def sh(a, di, dj, dk):
        return a[1 + di:a.shape[0] - 1 + di, 1 + dj:a.shape[1] - 1 + dj, 1 + dk:a.shape[2] - 1 + dk]

    # divergence at cell centres, also one cell outside the interior where needed
    def div_at(di, dj, dk):
        return ((sh(u, di, dj, dk) - sh(u, di - 1, dj, dk)) / dx + (sh(v, di, dj, dk) - sh(v, di, dj - 1, dk)) / dy
                + (sh(w, di, dj, dk) - sh(w, di, dj, dk - 1)) / dz)

    def lap(a):
        return ((sh(a, 1, 0, 0) - 2 * sh(a, 0, 0, 0) + sh(a, -1, 0, 0)) / dx ** 2 +
                (sh(a, 0, 1, 0) - 2 * sh(a, 0, 0, 0) + sh(a, 0, -1, 0)) / dy ** 2 +
                (sh(a, 0, 0, 1) - 2 * sh(a, 0, 0, 0) + sh(a, 0, 0, -1)) / dz ** 2)

    # u at (i+1/2,j,k): fluxes through the faces of its control volume
    adv_u = -(((sh(u, 0, 0, 0) + sh(u, 1, 0, 0)) ** 2 - (sh(u, -1, 0, 0) + sh(u, 0, 0, 0)) ** 2) / (4 * dx)
              + ((sh(v, 0, 0, 0) + sh(v, 1, 0, 0)) * (sh(u, 0, 0, 0) + sh(u, 0, 1, 0))
                 - (sh(v, 0, -1, 0) + sh(v, 1, -1, 0)) * (sh(u, 0, 0, 0) + sh(u, 0, -1, 0))) / (4 * dy)
              + ((sh(w, 0, 0, 0) + sh(w, 1, 0, 0)) * (sh(u, 0, 0, 0) + sh(u, 0, 0, 1))
                 - (sh(w, 0, 0, -1) + sh(w, 1, 0, -1)) * (sh(u, 0, 0, 0) + sh(u, 0, 0, -1))) / (4 * dz)) * rho
    ref_u = adv_u + mu * (lap(u) + (div_at(1, 0, 0) - div_at(0, 0, 0)) / dx)
    adv_w = -(((sh(u, 0, 0, 0) + sh(u, 0, 0, 1)) * (sh(w, 0, 0, 0) + sh(w, 1, 0, 0))
               - (sh(u, -1, 0, 0) + sh(u, -1, 0, 1)) * (sh(w, 0, 0, 0) + sh(w, -1, 0, 0))) / (4 * dx)
              + ((sh(v, 0, 0, 0) + sh(v, 0, 0, 1)) * (sh(w, 0, 0, 0) + sh(w, 0, 1, 0))
                 - (sh(v, 0, -1, 0) + sh(v, 0, -1, 1)) * (sh(w, 0, 0, 0) + sh(w, 0, -1, 0))) / (4 * dy)
              + ((sh(w, 0, 0, 0) + sh(w, 0, 0, 1)) ** 2 - (sh(w, 0, 0, -1) + sh(w, 0, 0, 0)) ** 2) / (4 * dz)) * rho
    ref_w = adv_w + mu * (lap(w) + (div_at(0, 0, 1) - div_at(0, 0, 0)) / dz)
    assert np.abs(du - ref_u).max() < 1e-11 * np.abs(ref_u).max()
    assert np.abs(dw - ref_w).max() < 1e-11 * np.abs(ref_w).max()


def test_mom_oth_pressure_gradient_and_gravity(rng):
    n = [6, 5, 4]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1])
    x = (np.arange(n[0] + 2) - 0.5) * dl[0]
    p = np.asfortranarray(np.broadcast_to((3.0 * x)[:, None, None], (n[0] + 2, n[1] + 2, n[2] + 2)).copy())
    psi = np.asfortranarray(rng.uniform(0, 1, p.shape))
    kappa = O.halo(n)
    du, dv, dw = O.interior(n), O.interior(n), O.interior(n)
    rho12 = [10.0, 2.0]
    L.or_mom_xyz_oth(I(n), D(dli), P(dzci), P(dzfi), C.c_double(1.0), D(rho12), D([0, 0, 0]), D([0, 0, -9.0]),
                     C.c_double(0.0), C.c_double(2.0), D([0, 0, 1.5]), P(p), P(psi), P(kappa), P(p), P(p), P(du),
                     P(dv), P(dw))
    assert np.allclose(du, -3.0, rtol=1e-12)
    assert np.abs(dv).max() < 1e-12
    rz = 2.0 + 8.0 * 0.5 * (psi[1:-1, 1:-1, 2:] + psi[1:-1, 1:-1, 1:-1])
    assert np.allclose(dw, (rz - 1.5) * -9.0, rtol=1e-12)


def test_acdi_rhs_is_divergence_of_its_fluxes(rng):
    n = [8, 7, 6]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1], gr=1.0)
    psi = periodic_fill(np.asfortranarray(rng.uniform(0, 1, (n[0] + 2, n[1] + 2, n[2] + 2))))
    phi = O.halo(n)
    seps = 0.05
    L.or_acdi_cmpt_phi(I(n), C.c_double(seps), P(psi), P(phi))
    u, v, w = [periodic_fill(np.asfortranarray(rng.standard_normal(psi.shape))) for _ in range(3)]
    nrm = [periodic_fill(np.asfortranarray(rng.standard_normal(psi.shape))) for _ in range(3)]
    dpsi = O.interior(n)
    fx, fy, fz = O.halo(n), O.halo(n), O.halo(n)
    L.or_acdi_transport_pf(I(n), D(dli), P(dzci), P(dzfi), C.c_double(0.7), C.c_double(seps), P(u), P(v), P(w),
                           P(nrm[0]), P(nrm[1]), P(nrm[2]), P(phi), P(psi), P(dpsi), P(fx), P(fy), P(fz))
    ref = -((fx[2:-1, 2:-1, 2:-1] - fx[1:-2, 2:-1, 2:-1]) * dli[0] + (fy[2:-1, 2:-1, 2:-1] - fy[2:-1, 1:-2, 2:-1]) * dli[1]
            + (fz[2:-1, 2:-1, 2:-1] - fz[2:-1, 2:-1, 1:-2]) * dzfi[None, None, 2:-1])
    assert np.abs(dpsi[1:, 1:, 1:] - ref).max() < 1e-12 * np.abs(ref).max()


def test_acdi_phi_inverts_the_logistic_profile():
    n = [20, 1, 1]
    seps = 0.03
    x = np.linspace(-0.2, 0.2, n[0] + 2)
    psi = O.halo(n)
    psi[:, :, :] = (1.0 / (1.0 + np.exp(-x / seps)))[:, None, None]
    phi = O.halo(n)
    L.or_acdi_cmpt_phi(I(n), C.c_double(seps), P(psi), P(phi))
    assert np.abs(phi[:, 1, 1] - x).max() < 1e-9


def test_youngs_linear_field_gives_exact_normal_and_zero_curvature():
    n = [7, 6, 5]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1.0, 0.9, 0.8])
    a = np.array([0.3, -0.5, 0.8])
    X = (np.arange(n[0] + 2) - 0.5) * dl[0]
    Y = (np.arange(n[1] + 2) - 0.5) * dl[1]
    phi = np.asfortranarray(a[0] * X[:, None, None] + a[1] * Y[None, :, None] + a[2] * zc[None, None, :])
    nx, ny, nz, ka = O.halo(n), O.halo(n), O.halo(n), O.halo(n)
    L.or_cmpt_norm_curv_youngs(I(n), D(dli), P(dzci), P(dzfi), P(phi), P(nx), P(ny), P(nz), P(ka))
    e = a / np.linalg.norm(a)
    for comp, val in zip((nx, ny, nz), e):
        assert np.abs(comp[1:-1, 1:-1, 1:-1] - val).max() < 1e-13
    assert np.abs(ka[1:-1, 1:-1, 1:-1]).max() < 1e-10


def test_youngs_sphere_curvature():
    n = [40, 40, 40]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1])
    X = (np.arange(n[0] + 2) - 0.5) * dl[0]
    R = 0.3
    r = np.sqrt((X[:, None, None] - 0.5) ** 2 + (X[None, :, None] - 0.5) ** 2 + (zc[None, None, :] - 0.5) ** 2)
    phi = np.asfortranarray(r - R)
    nx, ny, nz, ka = O.halo(n), O.halo(n), O.halo(n), O.halo(n)
    L.or_cmpt_norm_curv_youngs(I(n), D(dli), P(dzci), P(dzfi), P(phi), P(nx), P(ny), P(nz), P(ka))
    band = np.abs(phi[1:-1, 1:-1, 1:-1]) < 0.5 * dl[0]
    k = ka[1:-1, 1:-1, 1:-1][band]
    assert np.abs(k.mean() - (-2.0 / R)) < 0.02 * (2.0 / R)


def test_thinc_fluxes_bounded_by_velocity(rng):
    n = [6, 6, 6]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1])
    psi = np.asfortranarray(rng.uniform(0, 1, (8, 8, 8)))
    phi = O.halo(n)
    L.or_vof_thinc_cmpt_phi(I(n), C.c_double(1.0), P(psi), P(phi))
    u, v, w = [np.asfortranarray(rng.standard_normal(psi.shape)) for _ in range(3)]
    nrm = [np.asfortranarray(rng.uniform(-1, 1, psi.shape)) for _ in range(3)]
    dpsi, fx, fy, fz = O.interior(n), O.halo(n), O.halo(n), O.halo(n)
    L.or_vof_thinc_transport_psi(I(n), D(dli), P(dzfi), C.c_double(1.0), P(u), P(v), P(w), P(nrm[0]), P(nrm[1]),
                                 P(nrm[2]), P(phi), P(dpsi), P(fx), P(fy), P(fz))
    for f, vel in ((fx, u), (fy, v), (fz, w)):
        s = (slice(0, n[0] + 1), slice(0, n[1] + 1), slice(0, n[2] + 1))
        assert np.all(np.abs(f[s]) <= np.abs(vel[s]) * (1 + 1e-15))
        assert np.all(f[s] * vel[s] >= 0)


@pytest.mark.parametrize("cbc", ["PPPPNN", "NNPPNN", "NNNNDD", "PPDDPP"])
def test_projection_is_divergence_free(cbc, rng):
    # With periodic z the reference corrects both copies of the periodic face, w(0) with dzci(0) and
    # w(n3) with dzci(n3), then copies one over the other; dzc(0) and dzc(n3) agree only when the
    # single-precision face coordinate (k-0.)/(1.*n) of src/initgrid.f90:38 is exact, i.e. n3 = 2^m
    # on a uniform grid. Walls in z are checked on a stretched grid.
    periodic_z = cbc[4:] == "PP"
    n = [16, 12, 16 if periodic_z else 20]
    l = [1.0, 1.0, 2.0]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, l, gr=0.0 if periodic_z else 0.8)
    cbcvel = []
    for comp in range(3):
        for d in range(3):
            pair = cbc[2 * d:2 * d + 2]
            cbcvel += list(pair) if pair == "PP" else ["D", "D"]
    u, v, w = [O.halo(n) for _ in range(3)]
    for a in (u, v, w):
        a[1:-1, 1:-1, 1:-1] = rng.standard_normal(n)
    bcv = D([0.0] * 18)
    L.or_bounduvw(I(n), S(cbc), S(cbcvel), bcv, 0, D(dl), P(dzc), P(dzf), P(u), P(v), P(w))
    div0 = O.chkdiv(n, dli, dzfi, u, v, w)[1]
    sol = O.Solver(n, dli, dzci, dzfi, list(cbc))
    p = O.halo(n)
    dt, rho0 = 1e-2, 1.0
    L.or_fillps(I(n), D(dli), P(dzfi), C.c_double(1 / dt), C.c_double(rho0), P(u), P(v), P(w), P(p))
    sol.solve(p)
    L.or_boundp(I(n), S(cbc), S(cbc), D([0.0] * 6), D(dl), P(dzc), P(p))
    L.or_correc(I(n), D(dli), P(dzci), C.c_double(rho0), C.c_double(dt), P(p), P(u), P(v), P(w))
    L.or_bounduvw(I(n), S(cbc), S(cbcvel), bcv, 1, D(dl), P(dzc), P(dzf), P(u), P(v), P(w))
    div1 = O.chkdiv(n, dli, dzfi, u, v, w)[1]
    assert div1 < 1e-11 * div0, (div0, div1)


def test_boundary_conditions_semantics(rng):
    n = [5, 4, 6]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1], gr=1.0)
    p = rng.standard_normal((7, 6, 8)).copy(order="F")
    bc = [0.5, -0.25, 0.0, 0.0, 2.0, 3.0]
    L.or_boundp(I(n), S("NNPPDN"), S("NDPPDN"), D(bc), D(dl), P(dzc), P(p))
    c = slice(1, -1)  # edge rows are overwritten by the later directions
    assert np.allclose((p[1, c, c] - p[0, c, c]) / dl[0], 0.5)           # N at x0: dp/dx = bc
    assert np.allclose(0.5 * (p[-1, c, c] + p[-2, c, c]), -0.25)          # D at x1 (centred): mean = bc
    assert np.allclose(p[c, 0, c], p[c, -2, c]) and np.allclose(p[c, -1, c], p[c, 1, c])  # periodic y
    assert np.allclose(0.5 * (p[:, :, 0] + p[:, :, 1]), 2.0)             # D at z0 (applied last)
    assert np.allclose((p[:, :, -1] - p[:, :, -2]) / dzc[n[2]], 3.0)     # N at z1
    # staggered (face-normal) Dirichlet puts the value on the boundary face
    u = rng.standard_normal((7, 6, 8)).copy(order="F")
    v, w = u.copy(order="F"), u.copy(order="F")
    cbcvel = list("DDPPNN") + list("NNPPNN") + list("NNPPDD")
    bcvel = [0.0] * 18
    bcvel[1] = 0.7
    L.or_bounduvw(I(n), S("NNPPNN"), S(cbcvel), D(bcvel), 0, D(dl), P(dzc), P(dzf), P(u), P(v), P(w))
    assert np.allclose(u[0], 0.0) and np.allclose(u[n[0]], 0.7)
    assert np.allclose(w[:, :, 0], 0.0) and np.allclose(w[:, :, n[2]], 0.0)


def test_chkdt_uniform_flow_closed_form():
    n = [8, 8, 8]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(n, [1, 1, 1])
    u, v, w = O.halo(n, 2.0), O.halo(n, 0.0), O.halo(n, 0.0)
    dt = L.or_chkdt(I(n), D(dl), P(dzci), P(dzfi), 0, 1, 1, D([1, 1]), D([1, 1]), C.c_double(0), D([0, 0, 0]),
                    P(u), P(v), P(w), C.c_double(0), C.c_double(0))
    assert dt == pytest.approx(1.0 / (2.0 * 8), rel=1e-15)


def test_init2fl_sphere_volume():
    case = dict(ng=[32, 32, 64], l=[1, 1, 2], cbcvel=["P"] * 18, cbcpre=list("PPPPNN"), cbcpsi=list("PPPPNN"),
                cbcnor=["P"] * 18, inipsi="bub3", spheres=[(0.5, 0.5, 0.5, 0.25)], psi_thickness_factor=0.51)
    c = O.Case(case)
    c.init_fields()
    psi = c.field("psi")[1:-1, 1:-1, 1:-1]
    vol_gas = (1 - psi).sum() * (1 / 32) ** 3
    # volume of the smooth profile psi = (1 + tanh((r-R)/2eps))/2, eps = 0.51 dx (src/two_fluid.f90:503-513)
    eps, R = 0.51 / 32, 0.25
    s = np.linspace(0.0, 0.5, 200001)
    prof = 0.5 * (1 - np.tanh((s - R) / (2 * eps)))
    vol_ref = np.trapezoid(prof * 4 * np.pi * s ** 2, s)
    assert abs(vol_gas - vol_ref) < 2e-3 * vol_ref
    assert psi.min() >= 0 and psi.max() <= 1
