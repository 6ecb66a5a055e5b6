"""Pins of the oracle terms that the conservation and round-trip checks leave
open, each by a closed form or an asymptotic property that the mathematics
fixes, independently of the restated expressions:

- ACDI regularisation flux (src/acdi.f90:160-216): at the equilibrium profile
  psi = 1/(1+exp(-phi/eps)) with phi a signed distance and n its gradient, the
  diffusive flux eps*grad(psi) equals the sharpening flux
  1/4 (1 - tanh^2(phi/2eps)) n exactly, so the discrete flux and right-hand
  side vanish as O(h^2) under refinement;
- THINC/QQ face value (src/vof_thinc_qq.f90:66-116): for a planar profile the
  midpoint THINC reconstruction in the upwind cell evaluated at the face is
  the exact profile at the face, and the flux does not depend on the downwind
  cell;
- mom_xyz_oth (src/mom.f90:902-1008): with linear psi, kappa and pressures the
  Brackbill surface tension, gravity, body force and the extrapolated
  pressure split have closed forms at every face;
- chkdt (src/chkdt.f90:71-86): each limit alone in closed form (the viscous
  one is the FTCS bound dl^2/(6 nu) of the 3-D heat equation), and the
  combined limit is the positive root of x^2 - (C+V) x - (G^2+S^2) = 0;
- init2fl (src/two_fluid.f90:247-437, :503-513): the smooth step inverts to
  the signed distance through acdi_cmpt_phi's logit, for the film 'flm', the
  planar films 'bub1'/'drp1' and periodic droplets 'drp3' (images across the
  periodic boundary included), and the 'drp*' cases are one minus 'bub*'.

Every pin is a function of the oracle library. Its test passes on the oracle
and fails on liboracle_planted.so with each named planted mistake selected
(the PL_* list of oracle/fz_oracle.c): a restatement slip of that kind would
not go unnoticed."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

I, D, P, S = O.ivec, O.dvec, O.ptr, O.cstr


def planted(names):
    return pytest.mark.parametrize("mistake", names)


def grid(L, n, l, gr=0.0):
    arrs = [np.zeros(n[2] + 2) for _ in range(4)]
    L.or_initgrid(C.c_int(1), C.c_int(n[2]), C.c_double(gr), C.c_double(l[2]), *[P(a) for a in arrs])
    dzc, dzf, zc, zf = arrs
    dl = [l[d] / float(np.float32(n[d])) for d in range(3)]
    return dl, [1 / x for x in dl], dzc, dzf, zc, zf, 1 / dzc, 1 / dzf


def centres(n, dl, zc):
    x = (np.arange(n[0] + 2) - 0.5) * dl[0]
    y = (np.arange(n[1] + 2) - 0.5) * dl[1]
    return np.meshgrid(x, y, zc, indexing="ij")


# ---------------------------------------------------------------------------------------------
# ACDI: the regularisation flux vanishes at the equilibrium profile
def acdi_equilibrium(L, a, N, gr=0.0, seps=0.05):
    """max |flux| and max |dpsidt| for psi at equilibrium with the unit normal a
    (the axes along which a has a component get N cells, the others 4)."""
    a = np.asarray(a, float) / np.linalg.norm(a)
    n = [N if abs(a[d]) > 0 else 4 for d in range(3)]
    l = [1.0, 1.0, 1.0]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(L, n, l, gr)
    X, Y, Z = centres(n, dl, zc)
    phi = np.asfortranarray(a[0] * (X - 0.503) + a[1] * (Y - 0.497) + a[2] * (Z - 0.511))
    psi = np.asfortranarray(1.0 / (1.0 + np.exp(-phi / seps)))
    nrm = [np.asfortranarray(np.full(phi.shape, a[d])) for d in range(3)]
    zero = O.halo(n)
    dpsi = O.interior(n)
    fx, fy, fz = O.halo(n), O.halo(n), O.halo(n)
    L.or_acdi_transport_pf(I(n), D(dli), P(dzci), P(dzfi), C.c_double(1.0), C.c_double(seps), P(zero), P(zero),
                           P(zero), P(nrm[0]), P(nrm[1]), P(nrm[2]), P(phi), P(psi), P(dpsi), P(fx), P(fy), P(fz))
    inner = (slice(1, -1),) * 3
    flux = max(np.abs(f[inner]).max() for f in (fx, fy, fz))
    return flux, np.abs(dpsi).max()


ACDI_PROFILES = {"x": ((1, 0, 0), 0.0), "y": ((0, 1, 0), 0.0), "z_stretched": ((0, 0, 1), 1.2),
                 "oblique_xz": ((0.6, 0, 0.8), 0.0), "oblique_xy": ((-0.8, 0.6, 0), 0.0)}


def acdi_converges(L, a, gr):
    c = acdi_equilibrium(L, a, 64, gr)
    f = acdi_equilibrium(L, a, 128, gr)
    # second order: each halving of h cuts both residuals by about 4
    return c[0] / f[0] > 3.5 and c[1] / f[1] > 3.5 and f[0] < 1e-3, (c, f)


@pytest.mark.parametrize("profile", list(ACDI_PROFILES))
def test_acdi_regularisation_vanishes_at_equilibrium(profile):
    ok, res = acdi_converges(O.lib(), *ACDI_PROFILES[profile])
    assert ok, res


@planted(["ACDI_SHARP_SIGN", "ACDI_QUARTER", "ACDI_TANH_ARG", "ACDI_NORMAL_COMP", "ACDI_DZCI_M"])
def test_acdi_equilibrium_fails_on_planted(mistake):
    L = O.planted_lib(mistake)
    assert not all(acdi_converges(L, *p)[0] for p in ACDI_PROFILES.values())


# ---------------------------------------------------------------------------------------------
# THINC/QQ: the face value of a planar profile, and upwinding
def thinc_planar(L, axis, vsign, nsign, beta=1.0):
    """max |flux - vel*H(face)| / max|vel| for a planar profile along `axis`
    (phi in cell units, normal +-1) with velocities of one sign."""
    rng = np.random.default_rng(7 + axis)
    n = [6, 6, 6]
    n[axis] = 12
    dl = [1.0 / n[d] for d in range(3)]
    idx = np.indices((n[0] + 2, n[1] + 2, n[2] + 2))[axis].astype(float)  # cell centre i - 1/2 in cell units
    x0 = 6.3
    phi = np.asfortranarray(nsign * (idx - 0.5 - x0))
    nrm = [np.asfortranarray(rng.uniform(-1, 1, phi.shape)) for _ in range(3)]
    nrm[axis][...] = nsign
    vel = [np.asfortranarray(rng.uniform(0.2, 1.0, phi.shape) * rng.choice([-1, 1], phi.shape)) for _ in range(3)]
    vel[axis][...] = vsign * np.abs(vel[axis])
    dpsi, fx, fy, fz = O.interior(n), O.halo(n), O.halo(n), O.halo(n)
    dzfi = np.full(n[2] + 2, 1.0 / dl[2])
    L.or_vof_thinc_transport_psi(I(n), D([1 / dl[0], 1 / dl[1]] + [1 / dl[2]]), P(dzfi), C.c_double(beta),
                                 P(vel[0]), P(vel[1]), P(vel[2]), P(nrm[0]), P(nrm[1]), P(nrm[2]), P(phi), P(dpsi),
                                 P(fx), P(fy), P(fz))
    f = (fx, fy, fz)[axis]
    face = nsign * (idx - x0)  # the face i+1/2 of cell i, in cell units from the interface
    exact = vel[axis] / (1.0 + np.exp(-2.0 * beta * face))
    s = (slice(0, n[0] + 1), slice(0, n[1] + 1), slice(0, n[2] + 1))
    return np.abs(f[s] - exact[s]).max() / np.abs(vel[axis][s]).max()


THINC_CASES = [(ax, vs, ns) for ax in range(3) for vs in (1, -1) for ns in (1, -1)]


@pytest.mark.parametrize("axis,vsign,nsign", THINC_CASES)
def test_thinc_face_value_is_the_exact_profile_at_the_face(axis, vsign, nsign):
    assert thinc_planar(O.lib(), axis, vsign, nsign) < 1e-14


def thinc_downwind_independence(L):
    """Perturbing the downwind cell of an x face leaves its flux bitwise
    unchanged; perturbing the upwind cell changes it."""
    rng = np.random.default_rng(11)
    n = [8, 5, 5]
    shp = (n[0] + 2, n[1] + 2, n[2] + 2)
    base = dict(u=rng.uniform(-1, 1, shp), v=rng.uniform(-1, 1, shp), w=rng.uniform(-1, 1, shp),
                nx=rng.uniform(-1, 1, shp), ny=rng.uniform(-1, 1, shp), nz=rng.uniform(-1, 1, shp),
                phi=rng.uniform(-2, 2, shp))

    def flux_x(f):
        g = {k: np.asfortranarray(v) for k, v in f.items()}
        out = [O.interior(n), O.halo(n), O.halo(n), O.halo(n)]
        L.or_vof_thinc_transport_psi(I(n), D([1.0, 1.0, 1.0]), P(np.ones(n[2] + 2)), C.c_double(1.0), P(g["u"]),
                                     P(g["v"]), P(g["w"]), P(g["nx"]), P(g["ny"]), P(g["nz"]), P(g["phi"]),
                                     *[P(a) for a in out])
        return out[1]

    f0 = flux_x(base)
    ok = True
    for (i, j, k) in [(3, 2, 2), (4, 3, 1), (1, 1, 3), (6, 4, 4)]:
        up, down = (i, i + 1) if base["u"][i, j, k] > 0 else (i + 1, i)
        for field in ("phi", "nx", "ny", "nz"):
            f = {kk: vv.copy() for kk, vv in base.items()}
            f[field][down, j, k] += 0.37
            ok &= flux_x(f)[i, j, k] == f0[i, j, k]
        for field in ("phi", "nx"):
            f = {kk: vv.copy() for kk, vv in base.items()}
            f[field][up, j, k] += 0.37
            ok &= flux_x(f)[i, j, k] != f0[i, j, k]
    return ok


def test_thinc_flux_depends_on_the_upwind_cell_only():
    assert thinc_downwind_independence(O.lib())


@planted(["THINC_UPWIND", "THINC_OFFSET", "THINC_NORMAL_COMP"])
def test_thinc_pins_fail_on_planted(mistake):
    L = O.planted_lib(mistake)
    worst = max(thinc_planar(L, ax, vs, ns) for ax, vs, ns in THINC_CASES)
    assert worst > 1e-3 or not thinc_downwind_independence(L)


# ---------------------------------------------------------------------------------------------
# mom_xyz_oth: closed forms on linear fields
def mom_oth_linear(L, dt_r=0.37):
    """Relative error of mom_xyz_oth against the closed form on linear psi,
    kappa, p, pn, po (different slopes per field and direction), with gravity,
    a body force and a bulk density in every direction."""
    n = [6, 5, 8]
    l = [1.2, 0.9, 1.6]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(L, n, l)
    X, Y, Z = centres(n, dl, zc)
    lin = lambda c0, c: c0 + c[0] * X + c[1] * Y + c[2] * Z
    cpsi, ck = np.array([0.21, -0.17, 0.12]), np.array([3.0, -2.0, 1.5])
    gp, gn, go = np.array([2.0, -1.0, 0.5]), np.array([-0.7, 1.3, 2.2]), np.array([1.1, 0.4, -1.9])
    psi = np.asfortranarray(lin(0.4, cpsi))
    kappa = np.asfortranarray(lin(-5.0, ck))
    p, pn, po = (np.asfortranarray(lin(0.3, g)) for g in (gp, gn, go))
    rho12, sigma, rho0 = (7.0, 1.5), 0.8, 1.5
    bforce, gacc, rho_av = [0.3, -0.2, 0.1], [0.5, -0.25, -9.0], [2.0, 3.0, 4.0]
    out = [O.interior(n) for _ in range(3)]
    L.or_mom_xyz_oth(I(n), D(dli), P(dzci), P(dzfi), C.c_double(dt_r), D(rho12), D(bforce), D(gacc),
                     C.c_double(sigma), C.c_double(rho0), D(rho_av), P(p), P(psi), P(kappa), P(pn), P(po),
                     *[P(a) for a in out])
    xc, yc = (np.arange(1, n[0] + 1) - 0.5) * dl[0], (np.arange(1, n[1] + 1) - 0.5) * dl[1]
    zi = zc[1:-1]
    worst = 0.0
    for d in range(3):
        # face coordinates of the d-face above each interior cell
        fx = xc + (0.5 * dl[0] if d == 0 else 0.0)
        fy = yc + (0.5 * dl[1] if d == 1 else 0.0)
        fz = 0.5 * (zc[1:-1] + zc[2:]) if d == 2 else zi
        FX, FY, FZ = np.meshgrid(fx, fy, fz, indexing="ij")
        psi_f = 0.4 + cpsi[0] * FX + cpsi[1] * FY + cpsi[2] * FZ
        kap_f = -5.0 + ck[0] * FX + ck[1] * FY + ck[2] * FZ
        rho_f = rho12[1] + (rho12[0] - rho12[1]) * psi_f
        ge = (1.0 + dt_r) * gn[d] - dt_r * go[d]
        surf = sigma * 2.0 / (rho12[0] + rho12[1]) * kap_f * rho_f * cpsi[d]  # Brackbill normalisation
        ref = bforce[d] + (rho_f - rho_av[d]) * gacc[d] + surf - (ge + rho_f / rho0 * (gp[d] - ge))
        worst = max(worst, np.abs(out[d] - ref).max() / np.abs(ref).max())
    return worst


def test_mom_oth_closed_form_on_linear_fields():
    assert mom_oth_linear(O.lib()) < 1e-12


@planted(["OTH_BRACKBILL", "OTH_KAPPA_FACE", "OTH_EXTRAP", "OTH_SPLIT", "OTH_SURF_FACTOR"])
def test_mom_oth_closed_form_fails_on_planted(mistake):
    assert mom_oth_linear(O.planted_lib(mistake)) > 1e-6


# ---------------------------------------------------------------------------------------------
# chkdt: each limit alone, and their combination
def chkdt(L, n, dl, dzci, dzfi, u, rho12, mu12, sigma, gacc, track=0, gam=0.0, seps=0.0):
    z = O.halo(n)
    uu = O.halo(n, u)
    return L.or_chkdt(I(n), D(dl), P(dzci), P(dzfi), 1, track, 0, D(mu12), D(rho12), C.c_double(sigma), D(gacc),
                      P(uu), P(z), P(z), C.c_double(gam), C.c_double(seps))


def chkdt_errors(L):
    """Relative errors of chkdt against the closed forms; the grid's smallest
    spacing is a z cell of the stretched grid."""
    n = [8, 8, 16]
    l = [1.0, 1.0, 1.0]
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(L, n, l, gr=1.5)
    dlmin = min(dl[0], dl[1], (1.0 / dzfi).min())
    assert dlmin < 0.8 * dl[0]
    rho12, mu12 = (3.0, 0.5), (0.2, 0.4)
    nu = max(m / r for m, r in zip(mu12, rho12))
    errs = []
    # viscous alone: the explicit-diffusion (FTCS) bound of the 3-D heat equation
    dt = chkdt(L, n, dl, dzci, dzfi, 0.0, rho12, mu12, 0.0, [0, 0, 0])
    errs.append(dt * 6.0 * nu / dlmin ** 2 - 1.0)
    # capillary alone: sqrt(rho_min dl^3 / sigma)
    dt = chkdt(L, n, dl, dzci, dzfi, 0.0, rho12, (0.0, 0.0), 0.7, [0, 0, 0])
    errs.append(dt / np.sqrt(min(rho12) * dlmin ** 3 / 0.7) - 1.0)
    # gravity alone: dl/|g|max
    dt = chkdt(L, n, dl, dzci, dzfi, 0.0, rho12, (0.0, 0.0), 0.0, [0.3, -2.5, 1.0])
    errs.append(dt * 2.5 / dlmin - 1.0)
    # all together with a uniform flow U: 1/dt is the positive root of x^2 - (C+V) x - (G^2 + S^2)
    U = 1.7
    dt = chkdt(L, n, dl, dzci, dzfi, U, rho12, mu12, 0.7, [0.3, -2.5, 1.0])
    Cc, V = U / dl[0], nu * 6.0 / dlmin ** 2
    G, Sk = 2.5 / dlmin, np.sqrt(0.7 / min(rho12) / dlmin ** 3)
    x = 1.0 / dt
    errs.append((x * x - (Cc + V) * x - (G * G + Sk * Sk)) / (x * x))
    # the ACDI limit dl^2/(gam eps) when it is the smallest
    dt = chkdt(L, n, dl, dzci, dzfi, 0.0, rho12, (1e-9, 1e-9), 0.0, [0, 0, 0], track=1, gam=4.0, seps=0.3)
    errs.append(dt * 4.0 * 0.3 / dlmin ** 2 - 1.0)
    return max(abs(e) for e in errs)


def test_chkdt_limits_in_closed_form():
    assert chkdt_errors(O.lib()) < 1e-12


@planted(["DT_VISC", "DT_CAP_RHO", "DT_GRAV", "DT_DLMIN", "DT_COMBINE"])
def test_chkdt_closed_forms_fail_on_planted(mistake):
    assert chkdt_errors(O.planted_lib(mistake)) > 1e-3


# ---------------------------------------------------------------------------------------------
# init2fl: the smooth step inverts to the signed distance
def init_psi(L, inipsi, n, l, spheres, cbcpsi, seps, gr=0.0):
    case = dict(ng=n, l=l, cbcvel=["P"] * 18, cbcpre=list(cbcpsi), cbcpsi=list(cbcpsi), cbcnor=["P"] * 18,
                inipsi=inipsi, spheres=spheres)
    prm = O.OrParams.from_case(case)
    dl, dli, dzc, dzf, zc, zf, dzci, dzfi = grid(L, n, l, gr)
    psi = O.halo(n)
    L.or_init2fl(C.byref(prm), C.c_double(seps), P(zc), P(psi))
    return psi, dl, zc


def logit_distance(psi, seps):
    """acdi_cmpt_phi's inverse of the smooth step, eps*log(psi/(1-psi)) (infinite where psi is 0 or 1)."""
    with np.errstate(divide="ignore"):
        return seps * np.log(psi / (1.0 - psi))


def init2fl_errors(L):
    seps = 0.02
    errs = []
    band = lambda d: np.abs(d) < 6 * seps
    # 'flm': psi = H((z - lz/2)/2eps) on a stretched grid
    n, l = [4, 4, 40], [1.0, 1.0, 2.0]
    psi, dl, zc = init_psi(L, "flm", n, l, [], "PPPPNN", seps, gr=1.0)
    d = np.broadcast_to(zc[None, None, :] - 1.0, psi.shape)[1:-1, 1:-1, 1:-1]
    ph = logit_distance(psi[1:-1, 1:-1, 1:-1], seps)
    errs.append(np.abs(ph - d)[band(d)].max() / seps)
    # 'bub1': a gas film |z - zc| < r (psi ~ 0 inside), 'drp1' its complement
    n, l = [4, 4, 64], [1.0, 1.0, 1.0]
    sph = [(0.5, 0.5, 0.43, 0.12)]
    p1, dl, zc = init_psi(L, "bub1", n, l, sph, "PPPPNN", seps)
    p2, _, _ = init_psi(L, "drp1", n, l, sph, "PPPPNN", seps)
    d = np.broadcast_to(np.abs(zc[None, None, :] - 0.43) - 0.12, p1.shape)[1:-1, 1:-1, 1:-1]
    errs.append(np.abs(logit_distance(p1[1:-1, 1:-1, 1:-1], seps) - d)[band(d)].max() / seps)
    errs.append(np.abs(p2 + p1 - 1.0)[1:-1, 1:-1, 1:-1].max())
    # 'drp3': a droplet across the periodic x and y boundaries of a unit cube;
    # its distance field seen through the nearest periodic image
    n, l = [32, 32, 32], [1.0, 1.0, 1.0]
    sph = [(0.06, 0.95, 0.5, 0.2)]
    p3, dl, zc = init_psi(L, "drp3", n, l, sph, "PPPPPP", seps)
    X, Y, Z = centres(n, dl, zc)
    dx = np.minimum(np.abs(X - 0.06), 1.0 - np.abs(X - 0.06))
    dy = np.minimum(np.abs(Y - 0.95), 1.0 - np.abs(Y - 0.95))
    d = (np.sqrt(dx ** 2 + dy ** 2 + (Z - 0.5) ** 2) - 0.2)[1:-1, 1:-1, 1:-1]
    errs.append(np.abs(-logit_distance(p3[1:-1, 1:-1, 1:-1], seps) - d)[band(d)].max() / seps)
    # 'drp3' is one minus 'bub3' of the same spheres
    pb, _, _ = init_psi(L, "bub3", n, l, sph, "PPPPPP", seps)
    errs.append(np.abs(pb + p3 - 1.0)[1:-1, 1:-1, 1:-1].max())
    return max(errs)


def test_init2fl_profiles_invert_to_the_signed_distance():
    assert init2fl_errors(O.lib()) < 1e-8


@planted(["INIT_WIDTH", "INIT_PERIODIC", "INIT_FLM_INDEX"])
def test_init2fl_pins_fail_on_planted(mistake):
    assert init2fl_errors(O.planted_lib(mistake)) > 1e-3


def test_vof_cases_smooth_their_initial_droplets_with_half_a_cell():
    """A THINC/QQ case (psi_thickness_factor defaults to 0.50 for it,
    src/param.f90:148-151) takes eps = 0.5*dl from acdi_set_epsilon, which
    src/main.f90:301 calls in every build, for its 'drp3' droplets."""
    n = [24, 24, 24]
    spec = dict(ng=n, l=[1.0, 1.0, 1.0], cbcvel=["P"] * 18, cbcpre=["P"] * 6, cbcpsi=["P"] * 6, cbcnor=["P"] * 18,
                inipsi="drp3", spheres=[(0.5, 0.5, 0.5, 0.3)], vof=True, rho12=[2.0, 1.0], mu12=[0.01, 0.01])
    c = O.Case(spec)
    c.init_fields()
    # 0.5 dl, up to the single-precision face coordinates of initgrid (src/initgrid.f90:38)
    seps = c.scalar("seps")
    assert seps == pytest.approx(0.5 / 24, rel=1e-6)
    psi = c.field("psi")[1:-1, 1:-1, 1:-1]
    x = (np.arange(1, 25) - 0.5) / 24
    zc = O.initgrid(1, 24, 0.0, 1.0)[2][1:-1]  # the case's z centres (single-precision face coordinates)
    X, Y, Z = np.meshgrid(x, x, zc, indexing="ij")
    d = np.sqrt((X - 0.5) ** 2 + (Y - 0.5) ** 2 + (Z - 0.5) ** 2) - 0.3
    band = np.abs(d) < 6 * seps
    assert np.abs(-logit_distance(psi, seps) - d)[band].max() < 1e-8 * seps
