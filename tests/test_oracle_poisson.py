"""Pins the oracle's Poisson operator set-up and solve (initsolver,
src/initsolver.f90:15-228; solver, src/solver.f90:19-177) by discrete
eigenfunctions, by a dense solve of the assembled 7-point operator the
FFT solver inverts (src/debug.f90:104-107) and by the eps-regularised null
mode the reference leaves in p for singular (all periodic/Neumann) BCs."""
import numpy as np
import pytest

import oracle as O


def grid(n, l, gtype=1, gr=0.0):
    dzc, dzf, zc, zf = O.initgrid(gtype, n[2], gr, l[2])
    dl = [l[d] / float(np.float32(n[d])) for d in range(3)]
    return dl, [1.0 / x for x in dl], 1.0 / dzc, 1.0 / dzf


def eig1d(bc, n, dl):
    """1-D eigenvectors (columns) and eigenvalues of the cell-centred second difference."""
    i = np.arange(1, n + 1) - 0.5
    k = np.arange(n)
    if bc == "PP":
        V = np.cos(2 * np.pi * np.outer(i, k) / n + 0.3)  # any phase: cos and sin share the eigenvalue
        lam = -2 * (1 - np.cos(2 * np.pi * k / n))
    elif bc == "NN":
        V = np.cos(np.pi * np.outer(i, k) / n)
        lam = -2 * (1 - np.cos(np.pi * k / n))
    elif bc == "DD":
        V = np.sin(np.pi * np.outer(i, k + 1) / n)
        lam = -2 * (1 - np.cos(np.pi * (k + 1) / n))
    elif bc == "ND":
        V = np.cos(np.pi * np.outer(i, 2 * k + 1) / (2 * n))
        lam = -2 * (1 - np.cos(np.pi * (2 * k + 1) / (2 * n)))
    else:  # DN
        V = np.sin(np.pi * np.outer(i, 2 * k + 1) / (2 * n))
        lam = -2 * (1 - np.cos(np.pi * (2 * k + 1) / (2 * n)))
    return V, lam / dl ** 2


@pytest.mark.parametrize("cbc", ["PPPPNN", "NNPPNN", "DDNNDD", "PPDDPP", "NDDNNN", "DNPPDD", "NNNNNN", "PPPPPP"])
def test_discrete_eigenfunction_solved_exactly(cbc, rng):
    n = [12, 10, 8]
    l = [1.0, 0.7, 1.3]
    dl, dli, dzci, dzfi = grid(n, l)
    S = O.Solver(n, dli, dzci, dzfi, list(cbc))
    V = []
    L = []
    for d in range(3):
        v, lam = eig1d(cbc[2 * d:2 * d + 2], n[d], dl[d])
        V.append(v)
        L.append(lam)
    for trial in range(4):
        kx, ky, kz = [int(rng.integers(0, n[d])) for d in range(3)]
        lam = L[0][kx] + L[1][ky] + L[2][kz]
        if abs(lam) < 1e-9:
            continue
        pe = np.einsum("i,j,k->ijk", V[0][:, kx], V[1][:, ky], V[2][:, kz])
        f = O.halo(n)
        f[1:-1, 1:-1, 1:-1] = lam * pe
        p = f.copy(order="F")
        S.solve(p)
        pi = p[1:-1, 1:-1, 1:-1]
        singular = "D" not in cbc
        if singular:
            # the reference's null mode adds a roundoff-set constant (src/solver.f90:163-167)
            err = np.abs((pi - pi.mean()) - (pe - pe.mean())).max()
            assert err <= 1e-13 * max(1.0, np.abs(pi).max()), (cbc, kx, ky, kz, err)
        else:
            assert np.abs(pi - pe).max() <= 1e-12 * np.abs(pe).max(), (cbc, kx, ky, kz)


def dense_operator(n, dli, dzci, dzfi, cbc):
    """The 7-point operator of src/debug.f90:104-107 with ghost cells eliminated by the BCs."""
    N = n[0] * n[1] * n[2]
    A = np.zeros((N, N))

    def idx(i, j, k):
        return i + n[0] * (j + n[1] * k)

    coef = [dli[0] ** 2, dli[1] ** 2]
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                r = idx(i, j, k)
                ijk = [i, j, k]
                for d in range(3):
                    for side, s in ((0, -1), (1, 1)):
                        if d < 2:
                            w = coef[d]
                        else:
                            w = dzfi[k + 1] * (dzci[k + 1] if s == 1 else dzci[k])
                        A[r, r] -= w
                        nb = list(ijk)
                        nb[d] += s
                        if 0 <= nb[d] < n[d]:
                            A[r, idx(*nb)] += w
                        else:
                            c = cbc[2 * d + side]
                            if c == "P":
                                nb[d] %= n[d]
                                A[r, idx(*nb)] += w
                            elif c == "N":
                                A[r, r] += w
                            else:
                                A[r, r] -= w
    return A


@pytest.mark.parametrize("cbc", ["PPPPNN", "NNDDNN", "DDNNPP", "PPPPDD", "NNNNNN", "PPPPPP"])
def test_dense_solve_stretched_z(cbc, rng):
    n = [6, 4, 7]
    l = [1.0, 1.0, 1.0]
    dl, dli, dzci, dzfi = grid(n, l, gtype=1, gr=1.5)
    S = O.Solver(n, dli, dzci, dzfi, list(cbc))
    A = dense_operator(n, dli, dzci, dzfi, cbc)
    b = rng.standard_normal(n)
    singular = "D" not in cbc
    if singular:
        b -= b.mean()
    f = O.halo(n)
    f[1:-1, 1:-1, 1:-1] = b
    p = f.copy(order="F")
    S.solve(p)
    pi = p[1:-1, 1:-1, 1:-1]
    x = np.linalg.lstsq(A, b.flatten(order="F"), rcond=None)[0].reshape(n, order="F")
    if singular:
        err = np.abs((pi - pi.mean()) - (x - x.mean())).max()
        # roundoff relative to the null-mode constant the reference leaves in p
        assert err <= 1e-12 * max(1.0, np.abs(pi).max()), err
    else:
        assert np.abs(pi - x).max() <= 1e-12 * np.abs(x).max()


def test_tridmatrix_and_eigenvalues_closed_form():
    n = [8, 6, 5]
    l = [2.0, 1.0, 1.0]
    dl, dli, dzci, dzfi = grid(n, l)
    S = O.Solver(n, dli, dzci, dzfi, list("PPNNDN"))
    lx = -2 * (1 - np.cos(2 * np.pi * np.arange(8) / 8)) * dli[0] ** 2
    ly = -2 * (1 - np.cos(np.pi * np.arange(6) / 6)) * dli[1] ** 2
    assert np.allclose(S.lambdaxy, lx[:, None] + ly[None, :], rtol=1e-15, atol=1e-12)
    dz = 1.0 / 5
    a = np.full(5, 1 / dz ** 2)
    b = -2 * a
    b[0] += -1 * a[0]   # 'D' at z0: factor -1
    b[-1] += 1 * a[-1]  # 'N' at z1: factor +1
    assert np.allclose(S.a, a) and np.allclose(S.c, a) and np.allclose(S.b, b)
    assert S.normfft == pytest.approx(1.0 / (8 * 2 * 6))


def test_chk_poisson_residual_nonsingular(rng):
    n = [16, 12, 10]
    l = [1.0, 1.0, 2.0]
    dl, dli, dzci, dzfi = grid(n, l, gr=1.2)
    cbc = list("NNDDPP")
    S = O.Solver(n, dli, dzci, dzfi, cbc)
    f = O.halo(n)
    f[1:-1, 1:-1, 1:-1] = rng.standard_normal(n)
    p = f.copy(order="F")
    S.solve(p)
    # ghost cells of the homogeneous BCs (set_bc, src/bound.f90:178-211)
    p[0] = p[1]
    p[-1] = p[-2]
    p[:, 0] = -p[:, 1]
    p[:, -1] = -p[:, -2]
    p[:, :, 0] = p[:, :, -2]
    p[:, :, -1] = p[:, :, 1]
    r = O.chk_poisson(n, dli, dzci, dzfi, f, p)
    assert r < 1e-11 * np.abs(f).max()
