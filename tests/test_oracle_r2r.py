"""Pins the oracle's r2r transforms (the FFTW kinds planned by src/fft.f90:189-242)
against scipy golden vectors, brute-force sums of FFTW's definitions and the
round-trip normalisation normfft of src/fft.f90:96-139."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

KINDS = {"R2HC": O.R2HC, "HC2R": O.HC2R, "REDFT10": O.REDFT10, "REDFT01": O.REDFT01, "RODFT10": O.RODFT10,
         "RODFT01": O.RODFT01, "REDFT11": O.REDFT11, "RODFT11": O.RODFT11}


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "r2r.npz"))


def sizes(golden):
    return sorted(int(k.split("_")[1]) for k in golden.files if k.startswith("x_"))


@pytest.mark.parametrize("kind", list(KINDS))
def test_against_scipy_golden(golden, kind):
    for n in sizes(golden):
        x = golden[f"x_{n}"]
        y = O.r2r(KINDS[kind], x)
        ref = golden[f"{kind}_{n}"]
        err = np.abs(y - ref).max() / max(1.0, np.abs(ref).max())
        assert err < 2e-15, (kind, n, err)


def brute(kind, x):
    """FFTW's unnormalised definitions as O(n^2) sums."""
    n = x.shape[-1]
    j = np.arange(n)[:, None]
    k = np.arange(n)[None, :]
    if kind == "REDFT10":
        M = 2 * np.cos(np.pi * (2 * j + 1) * k / (2 * n))
    elif kind == "REDFT01":
        M = 2 * np.cos(np.pi * j * (2 * k + 1) / (2 * n))
        M[0, :] = 1.0
    elif kind == "RODFT10":
        M = 2 * np.sin(np.pi * (2 * j + 1) * (k + 1) / (2 * n))
    elif kind == "RODFT01":
        M = 2 * np.sin(np.pi * (j + 1) * (2 * k + 1) / (2 * n))
        M[n - 1, :] = (-1.0) ** np.arange(n)
    elif kind == "REDFT11":
        M = 2 * np.cos(np.pi * (2 * j + 1) * (2 * k + 1) / (4 * n))
    elif kind == "RODFT11":
        M = 2 * np.sin(np.pi * (2 * j + 1) * (2 * k + 1) / (4 * n))
    else:
        raise ValueError(kind)
    return x @ M


@pytest.mark.parametrize("kind", ["REDFT10", "REDFT01", "RODFT10", "RODFT01", "REDFT11", "RODFT11"])
@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 10, 15, 16])
def test_against_brute_force(kind, n, rng):
    x = rng.standard_normal((5, n))
    y = O.r2r(KINDS[kind], x)
    ref = brute(kind, x)
    assert np.abs(y - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("pair,norm", [(("R2HC", "HC2R"), 1), (("REDFT10", "REDFT01"), 2), (("RODFT10", "RODFT01"), 2),
                                       (("REDFT11", "REDFT11"), 2), (("RODFT11", "RODFT11"), 2)])
@pytest.mark.parametrize("n", [2, 5, 8, 12, 64])
def test_round_trip_equals_inverse_normfft(pair, norm, n, rng):
    x = rng.standard_normal((4, n))
    y = O.r2r(KINDS[pair[1]], O.r2r(KINDS[pair[0]], x))
    assert np.abs(y / (norm * n) - x).max() < 1e-14


def test_pairing_is_independent_of_neighbour(rng):
    # two real lines share one complex transform; each result must not depend on the other line
    n = 24
    a = rng.standard_normal((1, n))
    b = rng.standard_normal((1, n))
    ya = O.r2r(O.REDFT10, np.vstack([a, b]))[0]
    yb = O.r2r(O.REDFT10, np.vstack([a, 1e6 * b]))[0]
    assert np.abs(ya - yb).max() < 1e-9 * np.abs(ya).max()
