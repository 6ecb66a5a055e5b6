"""Writes the golden vectors under tests/golden/ (run from the repo root).

r2r.npz: scipy.fft results for every r2r kind the Poisson solver uses, with
FFTW's unnormalised definitions (REDFT10 = DCT-II, REDFT01 = DCT-III,
RODFT10 = DST-II, RODFT01 = DST-III, REDFT11 = DCT-IV, RODFT11 = DST-IV,
R2HC/HC2R halfcomplex). scipy's norm=None conventions equal FFTW's.

The reference-test fixtures c1g2l3s.txt (Hysing et al. 2009 bubble contour,
tests/rising_bubble), ghiau.txt and ghiav.txt (Ghia et al. 1982 cavity
profiles, tests/lid_driven_cavity) are data files copied from the reference's
own tests by copy_reference_fixtures(), and so are the input files of those
two tests (input.nml, and spheres.in of the rising bubble), under the names
rising_bubble.nml, rising_bubble_spheres.in and lid_driven_cavity.nml.
"""
import os
import shutil
import sys

import numpy as np
import scipy.fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))
SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 24, 26, 30, 32, 48, 64, 96, 128]


def halfcomplex(x):
    n = x.shape[-1]
    X = np.fft.rfft(x, axis=-1)
    out = np.zeros_like(x)
    out[..., :n // 2 + 1] = X.real
    for k in range(1, (n - 1) // 2 + 1):
        out[..., n - k] = X[..., k].imag
    return out


def hc2r(h):
    n = h.shape[-1]
    X = np.zeros(h.shape, complex)
    X[..., 0] = h[..., 0]
    for k in range(1, (n - 1) // 2 + 1):
        X[..., k] = h[..., k] + 1j * h[..., n - k]
        X[..., n - k] = np.conj(X[..., k])
    if n % 2 == 0:
        X[..., n // 2] = h[..., n // 2]
    return np.real(np.fft.ifft(X, axis=-1) * n)


def make_r2r():
    rng = np.random.default_rng(20261202)
    data = {}
    for n in SIZES:
        x = rng.standard_normal((3, n))
        data[f"x_{n}"] = x
        data[f"R2HC_{n}"] = halfcomplex(x)
        data[f"HC2R_{n}"] = hc2r(x)
        data[f"REDFT10_{n}"] = sf.dct(x, type=2, axis=-1)
        data[f"REDFT01_{n}"] = sf.dct(x, type=3, axis=-1)
        data[f"RODFT10_{n}"] = sf.dst(x, type=2, axis=-1)
        data[f"RODFT01_{n}"] = sf.dst(x, type=3, axis=-1)
        data[f"REDFT11_{n}"] = sf.dct(x, type=4, axis=-1)
        data[f"RODFT11_{n}"] = sf.dst(x, type=4, axis=-1)
    np.savez_compressed(os.path.join(HERE, "r2r.npz"), **data)


def copy_reference_fixtures(ref="/root/reference"):
    for rel, name in [("tests/rising_bubble/c1g2l3s.txt", None), ("tests/lid_driven_cavity/ghiau.txt", None),
                      ("tests/lid_driven_cavity/ghiav.txt", None),
                      ("tests/rising_bubble/input.nml", "rising_bubble.nml"),
                      ("tests/rising_bubble/spheres.in", "rising_bubble_spheres.in"),
                      ("tests/lid_driven_cavity/input.nml", "lid_driven_cavity.nml")]:
        src = os.path.join(ref, rel)
        if os.path.exists(src):
            shutil.copyfile(src, os.path.join(HERE, name or os.path.basename(rel)))


if __name__ == "__main__":
    make_r2r()
    if "--fixtures" in sys.argv:
        copy_reference_fixtures()
