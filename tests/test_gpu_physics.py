"""End-to-end physics against the reference's own test fixtures.

tests/rising_bubble of the reference: 2-D Hysing case 1 (64x2x128, density
ratio 10, Re=35, Eo=10), psi=0.5 contour at t=3 against the benchmark contour
c1g2l3s.txt (Hysing et al. 2009), directed Hausdorff distance below 0.1
(tests/rising_bubble/test.py:9-20), run entirely on the device.
"""
import os

import numpy as np
import pytest
from scipy.spatial.distance import directed_hausdorff

import fizzy_amd as F
from conftest import GOLDEN
from fizzy_amd import cases

pytestmark = pytest.mark.gpu


def contour_points(f, x, z, level=0.5):
    """Vertices of the level set of f(x_i, z_k) on the grid edges (marching squares)."""
    pts = []
    a = f - level
    # edges along x
    s = a[:-1, :] * a[1:, :] < 0
    i, k = np.nonzero(s)
    t = a[i, k] / (a[i, k] - a[i + 1, k])
    pts.append(np.stack([x[i] + t * (x[i + 1] - x[i]), z[k]], axis=1))
    # edges along z
    s = a[:, :-1] * a[:, 1:] < 0
    i, k = np.nonzero(s)
    t = a[i, k] / (a[i, k] - a[i, k + 1])
    pts.append(np.stack([x[i], z[k] + t * (z[k + 1] - z[k])], axis=1))
    return np.concatenate(pts)


def test_hysing_rising_bubble_contour():
    spec = cases.hysing_2d()
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        nstep, time_max = 6000, 3.0  # stop_type = T,T,F in tests/rising_bubble/input.nml
        done = 0
        while done < nstep and dc.scalar("time") < time_max - 1e-12:
            killed = dc.step(10)
            done += 10
            assert killed == 0, f"run stopped by the divergence/dt checks at step {done}"
        psi = dc.get("psi")[1:-1, 1, 1:-1]
        n = spec["ng"]
        dx = spec["l"][0] / n[0]
        x = (np.arange(1, n[0] + 1) - 0.5) * dx
        _, _, zc, _ = F.initgrid(1, n[2], 0.0, spec["l"][2])
        v = contour_points(psi, x, zc[1:-1])
        v_ref = np.loadtxt(os.path.join(GOLDEN, "c1g2l3s.txt"))
        err = directed_hausdorff(v, v_ref)[0]
        print(f"Hysing contour at t={dc.scalar('time'):.4f} after {int(dc.scalar('istep'))} steps: "
              f"directed Hausdorff {err:.4f}, divmax {dc.scalar('divmax'):.2e}")
        assert err < 1e-1
        assert dc.scalar("divmax") < 1e-9
    finally:
        dc.close()
