"""The Fortran side of the boundary: a main program written against the
reference's module interfaces (cans-fizzy_amd/fortran/fz_drive.F90, the call
sequence of src/main.f90:252-562 through mod_rk/mod_solver/mod_fillps/...)
must reproduce the device time loop exactly. The driver goes through the
per-routine entry points (mom_xyz_ad, rk, rk_2fl, clip_field, ... with array
copies for psio/pp/po/pn); fz_case_step uses fused kernels and pointer
rotations, so bitwise agreement also pins those fusions."""
import os
import subprocess

import numpy as np
import pytest

import fizzy_amd as F
from conftest import ROOT
from fizzy_amd import cases

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "cans-fizzy_amd", "fortran", "bin", "fz_drive")


@pytest.mark.parametrize("ng,nstep,rho,mu", [((32, 32, 64), 12, (1000.0, 100.0), (10.0, 1.0)),
                                             ((16, 24, 32), 21, (1000.0, 1.0), (10.0, 0.1))])
def test_fortran_main_matches_device_loop(tmp_path, ng, nstep, rho, mu):
    assert os.path.exists(DRIVER), "fz_drive is not built (run __graft_entry__.build())"
    (tmp_path / "spheres.in").write_text("0.5 0.5 0.5 0.25\n")
    out = tmp_path / "fields.bin"
    args = [DRIVER, *map(str, ng), str(nstep), *map(repr, rho), *map(repr, mu), str(out)]
    print("running", " ".join(args), flush=True)
    r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    nh = (ng[0] + 2) * (ng[1] + 2) * (ng[2] + 2)
    raw = np.fromfile(out, dtype=np.float64)
    assert raw.size == 5 * nh + 4
    fields = {name: raw[q * nh:(q + 1) * nh].reshape((ng[0] + 2, ng[1] + 2, ng[2] + 2), order="F")
              for q, name in enumerate(["u", "v", "w", "p", "psi"])}
    time, dt, gam, divmax = raw[5 * nh:]
    spec = cases.rising_bubble_3d(ng=ng, rho12=rho, mu12=mu)
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(nstep)
        for name, f in fields.items():
            ref = dc.get(name)
            assert np.array_equal(f, ref), (name, np.abs(f - ref).max())
        assert time == dc.scalar("time") and dt == dc.scalar("dt") and gam == dc.scalar("gam")
        assert divmax == dc.scalar("divmax")
    finally:
        dc.close()
