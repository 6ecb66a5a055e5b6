"""Field files on the device path: restart files in the load_one format
(src/load.f90:19-116), the checkpoint/restart cycle of src/main.f90:312-407
and :597-631, out3d sampling (src/output.f90:185-236) and the grid files of
src/main.f90:230-243, read back with the reference utilities' semantics."""
import os

import numpy as np
import pytest

import fizzy_amd as F
import oracle as O
from conftest import demean, rel_err
from fizzy_amd import cases
from fizzy_amd import io as fio

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "psi"]


def test_checkpoint_files_hold_the_interior_and_the_step(tmp_path):
    spec = cases.rising_bubble_3d(ng=(24, 20, 32))
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(7)
        prefix = str(tmp_path / "fld")
        dc.save(prefix)
        for name in FIELDS:
            data, time, istep = fio.read_restart_field(f"{prefix}_{name}.bin", spec["ng"])
            assert np.array_equal(data, dc.get(name)[1:-1, 1:-1, 1:-1]), name
            assert time == dc.scalar("time") and istep == 7
        # a truncated file is refused, as load_one does
        path = f"{prefix}_u.bin"
        with open(path, "r+b") as f:
            f.truncate(os.path.getsize(path) - 8)
        with pytest.raises(F.FizzyError, match="incorrect size"):
            dc.restart(prefix)
    finally:
        dc.close()


def test_restart_matches_the_oracle_restart(tmp_path):
    """Run, checkpoint, restart in a fresh case; the oracle restarts from the
    same files (the fields' interiors, then the reference's init sequence)
    and both then run on; the AB2 history starts over on both sides."""
    spec = cases.rising_bubble_3d(ng=(32, 32, 64))
    prefix = str(tmp_path / "fld")
    d1 = F.DeviceCase(spec)
    try:
        d1.init()
        d1.step(11)
        d1.save(prefix)
    finally:
        d1.close()
    d2 = F.DeviceCase(spec)
    oc = O.Case(spec)
    try:
        d2.restart(prefix)
        for name in FIELDS:
            data, time, istep = fio.read_restart_field(f"{prefix}_{name}.bin", spec["ng"])
            oc.field(name)[1:-1, 1:-1, 1:-1] = data
        oc.set_scalar("time", time)
        oc.set_scalar("istep", istep)
        oc.finish_init()
        for k in ("time", "dt", "dto", "gam", "istep", "is_first"):
            assert d2.scalar(k) == pytest.approx(oc.scalar(k), rel=1e-13, abs=0), k
        d2.step(12)
        oc.step(12)
        errs = {n: rel_err(d2.get(n), oc.field(n)) for n in ["u", "v", "w", "psi"]}
        errs["p"] = rel_err(demean(d2.get("p")[1:-1, 1:-1, 1:-1]), demean(oc.field("p")[1:-1, 1:-1, 1:-1]))
        print(errs)
        assert max(errs.values()) <= 1e-12, errs
    finally:
        d2.close()


def test_out3d_and_grid_files(tmp_path):
    spec = cases.rising_bubble_3d(ng=(16, 12, 24))
    spec["gr"] = 0.8
    spec["gtype"] = 1
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(3)
        datadir = str(tmp_path) + "/"
        dc.write_grid(datadir)
        fname = datadir + "psi_fld.bin"
        dc.out3d(fname, "psi", (2, 2, 3))
        data, xp, yp, zp, xu, yv, zw = fio.read_single_field_binary(fname, (2, 2, 3), datadir)
        psi = dc.get("psi")[1:-1, 1:-1, 1:-1]
        assert np.array_equal(data, psi[::2, ::2, ::3])
        _, _, zc, zf = F.initgrid(1, 24, 0.8, 2.0)
        assert np.array_equal(zp, zc[1:-1][::3]) and np.array_equal(zw, zf[1:-1][::3])
        ng, l = fio.read_geometry(datadir)
        assert list(ng) == [16, 12, 24] and list(l) == [1.0, 1.0, 2.0]
    finally:
        dc.close()
