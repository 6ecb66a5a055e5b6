"""Readers of the reference's field files, with the semantics of its own
utilities (utils/read_binary_data/python/read_single_field_binary.py:8-57):
geometry.out gives ng and l, grid.bin the stretched z grid, and a field file
is float64 in Fortran order. Restart files (load_one, src/load.f90:19-116)
carry the interior of one field followed by [time, istep]."""
import os

import numpy as np


def read_geometry(datadir="."):
    geo = np.loadtxt(os.path.join(datadir, "geometry.out"), comments="!", max_rows=2)
    return geo[0, :].astype(int), geo[1, :]


def read_grid(datadir=".", non_uniform_grid=True):
    """Cell-centre and face coordinates (xp, yp, zp, xu, yv, zw) of the grid."""
    ng, l = read_geometry(datadir)
    dl = l / (1.0 * ng)
    xp = np.arange(dl[0] / 2.0, l[0], dl[0])
    yp = np.arange(dl[1] / 2.0, l[1], dl[1])
    zp = np.arange(dl[2] / 2.0, l[2], dl[2])
    xu, yv, zw = xp + dl[0] / 2.0, yp + dl[1] / 2.0, zp + dl[2] / 2.0
    if non_uniform_grid:
        g = np.fromfile(os.path.join(datadir, "grid.bin"), dtype=np.float64).reshape((ng[2], 4), order="F")
        zp, zw = g[:, 2].copy(), g[:, 3].copy()
    return xp, yp, zp, xu, yv, zw


def read_single_field_binary(filename, iskip=(1, 1, 1), datadir="."):
    """A 3-D output field (out3d) and its grid, sampled every iskip points."""
    ng, _ = read_geometry(datadir)
    iskip = np.asarray(iskip, dtype=int)
    n = (ng // iskip).astype(int)
    data = np.fromfile(filename, dtype=np.float64).reshape(tuple(n), order="F")
    xp, yp, zp, xu, yv, zw = read_grid(datadir)
    return (data, xp[0:ng[0]:iskip[0]], yp[0:ng[1]:iskip[1]], zp[0:ng[2]:iskip[2]], xu[0:ng[0]:iskip[0]],
            yv[0:ng[1]:iskip[1]], zw[0:ng[2]:iskip[2]])


def read_restart_field(filename, ng):
    """One restart file: (interior field, time, istep); refuses a wrong size as load_one does."""
    n = int(np.prod(ng))
    size = os.path.getsize(filename)
    if size != (n + 2) * 8:
        raise ValueError(f"checkpoint file with incorrect size: {filename} | expected size: {(n + 2) * 8} "
                         f"| actual size: {size}")
    raw = np.fromfile(filename, dtype=np.float64)
    return raw[:n].reshape(tuple(ng), order="F"), float(raw[n]), int(round(raw[n + 1]))


def write_restart_field(filename, field, time, istep):
    """The inverse of read_restart_field (the load_one 'w' layout)."""
    a = np.asarray(field, dtype=np.float64)
    with open(filename, "wb") as f:
        f.write(np.asfortranarray(a).tobytes(order="F"))
        f.write(np.array([time, float(istep)], dtype=np.float64).tobytes())
