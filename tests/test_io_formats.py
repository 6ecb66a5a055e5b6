"""The field-file formats on the host side (no GPU): restart files in the
load_one layout (src/load.f90:19-116: interior in Fortran order, then
[time, istep], exact size) and the reader semantics of the reference's
utils/read_binary_data/python/read_single_field_binary.py:8-57
(geometry.out, grid.bin with columns dzc, dzf, zc, zf)."""
import numpy as np
import pytest

from fizzy_amd import io as fio


def test_restart_field_round_trip_and_size_check(tmp_path, rng):
    ng = (5, 4, 3)
    a = np.asfortranarray(rng.standard_normal(ng))
    path = str(tmp_path / "fld_u.bin")
    fio.write_restart_field(path, a, 1.25, 40)
    raw = np.fromfile(path, dtype=np.float64)
    assert raw.size == a.size + 2 and raw[-2] == 1.25 and raw[-1] == 40.0
    assert np.array_equal(raw[:a.size], a.ravel(order="F"))
    b, t, s = fio.read_restart_field(path, ng)
    assert np.array_equal(a, b) and t == 1.25 and s == 40
    with pytest.raises(ValueError, match="incorrect size"):
        fio.read_restart_field(path, (5, 4, 4))


def test_single_field_reader_with_stretched_grid(tmp_path, rng):
    ng = (4, 6, 8)
    l = (1.0, 1.5, 2.0)
    (tmp_path / "geometry.out").write_text(f" {ng[0]} {ng[1]} {ng[2]}\n {l[0]} {l[1]} {l[2]}\n")
    zc, zf = np.sort(rng.uniform(0, 2, ng[2])), np.sort(rng.uniform(0, 2, ng[2]))
    cols = np.stack([rng.uniform(size=ng[2]), rng.uniform(size=ng[2]), zc, zf], axis=1)
    cols.ravel(order="F").tofile(str(tmp_path / "grid.bin"))
    f = np.asfortranarray(rng.standard_normal(ng))
    f[::2, ::3, ::2].ravel(order="F").tofile(str(tmp_path / "psi.bin"))
    data, xp, yp, zp, xu, yv, zw = fio.read_single_field_binary(str(tmp_path / "psi.bin"), (2, 3, 2), str(tmp_path))
    assert np.array_equal(data, f[::2, ::3, ::2])
    assert np.allclose(xp, [0.125, 0.625]) and np.allclose(xu, [0.25, 0.75])
    assert np.array_equal(zp, zc[::2]) and np.array_equal(zw, zf[::2])
