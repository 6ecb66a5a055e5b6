"""The C-ABI library loads and exports every entry point include/fizzy.h
declares (no GPU needed: nothing is called)."""
import ctypes as C
import os

import fizzy_amd as F
from fizzy_amd import _lib
from conftest import ROOT


def test_library_exports_every_declared_symbol():
    L = F.lib()
    declared = F.declared_symbols()
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(_lib.SIGNATURES) == F.declared_symbols()


def test_params_layout_matches_oracle():
    import oracle as O
    a, b = F.FzParams, O.OrParams
    assert C.sizeof(a) == C.sizeof(b)
    for (na, ta), (nb, tb) in zip(a._fields_, b._fields_):
        assert na == nb and getattr(a, na).offset == getattr(b, nb).offset and C.sizeof(ta) == C.sizeof(tb)


def test_library_is_gfx950_code_object():
    # the offload bundle in .hip_fatbin names its target
    blob = open(F.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_host_only_entry_points_run_without_gpu():
    import numpy as np
    dzc, dzf, zc, zf = F.initgrid(1, 8, 0.0, 2.0)
    assert np.allclose(dzf[1:-1], 0.25) and np.isclose(zc[1], 0.125)
    assert os.path.exists(F.LIB_PATH)


def test_fortran_modules_bind_the_library():
    """The drop-in Fortran modules and their driver are built and link libfizzy."""
    import subprocess
    drv = os.path.join(ROOT, "cans-fizzy_amd", "fortran", "bin", "fz_drive")
    assert os.path.exists(drv), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", drv], capture_output=True, text=True).stdout
    assert "libfizzy.so" in out and "not found" not in out
    mods = os.listdir(os.path.join(ROOT, "cans-fizzy_amd", "fortran", "bin"))
    for m in ["mod_solver", "mod_initsolver", "mod_fillps", "mod_correc", "mod_updatep", "mod_acdi",
              "mod_vof_thinc_qq", "mod_two_fluid", "mod_mom", "mod_rk", "mod_bound", "mod_chkdiv", "mod_chkdt",
              "mod_forcing", "mod_utils"]:
        assert m + ".mod" in mods, m


def test_mixed_precision_refuses_what_it_does_not_cover():
    """Argument checks of fz_case_create_ex run on the host before any device
    allocation: unknown precision, a refinement count out of range and a
    single-phase case in mixed precision are refused with a message."""
    import ctypes as C
    from fizzy_amd import cases
    from fizzy_amd._lib import FzParams, FZ_OK
    L = F.lib()
    for spec, prec, nref, msg in [(cases.rising_bubble_3d(ng=(16, 16, 16)), 7, 1, b"unknown precision"),
                                  (cases.rising_bubble_3d(ng=(16, 16, 16)), 1, 9, b"refinements"),
                                  (dict(cases.rising_bubble_3d(ng=(16, 16, 16)), is_track_interface=False), 1, 1,
                                   b"two-fluid")]:
        prm = cases.fill_params(FzParams(), spec)
        h = C.c_void_p()
        rc = L.fz_case_create_ex(C.byref(h), C.byref(prm), None, prec, nref)
        assert rc != FZ_OK and msg in L.fz_last_error(), L.fz_last_error()
