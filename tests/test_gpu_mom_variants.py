"""Every variant of the momentum kernel (staged loads, plain, waves-per-SIMD
caps, the stencil fields staged in LDS; a context reads FZ_MOM_WPE when it
is made) gives the default variant's fields bit for bit, in FP64 and in
mixed precision, on a grid whose sizes are not multiples of the blocks."""
import numpy as np
import pytest

import fizzy_amd as F
from fizzy_amd import cases

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "psi", "kappa"]
BUBBLE = cases.rising_bubble_3d(ng=(72, 20, 40))
SPECS = {"fp64": BUBBLE, "mixed": dict(BUBBLE, precision="mixed", refine=1)}
_reference = {}


def run(spec, nsteps, variant, monkeypatch):
    if variant is None:
        monkeypatch.delenv("FZ_MOM_WPE", raising=False)
    else:
        monkeypatch.setenv("FZ_MOM_WPE", str(variant))
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        dc.step(nsteps)
        return {k: dc.get(k) for k in FIELDS}, dc.scalars()
    finally:
        dc.close()


@pytest.mark.parametrize("variant", [1, 2, 4, 5, 8])
@pytest.mark.parametrize("name", list(SPECS))
def test_momentum_variants_agree_bitwise(name, variant, monkeypatch):
    if name not in _reference:
        _reference[name] = run(SPECS[name], 8, 0, monkeypatch)
    a, sa = _reference[name]
    b, sb = run(SPECS[name], 8, variant, monkeypatch)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(sa, sb)


@pytest.mark.parametrize("env", ["FZ_ACDI_VARIANT=1", "FZ_ACDI_VARIANT=2", "FZ_ACDI_VARIANT=3", "FZ_YOUNGS_VARIANT=1",
                                 "FZ_YOUNGS_VARIANT=2", "FZ_YOUNGS_VARIANT=3", "FZ_CORREC_KZ=0",
                                 "FZ_CORREC_KZ=8", "FZ_CORREC_MY=8", "FZ_CORREC_MY=16", "FZ_CORREC_MY=5",
                                 "FZ_P2_XDIV=1", "FZ_TILE_Y=0", "FZ_TILE_Y=1", "FZ_TILE_Y=3"])
@pytest.mark.parametrize("name", list(SPECS))
def test_acdi_and_correction_variants_agree_bitwise(name, env, monkeypatch):
    """The ACDI step with six faces per cell staged in LDS, one cell per
    thread or each face once, the Youngs kernel marching in z, one cell per
    thread or the check-free pair, the correction step one cell per thread,
    marching in z or with other tile heights (5 is not one it is built for and
    falls back to 4), x-line transforms with twice the default's lines per
    block, and the stencils' block order plain or in pencils of 1 or 3 block
    rows, give the defaults' fields."""
    if name not in _reference:
        _reference[name] = run(SPECS[name], 8, 0, monkeypatch)
    a, sa = _reference[name]
    key, val = env.split("=")
    monkeypatch.setenv(key, val)
    b, sb = run(SPECS[name], 8, None, monkeypatch)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(sa, sb)
