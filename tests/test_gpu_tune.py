"""The autotuner (fz_case_autotune) changes launch parameters only: a tuned
case gives the same fields bit for bit as an untuned one, its choice comes
back as a report, and a second case of the same shape reuses it."""
import re

import numpy as np
import pytest

import fizzy_amd as F
from fizzy_amd import cases

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "w", "p", "psi", "kappa"]


def run(spec, nsteps, tune):
    dc = F.DeviceCase(spec)
    try:
        dc.init()
        rep = dc.autotune(2) if tune else None
        dc.step(nsteps)
        return {k: dc.get(k) for k in FIELDS}, dc.scalars(), rep
    finally:
        dc.close()


@pytest.mark.parametrize("spec", [cases.rising_bubble_3d(ng=(32, 32, 64)),
                                  dict(cases.rising_bubble_3d(ng=(32, 32, 64)), precision="mixed", refine=1)],
                         ids=["fp64", "mixed"])
def test_tuned_case_gives_the_same_fields(spec):
    a, sa, _ = run(spec, 25, False)
    b, sb, rep = run(spec, 25, True)
    print(rep)
    m = re.match(r"tile_y=(\d+) mom_variant=(\d+) thomas_rows=(\d+)(?:x2|rc)? ms_per_step=([0-9.]+)", rep)
    assert m and int(m.group(1)) in (0, 2, 4, 8, 16) and int(m.group(2)) in (0, 1, 2, 4, 5, 8)
    assert int(m.group(3)) in (8, 16, 32) and float(m.group(4)) > 0
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(sa, sb)
    _, _, rep2 = run(spec, 1, True)
    assert rep2.endswith("(cached)") and rep2.split(" ms")[0] == rep.split(" ms")[0]
