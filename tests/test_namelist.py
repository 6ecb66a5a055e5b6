"""The reader of the reference's input files (fizzy_amd.namelist) against
the input files of the reference's own tests (tests/rising_bubble,
tests/lid_driven_cavity; copies under tests/golden/) and the cases this
package defines from them by hand, and on the namelist forms the reference's
reader accepts."""
import os

import pytest

from conftest import GOLDEN
from fizzy_amd import cases, namelist

KEYS = ["ng", "l", "cfl", "dtmax", "dt_f", "icheck", "cbcvel", "bcvel", "cbcpre", "cbcpsi", "cbcnor", "gacc",
        "rho12", "mu12", "sigma", "inipsi", "psi_thickness_factor", "vof"]


def read(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def test_rising_bubble_input_gives_the_hysing_case():
    case = namelist.case_from_input(read("rising_bubble.nml"), read("rising_bubble_spheres.in"))
    ref = cases.hysing_2d()
    for k in KEYS:
        if k in ref:
            assert case[k] == ref[k], k
    assert case["spheres"] == ref["spheres"]
    assert case["nstep"] == 6000 and case["time_max"] == 3.0 and case["stop_type"] == [True, True, False]


def test_lid_driven_cavity_input_gives_the_cavity_case():
    case = namelist.case_from_input(read("lid_driven_cavity.nml"))
    ref = cases.lid_driven_cavity()
    for k in KEYS:
        if k in ref:
            assert case[k] == ref[k], k
    assert case["bcvel"][11] == 1.0 and sum(case["bcvel"]) == 1.0  # bcvel(1,3,2) = 1: v on the lid


def test_namelist_forms():
    text = """
    ! a comment line
    &dns
    ng(1:3) = 16, 8, 4   ! trailing comment
    l = 1.d0, 2.D0, 3.e0
    is_solve_ns = .false., is_forced_hit = .T.
    cbcvel = 18*'D'
    cbcvel(0:1,3,2) = 'N','N'
    bcvel(1,1:3,3) = 3*2.5
    gacc(3) = -9.81
    inivel = 'zer'
    /
    &cudecomp
    cudecomp_t_comm_backend = 0
    /
    &two_fluid
    inipsi = 'dis3', rho12 = 2*3., sigma = 1e-2
    /
    """
    v = namelist.parse(text)
    assert v["ng"] == [16, 8, 4] and v["l"] == [1.0, 2.0, 3.0]
    assert v["is_solve_ns"] is False and v["is_forced_hit"] is True
    # cbcvel(0:1,1:3,1:3): the section (0:1,3,2) is elements 4, 5 of component 2 (offset 6)
    assert v["cbcvel"] == ["D"] * 10 + ["N", "N"] + ["D"] * 6
    assert v["bcvel"] == [0.0] * 12 + [0.0, 2.5, 0.0, 2.5, 0.0, 2.5]
    assert v["gacc"] == [0.0, 0.0, -9.81]
    assert v["rho12"] == [3.0, 3.0] and v["sigma"] == 1e-2
    assert v["cfl"] == 0.95 and v["mu12"] == [0.01, 0.01]  # defaults kept (src/param.f90:123-151)
    assert namelist.case_from_input(text)["inipsi"] == "drp3"
    assert namelist.parse("&two_fluid\n/\n", vof=True)["psi_thickness_factor"] == 0.50


@pytest.mark.parametrize("text,msg", [("&dns\nnx = 3\n/\n", "unknown variable"),
                                      ("&dns\ncbcpre(0:2,1) = 'P','P','P'\n/\n", "outside"),
                                      ("&dns\nng = 1, 2, 3, 4\n/\n", "values for"),
                                      ("&dns\ninivel = 'log'\n/\n", "only a fluid at rest")])
def test_namelist_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        namelist.case_from_input(text)
