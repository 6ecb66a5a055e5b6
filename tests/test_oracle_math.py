"""Pins the oracle's exp/expm1/log/tanh/sin/cos (oracle/fz_oracle_math.h, the
method the device kernels share) against the C library evaluated in x87
extended precision (numpy longdouble): at most two ulps (tanh 2.5) over the argument
ranges the hot path uses, exact special values."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

FN = {"exp": 0, "expm1": 1, "log": 2, "tanh": 3, "sin": 4, "cos": 5}
REF = {"exp": np.exp, "expm1": np.expm1, "log": np.log, "tanh": np.tanh, "sin": np.sin, "cos": np.cos}


def ev(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    O.lib().or_math_eval(C.c_int(FN[fn]), C.c_long(x.size), x.ctypes.data_as(C.POINTER(C.c_double)),
                         y.ctypes.data_as(C.POINTER(C.c_double)))
    return y


def ulps(y, ref):
    ref = np.asarray(ref, dtype=np.longdouble)
    sp = np.spacing(np.abs(ref.astype(np.float64))).astype(np.longdouble)
    return np.abs((y.astype(np.longdouble) - ref) / sp)


RANGES = {
    "exp": [(-745.0, 709.0), (-40.0, 40.0), (-1.0, 1.0), (-1e-8, 1e-8)],
    "expm1": [(-40.0, 40.0), (-0.4, 0.4), (-1e-10, 1e-10)],
    "log": [(1e-300, 1e-200), (1e-20, 1.0), (0.5, 2.0), (1.0, 1e10), (2.2e-16, 1e-8)],
    "tanh": [(-30.0, 30.0), (-1.0, 1.0), (-1e-3, 1e-3), (15.0, 25.0)],
    "sin": [(-10.0, 10.0), (-1.0, 1.0), (0.0, 1e3)],
    "cos": [(-10.0, 10.0), (-1.0, 1.0), (0.0, 1e3)],
}


@pytest.mark.parametrize("fn", list(FN))
def test_accuracy_in_ulps(fn, rng):
    if np.finfo(np.longdouble).nmant < 60:
        pytest.skip("no extended precision on this platform")
    worst = 0.0
    for lo, hi in RANGES[fn]:
        x = rng.uniform(lo, hi, 200000)
        y = ev(fn, x)
        ref = REF[fn](x.astype(np.longdouble))
        e = ulps(y, ref)
        worst = max(worst, float(e.max()))
    assert worst <= (2.5 if fn == "tanh" else 2.0), (fn, worst)


def test_special_values():
    assert ev("exp", [0.0])[0] == 1.0 and ev("exp", [1000.0])[0] == np.inf and ev("exp", [-1000.0])[0] == 0.0
    assert ev("log", [1.0])[0] == 0.0 and ev("log", [0.0])[0] == -np.inf and np.isnan(ev("log", [-1.0])[0])
    assert ev("log", [5e-324])[0] == pytest.approx(np.log(5e-324), rel=1e-15)
    assert ev("tanh", [0.0])[0] == 0.0 and ev("tanh", [40.0])[0] == 1.0 and ev("tanh", [-40.0])[0] == -1.0
    t = ev("tanh", np.linspace(-5, 5, 1001))
    assert np.all(np.diff(t) >= 0)  # monotone
    assert np.all(np.isnan(ev("tanh", [np.nan])))
