"""The kernels' FP64 routines on the device against the oracle's copies of
the same method (oracle/fz_oracle_math.h), bit for bit: exp, expm1, log, tanh,
sin and cos over the argument ranges of the ACDI, THINC/QQ and forcing
kernels, and the tanh's in-range division (the device's division sequence
without its range scaling and fix-up) against IEEE division over operands
whose magnitudes span [2^-300, 2^300]."""
import ctypes as C

import numpy as np
import pytest

import fizzy_amd as F
import oracle as O
from fizzy_amd.device import DArray

pytestmark = pytest.mark.gpu

FN = {"exp": 0, "expm1": 1, "log": 2, "tanh": 3, "sin": 4, "cos": 5}
RANGES = {
    "exp": [(-745.0, 709.0), (-40.0, 40.0), (-1e-8, 1e-8)],
    "expm1": [(-40.0, 40.0), (-0.4, 0.4)],
    "log": [(1e-300, 1e-200), (1e-20, 1.0), (0.5, 2.0), (1.0, 1e10)],
    # the ACDI argument 0.25 (phi1 + phi2) / eps lies within about 18.5 of 0
    "tanh": [(-30.0, 30.0), (-1.0, 1.0), (-1e-3, 1e-3), (-1e-8, 1e-8), (15.0, 25.0), (-0.2, 0.2)],
    "sin": [(-10.0, 10.0), (0.0, 1e3)],
    "cos": [(-10.0, 10.0), (0.0, 1e3)],
}


@pytest.fixture(scope="module")
def ctx():
    c = F.Context([8, 8, 8], [1.0, 1.0, 1.0], "PPPPPP")
    yield c
    c.close()


def device_eval(ctx, fn, x, y=None):
    x = np.ascontiguousarray(x, dtype=np.float64)
    dx, dout = DArray(ctx, x.shape), DArray(ctx, x.shape)
    dy = DArray(ctx, x.shape) if y is not None else None
    try:
        dx.upload(x)
        if dy is not None:
            dy.upload(np.ascontiguousarray(y, dtype=np.float64))
        F.call("fz_math_eval", ctx.h, fn, C.c_long(x.size), dx.ptr, dy.ptr if dy else None, dout.ptr)
        ctx.sync()
        return dout.download()
    finally:
        for d in (dx, dout, dy):
            if d is not None:
                d.free()


def oracle_eval(fn, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    O.lib().or_math_eval(C.c_int(FN[fn]), C.c_long(x.size), x.ctypes.data_as(C.POINTER(C.c_double)),
                         y.ctypes.data_as(C.POINTER(C.c_double)))
    return y


@pytest.mark.parametrize("fn", list(FN))
def test_device_math_equals_the_oracle_bit_for_bit(ctx, fn, rng):
    x = np.concatenate([rng.uniform(lo, hi, 400000) for lo, hi in RANGES[fn]])
    if fn == "tanh":
        x = np.concatenate([x, [0.0, -0.0, 22.0, -22.0, np.nextafter(22.0, 30.0), 3.725290298461914e-09,
                                np.nextafter(3.725290298461914e-09, 0.0), 0.5 * np.log(2.0), 40.0, -40.0]])
    got, ref = device_eval(ctx, FN[fn], x), oracle_eval(fn, x)
    same = (got.view(np.uint64) == ref.view(np.uint64)) | (np.isnan(got) & np.isnan(ref))
    assert same.all(), (fn, x[~same][:5], got[~same][:5], ref[~same][:5])


def test_inrange_division_is_ieee_division(ctx, rng):
    """fn 6 (the in-range division) and fn 7 (the compiler's division) against
    numpy's IEEE division, 4 million quotients with both operands' magnitudes
    log-uniform in [2^-300, 2^300] and random signs, plus operands of the
    tanh's range (numerator in [7e-9, 2^64], denominator numerator + 2)."""
    m = 2_000_000
    a = np.exp2(rng.uniform(-300.0, 300.0, m)) * rng.choice([-1.0, 1.0], m)
    b = np.exp2(rng.uniform(-300.0, 300.0, m)) * rng.choice([-1.0, 1.0], m)
    t = np.exp2(rng.uniform(np.log2(7e-9), 64.0, m))
    x, y = np.concatenate([a, t]), np.concatenate([b, t + 2.0])
    ref = x / y
    for fn in (6, 7):
        got = device_eval(ctx, fn, x, y)
        same = got.view(np.uint64) == ref.view(np.uint64)
        assert same.all(), (fn, x[~same][:5], y[~same][:5])
