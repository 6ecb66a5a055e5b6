"""fizzy_amd: Python access to libfizzy.so, the MI355X-native CaNS-Fizzy hot path.

The compute runs in the HIP library (cans-fizzy_amd/csrc, C ABI in
include/fizzy.h); this package only binds it with ctypes.
"""
from ._lib import FizzyError, FzParams, LIB_PATH, call, declared_symbols, lib  # noqa: F401
from .cases import CaseBatch, DeviceCase, FIELD_IDS, SCALAR_IDS, make_params  # noqa: F401
from .device import Context, DArray, initgrid  # noqa: F401
