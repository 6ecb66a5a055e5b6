"""Per-routine access to libfizzy: a context plus device arrays.

Mirrors how the reference's routines are called from src/main.f90: the
caller owns the fields (here device arrays), the routine works in place.
"""
import ctypes as C

import numpy as np

from ._lib import call, lib, vp


def initgrid(gtype, n, gr, lz):
    """initgrid, src/initgrid.f90:14-78: (dzc, dzf, zc, zf), each n+2 long."""
    arrs = [np.zeros(n + 2) for _ in range(4)]
    call("fz_initgrid", int(gtype), int(n), float(gr), float(lz),
         *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in arrs])
    return arrs


def dvec(v):
    return (C.c_double * len(v))(*[float(x) for x in v])


class DArray:
    """A device buffer holding a Fortran-ordered float64 array."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = vp()
        call("fz_malloc", ctx.h, C.c_size_t(self.nbytes), C.byref(p))
        self.ptr = p

    def upload(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        call("fz_upload", self.ctx.h, self.ptr, a.ctypes.data_as(vp), C.c_size_t(self.nbytes))
        return self

    def download(self):
        out = np.empty(self.shape, dtype=np.float64, order="F")
        call("fz_download", self.ctx.h, out.ctypes.data_as(vp), self.ptr, C.c_size_t(self.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().fz_free(self.ctx.h, self.ptr)
            self.ptr = None


class Context:
    """fz_ctx over a grid: interior size n, box l, pressure BCs cbcpre."""

    def __init__(self, n, l, cbcpre, gtype=1, gr=0.0, stream=None):
        self.n = [int(x) for x in n]
        self.l = [float(x) for x in l]
        self.cbcpre = list(cbcpre)
        self.dzc, self.dzf, self.zc, self.zf = initgrid(gtype, self.n[2], gr, self.l[2])
        h = vp()
        dp = C.POINTER(C.c_double)
        call("fz_ctx_create", C.byref(h), (C.c_int * 3)(*self.n), dvec(self.l),
             self.dzc.ctypes.data_as(dp), self.dzf.ctypes.data_as(dp), self.zc.ctypes.data_as(dp),
             self.zf.ctypes.data_as(dp), "".join(self.cbcpre).encode(), vp(stream) if stream else None)
        self.h = h
        self.dl = [self.l[d] / float(np.float32(self.n[d])) for d in range(3)]
        self.dli = [1.0 / x for x in self.dl]
        self.dzci = 1.0 / self.dzc
        self.dzfi = 1.0 / self.dzf
        self._arrays = []

    def halo(self, a=None):
        d = DArray(self, (self.n[0] + 2, self.n[1] + 2, self.n[2] + 2))
        self._arrays.append(d)
        if a is not None:
            d.upload(a)
        return d

    def interior(self, a=None):
        d = DArray(self, tuple(self.n))
        self._arrays.append(d)
        if a is not None:
            d.upload(a)
        return d

    def sync(self):
        call("fz_ctx_sync", self.h)

    def close(self):
        for d in self._arrays:
            d.free()
        self._arrays = []
        if self.h:
            lib().fz_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
