"""Serving many cases on one device (fz_server_* of libfizzy).

A request is a case (the dict of cases.py) and a number of time steps. The
server admits requests in arrival order while their HBM footprint fits the
budget, steps every admitted case once per round on the case's own stream,
and keeps finished cases until they are released. `run_stream` drives it
from a wall-clock arrival schedule, submitting requests as they arrive
while others are running (continuous batching).
"""
import ctypes as C
import math
import time

import numpy as np

from ._lib import FzParams, call, vp
from .cases import PRECISIONS, DeviceCase, fill_params, splitmix64


def footprint(spec):
    """Device bytes a case needs (fz_case_footprint_ex, in the spec's precision)."""
    prm = fill_params(FzParams(), spec)
    b = C.c_size_t()
    call("fz_case_footprint_ex", C.byref(prm), PRECISIONS[spec.get("precision", "fp64")], C.byref(b))
    return b.value


class Server:
    def __init__(self, hbm_budget):
        h = vp()
        call("fz_server_create", C.byref(h), C.c_size_t(int(hbm_budget)))
        self._h = h
        self.specs = {}

    def close(self):
        if getattr(self, "_h", None):
            call("fz_server_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, spec, nsteps):
        prm = fill_params(FzParams(), spec)
        rid = C.c_int()
        call("fz_server_submit_ex", self._h, C.byref(prm), int(nsteps), PRECISIONS[spec.get("precision", "fp64")],
             int(spec.get("refine", 1)), C.byref(rid))
        self.specs[rid.value] = spec
        return rid.value

    def round(self):
        """Admit what fits, then one step of every running case; (running, queued)."""
        na, nq = C.c_int(), C.c_int()
        call("fz_server_round", self._h, C.byref(na), C.byref(nq))
        return na.value, nq.value

    QUEUED, RUNNING, FINISHED, FAILED = 0, 1, 2, 3

    def status(self, rid):
        """(state, killed): state 0 queued, 1 running, 2 finished, 3 failed, -1 unknown/released."""
        st, k = C.c_int(), C.c_int()
        call("fz_server_status", self._h, int(rid), C.byref(st), C.byref(k))
        return st.value, k.value

    def error(self, rid):
        """The error message of a failed request ('' otherwise)."""
        buf = C.create_string_buffer(1024)
        call("fz_server_error", self._h, int(rid), buf, len(buf))
        return buf.value.decode(errors="replace")

    def case(self, rid):
        """The finished case as a DeviceCase view (owned by the server)."""
        h = vp()
        call("fz_server_case", self._h, int(rid), C.byref(h))
        return DeviceCase.view(h, self.specs[rid])

    def release(self, rid):
        call("fz_server_release", self._h, int(rid))

    def step_times(self, rid=0):
        """Durations (ms) of every step of request rid, or of all steps (rid 0)."""
        n = C.c_int()
        call("fz_server_step_times", self._h, int(rid), None, 0, C.byref(n))
        buf = (C.c_float * max(n.value, 1))()
        call("fz_server_step_times", self._h, int(rid), buf, n.value, C.byref(n))
        return np.array(buf[:n.value], dtype=np.float64)

    def memory(self):
        used, budget = C.c_size_t(), C.c_size_t()
        call("fz_server_memory", self._h, C.byref(used), C.byref(budget))
        return used.value, budget.value


def run_stream(server, requests, release=True, on_finish=None):
    """Submit (t_arrival_s, spec, nsteps) requests at their arrival times
    (relative to the start) while rounds run; returns the wall time and the
    request ids in submission order. A request that fails (state 3) ends like
    a finished one: on_finish sees it, and server.error(rid) gives why."""
    t0 = time.perf_counter()
    pending = sorted(requests, key=lambda r: r[0])
    ids, live = [], set()
    q = 0
    while q < len(pending) or live:
        now = time.perf_counter() - t0
        while q < len(pending) and pending[q][0] <= now:
            rid = server.submit(pending[q][1], pending[q][2])
            ids.append(rid)
            live.add(rid)
            q += 1
        if not live:
            time.sleep(max(0.0, min(pending[q][0] - now, 0.01)))
            continue
        server.round()
        for rid in list(live):
            st, _ = server.status(rid)
            if st in (Server.FINISHED, Server.FAILED):
                live.discard(rid)
                if on_finish:
                    on_finish(rid)
                if release:
                    server.release(rid)
    return time.perf_counter() - t0, ids


def bubble_case(n, nz=None, two_d=False, density_ratio=10.0, mu12=(10.0, 1.0)):
    """A rising bubble of the serving mix: periodic x/y, walls in z, one bubble
    of radius 0.25 at the centre of a unit box (2-D: ng = (n, 2, n), the
    reference's 2-D convention of two cells in y, tests/rising_bubble/input.nml:2-3)."""
    from .cases import bc3
    if two_d:
        ng, l, inipsi, sph = [n, 2, n], [1.0, 2.0 / n, 1.0], "bub2", [(0.5, 1.0 / n, 0.5, 0.25)]
    else:
        nz = nz or n
        ng, l, inipsi, sph = [n, n, nz], [1.0, 1.0, nz / n], "bub3", [(0.5, 0.5, 0.5, 0.25)]
    return dict(
        ng=ng, l=l, cfl=0.25, dtmax=1e9, dt_f=-1.0, icheck=10,
        cbcvel=bc3("PP", "PP", "DD") * 3, cbcpre=bc3("PP", "PP", "NN"), cbcpsi=bc3("PP", "PP", "NN"),
        cbcnor=bc3("PP", "PP", "NN") + bc3("PP", "PP", "NN") + bc3("PP", "PP", "DD"),
        gacc=[0.0, 0.0, -0.98], rho12=[density_ratio, 1.0], mu12=list(mu12), sigma=24.5,
        inipsi=inipsi, spheres=sph, psi_thickness_factor=1.0, vof=False)


GRIDS = [(64, None, True), (128, None, True), (256, None, True), (64, 128, False), (128, 256, False)]


def serving_requests(n, mean_interarrival=0.05, nsteps=200, seed=20261202, precision="fp64", refine=1,
                     grids=None):
    """The serving mix: Poisson arrivals (exponential gaps, SplitMix64), grids
    uniform over {64^2, 128^2, 256^2, 64^2x128, 128^2x256}, density ratio
    uniform in [2, 1000], nsteps steps each, in the given precision mode
    (mixed: refine FP64-residual refinements of an FP32 solve, or -1 for the
    FP64 projection); grids: a subset of GRIDS to draw from instead."""
    g = splitmix64(seed)
    t, out = 0.0, []
    for _ in range(n):
        t += -mean_interarrival * math.log(1.0 - next(g))
        pool = grids or GRIDS
        gi = min(int(next(g) * len(pool)), len(pool) - 1)
        ratio = 2.0 + 998.0 * next(g)
        nxy, nz, two_d = pool[gi]
        spec = bubble_case(nxy, nz, two_d, ratio)
        if precision != "fp64":
            spec = dict(spec, precision=precision, refine=refine)
        out.append((t, spec, nsteps))
    return out
