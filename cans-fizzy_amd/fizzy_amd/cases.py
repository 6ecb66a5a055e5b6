"""Case definitions and a Python handle over the device-resident case.

A case is a dict with the reference's namelist names (src/param.f90:91-115):
ng, l, cfl, dtmax, dt_f, cbcvel (18 codes: u then v then w, each x0 x1 y0 y1
z0 z1), cbcpre, cbcpsi, cbcnor (18), rho12, mu12, sigma, gacc, inipsi,
spheres ([(x, y, z, r), ...], the records of spheres.in), vof, ...
"""
import ctypes as C
import math

import numpy as np

from ._lib import FzParams, call, lib, vp

INIPSI = {"uni": 0, "zer": 1, "bub3": 2, "bub2": 3, "bub1": 4, "drp3": 5, "drp2": 6, "drp1": 7, "flm": 8}
FIELD_IDS = {"u": 0, "v": 1, "w": 2, "p": 3, "pn": 4, "po": 5, "psi": 6, "phi": 7, "kappa": 8,
             "normx": 9, "normy": 10, "normz": 11, "psiflx_x": 12, "psiflx_y": 13, "psiflx_z": 14,
             "dpsidt_old": 15, "dudt_old": 16, "dvdt_old": 17, "dwdt_old": 18}
SCALAR_IDS = {"time": 0, "dt": 1, "dto": 2, "gam": 3, "seps": 4, "rho0": 5, "divtot": 6, "divmax": 7,
              "dt_cfl": 8, "istep": 9, "is_first": 10}


def bc3(x, y, z):
    """Six boundary codes from three pairs, e.g. bc3('PP','PP','DD')."""
    return list(x) + list(y) + list(z)


def rising_bubble_3d(ng=(64, 64, 128), rho12=(1000.0, 100.0), mu12=(10.0, 1.0), cfl=0.25, thickness=1.0):
    """3-D rising bubble: periodic x/y, no-slip walls in z, one sphere r=0.25 at
    (0.5,0.5,0.5) in a 1x1x2 box, Re=35, Eo=10 (sigma=24.5, g=-0.98), ACDI."""
    return dict(
        ng=list(ng), l=[1.0, 1.0, 2.0], cfl=cfl, dtmax=1e9, dt_f=-1.0, icheck=10,
        cbcvel=bc3("PP", "PP", "DD") * 3, cbcpre=bc3("PP", "PP", "NN"), cbcpsi=bc3("PP", "PP", "NN"),
        cbcnor=bc3("PP", "PP", "NN") + bc3("PP", "PP", "NN") + bc3("PP", "PP", "DD"),
        gacc=[0.0, 0.0, -0.98], rho12=list(rho12), mu12=list(mu12), sigma=24.5,
        inipsi="bub3", spheres=[(0.5, 0.5, 0.5, 0.25)], psi_thickness_factor=thickness, vof=False)


def config0():
    """64x64x128 plain FP64 rising bubble, density and viscosity ratio 10, 100 steps."""
    return rising_bubble_3d()


def config2():
    """256x256x512 ACDI rising bubble, density ratio 1000 and viscosity ratio 100."""
    return rising_bubble_3d(ng=(256, 256, 512), rho12=(1000.0, 1.0), mu12=(10.0, 0.1))


def hysing_2d(ng=(64, 2, 128), cfl=0.95, dtmax=5e-4):
    """The reference's own test tests/rising_bubble/input.nml: Hysing case 1 in 2-D."""
    return dict(
        ng=list(ng), l=[1.0, 0.03125 * (ng[1] / 2.0) * (64.0 / ng[0]), 2.0], cfl=cfl, dtmax=dtmax, dt_f=-1.0,
        icheck=10,
        cbcvel=bc3("DD", "PP", "DD") + bc3("NN", "PP", "DD") + bc3("NN", "PP", "DD"),
        cbcpre=bc3("NN", "PP", "NN"), cbcpsi=bc3("NN", "PP", "NN"),
        cbcnor=bc3("DD", "PP", "NN") + bc3("NN", "PP", "NN") + bc3("NN", "PP", "DD"),
        gacc=[0.0, 0.0, -0.98], rho12=[1000.0, 100.0], mu12=[10.0, 1.0], sigma=24.5,
        inipsi="bub2", spheres=[(0.5, 1.0, 0.5, 0.25)], psi_thickness_factor=0.51, vof=False)


def lid_driven_cavity(ng=(2, 64, 64)):
    """The reference's test tests/lid_driven_cavity/input.nml (single phase, lid v=1 at z=1)."""
    bcvel = [0.0] * 18
    bcvel[5 + 6] = 1.0  # bcvel(1,3,2) = 1: v on the top wall
    return dict(
        ng=list(ng), l=[0.03125 * ng[0] / 2.0, 1.0, 1.0], cfl=0.95, dtmax=1e9, dt_f=-1.0, icheck=10,
        cbcvel=bc3("PP", "DD", "DD") * 3, bcvel=bcvel, cbcpre=bc3("PP", "NN", "NN"),
        cbcpsi=bc3("PP", "NN", "NN"),
        cbcnor=bc3("PP", "NN", "NN") + bc3("PP", "DD", "NN") + bc3("PP", "NN", "DD"),
        gacc=[0.0, 0.0, 0.0], rho12=[1.0, 1.0], mu12=[0.01, 0.01], sigma=0.0, inipsi="uni", spheres=[], vof=False)


def vof_droplets(ng=(64, 64, 64), ndrop=4, r=0.3, seed=20261202, rho12=(50.0, 1.0), mu12=(0.02, 0.002), sigma=0.1):
    """Triply periodic THINC/QQ droplets with the reference's Taylor-Green forcing
    (src/forcing.f90:48-63), droplet centres from SplitMix64 rejection sampling."""
    L = 2.0 * math.pi
    cents = seeded_centres(ndrop, r, L, 2.0 * r + 2.0 * L / ng[0], seed)
    return dict(
        ng=list(ng), l=[L, L, L], cfl=0.5, dtmax=1e9, dt_f=-1.0, icheck=10, is_forced_hit=True,
        cbcvel=bc3("PP", "PP", "PP") * 3, cbcpre=bc3("PP", "PP", "PP"), cbcpsi=bc3("PP", "PP", "PP"),
        cbcnor=bc3("PP", "PP", "PP") * 3, gacc=[0.0, 0.0, 0.0], rho12=list(rho12), mu12=list(mu12), sigma=sigma,
        inipsi="drp3", spheres=[(x, y, z, r) for (x, y, z) in cents], vof=True, psi_thickness_factor=0.5)


def config3(ng=256, mixed=False):
    """THINC/QQ droplets in forced turbulence: triply periodic box of side
    2*pi, 32 droplets of radius 0.3 at seeded non-overlapping centres, density
    ratio 50, Weber number 5 on the carrier density, a unit velocity and the
    droplet diameter (sigma = 1*1*0.6/5), the reference's Taylor-Green
    large-scale forcing (src/forcing.f90; the reference has no random
    solenoidal forcing), fluid initially at rest."""
    spec = vof_droplets(ng=(ng, ng, ng), ndrop=32, r=0.3, rho12=(50.0, 1.0), mu12=(0.02, 0.002),
                        sigma=1.0 * 1.0 * 0.6 / 5.0)
    if mixed:
        spec = dict(spec, precision="mixed", refine=1)
    return spec


def splitmix64(seed):
    """SplitMix64 generator of 53-bit uniforms in [0,1)."""
    s = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z = z ^ (z >> 31)
        yield (z >> 11) * (1.0 / 9007199254740992.0)


def seeded_centres(n, r, L, dmin, seed):
    g = splitmix64(seed)
    out = []
    while len(out) < n:
        c = (next(g) * L, next(g) * L, next(g) * L)
        ok = True
        for o in out:
            d2 = 0.0
            for a, b in zip(c, o):
                d = abs(a - b)
                d = min(d, L - d)
                d2 += d * d
            if d2 < dmin * dmin:
                ok = False
                break
        if ok:
            out.append(c)
    return out


def seeded_uniform(n, seed=20261202):
    """n SplitMix64 uniforms in [-1,1), vectorised (same stream as splitmix64)."""
    idx = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        s = np.uint64(seed) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = s
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def fill_params(p, case):
    p.ng[:] = [int(x) for x in case["ng"]]
    p.l[:] = [float(x) for x in case["l"]]
    p.gtype = int(case.get("gtype", 1))
    p.gr = float(case.get("gr", 0.0))
    p.cfl = float(case.get("cfl", 0.95))
    p.dtmax = float(case.get("dtmax", 1e9))
    p.dt_f = float(case.get("dt_f", -1.0))
    p.is_solve_ns = int(case.get("is_solve_ns", True))
    p.is_track_interface = int(case.get("is_track_interface", True))
    p.is_forced_hit = int(case.get("is_forced_hit", False))
    p.icheck = int(case.get("icheck", 10))
    p.vof = int(case.get("vof", False))
    p.cbcvel = "".join(case["cbcvel"]).encode()
    p.cbcpre = "".join(case["cbcpre"]).encode()
    p.cbcpsi = "".join(case["cbcpsi"]).encode()
    p.cbcnor = "".join(case["cbcnor"]).encode()
    p.bcvel[:] = list(case.get("bcvel", [0.0] * 18))
    p.bcpre[:] = list(case.get("bcpre", [0.0] * 6))
    p.bcpsi[:] = list(case.get("bcpsi", [0.0] * 6))
    p.bcnor[:] = list(case.get("bcnor", [0.0] * 18))
    p.bforce[:] = list(case.get("bforce", [0.0] * 3))
    p.gacc[:] = list(case.get("gacc", [0.0] * 3))
    p.sigma = float(case.get("sigma", 0.0))
    p.rho12[:] = list(case.get("rho12", [1.0, 1.0]))
    p.mu12[:] = list(case.get("mu12", [0.01, 0.01]))
    p.beta12[:] = list(case.get("beta12", [1.0, 1.0]))
    p.psi_thickness_factor = float(case.get("psi_thickness_factor", 0.50 if case.get("vof") else 0.51))
    p.acdi_gam_factor = float(case.get("acdi_gam_factor", 1.0))
    p.acdi_gam_min = float(case.get("acdi_gam_min", 1e-12))
    sph = case.get("spheres", [])
    if len(sph) > 64:
        raise ValueError("at most 64 spheres")
    p.nspheres = len(sph)
    flat = [float(x) for s in sph for x in s]
    p.spheres[:len(flat)] = flat
    p.inipsi = INIPSI[case.get("inipsi", "uni")]
    return p


def make_params(case):
    return fill_params(FzParams(), case)


PRECISIONS = {"fp64": 0, "mixed": 1}
# fields a mixed-precision case holds in FP32
F32_FIELDS = {"psi", "phi", "kappa", "normx", "normy", "normz", "psiflx_x", "psiflx_y", "psiflx_z", "dpsidt_old",
              "dudt_old", "dvdt_old", "dwdt_old"}


class CaseBatch:
    """fz_case_batch_*: same-shaped cases stepped together, one launch per
    stage of the step for all of them (each case's results are those of its
    own steps bit for bit)."""

    def __init__(self, cases):
        self.cases = list(cases)
        arr = (vp * len(self.cases))(*[c._h.value for c in self.cases])
        h = vp()
        call("fz_case_batch_create", arr, len(self.cases), C.byref(h))
        self._h = h

    def step(self, nsteps=1):
        """nsteps steps of every case; returns the cases a check stopped."""
        killed = (C.c_int * len(self.cases))()
        call("fz_case_batch_step", self._h, int(nsteps), killed)
        return [c for c, k in zip(self.cases, killed) if k]

    def close(self):
        if getattr(self, "_h", None):
            lib().fz_case_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceCase:
    """The device-resident case (fz_case_* of include/fizzy.h)."""

    def __init__(self, case, stream=None):
        self.spec = case
        self.prm = make_params(case)
        self.n = [int(x) for x in case["ng"]]
        h = vp()
        prec = PRECISIONS[case.get("precision", "fp64")]
        call("fz_case_create_ex", C.byref(h), C.byref(self.prm), vp(stream) if stream else None, prec,
             int(case.get("refine", 1)))
        self._h = h
        c = vp()
        call("fz_case_ctx", self._h, C.byref(c))
        self.ctx = c

    @classmethod
    def view(cls, handle, case):
        """A DeviceCase over a case another owner created (e.g. the server); not destroyed by close()."""
        self = cls.__new__(cls)
        self.spec = case
        self.prm = make_params(case)
        self.n = [int(x) for x in case["ng"]]
        self._h = handle if isinstance(handle, vp) else vp(handle)
        self._owned = False
        c = vp()
        call("fz_case_ctx", self._h, C.byref(c))
        self.ctx = c
        return self

    def close(self):
        if getattr(self, "_h", None):
            if getattr(self, "_owned", True):
                lib().fz_case_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _fshape(self, fid):
        n = self.n
        return (n[0] + 2, n[1] + 2, n[2] + 2) if fid < 15 else (n[0], n[1], n[2])

    def autotune(self, trial_steps=3):
        """fz_case_autotune: launch parameters picked by timing; returns the report."""
        buf = C.create_string_buffer(256)
        call("fz_case_autotune", self._h, int(trial_steps), buf, len(buf))
        return buf.value.decode()

    def set_graphs(self, on=True):
        call("fz_case_set_graphs", self._h, 1 if on else 0)

    def precision(self):
        """('fp64' | 'mixed', refinements)."""
        pr, nr = C.c_int(), C.c_int()
        call("fz_case_precision", self._h, C.byref(pr), C.byref(nr))
        return {v: k for k, v in PRECISIONS.items()}[pr.value], nr.value

    def field_dtype(self, name):
        mixed = self.spec.get("precision", "fp64") == "mixed"
        return np.float32 if mixed and name in F32_FIELDS else np.float64

    def field_ptr(self, name):
        fid = FIELD_IDS[name]
        p = vp()
        cnt = C.c_size_t()
        fn = "fz_case_field_f32" if self.field_dtype(name) == np.float32 else "fz_case_field"
        call(fn, self._h, fid, C.byref(p), C.byref(cnt))
        return p.value, cnt.value

    def get(self, name):
        """The field as float64 (FP32 fields of a mixed case are widened)."""
        fid = FIELD_IDS[name]
        ptr, cnt = self.field_ptr(name)
        dt = self.field_dtype(name)
        out = np.empty(self._fshape(fid), dtype=dt, order="F")
        call("fz_download", self.ctx, out.ctypes.data_as(vp), vp(ptr), C.c_size_t(cnt * out.itemsize))
        return out.astype(np.float64, order="F")

    def set(self, name, arr):
        fid = FIELD_IDS[name]
        ptr, cnt = self.field_ptr(name)
        a = np.asfortranarray(arr, dtype=self.field_dtype(name))
        assert a.shape == self._fshape(fid) and a.size == cnt
        call("fz_upload", self.ctx, vp(ptr), a.ctypes.data_as(vp), C.c_size_t(cnt * a.itemsize))

    def scalars(self):
        out = (C.c_double * 12)()
        call("fz_case_get_scalars", self._h, out)
        return np.array(out[:])

    def set_scalars(self, vals):
        arr = (C.c_double * 12)(*[float(x) for x in vals])
        call("fz_case_set_scalars", self._h, arr)

    def scalar(self, name):
        return float(self.scalars()[SCALAR_IDS[name]])

    def init(self):
        call("fz_case_init", self._h)

    def step(self, nsteps=1):
        killed = C.c_int(0)
        call("fz_case_step", self._h, int(nsteps), C.byref(killed))
        return killed.value

    def sync(self):
        call("fz_ctx_sync", self.ctx)

    def save(self, prefix):
        """Checkpoint: prefix_u.bin, _v, _w, _p, _psi in the load_one format."""
        call("fz_case_save", self._h, prefix.encode())

    def restart(self, prefix):
        """Restart from save()'s files (src/main.f90:312-407 with restart = T)."""
        call("fz_case_restart", self._h, prefix.encode())

    def write_grid(self, datadir):
        call("fz_write_grid", self.ctx, datadir.encode())

    def out3d(self, filename, name, nskip=(1, 1, 1)):
        ptr, _ = self.field_ptr(name)
        call("fz_out3d", self.ctx, filename.encode(), (C.c_int * 3)(*nskip), vp(ptr))
