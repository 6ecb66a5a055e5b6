"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg. The product path never imports it.

Arrays follow the reference layout (column-major, one-cell halo); numpy arrays
passed here are Fortran-ordered with shape (n1+2, n2+2, n3+2) for halo'd
fields and (n1, n2, n3) for interior-only ones.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

R2HC, HC2R, REDFT00, REDFT01, REDFT10, REDFT11, RODFT00, RODFT01, RODFT10, RODFT11 = 0, 1, 3, 4, 5, 6, 7, 8, 9, 10

FIELD_IDS = {"u": 0, "v": 1, "w": 2, "p": 3, "pn": 4, "po": 5, "psi": 6, "phi": 7, "kappa": 8,
             "normx": 9, "normy": 10, "normz": 11, "psiflx_x": 12, "psiflx_y": 13, "psiflx_z": 14,
             "dpsidt_old": 15, "dudt_old": 16, "dvdt_old": 17, "dwdt_old": 18}
SCALAR_IDS = {"time": 0, "dt": 1, "dto": 2, "gam": 3, "seps": 4, "rho0": 5, "divtot": 6, "divmax": 7,
              "dt_cfl": 8, "istep": 9, "is_first": 10}
INIPSI = {"uni": 0, "zer": 1, "bub3": 2, "bub2": 3, "bub1": 4, "drp3": 5, "drp2": 6, "drp1": 7, "flm": 8}


class OrParams(C.Structure):
    _fields_ = [
        ("ng", C.c_int * 3), ("l", C.c_double * 3), ("gtype", C.c_int), ("gr", C.c_double),
        ("cfl", C.c_double), ("dtmax", C.c_double), ("dt_f", C.c_double),
        ("is_solve_ns", C.c_int), ("is_track_interface", C.c_int), ("is_forced_hit", C.c_int),
        ("icheck", C.c_int), ("vof", C.c_int),
        ("cbcvel", C.c_char * 18), ("cbcpre", C.c_char * 6), ("cbcpsi", C.c_char * 6), ("cbcnor", C.c_char * 18),
        ("bcvel", C.c_double * 18), ("bcpre", C.c_double * 6), ("bcpsi", C.c_double * 6), ("bcnor", C.c_double * 18),
        ("bforce", C.c_double * 3), ("gacc", C.c_double * 3), ("sigma", C.c_double),
        ("rho12", C.c_double * 2), ("mu12", C.c_double * 2), ("beta12", C.c_double * 2),
        ("psi_thickness_factor", C.c_double), ("acdi_gam_factor", C.c_double), ("acdi_gam_min", C.c_double),
        ("nspheres", C.c_int), ("inipsi", C.c_int), ("spheres", C.c_double * 256),
    ]

    @classmethod
    def from_case(cls, case):
        p = cls()
        fill_struct(p, case)
        return p


def fill_struct(p, case):
    """Fill a params struct from a case dict (keys as in fizzy_amd.cases)."""
    p.ng[:] = list(case["ng"])
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
    p.nspheres = len(sph)
    flat = [float(x) for s in sph for x in s]
    p.spheres[:len(flat)] = flat
    p.inipsi = INIPSI[case.get("inipsi", "uni")]


PLANTED_PATH = os.path.join(HERE, "liboracle_planted.so")
# the numbered planted mistakes of liboracle_planted.so (the PL_* enum of fz_oracle.c)
PLANTED = ["ACDI_SHARP_SIGN", "ACDI_QUARTER", "ACDI_TANH_ARG", "ACDI_NORMAL_COMP", "ACDI_DZCI_M", "THINC_UPWIND",
           "THINC_OFFSET", "THINC_NORMAL_COMP", "OTH_BRACKBILL", "OTH_KAPPA_FACE", "OTH_EXTRAP", "OTH_SPLIT",
           "OTH_SURF_FACTOR", "DT_VISC", "DT_CAP_RHO", "DT_GRAV", "DT_DLMIN", "DT_COMBINE", "INIT_WIDTH",
           "INIT_PERIODIC", "INIT_FLM_INDEX"]

_lib = None
_planted = None


def planted_lib(name):
    """liboracle_planted.so with the planted mistake `name` selected (a name
    of PLANTED): the restatement with one deliberate slip, for showing that a
    pin test fails on it. Never the checker of anything."""
    global _planted
    if _planted is None:
        if not os.path.exists(PLANTED_PATH):
            raise RuntimeError(f"oracle library missing: {PLANTED_PATH} (run make -C oracle)")
        _planted = _setup(C.CDLL(PLANTED_PATH))
    C.c_int.in_dll(_planted, "or_planted").value = PLANTED.index(name) + 1
    return _planted


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = _setup(C.CDLL(LIB_PATH))
    return _lib


def _setup(L):
    dp = C.POINTER(C.c_double)
    L.or_case_create.restype = C.c_void_p
    L.or_case_create.argtypes = [C.POINTER(OrParams)]
    L.or_case_destroy.argtypes = [C.c_void_p]
    L.or_case_field.restype = dp
    L.or_case_field.argtypes = [C.c_void_p, C.c_int]
    L.or_case_scalars.restype = dp
    L.or_case_scalars.argtypes = [C.c_void_p]
    L.or_case_init_fields.argtypes = [C.c_void_p]
    L.or_case_step.argtypes = [C.c_void_p, C.c_int]
    L.or_case_step.restype = C.c_int
    for name in ["or_chk_poisson", "or_acdi_set_gamma", "or_acdi_set_epsilon", "or_chkdt", "or_bulk_mean_12_stag"]:
        getattr(L, name).restype = C.c_double
    return L


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1), "need float64 Fortran-ordered arrays"
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ivec(v):
    return (C.c_int * len(v))(*[int(x) for x in v])


def dvec(v):
    return (C.c_double * len(v))(*[float(x) for x in v])


def cstr(s):
    return C.c_char_p("".join(s).encode())


def halo(n, fill=0.0):
    return np.full((n[0] + 2, n[1] + 2, n[2] + 2), fill, dtype=np.float64, order="F")


def interior(n, fill=0.0):
    return np.full((n[0], n[1], n[2]), fill, dtype=np.float64, order="F")


# --------------------------------------------------------------------------------------------
def r2r(kind, x):
    """Transform the rows of a 2-D C-ordered array (each row one line) in place copy."""
    x = np.ascontiguousarray(x, dtype=np.float64).copy()
    if x.ndim == 1:
        x = x[None, :]
    lib().or_r2r(C.c_int(kind), C.c_int(x.shape[1]), C.c_int(x.shape[0]), x.ctypes.data_as(C.POINTER(C.c_double)))
    return x


def initgrid(gtype, n, gr, lz):
    arrs = [np.zeros(n + 2) for _ in range(4)]
    lib().or_initgrid(C.c_int(gtype), C.c_int(n), C.c_double(gr), C.c_double(lz), *[ptr(a) for a in arrs])
    return arrs  # dzc, dzf, zc, zf


class Solver:
    """initsolver + solver for a cell-centred grid (src/initsolver.f90, src/solver.f90)."""

    def __init__(self, ng, dli, dzci, dzfi, cbc, bc=(0.0,) * 6):
        self.n = list(ng)
        self.cbc = list(cbc)
        n1, n2, n3 = ng
        self.lambdaxy = np.zeros((n1, n2), order="F")
        self.a, self.b, self.c = np.zeros(n3), np.zeros(n3), np.zeros(n3)
        self.rhsbx = np.zeros((n2, n3, 2), order="F")
        self.rhsby = np.zeros((n1, n3, 2), order="F")
        self.rhsbz = np.zeros((n1, n2, 2), order="F")
        nf = C.c_double()
        lib().or_initsolver(ivec(ng), dvec(dli), ptr(np.asarray(dzci, float)), ptr(np.asarray(dzfi, float)),
                            cstr(cbc), dvec(bc), ptr(self.lambdaxy), ptr(self.a), ptr(self.b), ptr(self.c),
                            C.byref(nf), ptr(self.rhsbx), ptr(self.rhsby), ptr(self.rhsbz))
        self.normfft = nf.value

    def solve(self, p):
        lib().or_solver(ivec(self.n), cstr(self.cbc), C.c_double(self.normfft), ptr(self.lambdaxy), ptr(self.a),
                        ptr(self.b), ptr(self.c), ptr(p))
        return p


def chk_poisson(n, dli, dzci, dzfi, fp, fpp):
    return lib().or_chk_poisson(ivec(n), dvec(dli), ptr(dzci), ptr(dzfi), ptr(fp), ptr(fpp))


def chkdiv(n, dli, dzfi, u, v, w):
    out = np.zeros(2)
    lib().or_chkdiv(ivec(n), dvec(dli), ptr(dzfi), ptr(u), ptr(v), ptr(w), ptr(out))
    return out


class Case:
    """A whole reference time loop (src/main.f90:414-562) on the CPU."""

    def __init__(self, case):
        self.spec = case
        self.prm = OrParams.from_case(case)
        self._h = lib().or_case_create(C.byref(self.prm))
        self.n = list(case["ng"])

    def __del__(self):
        try:
            if self._h:
                lib().or_case_destroy(C.c_void_p(self._h))
                self._h = None
        except Exception:
            pass

    def field(self, name):
        """A live numpy view of one oracle field."""
        fid = FIELD_IDS[name]
        n = self.n
        shape = (n[0] + 2, n[1] + 2, n[2] + 2) if fid < 15 else (n[0], n[1], n[2])
        p = lib().or_case_field(C.c_void_p(self._h), C.c_int(fid))
        return np.ctypeslib.as_array(p, shape=(int(np.prod(shape)),)).reshape(shape, order="F")

    def scalars(self):
        p = lib().or_case_scalars(C.c_void_p(self._h))
        return np.ctypeslib.as_array(p, shape=(12,))

    def scalar(self, name):
        return float(self.scalars()[SCALAR_IDS[name]])

    def set_scalar(self, name, value):
        self.scalars()[SCALAR_IDS[name]] = value

    def init_fields(self):
        lib().or_case_init_fields(C.c_void_p(self._h))

    def finish_init(self):
        """The initialisation after the fields are set, as on a restart."""
        lib().or_case_finish_init(C.c_void_p(self._h))

    def step(self, nsteps=1):
        return lib().or_case_step(C.c_void_p(self._h), C.c_int(nsteps))
