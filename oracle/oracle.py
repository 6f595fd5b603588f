"""ctypes wrapper around the CPU restatement (liboracle.so) and, where it was
built, the reference's own scheme objects (_ref/libnucref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product path never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)

WENO5JS, WENO5Z, WENO7JS, CRWENO5, HCCS, WENO3JS, UPWIND1ST, HCCS78, HCCS77 = range(9)
CD6, CD8, CD2 = 0, 1, 2
LF, AUSM, AUSMP = 0, 1, 2
EULER, NS = 0, 1
SUTHERLAND, CONSTANT = 0, 1
FORCE_NONE, FORCE_CONST, FORCE_ABC, FORCE_TG = 0, 1, 2, 3
SCHEME_NAMES = {"weno5js": WENO5JS, "weno5z": WENO5Z, "weno7js": WENO7JS, "crweno5": CRWENO5, "hccs": HCCS,
                "weno3js": WENO3JS, "upwind1st": UPWIND1ST, "hccs78": HCCS78, "hccs77": HCCS77}


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("npx", ctypes.c_int), ("npy", ctypes.c_int), ("npz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("scheme", ctypes.c_int * 3), ("vis_scheme", ctypes.c_int * 3),
        ("riemann", ctypes.c_int), ("model", ctypes.c_int), ("viscosity", ctypes.c_int),
        ("forcing", ctypes.c_int),
        ("gamma", ctypes.c_double), ("mach", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("t_ref", ctypes.c_double), ("t_inf", ctypes.c_double),
        ("forcing_amp", ctypes.c_double), ("forcing_amp_hel", ctypes.c_double),
        ("forcing_k", ctypes.c_double), ("cfl", ctypes.c_double), ("nthreads", ctypes.c_int),
        ("coriolis", ctypes.c_int), ("centrifugal", ctypes.c_int),
        ("rossby", ctypes.c_double), ("x_center", ctypes.c_double), ("y_center", ctypes.c_double),
        ("well_rad", ctypes.c_double), ("origin", ctypes.c_double * 3),
    ]


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        # NUC_ORACLE_LIB: a throwaway build with a planted mistake (tests only)
        path = os.environ.get("NUC_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run: make -f oracle/Makefile oracle)")
        L = ctypes.CDLL(path)
        L.orc_scheme_halo.restype = ctypes.c_int
        L.orc_line_reconstruct.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_line_diff.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_field_reconstruct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.orc_field_diff.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        L.orc_rk3.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_long]
        L.orc_con2prim.argtypes = [_dp, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_split.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_viscosity.argtypes = [ctypes.c_int] + [ctypes.c_double] * 7 + [_dp, _dp]
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_state.argtypes = [ctypes.c_void_p, _dp]
        L.orc_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_init_tgv.argtypes = [ctypes.POINTER(OrcConfig), _dp]
        L.orc_init_default.argtypes = [ctypes.POINTER(OrcConfig), _dp]
        L.orc_init_ivc.argtypes = [ctypes.POINTER(OrcConfig), _dp]
        L.orc_stats.argtypes = [ctypes.c_void_p, _dp]
        L.orc_hccs_row.argtypes = [_dp, _dp]
        L.orc_post_create.restype = ctypes.c_void_p
        L.orc_post_create.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_post_destroy.argtypes = [ctypes.c_void_p]
        L.orc_post_update.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.orc_post_field.restype = ctypes.POINTER(ctypes.c_double)
        L.orc_post_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_post_steps.restype = ctypes.c_long
        L.orc_post_steps.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def ref_lib():
    """The reference's scheme objects, or None where oracle/_ref was not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libnucref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        R.ref_reconstruct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        R.ref_diff.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, _dp, _dp, ctypes.c_int, _dp]
        R.ref_rk3.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int]
        _ref = R
    return _ref


# --------------------------------------------------------------------------- operators
def halo(scheme):
    return lib().orc_scheme_halo(scheme)


def line_reconstruct(scheme, upwind, f_ext, n, ts, bf):
    """f_ext: line of n cells with bf ghost cells each side (length n+2bf)."""
    f_ext = np.ascontiguousarray(f_ext, dtype=np.float64)
    out = np.zeros(n + 1)
    base = f_ext.ctypes.data + 8 * bf
    lib().orc_line_reconstruct(scheme, upwind, ctypes.cast(base, _dp), n, ts, _ptr(out))
    return out


def line_diff(scheme, f_ext, n, ts, h, bf):
    f_ext = np.ascontiguousarray(f_ext, dtype=np.float64)
    out = np.zeros(n)
    base = f_ext.ctypes.data + 8 * bf
    lib().orc_line_diff(scheme, ctypes.cast(base, _dp), n, ts, h, _ptr(out))
    return out


def field_reconstruct(scheme, upwind, ts, fin, gL, gR, use_ref=False):
    """fin: array shaped (n3, n2, n1) (C order = x fastest); ghosts (n3, n2, bf)."""
    n3, n2, n1 = fin.shape
    bf = gL.shape[2]
    out = np.zeros((n3, n2, n1 + 1))
    args = (scheme, upwind, ts, _ptr(np.ascontiguousarray(fin)), n1, n2, n3,
            _ptr(np.ascontiguousarray(gL)), _ptr(np.ascontiguousarray(gR)), bf, _ptr(out))
    if use_ref:
        rc = ref_lib().ref_reconstruct(*args)
        assert rc == 0
    else:
        lib().orc_field_reconstruct(*args)
    return out


def field_diff(scheme, ts, fin, h, gL, gR, use_ref=False):
    n3, n2, n1 = fin.shape
    bf = gL.shape[2]
    out = np.zeros((n3, n2, n1))
    args = (scheme, ts, _ptr(np.ascontiguousarray(fin)), float(h), n1, n2, n3,
            _ptr(np.ascontiguousarray(gL)), _ptr(np.ascontiguousarray(gR)), bf, _ptr(out))
    if use_ref:
        assert ref_lib().ref_diff(*args) == 0
    else:
        lib().orc_field_diff(*args)
    return out


def rk3(stage, rhs, un, ui, dt, use_ref=False):
    ui = np.array(ui, dtype=np.float64, copy=True)
    if use_ref:
        nvar, n = rhs.shape
        ref_lib().ref_rk3(stage, _ptr(np.ascontiguousarray(rhs)), _ptr(np.ascontiguousarray(un)), _ptr(ui),
                          float(dt), nvar, n)
    else:
        lib().orc_rk3(stage, _ptr(np.ascontiguousarray(rhs)), _ptr(np.ascontiguousarray(un)), _ptr(ui),
                      float(dt), ui.size)
    return ui


def hccs_row(f5):
    """(lo, di, hi, rhs) of one interior hccs row from its 5-point stencil."""
    f5 = np.ascontiguousarray(f5, dtype=np.float64)
    out = np.zeros(4)
    lib().orc_hccs_row(_ptr(f5), _ptr(out))
    return out


def con2prim(q, gamma, mach):
    q = np.ascontiguousarray(q, dtype=np.float64)
    w = np.zeros(8)
    lib().orc_con2prim(_ptr(q), gamma, mach, _ptr(w))
    return w


def split(riemann, d, q, w):
    fp = np.zeros(5)
    fm = np.zeros(5)
    lib().orc_split(riemann, d, _ptr(np.ascontiguousarray(q, dtype=np.float64)),
                    _ptr(np.ascontiguousarray(w, dtype=np.float64)), _ptr(fp), _ptr(fm))
    return fp, fm


def viscosity(law, T, re, mach, pr, gamma, t_ref, t_inf):
    m = ctypes.c_double()
    c = ctypes.c_double()
    lib().orc_viscosity(law, T, re, mach, pr, gamma, t_ref, t_inf, ctypes.byref(m), ctypes.byref(c))
    return m.value, c.value


# --------------------------------------------------------------------------- solver
def make_config(n=(16, 16, 16), npr=(1, 1, 1), L=None, scheme=WENO5JS, vis=CD6, riemann=LF, model=NS,
                viscosity=SUTHERLAND, forcing=FORCE_NONE, gamma=1.4, mach=0.1, reynolds=1600.0,
                prandtl=0.72, t_ref=110.4, t_inf=288.0, forcing_amp=0.01, forcing_amp_hel=0.01,
                forcing_k=1.0, cfl=0.5, nthreads=1, coriolis=False, centrifugal=False, rossby=0.0,
                x_center=3.1415926, y_center=3.1415926, well_rad=1.0, origin=(0.0, 0.0, 0.0)):
    c = OrcConfig()
    c.nx, c.ny, c.nz = n
    c.npx, c.npy, c.npz = npr
    if L is None:
        L = (2 * np.pi,) * 3
    c.lx, c.ly, c.lz = L
    sch = scheme if isinstance(scheme, (tuple, list)) else (scheme,) * 3
    vs = vis if isinstance(vis, (tuple, list)) else (vis,) * 3
    for d in range(3):
        c.scheme[d] = sch[d]
        c.vis_scheme[d] = vs[d]
    c.riemann, c.model, c.viscosity, c.forcing = riemann, model, viscosity, forcing
    c.gamma, c.mach, c.reynolds, c.prandtl = gamma, mach, reynolds, prandtl
    c.t_ref, c.t_inf = t_ref, t_inf
    c.forcing_amp, c.forcing_amp_hel, c.forcing_k = forcing_amp, forcing_amp_hel, forcing_k
    c.cfl = cfl
    c.nthreads = nthreads
    c.coriolis, c.centrifugal = int(bool(coriolis)), int(bool(centrifugal))
    c.rossby, c.x_center, c.y_center, c.well_rad = rossby, x_center, y_center, well_rad
    for d in range(3):
        c.origin[d] = origin[d]
    return c


def init_tgv(cfg):
    N = cfg.nx * cfg.ny * cfg.nz
    q = np.zeros(5 * N)
    lib().orc_init_tgv(ctypes.byref(cfg), _ptr(q))
    return q


def init_default(cfg):
    N = cfg.nx * cfg.ny * cfg.nz
    q = np.zeros(5 * N)
    lib().orc_init_default(ctypes.byref(cfg), _ptr(q))
    return q


def init_ivc(cfg):
    N = cfg.nx * cfg.ny * cfg.nz
    q = np.zeros(5 * N)
    lib().orc_init_ivc(ctypes.byref(cfg), _ptr(q))
    return q


class Solver:
    def __init__(self, cfg):
        self.cfg = cfg
        self.N = cfg.nx * cfg.ny * cfg.nz
        self.h = lib().orc_create(ctypes.byref(cfg))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_state(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == 5 * self.N
        lib().orc_set_state(self.h, _ptr(q))

    def get_state(self):
        q = np.zeros(5 * self.N)
        lib().orc_get_state(self.h, _ptr(q))
        return q

    def rhs(self):
        r = np.zeros(5 * self.N)
        dt = ctypes.c_double()
        lib().orc_rhs(self.h, _ptr(r), ctypes.byref(dt))
        return r, dt.value

    def step(self, nsteps=1):
        dt = ctypes.c_double()
        res = ctypes.c_double()
        t = ctypes.c_double()
        lib().orc_step(self.h, nsteps, ctypes.byref(dt), ctypes.byref(res), ctypes.byref(t))
        return dt.value, res.value, t.value

    def stats(self):
        out = np.zeros(25)
        lib().orc_stats(self.h, _ptr(out))
        return out


POST_COUNTS = (10, 21, 28, 49, 45)   # primary, grads, derived, TKE, budget (src/postproc.hpp:49-121)


class Post:
    """Time-averaged fields and TKE budget of postproc::solveAveraged0/1/2
    (src/postproc.cpp:1186-2110) over a Solver's successive states."""

    def __init__(self, solver, avg_type=3):
        self.solver = solver
        self.N = solver.N
        self.h = lib().orc_post_create(solver.h, int(avg_type))
        if not self.h:
            raise ValueError("averaging type must be 1, 2 or 3")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_post_destroy(self.h)
            self.h = None

    def update(self):
        lib().orc_post_update(self.h, self.solver.h)

    def steps(self):
        return lib().orc_post_steps(self.h)

    def field(self, group, index):
        ptr = lib().orc_post_field(self.h, int(group), int(index))
        if not ptr:
            raise ValueError(f"no field {group}/{index} at this averaging type")
        return np.ctypeslib.as_array(ptr, shape=(self.N,)).copy()

    def group(self, group):
        return np.stack([self.field(group, i) for i in range(POST_COUNTS[group])])
