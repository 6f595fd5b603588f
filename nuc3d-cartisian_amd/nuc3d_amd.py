"""Python binding of libnuc3d_amd.so (include/nuc3d_gpu.h, include/nuc3d_serve.h)
through ctypes.

Loads the in-tree library and fails loudly when it is missing; there is no
CPU fallback.  Arrays are numpy float64, x-fastest, conserved state as
5*nx*ny*nz values (rho, rho*u, rho*v, rho*w, rho*E).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnuc3d_amd.so")
_dp = ctypes.POINTER(ctypes.c_double)

WENO5JS, WENO5Z, WENO7JS, CRWENO5, HCCS, WENO3JS, UPWIND1ST, HCCS78, HCCS77 = range(9)
CD6, CD8, CD2 = 0, 1, 2
LF, AUSM, AUSMP = 0, 1, 2
EULER3D, NAVIERSTOKES3D = 0, 1
SUTHERLAND, CONSTANT = 0, 1
FORCING_NONE, FORCING_CONST, FORCING_ABC, FORCING_TG = 0, 1, 2, 3
NSTATS_EX = 27
FP64, MIXED, FP32 = 0, 1, 2
SCHEMES = {"weno5js": WENO5JS, "weno5z": WENO5Z, "weno7js": WENO7JS, "crweno5": CRWENO5, "hccs": HCCS}
STATUS = {0: "NUC_OK", 1: "NUC_EINVAL", 2: "NUC_ENOMEM", 3: "NUC_EHIP", 4: "NUC_ENONFINITE", 5: "NUC_ENODEV",
          6: "NUC_EBUSY"}
NUC_OK, NUC_EBUSY = 0, 6
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = range(4)
IC_STATE, IC_TGV, IC_HIT, IC_IVC, IC_DEFAULT = range(5)
SPEC_DENSITY, SPEC_PRESSURE, SPEC_ENSTROPHY, SPEC_HELICITY = range(4)

# every entry point include/nuc3d_gpu.h and include/nuc3d_serve.h declare
EXPORTS = ["nuc_create", "nuc_destroy", "nuc_last_error", "nuc_footprint", "nuc_upload_state",
           "nuc_download_state", "nuc_rhs", "nuc_rhs_noalloc", "nuc_step", "nuc_step_async", "nuc_set_graph", "nuc_query", "nuc_sync",
           "nuc_stats", "nuc_stats_ex", "nuc_op_reconstruct", "nuc_op_diff", "nuc_op_rk3", "nuc_config_from_inp",
           "nuc_init_tgv", "nuc_init_hit", "nuc_init_shu_osher", "nuc_profile", "nuc_profile_read",
           "nuc_server_create", "nuc_server_destroy", "nuc_server_submit", "nuc_server_result", "nuc_server_drain",
           "nuc_server_stats_get", "nuc_server_now", "nuc_server_job_error",
           "nuc_pack_compatible", "nuc_pack_create", "nuc_pack_step_async", "nuc_pack_destroy",
           "nuc_pack_last_error", "nuc_autotune", "nuc_tile_shape", "nuc_set_tile_shape", "nuc_rk3_stage",
           "nuc_average_begin", "nuc_average_update", "nuc_average_steps", "nuc_average_field",
           "nuc_average_means", "nuc_average_end", "nuc_init_tgv_device", "nuc_init_hit_device",
           "nuc_server_submit_ic", "nuc_spectra_kinetic", "nuc_spectrum_scalar", "nuc_init_default",
           "nuc_init_ivc", "nuc_init_default_device", "nuc_init_ivc_device", "nuc_average_restore",
           "nuc_write_restart", "nuc_read_restart", "nuc_write_tecplot", "nuc_average_write", "nuc_average_read",
           "nuc_average_write_tecplot", "nuc_q_fields", "nuc_write_q_tecplot"]


class NucConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("npx", ctypes.c_int), ("npy", ctypes.c_int), ("npz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("scheme", ctypes.c_int * 3), ("vis_scheme", ctypes.c_int * 3),
        ("riemann", ctypes.c_int), ("model", ctypes.c_int), ("viscosity", ctypes.c_int),
        ("forcing", ctypes.c_int), ("precision", ctypes.c_int),
        ("gamma", ctypes.c_double), ("mach", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("t_ref", ctypes.c_double), ("t_inf", ctypes.c_double),
        ("forcing_amp", ctypes.c_double), ("forcing_amp_hel", ctypes.c_double),
        ("forcing_k", ctypes.c_double), ("cfl", ctypes.c_double),
        ("coriolis", ctypes.c_int), ("centrifugal", ctypes.c_int),
        ("rossby", ctypes.c_double), ("x_center", ctypes.c_double), ("y_center", ctypes.c_double),
        ("well_rad", ctypes.c_double), ("origin", ctypes.c_double * 3),
    ]


class StepInfo(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("time", ctypes.c_double), ("res", ctypes.c_double),
                ("step", ctypes.c_long)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_longlong), ("max_active", ctypes.c_int),
                ("steps_per_slice", ctypes.c_int), ("use_graphs", ctypes.c_int), ("pack_cells", ctypes.c_longlong),
                ("autotune", ctypes.c_int)]


class JobResult(ctypes.Structure):
    _fields_ = [("info", StepInfo), ("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("state", ctypes.c_int), ("error", ctypes.c_int)]


class ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in ("completed", "failed", "queued", "active", "peak_active",
                                                  "bytes_in_use", "peak_bytes", "single_slices",
                                                  "pack_slices")]


class NucError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                               "(make -C nuc3d-cartisian_amd)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.nuc_create.argtypes = [ctypes.POINTER(NucConfig), ctypes.POINTER(vp)]
        L.nuc_destroy.argtypes = [vp]
        L.nuc_last_error.argtypes = [vp]
        L.nuc_last_error.restype = ctypes.c_char_p
        L.nuc_footprint.argtypes = [ctypes.POINTER(NucConfig)]
        L.nuc_footprint.restype = ctypes.c_longlong
        L.nuc_upload_state.argtypes = [vp, _dp]
        L.nuc_download_state.argtypes = [vp, _dp]
        L.nuc_rhs.argtypes = [vp, _dp, _dp]
        L.nuc_rhs_noalloc.argtypes = [ctypes.POINTER(NucConfig), vp, vp, vp, vp]
        L.nuc_step.argtypes = [vp, ctypes.c_int, ctypes.POINTER(StepInfo)]
        L.nuc_step_async.argtypes = [vp, ctypes.c_int]
        L.nuc_set_graph.argtypes = [vp, ctypes.c_int]
        L.nuc_query.argtypes = [vp]
        L.nuc_sync.argtypes = [vp, ctypes.POINTER(StepInfo)]
        L.nuc_stats.argtypes = [vp, _dp]
        L.nuc_stats_ex.argtypes = [vp, _dp, ctypes.c_int]
        L.nuc_op_reconstruct.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.nuc_op_diff.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int]
        L.nuc_op_rk3.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, ctypes.c_long]
        L.nuc_config_from_inp.argtypes = [ctypes.c_char_p, ctypes.POINTER(NucConfig),
                                          ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.nuc_init_tgv.argtypes = [ctypes.POINTER(NucConfig), _dp]
        L.nuc_init_default.argtypes = [ctypes.POINTER(NucConfig), _dp]
        L.nuc_init_ivc.argtypes = [ctypes.POINTER(NucConfig), _dp]
        L.nuc_init_hit.argtypes = [ctypes.POINTER(NucConfig), ctypes.c_double, ctypes.c_double,
                                   ctypes.c_ulonglong, _dp]
        L.nuc_init_shu_osher.argtypes = [ctypes.POINTER(NucConfig), ctypes.c_double, ctypes.c_ulonglong, _dp]
        L.nuc_profile.argtypes = [vp, ctypes.c_int]
        L.nuc_profile_read.argtypes = [vp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_long),
                                       ctypes.POINTER(ctypes.c_char_p)]
        L.nuc_server_create.argtypes = [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        L.nuc_server_destroy.argtypes = [vp]
        L.nuc_server_submit.argtypes = [vp, ctypes.POINTER(NucConfig), _dp, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_longlong)]
        L.nuc_server_submit_ic.argtypes = [vp, ctypes.POINTER(NucConfig), ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_ulonglong, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_longlong)]
        L.nuc_server_result.argtypes = [vp, ctypes.c_longlong, _dp, ctypes.POINTER(JobResult)]
        L.nuc_server_drain.argtypes = [vp, ctypes.c_double]
        L.nuc_server_stats_get.argtypes = [vp, ctypes.POINTER(ServerStats)]
        L.nuc_server_now.argtypes = [vp]
        L.nuc_server_now.restype = ctypes.c_double
        L.nuc_server_job_error.argtypes = [vp, ctypes.c_longlong]
        L.nuc_server_job_error.restype = ctypes.c_char_p
        L.nuc_autotune.argtypes = [vp]
        L.nuc_spectra_kinetic.argtypes = [vp, _dp, ctypes.c_int]
        L.nuc_spectrum_scalar.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int]
        L.nuc_init_tgv_device.argtypes = [vp]
        L.nuc_init_default_device.argtypes = [vp]
        L.nuc_init_ivc_device.argtypes = [vp]
        L.nuc_init_hit_device.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong]
        L.nuc_average_begin.argtypes = [vp, ctypes.c_int]
        L.nuc_average_update.argtypes = [vp]
        L.nuc_average_steps.argtypes = [vp]
        L.nuc_average_steps.restype = ctypes.c_longlong
        L.nuc_average_field.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp]
        L.nuc_average_means.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int]
        L.nuc_average_end.argtypes = [vp]
        L.nuc_average_restore.argtypes = [vp, ctypes.c_int, ctypes.c_longlong, _dp]
        L.nuc_write_restart.argtypes = [ctypes.POINTER(NucConfig), ctypes.c_char_p, ctypes.c_int, ctypes.c_double, _dp]
        L.nuc_read_restart.argtypes = [ctypes.POINTER(NucConfig), ctypes.c_char_p, ctypes.c_int, _dp, _dp]
        L.nuc_write_tecplot.argtypes = [ctypes.POINTER(NucConfig), ctypes.c_char_p, ctypes.c_int, _dp, _dp]
        L.nuc_average_write.argtypes = [vp, ctypes.POINTER(NucConfig), ctypes.c_int, ctypes.c_char_p]
        L.nuc_average_read.argtypes = [vp, ctypes.POINTER(NucConfig), ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.nuc_average_write_tecplot.argtypes = [vp, ctypes.POINTER(NucConfig), ctypes.c_int, ctypes.c_char_p,
                                                ctypes.c_double, _dp]
        L.nuc_q_fields.argtypes = [vp, _dp]
        L.nuc_write_q_tecplot.argtypes = [vp, ctypes.POINTER(NucConfig), ctypes.c_char_p, ctypes.c_int,
                                          ctypes.c_double, _dp]
        L.nuc_rk3_stage.argtypes = [vp, ctypes.c_int]
        L.nuc_tile_shape.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
        L.nuc_set_tile_shape.argtypes = [vp, ctypes.c_int]
        L.nuc_pack_compatible.argtypes = [vp, vp]
        L.nuc_pack_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
        L.nuc_pack_step_async.argtypes = [vp, ctypes.c_int]
        L.nuc_pack_destroy.argtypes = [vp]
        L.nuc_pack_last_error.argtypes = [vp]
        L.nuc_pack_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _check(rc, what, ctx=None):
    if rc != 0:
        msg = what
        if ctx is not None:
            msg += ": " + lib().nuc_last_error(ctx).decode()
        raise NucError(rc, msg)


def make_config(n=(16, 16, 16), npr=(1, 1, 1), L=None, scheme=WENO5JS, vis=CD6, riemann=LF,
                model=NAVIERSTOKES3D, viscosity=SUTHERLAND, forcing=FORCING_NONE, precision=FP64,
                gamma=1.4, mach=0.1, reynolds=1600.0, prandtl=0.72, t_ref=110.4, t_inf=288.0,
                forcing_amp=0.01, forcing_amp_hel=0.01, forcing_k=1.0, cfl=0.5, coriolis=False, centrifugal=False,
                rossby=0.0, x_center=3.1415926, y_center=3.1415926, well_rad=1.0, origin=(0.0, 0.0, 0.0)):
    c = NucConfig()
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
    c.riemann, c.model, c.viscosity, c.forcing, c.precision = riemann, model, viscosity, forcing, precision
    c.gamma, c.mach, c.reynolds, c.prandtl = gamma, mach, reynolds, prandtl
    c.t_ref, c.t_inf = t_ref, t_inf
    c.forcing_amp, c.forcing_amp_hel, c.forcing_k = forcing_amp, forcing_amp_hel, forcing_k
    c.cfl = cfl
    c.coriolis, c.centrifugal = int(bool(coriolis)), int(bool(centrifugal))
    c.rossby, c.x_center, c.y_center, c.well_rad = rossby, x_center, y_center, well_rad
    for d in range(3):
        c.origin[d] = origin[d]
    return c


def config_from_inp(directory):
    c = NucConfig()
    s = ctypes.c_int()
    e = ctypes.c_int()
    _check(lib().nuc_config_from_inp(directory.encode(), ctypes.byref(c), ctypes.byref(s), ctypes.byref(e)),
           "nuc_config_from_inp")
    return c, s.value, e.value


def init_tgv(cfg):
    q = np.zeros(5 * cfg.nx * cfg.ny * cfg.nz)
    _check(lib().nuc_init_tgv(ctypes.byref(cfg), _ptr(q)), "nuc_init_tgv")
    return q


def write_restart(cfg, directory, step, time, q):
    """Per-rank restart files NUC3d_step_<step>_ID_<rank>.bin (nuc_write_restart)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    _check(lib().nuc_write_restart(ctypes.byref(cfg), str(directory).encode(), int(step), float(time), _ptr(q)),
           "nuc_write_restart")


def read_restart(cfg, directory, step):
    """(time, q) assembled from the per-rank restart files (nuc_read_restart)."""
    q = np.zeros(5 * cfg.nx * cfg.ny * cfg.nz)
    t = ctypes.c_double()
    _check(lib().nuc_read_restart(ctypes.byref(cfg), str(directory).encode(), int(step), ctypes.byref(t), _ptr(q)),
           "nuc_read_restart")
    return t.value, q


def write_tecplot(cfg, directory, step, q, origin=None):
    """Per-rank Tecplot files NUC3d_step_<step>_id_<rank>.dat (nuc_write_tecplot)."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    o = None if origin is None else np.ascontiguousarray(origin, dtype=np.float64)
    _check(lib().nuc_write_tecplot(ctypes.byref(cfg), str(directory).encode(), int(step),
                                   None if o is None else _ptr(o), _ptr(q)), "nuc_write_tecplot")


def init_default(cfg):
    """Fluid at rest, p = 1/(gamma M^2) (nuc_init_default)."""
    q = np.zeros(5 * cfg.nx * cfg.ny * cfg.nz)
    _check(lib().nuc_init_default(ctypes.byref(cfg), _ptr(q)), "nuc_init_default")
    return q


def init_ivc(cfg):
    """Isentropic vortex at (5, 5) carried by u = 0.5 (nuc_init_ivc)."""
    q = np.zeros(5 * cfg.nx * cfg.ny * cfg.nz)
    _check(lib().nuc_init_ivc(ctypes.byref(cfg), _ptr(q)), "nuc_init_ivc")
    return q


def init_hit(cfg, k0=4.0, u_rms=1.0, seed=20261202):
    q = np.zeros(5 * cfg.nx * cfg.ny * cfg.nz)
    _check(lib().nuc_init_hit(ctypes.byref(cfg), float(k0), float(u_rms), int(seed), _ptr(q)), "nuc_init_hit")
    return q


def init_shu_osher(cfg, pert=0.05, seed=20261602):
    """Shu-Osher profile along x (box centred on x = 0) with a seeded solenoidal
    velocity perturbation of rms pert * 2.629369 (nuc_init_shu_osher)."""
    q = np.empty(5 * cfg.nx * cfg.ny * cfg.nz)
    _check(lib().nuc_init_shu_osher(ctypes.byref(cfg), float(pert), int(seed), _ptr(q)), "nuc_init_shu_osher")
    return q


def footprint(cfg):
    return lib().nuc_footprint(ctypes.byref(cfg))


class Context:
    """One device context: state resident in HBM, the RHS / RK3 hot path on the GPU."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.N = cfg.nx * cfg.ny * cfg.nz
        h = ctypes.c_void_p()
        rc = lib().nuc_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != 0:
            raise NucError(rc, lib().nuc_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nuc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == 5 * self.N
        _check(lib().nuc_upload_state(self.h, _ptr(q)), "nuc_upload_state", self.h)

    def download(self):
        q = np.zeros(5 * self.N)
        _check(lib().nuc_download_state(self.h, _ptr(q)), "nuc_download_state", self.h)
        return q

    def rhs(self):
        r = np.zeros(5 * self.N)
        dt = ctypes.c_double()
        _check(lib().nuc_rhs(self.h, _ptr(r), ctypes.byref(dt)), "nuc_rhs", self.h)
        return r, dt.value

    def step(self, n=1):
        info = StepInfo()
        _check(lib().nuc_step(self.h, n, ctypes.byref(info)), "nuc_step", self.h)
        return info.dt, info.res, info.time

    def step_async(self, n):
        _check(lib().nuc_step_async(self.h, n), "nuc_step_async", self.h)

    def q_fields(self):
        """(16, N): solveQ's per-point fields of the current state (nuc_q_fields)."""
        out = np.zeros(16 * self.cfg.nx * self.cfg.ny * self.cfg.nz)
        _check(lib().nuc_q_fields(self.h, _ptr(out)), "nuc_q_fields", self.h)
        return out.reshape(16, -1)

    def write_q_tecplot(self, directory, step, time, origin=None):
        """Per-rank Tecplot files Q_tec_step_<step>_id_<rank>.dat of the 16 Q fields."""
        o = None if origin is None else np.ascontiguousarray(origin, dtype=np.float64)
        _check(lib().nuc_write_q_tecplot(self.h, ctypes.byref(self.cfg), str(directory).encode(), int(step),
                                         float(time), None if o is None else _ptr(o)), "nuc_write_q_tecplot", self.h)

    def spectra_kinetic(self):
        """(5, kmax): total, compressive, solenoidal, left, right kinetic energy per shell."""
        kmax = min(self.cfg.nx, self.cfg.ny, self.cfg.nz) // 2
        out = np.zeros(5 * kmax)
        _check(lib().nuc_spectra_kinetic(self.h, _ptr(out), kmax), "nuc_spectra_kinetic", self.h)
        return out.reshape(5, kmax)

    def spectrum_scalar(self, field):
        """|f^|^2 per shell of SPEC_DENSITY, SPEC_PRESSURE, SPEC_ENSTROPHY or SPEC_HELICITY."""
        kmax = min(self.cfg.nx, self.cfg.ny, self.cfg.nz) // 2
        out = np.zeros(kmax)
        _check(lib().nuc_spectrum_scalar(self.h, int(field), _ptr(out), kmax), "nuc_spectrum_scalar", self.h)
        return out

    def init_tgv_device(self):
        """The Taylor-Green field generated on the device into the state."""
        _check(lib().nuc_init_tgv_device(self.h), "nuc_init_tgv_device", self.h)

    def init_default_device(self):
        """init_default's field generated on the device into the state."""
        _check(lib().nuc_init_default_device(self.h), "nuc_init_default_device", self.h)

    def init_ivc_device(self):
        """init_ivc's field generated on the device into the state."""
        _check(lib().nuc_init_ivc_device(self.h), "nuc_init_ivc_device", self.h)

    def init_hit_device(self, k0=4.0, u_rms=1.0, seed=20261202):
        """init_hit's seeded field synthesised on the device into the state."""
        _check(lib().nuc_init_hit_device(self.h, float(k0), float(u_rms), int(seed)), "nuc_init_hit_device", self.h)

    # time-averaged statistics and TKE budget (nuc_average_*)
    AVG_COUNTS = (10, 21, 28, 49, 45)

    def average_begin(self, avg_type=3):
        _check(lib().nuc_average_begin(self.h, int(avg_type)), "nuc_average_begin", self.h)

    def average_update(self):
        _check(lib().nuc_average_update(self.h), "nuc_average_update", self.h)

    def average_steps(self):
        return lib().nuc_average_steps(self.h)

    def average_field(self, group, index):
        out = np.zeros(self.N)
        _check(lib().nuc_average_field(self.h, int(group), int(index), _ptr(out)), "nuc_average_field", self.h)
        return out

    def average_group(self, group):
        return np.stack([self.average_field(group, i) for i in range(self.AVG_COUNTS[group])])

    def average_means(self, group):
        out = np.zeros(self.AVG_COUNTS[group])
        _check(lib().nuc_average_means(self.h, int(group), _ptr(out), out.size), "nuc_average_means", self.h)
        return out

    def average_restore(self, avg_type, steps, fields):
        """Start an average from saved fields (every field of the type's groups, in order)."""
        f = np.ascontiguousarray(fields, dtype=np.float64)
        _check(lib().nuc_average_restore(self.h, int(avg_type), int(steps), _ptr(f)), "nuc_average_restore", self.h)

    def average_write(self, avg_type, directory):
        """Per-rank averaged-statistics files Averaged_step_<n>_id_<rank>.bin."""
        _check(lib().nuc_average_write(self.h, ctypes.byref(self.cfg), int(avg_type), str(directory).encode()),
               "nuc_average_write", self.h)

    def average_write_tecplot(self, avg_type, directory, time, origin=None):
        """Per-rank Tecplot files of the averaged fields, Averaged_tec_step_<n>_id_<rank>.dat."""
        o = None if origin is None else np.ascontiguousarray(origin, dtype=np.float64)
        _check(lib().nuc_average_write_tecplot(self.h, ctypes.byref(self.cfg), int(avg_type), str(directory).encode(),
                                               float(time), None if o is None else _ptr(o)),
               "nuc_average_write_tecplot", self.h)

    def average_read(self, avg_type, directory, step):
        """Restore an average from the per-rank files of averaging step `step`."""
        _check(lib().nuc_average_read(self.h, ctypes.byref(self.cfg), int(avg_type), str(directory).encode(),
                                      int(step)), "nuc_average_read", self.h)

    def average_end(self):
        _check(lib().nuc_average_end(self.h), "nuc_average_end", self.h)

    def rk3_stage(self, stage):
        """One RK stage (nuc_rk3_stage); stages run in order 0, 1, 2."""
        _check(lib().nuc_rk3_stage(self.h, int(stage)), "nuc_rk3_stage", self.h)

    def autotune(self):
        """Pick the gradient kernel's tile shape by timing (nuc_autotune); returns it."""
        _check(lib().nuc_autotune(self.h), "nuc_autotune", self.h)
        return self.tile_shape()

    def set_tile_shape(self, grad_planes):
        _check(lib().nuc_set_tile_shape(self.h, int(grad_planes)), "nuc_set_tile_shape", self.h)

    def tile_shape(self):
        g = ctypes.c_int()
        _check(lib().nuc_tile_shape(self.h, ctypes.byref(g)), "nuc_tile_shape", self.h)
        return g.value

    def set_graph(self, enable=True):
        """Replay one captured hipGraph per RK3 step instead of per-kernel launches."""
        _check(lib().nuc_set_graph(self.h, 1 if enable else 0), "nuc_set_graph", self.h)

    def query(self):
        """True when all enqueued work has finished (never blocks)."""
        rc = lib().nuc_query(self.h)
        if rc == NUC_EBUSY:
            return False
        _check(rc, "nuc_query", self.h)
        return True

    def sync(self):
        info = StepInfo()
        _check(lib().nuc_sync(self.h, ctypes.byref(info)), "nuc_sync", self.h)
        return info.dt, info.res, info.time

    def profile(self, enable=True):
        _check(lib().nuc_profile(self.h, 1 if enable else 0), "nuc_profile", self.h)

    def profile_read(self):
        """{kernel: (total_ms, launches)} accumulated since profile(True)."""
        out = {}
        for slot in range(8):
            ms = ctypes.c_double()
            n = ctypes.c_long()
            name = ctypes.c_char_p()
            _check(lib().nuc_profile_read(self.h, slot, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(name)),
                   "nuc_profile_read", self.h)
            out[name.value.decode()] = (ms.value, n.value)
        return out

    def stats(self):
        out = np.zeros(25)
        _check(lib().nuc_stats(self.h, _ptr(out)), "nuc_stats", self.h)
        return out

    def stats_ex(self):
        """nuc_stats values followed by the solenoidal and dilatational dissipation."""
        out = np.zeros(NSTATS_EX)
        _check(lib().nuc_stats_ex(self.h, _ptr(out), NSTATS_EX), "nuc_stats_ex", self.h)
        return out


def op_reconstruct(scheme, upwind, tilesize, fin, gL=None, gR=None, zones=7, out=None):
    """fin shaped (n3, n2, n1); ghosts (n3, n2, bf).  Returns (n3, n2, n1+1)."""
    n3, n2, n1 = fin.shape
    bf = gL.shape[2] if gL is not None else (gR.shape[2] if gR is not None else 0)
    if out is None:
        out = np.zeros((n3, n2, n1 + 1))
    fin = np.ascontiguousarray(fin, dtype=np.float64)
    gL = None if gL is None else np.ascontiguousarray(gL, dtype=np.float64)
    gR = None if gR is None else np.ascontiguousarray(gR, dtype=np.float64)
    _check(lib().nuc_op_reconstruct(scheme, upwind, tilesize, _ptr(fin), n1, n2, n3, _ptr(gL), _ptr(gR), bf,
                                    _ptr(out), zones), "nuc_op_reconstruct")
    return out


def op_diff(scheme, tilesize, fin, h, gL=None, gR=None, zones=7, out=None):
    n3, n2, n1 = fin.shape
    bf = gL.shape[2] if gL is not None else (gR.shape[2] if gR is not None else 0)
    if out is None:
        out = np.zeros((n3, n2, n1))
    fin = np.ascontiguousarray(fin, dtype=np.float64)
    gL = None if gL is None else np.ascontiguousarray(gL, dtype=np.float64)
    gR = None if gR is None else np.ascontiguousarray(gR, dtype=np.float64)
    _check(lib().nuc_op_diff(scheme, tilesize, _ptr(fin), float(h), n1, n2, n3, _ptr(gL), _ptr(gR), bf, _ptr(out),
                             zones), "nuc_op_diff")
    return out


def op_rk3(stage, rhs, un, ui, dt):
    ui = np.array(ui, dtype=np.float64, copy=True).ravel()
    _check(lib().nuc_op_rk3(stage, _ptr(np.ascontiguousarray(rhs, dtype=np.float64).ravel()),
                            _ptr(np.ascontiguousarray(un, dtype=np.float64).ravel()), _ptr(ui), float(dt), ui.size),
           "nuc_op_rk3")
    return ui


def pack_compatible(a, b):
    """True when two contexts can share packed launches (nuc_pack_compatible)."""
    return bool(lib().nuc_pack_compatible(a.h, b.h))


class Pack:
    """Several contexts advanced by one launch per kernel (nuc_pack_*).  Each
    member's result equals nuc_step on that member alone; close the pack before
    its members."""

    def __init__(self, contexts, use_graph=False):
        self.members = list(contexts)
        arr = (ctypes.c_void_p * len(self.members))(*[c.h for c in self.members])
        h = ctypes.c_void_p()
        rc = lib().nuc_pack_create(arr, len(self.members), 1 if use_graph else 0, ctypes.byref(h))
        if rc != 0:
            raise NucError(rc, "nuc_pack_create: " + lib().nuc_pack_last_error(None).decode())
        self.h = h

    def step_async(self, n=1):
        rc = lib().nuc_pack_step_async(self.h, n)
        if rc != 0:
            raise NucError(rc, "nuc_pack_step_async: " + lib().nuc_pack_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().nuc_pack_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """The serving loop (include/nuc3d_serve.h): submit jobs, fetch results."""

    def __init__(self, hbm_budget=64 << 30, max_active=16, steps_per_slice=0, use_graphs=True, pack_cells=0,
                 autotune=False):
        c = ServerConfig(int(hbm_budget), int(max_active), int(steps_per_slice), 1 if use_graphs else 0,
                         int(pack_cells), 1 if autotune else 0)
        h = ctypes.c_void_p()
        _check(lib().nuc_server_create(ctypes.byref(c), ctypes.byref(h)), "nuc_server_create")
        self.h = h
        self._shapes = {}

    def close(self):
        if self.h:
            lib().nuc_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def now(self):
        return lib().nuc_server_now(self.h)

    def submit(self, cfg, q0, nsteps):
        q0 = np.ascontiguousarray(q0, dtype=np.float64)
        jid = ctypes.c_longlong()
        _check(lib().nuc_server_submit(self.h, ctypes.byref(cfg), _ptr(q0), int(nsteps), ctypes.byref(jid)),
               "nuc_server_submit")
        self._shapes[jid.value] = q0.size
        return jid.value

    def submit_ic(self, cfg, ic, nsteps, k0=4.0, u_rms=1.0, seed=0):
        """A job whose initial state (IC_TGV, IC_HIT, IC_IVC or IC_DEFAULT) is generated on the device."""
        jid = ctypes.c_longlong()
        _check(lib().nuc_server_submit_ic(self.h, ctypes.byref(cfg), int(ic), float(k0), float(u_rms), int(seed),
                                          int(nsteps), ctypes.byref(jid)), "nuc_server_submit_ic")
        self._shapes[jid.value] = 5 * cfg.nx * cfg.ny * cfg.nz
        return jid.value

    def result(self, jid, want_state=True):
        """None while the job is queued or running; else (state array or None, JobResult)."""
        r = JobResult()
        q = np.empty(self._shapes[jid]) if want_state else None
        rc = lib().nuc_server_result(self.h, jid, _ptr(q), ctypes.byref(r))
        if rc == NUC_EBUSY:
            return None
        _check(rc, "nuc_server_result")
        self._shapes.pop(jid, None)
        return (q if r.state == JOB_DONE else None), r

    def drain(self, timeout_s=-1.0):
        return lib().nuc_server_drain(self.h, float(timeout_s)) == NUC_OK

    def stats(self):
        st = ServerStats()
        _check(lib().nuc_server_stats_get(self.h, ctypes.byref(st)), "nuc_server_stats_get")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
