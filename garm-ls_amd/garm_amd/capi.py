"""ctypes binding of libgarm_gpu.so (the C-ABI declared in include/garm_gpu.h).

The shared library is built in-tree (garm-ls_amd/lib/).  There is no CPU
fallback: if the library or a gfx950 device is missing, every call raises.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libgarm_gpu.so")
HEADER_PATH = os.path.join(os.path.dirname(_PKG), "include", "garm_gpu.h")

# garm_field
PHI, PHI_GRAD, REFMAP, REFMAP_GRAD, PHI_TMP, PHI_ORI, PHI_REF, PHI_REF_GRAD, ADV_REGION, VELOCITY, \
    CELL_DIST, NODE_DIST, FRACTION, SOLID_VELOCITY, NORMAL, PRESSURE, RHS, UNKNOWNS = range(18)
FIELD_NAMES = ["phi", "phi_grad", "refmap", "refmap_grad", "phi_tmp", "phi_ori", "phi_ref", "phi_ref_grad",
               "adv_region", "velocity", "cell_dist", "node_dist", "fraction", "solid_velocity", "normal", "pressure",
               "rhs", "unknowns"]

# garm_kernel (instrumentation)
K_ADVECT_REFMAP, K_ADVECT_VELOCITY, K_COMPOSE, K_REINIT_STAGE, K_PCG_SPMV, K_PCG_UPDATE, K_PCG32_SPMV, K_PCG32_UPDATE, \
    K_COPY = range(9)
KERNEL_NAMES = ["advect_refmap", "advect_velocity", "compose", "reinit_stage", "pcg_spmv", "pcg_update", "pcg32_spmv",
                "pcg32_update", "copy"]

# garm_velocity_field
VELOCITY_ENRIGHT = 0

# garm_volume_order
VOLUME_REFERENCE, VOLUME_CORRECTED = 0, 1

# garm_pcg_mode
PCG_FP64, PCG_MIXED, PCG_LEARNED, PCG_MULTIGRID = 0, 1, 2, 3
# garm_restart_mode
RESTART_PLAIN, RESTART_NEWTON = 0, 1
FP64, FP32 = 0, 1

GARM_OK, GARM_EINVAL, GARM_EHIP, GARM_ENOCONV, GARM_ENOMEM, GARM_ENODEV = 0, -1, -2, -3, -4, -5
STATUS_NAMES = {0: "GARM_OK", -1: "GARM_EINVAL", -2: "GARM_EHIP", -3: "GARM_ENOCONV", -4: "GARM_ENOMEM",
                -5: "GARM_ENODEV"}


class PcgOpts(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("check_interval", ctypes.c_int), ("fallback_mode", ctypes.c_int)]


class PcgStats(ctypes.Structure):
    _fields_ = [("unknowns", ctypes.c_int), ("iterations", ctypes.c_int), ("outer_iterations", ctypes.c_int),
                ("fell_back", ctypes.c_int), ("relative_residual", ctypes.c_double),
                ("preconditioner_applications", ctypes.c_int), ("hbm_bytes", ctypes.c_double),
                ("solver_cells", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class SurfaceTensionOpts(ctypes.Structure):
    _fields_ = [("coeff", ctypes.c_double), ("delta_min", ctypes.c_double), ("delta_max", ctypes.c_double),
                ("delta_crit_time", ctypes.c_double), ("tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("little_drop", ctypes.c_int)]


class SurfaceTensionStats(ctypes.Structure):
    _fields_ = [("band", ctypes.c_double), ("capillary", PcgStats), ("reproject", PcgStats)]

    def as_dict(self):
        return {"band": self.band, "capillary": self.capillary.as_dict(), "reproject": self.reproject.as_dict()}


class StepOpts(ctypes.Structure):
    _fields_ = [("reinit_interval", ctypes.c_int), ("reinit_iterations", ctypes.c_int), ("reinit_cfl", ctypes.c_double),
                ("advect_precision", ctypes.c_int), ("pcg", PcgOpts), ("restart_mode", ctypes.c_int),
                ("surface_tension", ctypes.POINTER(SurfaceTensionOpts))]


class StepStats(ctypes.Structure):
    _fields_ = [("t_advect_ms", ctypes.c_double), ("t_restart_ms", ctypes.c_double), ("t_project_ms", ctypes.c_double),
                ("t_extrap_ms", ctypes.c_double), ("t_total_ms", ctypes.c_double), ("restarted", ctypes.c_int),
                ("reinitialized", ctypes.c_int), ("pcg", PcgStats), ("surface_tension", SurfaceTensionStats),
                ("max_speed", ctypes.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in ("pcg", "surface_tension")}
        d["pcg"] = self.pcg.as_dict()
        d["surface_tension"] = self.surface_tension.as_dict()
        return d


class GarmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libgarm_gpu.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libgarm_gpu.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, d, i, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
        dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        sig = {
            "garm_create": (i, [ctypes.POINTER(vp), i, ip, d, dp, i]),
            "garm_destroy": (None, [vp]),
            "garm_last_error": (ctypes.c_char_p, [vp]),
            "garm_create_error": (ctypes.c_char_p, []),
            "garm_field_count": (sz, [vp, i]),
            "garm_upload": (i, [vp, i, vp, sz]),
            "garm_download": (i, [vp, i, vp, sz]),
            "garm_union_box": (i, [vp, dp, dp]),
            "garm_union_sphere": (i, [vp, dp, d]),
            "garm_union_plane": (i, [vp, dp, dp]),
            "garm_initialize": (i, [vp, ctypes.POINTER(StepOpts)]),
            "garm_advect_refmap": (i, [vp, d]),
            "garm_advect_velocity": (i, [vp, d]),
            "garm_compose_levelset": (i, [vp]),
            "garm_reinit_weno5": (i, [vp, i, d]),
            "garm_extrapolate_collider": (i, [vp]),
            "garm_extrapolate_air": (i, [vp]),
            "garm_enforce_boundary": (i, [vp]),
            "garm_apply_gravity": (i, [vp, d]),
            "garm_check_refmap": (i, [vp, d, d, d, ip]),
            "garm_restart": (i, [vp]),
            "garm_restart_newton": (i, [vp]),
            "garm_project": (i, [vp, ctypes.POINTER(PcgOpts), ctypes.POINTER(PcgStats)]),
            "garm_step": (i, [vp, d, i, ctypes.POINTER(StepOpts), ctypes.POINTER(StepStats)]),
            "garm_apply_surface_tension": (i, [vp, d, ctypes.POINTER(SurfaceTensionOpts), ctypes.POINTER(PcgOpts),
                                               ctypes.POINTER(SurfaceTensionStats)]),
            "garm_set_elapsed_time": (i, [vp, d]),
            "garm_contour": (i, [vp, vp, vp, sz, vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz)]),
            "garm_write_mesh": (i, [vp, ctypes.c_char_p, ctypes.c_char_p]),
            "garm_get_elapsed_time": (i, [vp, dp]),
            "garm_max_speed": (i, [vp, dp]),
            "garm_apply_pressure_system": (i, [vp, vp, vp]),
            "garm_volume": (i, [vp, i, dp]),
            "garm_volume_device": (i, [i, ip, d, vp, i, vp, vp]),
            "garm_set_precision": (i, [vp, i]),
            "garm_set_velocity_field": (i, [vp, i, d, d]),
            "garm_save_frame": (i, [vp, ctypes.c_char_p]),
            "garm_load_frame": (i, [vp, ctypes.c_char_p]),
            "garm_synchronize": (i, [vp]),
            "garm_time_kernel": (i, [vp, i, i, dp, dp]),
            "garm_unet_create": (i, [ctypes.POINTER(vp), ctypes.c_ulonglong]),
            "garm_unet_destroy": (None, [vp]),
            "garm_unet_last_error": (ctypes.c_char_p, [vp]),
            "garm_unet_create_error": (ctypes.c_char_p, []),
            "garm_unet_param_count": (sz, []),
            "garm_unet_get_params": (i, [vp, vp, sz]),
            "garm_unet_set_params": (i, [vp, vp, sz]),
            "garm_unet_forward": (i, [vp, ip, vp, vp, vp]),
            "garm_unet_loss_grad": (i, [vp, ip, vp, vp, vp, dp, vp, sz]),
            "garm_unet_train": (i, [vp, ip, i, vp, vp, vp, i, ctypes.c_float, ctypes.c_ulonglong, vp]),
            "garm_unet_time_forward": (i, [vp, ip, i, dp, dp]),
            "garm_attach_preconditioner": (i, [vp, vp]),
            "garm_unet_train_on_system": (i, [vp, vp, i, ctypes.c_float, d, ctypes.c_ulonglong, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols(header=HEADER_PATH):
    """Names of the functions declared in include/garm_gpu.h."""
    import re
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s*(garm_[a-z0-9_]+)\s*\(", text, flags=re.M)))


def _d3(v):
    v = list(v) + [0.0] * (3 - len(v))
    return (ctypes.c_double * 3)(*v)


def surface_tension_opts(coeff=7.28e-2 / 1e3, delta_min=2.0, delta_max=8.0, delta_crit_time=1.0, tolerance=1e-5,
                         max_iterations=2000, little_drop=True):
    """LevelSetLiquid's surface-tension settings (LevelSetLiquid.h:86-91
    defaults; coeff is _surfaceTensionCoeff / _density)."""
    return SurfaceTensionOpts(coeff, delta_min, delta_max, delta_crit_time, tolerance, max_iterations, int(bool(little_drop)))


def step_opts(reinit_interval=1, reinit_iterations=30, reinit_cfl=0.9, precision=FP64, mode=PCG_FP64, tol=1e-5,
              max_iterations=2000, check_interval=8, fallback_mode=0, restart_mode=RESTART_PLAIN, surface_tension=None):
    o = StepOpts()
    if surface_tension is not None:
        o._surface_tension = surface_tension   # keeps the pointed-to settings alive with o
        o.surface_tension = ctypes.pointer(surface_tension)
    o.restart_mode = restart_mode
    o.reinit_interval = reinit_interval
    o.reinit_iterations = reinit_iterations
    o.reinit_cfl = reinit_cfl
    o.advect_precision = precision
    o.pcg.mode = mode
    o.pcg.tolerance = tol
    o.pcg.max_iterations = max_iterations
    o.pcg.check_interval = check_interval
    o.pcg.fallback_mode = fallback_mode
    return o


class Liquid:
    """One device-resident LevelSetLiquid<Dim> (LevelSetLiquid.h:17-121)."""

    def __init__(self, dim, res, dx, origin, bw=2):
        L = lib()
        r = list(res) + [1] * (3 - len(res))
        h = ctypes.c_void_p()
        rc = L.garm_create(ctypes.byref(h), dim, (ctypes.c_int * 3)(*r), dx, _d3(origin), bw)
        if rc != GARM_OK:
            raise GarmError(rc, L.garm_create_error().decode())
        self.h = h
        self.dim = dim

    def close(self):
        if getattr(self, "h", None):
            lib().garm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != GARM_OK:
            raise GarmError(rc, lib().garm_last_error(self.h).decode())
        return rc

    def count(self, field):
        return lib().garm_field_count(self.h, field)

    def get(self, field):
        n = self.count(field)
        a = np.empty(n, dtype=np.uint8 if field == ADV_REGION else np.float64)
        self._check(lib().garm_download(self.h, field, a.ctypes.data, n))
        return a

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint8 if field == ADV_REGION else np.float64).ravel()
        self._check(lib().garm_upload(self.h, field, a.ctypes.data, a.size))

    def union_box(self, mn, lengths):
        self._check(lib().garm_union_box(self.h, _d3(mn), _d3(lengths)))

    def union_sphere(self, c, r):
        self._check(lib().garm_union_sphere(self.h, _d3(c), r))

    def union_plane(self, p, n):
        self._check(lib().garm_union_plane(self.h, _d3(p), _d3(n)))

    def initialize(self, opts):
        self._check(lib().garm_initialize(self.h, ctypes.byref(opts)))

    def advect_refmap(self, dt):
        self._check(lib().garm_advect_refmap(self.h, dt))

    def advect_velocity(self, dt):
        self._check(lib().garm_advect_velocity(self.h, dt))

    def compose(self):
        self._check(lib().garm_compose_levelset(self.h))

    def reinit(self, iters=30, cfl=0.9):
        self._check(lib().garm_reinit_weno5(self.h, iters, cfl))

    def extrapolate_collider(self):
        self._check(lib().garm_extrapolate_collider(self.h))

    def extrapolate_air(self):
        self._check(lib().garm_extrapolate_air(self.h))

    def enforce_boundary(self):
        self._check(lib().garm_enforce_boundary(self.h))

    def gravity(self, dt):
        self._check(lib().garm_apply_gravity(self.h, dt))

    def check_refmap(self, threshold=0.883022221559489, band=4.0, ratio=1e-5):
        ok = ctypes.c_int()
        self._check(lib().garm_check_refmap(self.h, threshold, band, ratio, ctypes.byref(ok)))
        return bool(ok.value)

    def restart(self):
        self._check(lib().garm_restart(self.h))

    def restart_newton(self):
        self._check(lib().garm_restart_newton(self.h))

    def apply_surface_tension(self, dt, opts, mode=PCG_FP64, tol=1e-5, max_iterations=2000, check_interval=8, fallback_mode=0):
        """garm_apply_surface_tension: opts from surface_tension_opts(), the
        reprojection solved with the given PCG settings."""
        o = PcgOpts(mode, tol, max_iterations, check_interval, fallback_mode)
        st = SurfaceTensionStats()
        self._check(lib().garm_apply_surface_tension(self.h, dt, ctypes.byref(opts), ctypes.byref(o), ctypes.byref(st)))
        return st.as_dict()

    def contour(self):
        """garm_contour of phi_tmp: (positions (nv, dim) float32, normals
        (nv, 3) float32 (3D; empty in 2D), triangles (nt, 3) uint32)."""
        nv, ni = ctypes.c_size_t(), ctypes.c_size_t()
        rc = lib().garm_contour(self.h, None, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(ni))
        if rc not in (GARM_OK, GARM_EINVAL) or (rc == GARM_EINVAL and nv.value == 0 and ni.value == 0):
            self._check(rc)
        pos = np.zeros((nv.value, self.dim), np.float32)
        nrm = np.zeros((nv.value if self.dim == 3 else 0, 3), np.float32)
        idx = np.zeros(ni.value, np.uint32)
        self._check(lib().garm_contour(self.h, pos.ctypes.data, nrm.ctypes.data, nv.value, idx.ctypes.data, ni.value,
                                       ctypes.byref(nv), ctypes.byref(ni)))
        return pos, nrm, idx.reshape(-1, 3)

    def write_mesh(self, mesh_path=None, obj_path=None):
        """garm_write_mesh: phi_tmp's surface as a liquid.mesh file and/or an
        OBJ file."""
        self._check(lib().garm_write_mesh(self.h, os.fsencode(mesh_path) if mesh_path else None,
                                          os.fsencode(obj_path) if obj_path else None))

    @property
    def elapsed_time(self):
        t = ctypes.c_double()
        self._check(lib().garm_get_elapsed_time(self.h, ctypes.byref(t)))
        return t.value

    @elapsed_time.setter
    def elapsed_time(self, t):
        self._check(lib().garm_set_elapsed_time(self.h, t))

    def project(self, mode=PCG_FP64, tol=1e-5, max_iterations=2000, check_interval=8, raise_on_noconv=True,
                fallback_mode=0):
        o = PcgOpts(mode, tol, max_iterations, check_interval, fallback_mode)
        st = PcgStats()
        rc = lib().garm_project(self.h, ctypes.byref(o), ctypes.byref(st))
        if rc != GARM_OK and (raise_on_noconv or rc != GARM_ENOCONV):
            self._check(rc)
        d = st.as_dict()
        d["rc"] = rc
        return d

    def step(self, dt, step_index, opts):
        st = StepStats()
        self._check(lib().garm_step(self.h, dt, step_index, ctypes.byref(opts), ctypes.byref(st)))
        return st.as_dict()

    def set_velocity_field(self, field, time, period):
        """Analytic face velocity (VELOCITY_ENRIGHT) at time t with period T."""
        self._check(lib().garm_set_velocity_field(self.h, field, time, period))

    def save_frame(self, frame_dir):
        """Writes the reference's .sav files of this state into frame_dir."""
        self._check(lib().garm_save_frame(self.h, os.fsencode(frame_dir)))

    def load_frame(self, frame_dir):
        """Reads the reference's .sav files from frame_dir into this state."""
        self._check(lib().garm_load_frame(self.h, os.fsencode(frame_dir)))

    def set_precision(self, precision):
        """FP64 or FP32 compute for the per-operator advection / reinit calls."""
        self._check(lib().garm_set_precision(self.h, precision))

    def volume(self, corrected=False):
        """Liquid volume of phi_tmp (3D: MC-style fractions in the reference's
        vertex order unless corrected; 2D: filled-contour area)."""
        v = ctypes.c_double()
        self._check(lib().garm_volume(self.h, 1 if corrected else 0, ctypes.byref(v)))
        return v.value

    def apply_pressure_system(self, x):
        """b = A x with the operator of the last projection (cell fields)."""
        xx = np.ascontiguousarray(x, dtype=np.float64).ravel()
        b = np.empty_like(xx)
        if xx.size != self.count(PHI):
            raise ValueError("x must hold one value per cell")
        self._check(lib().garm_apply_pressure_system(self.h, xx.ctypes.data, b.ctypes.data))
        return b

    def max_speed(self):
        v = ctypes.c_double()
        self._check(lib().garm_max_speed(self.h, ctypes.byref(v)))
        return v.value

    def time_kernel(self, which, reps=10):
        """(mean ms per launch, algorithmic HBM bytes per launch)."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        self._check(lib().garm_time_kernel(self.h, which, reps, ctypes.byref(ms), ctypes.byref(by)))
        return ms.value, by.value

    def synchronize(self):
        self._check(lib().garm_synchronize(self.h))

    def train_network(self, net, steps, lr=1e-3, operator_fraction=0.75, seed=0):
        """Trains `net` on pairs made from this context's last pressure system
        (garm_unet_train_on_system); returns the per-step losses."""
        hist = np.zeros(max(steps, 1), dtype=np.float32)
        self._check(lib().garm_unet_train_on_system(net.h, self.h, steps, lr, operator_fraction, seed, hist.ctypes.data))
        return hist[:steps]

    def attach_preconditioner(self, net):
        """Use `net` (a Unet, or None) for PCG_LEARNED solves."""
        self._check(lib().garm_attach_preconditioner(self.h, net.h if net is not None else None))
        self._net = net   # keep the network alive while attached


def volume_device(dim, res, dx, phi_ptr, out_ptr, corrected=False, stream=None):
    """garm_volume_device on device memory (phi_ptr, out_ptr: device addresses,
    stream: a hipStream_t value or None); enqueues and returns without waiting."""
    r = list(res) + [1] * (3 - len(res))
    rc = lib().garm_volume_device(dim, (ctypes.c_int * 3)(*r), dx, ctypes.c_void_p(phi_ptr), 1 if corrected else 0,
                                  ctypes.c_void_p(out_ptr), ctypes.c_void_p(stream) if stream else None)
    if rc != GARM_OK:
        raise GarmError(rc, lib().garm_create_error().decode())


class Unet:
    """The learned pressure preconditioner (garm_unet_* in include/garm_gpu.h).

    Grids are numpy arrays of shape (nz, ny, nx), each size a multiple of 4.
    """

    def __init__(self, seed=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.garm_unet_create(ctypes.byref(h), seed)
        if rc != GARM_OK:
            raise GarmError(rc, L.garm_unet_create_error().decode())
        self.h = h

    @staticmethod
    def param_count():
        return lib().garm_unet_param_count()

    def close(self):
        if getattr(self, "h", None):
            lib().garm_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != GARM_OK:
            raise GarmError(rc, lib().garm_unet_last_error(self.h).decode())

    def get_params(self):
        a = np.empty(self.param_count(), dtype=np.float32)
        self._check(lib().garm_unet_get_params(self.h, a.ctypes.data, a.size))
        return a

    def set_params(self, flat):
        a = np.ascontiguousarray(flat, dtype=np.float32).ravel()
        self._check(lib().garm_unet_set_params(self.h, a.ctypes.data, a.size))

    @staticmethod
    def _dims(shape):
        nz, ny, nx = shape
        return (ctypes.c_int * 3)(nx, ny, nz)

    def forward(self, r, mask):
        r = np.ascontiguousarray(r, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        y = np.empty_like(r)
        self._check(lib().garm_unet_forward(self.h, self._dims(r.shape), r.ctypes.data, m.ctypes.data, y.ctypes.data))
        return y

    def time_forward(self, shape, reps=10):
        """(average ms, FLOPs) of one forward pass on a (nz, ny, nx) grid of synthetic input."""
        ms, fl = ctypes.c_double(), ctypes.c_double()
        self._check(lib().garm_unet_time_forward(self.h, self._dims(shape), reps, ctypes.byref(ms), ctypes.byref(fl)))
        return ms.value, fl.value

    def loss_grad(self, r, mask, target):
        r = np.ascontiguousarray(r, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        t = np.ascontiguousarray(target, dtype=np.float32)
        g = np.empty(self.param_count(), dtype=np.float32)
        loss = ctypes.c_double()
        self._check(lib().garm_unet_loss_grad(self.h, self._dims(r.shape), r.ctypes.data, m.ctypes.data, t.ctypes.data,
                                              ctypes.byref(loss), g.ctypes.data, g.size))
        return loss.value, g

    def train(self, rhs, pressure, mask, steps, lr=1e-3, seed=0):
        """rhs, pressure, mask: (pairs, nz, ny, nx).  Returns the per-step losses."""
        r = np.ascontiguousarray(rhs, dtype=np.float32)
        p = np.ascontiguousarray(pressure, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        hist = np.zeros(max(steps, 1), dtype=np.float32)
        self._check(lib().garm_unet_train(self.h, self._dims(r.shape[1:]), r.shape[0], r.ctypes.data, p.ctypes.data,
                                          m.ctypes.data, steps, lr, seed, hist.ctypes.data))
        return hist[:steps]


def read_mesh(path, dim=3):
    """A liquid.mesh file (SurfaceMesh::write, SurfaceMesh.cpp:103-124):
    (positions (nv, dim), normals (nv, 3) (empty in 2D), triangles (nt, 3))."""
    raw = open(path, "rb").read()
    nv = int(np.frombuffer(raw, np.uint32, 1, 0)[0])
    off = 4
    pos = np.frombuffer(raw, np.float32, nv * dim, off).reshape(nv, dim)
    off += nv * 4 * dim
    if dim == 3:
        nrm = np.frombuffer(raw, np.float32, nv * 3, off).reshape(nv, 3)
        off += nv * 12
    else:
        nrm = np.zeros((0, 3), np.float32)
    ni = int(np.frombuffer(raw, np.uint32, 1, off)[0])
    off += 4
    idx = np.frombuffer(raw, np.uint32, ni, off)
    if off + ni * 4 != len(raw):
        raise ValueError(f"{path}: {len(raw) - off - ni * 4} trailing bytes")
    return pos, nrm, idx.reshape(-1, 3)
