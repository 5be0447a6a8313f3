"""ctypes binding of the CPU oracle (oracle/lib/libgarm_oracle.so).

ORACLE — test infrastructure only.  Imported by tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg, never by the product package.  Method names
follow the reference operators they restate (see garm_oracle.cpp).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgarm_oracle.so")

# field ids (same numbering as include/garm_gpu.h)
PHI, PHI_GRAD, REFMAP, REFMAP_GRAD, PHI_TMP, PHI_ORI, PHI_REF, PHI_REF_GRAD, ADV_REGION, VELOCITY, \
    CELL_DIST, NODE_DIST, FRACTION, SOLID_VELOCITY, NORMAL, PRESSURE = range(16)

PRECOND_NONE, PRECOND_JACOBI, PRECOND_IC0 = 0, 1, 2

_lib = None


def build():
    """Compile the oracle with its Makefile (g++ + OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        vp, d, i, sz = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_size_t
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        llp = ctypes.POINTER(ctypes.c_longlong)
        sig = {
            "ora_create": (vp, [i, ip, d, dp, i]),
            "ora_destroy": (None, [vp]),
            "ora_set_threads": (None, [i]),
            "ora_get_threads": (i, []),
            "ora_field_count": (sz, [vp, i]),
            "ora_set": (i, [vp, i, vp, sz]),
            "ora_get": (i, [vp, i, vp, sz]),
            "ora_union_box": (None, [vp, dp, dp]),
            "ora_union_sphere": (None, [vp, dp, d]),
            "ora_union_plane": (None, [vp, dp, dp]),
            "ora_solid_plane": (None, [vp, dp, dp]),
            "ora_initialize": (i, [vp, i, d, i, d, i]),
            "ora_advect_refmap": (None, [vp, d]),
            "ora_advect_velocity": (None, [vp, d]),
            "ora_compose": (None, [vp]),
            "ora_reinit": (None, [vp, i, d]),
            "ora_extrapolate_collider": (None, [vp]),
            "ora_extrapolate_air": (None, [vp]),
            "ora_enforce_boundary": (None, [vp]),
            "ora_gravity": (None, [vp, d]),
            "ora_check_refmap": (i, [vp, d, d, d, llp, llp]),
            "ora_restart": (None, [vp]),
            "ora_restart_newton": (None, [vp]),
            "ora_set_restart_newton": (None, [vp, i]),
            "ora_set_surface_tension": (None, [vp, i, d, d, d, d, d, i, i]),
            "ora_apply_surface_tension": (i, [vp, d, i, d, i, ip, dp, ip]),
            "ora_set_surface_tension_solvers": (None, [vp, i, i, i]),
            "ora_capillary_stats": (None, [vp, ip, ip, dp, dp]),
            "ora_set_elapsed_time": (None, [vp, d]),
            "ora_capillary": (i, [vp, d, d, d, d, i, i]),
            "ora_contour3": (i, [ip, dp, d, dp, vp, vp, sz, vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz), vp]),
            "ora_mc_table": (i, [ip, ip]),
            "ora_contour2": (i, [ip, dp, d, dp, vp, sz, vp, sz, ctypes.POINTER(sz), ctypes.POINTER(sz)]),
            "ora_curvature": (d, [vp, dp]),
            "ora_little_drop": (None, [vp, d]),
            "ora_get_elapsed_time": (d, [vp]),
            "ora_project": (i, [vp, i, d, i, ip, dp, ip]),
            "ora_step": (i, [vp, d, i, i, d, i, d, i, ip, ip, dp, ip]),
            "ora_max_speed": (d, [vp]),
            "ora_volume": (d, [vp, ctypes.c_int]),
            "ora_mc_fraction": (d, [ctypes.POINTER(ctypes.c_double)]),
            "ora_filled_area": (d, [ctypes.POINTER(ctypes.c_double)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d3(v):
    v = list(v) + [0.0] * (3 - len(v))
    return (ctypes.c_double * 3)(*v)


class Oracle:
    """One CPU LevelSetLiquid<Dim> (LevelSetLiquid.h:17-121) restated."""

    def __init__(self, dim, res, dx, origin, bw=2):
        L = lib()
        r = list(res) + [1] * (3 - len(res))
        self.dim = dim
        self.h = L.ora_create(dim, (ctypes.c_int * 3)(*r), dx, _d3(origin), bw)
        if not self.h:
            raise ValueError("ora_create failed")

    def close(self):
        if self.h:
            lib().ora_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def set_threads(n):
        lib().ora_set_threads(int(n))

    @staticmethod
    def threads():
        return lib().ora_get_threads()

    def count(self, field):
        return lib().ora_field_count(self.h, field)

    def get(self, field):
        n = self.count(field)
        a = np.empty(n, dtype=np.uint8 if field == ADV_REGION else np.float64)
        if lib().ora_get(self.h, field, a.ctypes.data, n) != 0:
            raise ValueError(f"ora_get({field}) failed")
        return a

    def set(self, field, arr):
        a = np.ascontiguousarray(arr, dtype=np.uint8 if field == ADV_REGION else np.float64).ravel()
        if lib().ora_set(self.h, field, a.ctypes.data, a.size) != 0:
            raise ValueError(f"ora_set({field}) size {a.size} != {self.count(field)}")

    def union_box(self, mn, lengths):
        lib().ora_union_box(self.h, _d3(mn), _d3(lengths))

    def union_sphere(self, c, r):
        lib().ora_union_sphere(self.h, _d3(c), r)

    def union_plane(self, p, n):
        lib().ora_union_plane(self.h, _d3(p), _d3(n))

    def solid_plane(self, p, n):
        """Domain box plus the solid half-space (x - p).n < 0 (Boundary::unions
        with an ImplicitPlane, then finish)."""
        lib().ora_solid_plane(self.h, _d3(p), _d3(n))

    def initialize(self, reinit_iters=30, cfl=0.9, precond=PRECOND_JACOBI, tol=1e-10, max_it=20000):
        return lib().ora_initialize(self.h, reinit_iters, cfl, precond, tol, max_it)

    def advect_refmap(self, dt):
        lib().ora_advect_refmap(self.h, dt)

    def advect_velocity(self, dt):
        lib().ora_advect_velocity(self.h, dt)

    def compose(self):
        lib().ora_compose(self.h)

    def reinit(self, iters=30, cfl=0.9):
        lib().ora_reinit(self.h, iters, cfl)

    def extrapolate_collider(self):
        lib().ora_extrapolate_collider(self.h)

    def extrapolate_air(self):
        lib().ora_extrapolate_air(self.h)

    def enforce_boundary(self):
        lib().ora_enforce_boundary(self.h)

    def gravity(self, dt):
        lib().ora_gravity(self.h, dt)

    def check_refmap(self, threshold=0.883022221559489, band=4.0, ratio=1e-5):
        tot, cnt = ctypes.c_longlong(), ctypes.c_longlong()
        ok = lib().ora_check_refmap(self.h, threshold, band, ratio, ctypes.byref(tot), ctypes.byref(cnt))
        return bool(ok), tot.value, cnt.value

    def restart(self):
        lib().ora_restart(self.h)

    def restart_newton(self):
        """LevelSetLiquid::restart of the default build: Hybrid Newton with
        characteristic steps (enable-fmm=n)."""
        lib().ora_restart_newton(self.h)

    def set_restart_newton(self, on):
        """step() restarts with Hybrid Newton when on, the plain branch otherwise."""
        lib().ora_set_restart_newton(self.h, int(bool(on)))

    def set_surface_tension(self, on, coeff=7.28e-2 / 1e3, delta_min=2.0, delta_max=8.0, crit_time=1.0, tol=1e-5, max_it=2000,
                            little_drop=True):
        """step() applies surface tension after the projection when on
        (LevelSetLiquid.h:86-91 settings; the capillary solve uses IC(0))."""
        lib().ora_set_surface_tension(self.h, int(bool(on)), coeff, delta_min, delta_max, crit_time, tol, max_it, int(bool(little_drop)))

    def set_surface_tension_solvers(self, capillary_precond=PRECOND_IC0, capillary_guess=True, reproject_warm=True):
        """How the capillary solve and the reprojection start: the reference's
        by default (IC(0) from the velocity guess, the reprojection from the
        last pressure); (PRECOND_JACOBI, False, False) is the GPU's."""
        lib().ora_set_surface_tension_solvers(self.h, capillary_precond, int(bool(capillary_guess)), int(bool(reproject_warm)))

    def apply_surface_tension(self, dt, precond=PRECOND_JACOBI, tol=1e-10, max_it=20000):
        """applySurfaceTension + trickLittleDrop with the settings of
        set_surface_tension; precond/tol/max_it are the reprojection's."""
        it, unk, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        rc = lib().ora_apply_surface_tension(self.h, dt, precond, tol, max_it, ctypes.byref(it), ctypes.byref(err), ctypes.byref(unk))
        return dict(rc=rc, iterations=it.value, residual=err.value, unknowns=unk.value, capillary=self.capillary_stats())

    def capillary_stats(self):
        f, it, err, band = ctypes.c_int(), ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
        lib().ora_capillary_stats(self.h, ctypes.byref(f), ctypes.byref(it), ctypes.byref(err), ctypes.byref(band))
        return dict(faces=f.value, iterations=it.value, residual=err.value, band=band.value)

    def capillary(self, coeff, band, dt, tol=1e-12, max_it=20000, precond=PRECOND_IC0):
        """SurfaceTension::solve alone on the current phi and velocity."""
        rc = lib().ora_capillary(self.h, coeff, band, dt, tol, max_it, precond)
        return dict(rc=rc, **self.capillary_stats())

    def curvature(self, pos):
        """SurfaceTension::getCurvature at pos with the cubic view of phi."""
        return lib().ora_curvature(self.h, _d3(pos))

    def little_drop(self, dt):
        lib().ora_little_drop(self.h, dt)

    @property
    def elapsed_time(self):
        return lib().ora_get_elapsed_time(self.h)

    @elapsed_time.setter
    def elapsed_time(self, t):
        lib().ora_set_elapsed_time(self.h, t)

    def project(self, precond=PRECOND_JACOBI, tol=1e-10, max_it=20000):
        it, unk, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        rc = lib().ora_project(self.h, precond, tol, max_it, ctypes.byref(it), ctypes.byref(err), ctypes.byref(unk))
        return dict(rc=rc, iterations=it.value, residual=err.value, unknowns=unk.value)

    def step(self, dt, do_reinit=True, reinit_iters=30, cfl=0.9, precond=PRECOND_JACOBI, tol=1e-10, max_it=20000):
        rs, it, unk, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        rc = lib().ora_step(self.h, dt, int(do_reinit), reinit_iters, cfl, precond, tol, max_it,
                            ctypes.byref(rs), ctypes.byref(it), ctypes.byref(err), ctypes.byref(unk))
        return dict(rc=rc, restarted=rs.value, iterations=it.value, residual=err.value, unknowns=unk.value)

    def max_speed(self):
        return lib().ora_max_speed(self.h)

    def volume(self, corrected=False):
        """LevelSetLiquid::calculateVolume on phi_tmp (reference vertex order
        unless corrected)."""
        return lib().ora_volume(self.h, 1 if corrected else 0)


def mc_fraction(v):
    """Fraction::get_mc_vol of 8 corner values in the fraction's own order."""
    a = (ctypes.c_double * 8)(*[float(x) for x in v])
    return lib().ora_mc_fraction(a)


def filled_area(v):
    """Filled-contour area of one cell, nodes (0,0), (1,0), (0,1), (1,1)."""
    a = (ctypes.c_double * 4)(*[float(x) for x in v])
    return lib().ora_filled_area(a)


def contour3(field, origin, dx, fp64_positions=False):
    """Contourer<3>::contour at 0 of a node field of shape (nz, ny, nx):
    (positions (nv, 3) float32, normals (nv, 3) float32, triangles (nt, 3) uint32),
    positions in float64 when fp64_positions."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    n = (ctypes.c_int * 3)(f.shape[2], f.shape[1], f.shape[0])
    nv, ni = ctypes.c_size_t(), ctypes.c_size_t()
    args = (n, _d3(origin), dx, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    lib().ora_contour3(*args, None, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(ni), None)
    pos = np.zeros((nv.value, 3), np.float32)
    nrm = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros(ni.value, np.uint32)
    p64 = np.zeros((nv.value, 3), np.float64)
    rc = lib().ora_contour3(*args, pos.ctypes.data, nrm.ctypes.data, nv.value, idx.ctypes.data, ni.value,
                            ctypes.byref(nv), ctypes.byref(ni), p64.ctypes.data)
    assert rc == 0
    return (p64 if fp64_positions else pos), nrm, idx.reshape(-1, 3)


def mc_table():
    """The oracle's triangle table: (ntri[256], tri[256, 24] edge ords, -1 padded)."""
    ntri = np.zeros(256, np.int32)
    tri = np.zeros((256, 24), np.int32)
    assert lib().ora_mc_table(ntri.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), tri.ctypes.data_as(ctypes.POINTER(ctypes.c_int))) == 0
    return ntri, tri


def contour2(field, origin, dx):
    """Contourer<2, true>::contour at 0 of a node field of shape (ny, nx):
    (positions (nv, 2) float64, triangles (nt, 3) uint32)."""
    f = np.ascontiguousarray(field, dtype=np.float64)
    n = (ctypes.c_int * 3)(f.shape[1], f.shape[0], 1)
    nv, ni = ctypes.c_size_t(), ctypes.c_size_t()
    args = (n, _d3(origin), dx, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    lib().ora_contour2(*args, None, 0, None, 0, ctypes.byref(nv), ctypes.byref(ni))
    pos = np.zeros((nv.value, 2), np.float64)
    idx = np.zeros(ni.value, np.uint32)
    assert lib().ora_contour2(*args, pos.ctypes.data, nv.value, idx.ctypes.data, ni.value, ctypes.byref(nv), ctypes.byref(ni)) == 0
    return pos, idx.reshape(-1, 3)
