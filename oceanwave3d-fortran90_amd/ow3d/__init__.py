"""Python host interface to libow3d.so, the MI355X (gfx950) implementation of OceanWave3D's
Laplace-solve-per-RK4-stage path.

The entry points mirror the reference subroutines they replace (see include/ow3d.h):

    Tank.laplace_apply   BuildLinearSystem            (functions/BuildLinearSystem.f90)
    Tank.precond_apply   MultiGridSolveLinearSystem   (multigrid/MultiGridSolveLinearSystem.f90)
    Tank.solve           iterative_solution           (wrappers/iterative_solution.f90)
    Tank.set_surface     DifferentiationsFreeSurfacePlane + DetermineTransformationConstantsArray
    Tank.vertical_velocity  VerticalFreeSurfaceVelocity
    Tank.rhs_free_surface   rhsFreeSurface3D / rhsLinearFreeSurface3D
    Tank.rk4_steps       Runge_Kutta_4 driven by OceanWave3DTakeATimeStep

and, with no counterpart in the reference, many tanks at once: batch_rk4_steps (tanks advanced
concurrently, each RK4 step one hipGraph replay on the tank's stream), Tank.rk4_steps_async /
query / wait for request loops (ow3d/serve.py), and the device-memory figures used for admission.

Arrays follow the reference layout: 3D fields have shape (nyg, nxg, nzg) (sigma fastest),
plane fields (nyg, nxg).  The library must be built (``python build.py``); nothing here falls
back to a CPU implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libow3d.so")
_lib = None


class Config(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("alpha", ctypes.c_int), ("beta", ctypes.c_int), ("gamma", ctypes.c_int),
        ("gridz", ctypes.c_int),
        ("lx", ctypes.c_double), ("ly", ctypes.c_double), ("lz", ctypes.c_double),
        ("g", ctypes.c_double),
        ("nonlinear", ctypes.c_int), ("solver", ctypes.c_int), ("precond", ctypes.c_int),
        ("gmres_m", ctypes.c_int), ("reltol", ctypes.c_double), ("abstol", ctypes.c_double),
        ("maxit", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
        ("mg_maxgrids", ctypes.c_int), ("extrapolation", ctypes.c_int),
        ("mg_precision", ctypes.c_int), ("pressure_term", ctypes.c_int), ("pressure_damping", ctypes.c_int),
        ("gmres_side", ctypes.c_int),
    ]


DEFAULTS = dict(alpha=3, beta=3, gamma=3, gridz=1, lx=1.0, ly=1.0, lz=1.0, g=9.81, nonlinear=1,
                solver=1, precond=3, gmres_m=23, reltol=1e-8, abstol=0.0, maxit=1, nu1=1, nu2=1,
                mg_maxgrids=20, extrapolation=1, mg_precision=0, pressure_term=0, pressure_damping=0,
                gmres_side=0)

EXPORTS = ["ow3d_create", "ow3d_destroy", "ow3d_dims", "ow3d_last_error", "ow3d_laplace_apply",
           "ow3d_precond_apply", "ow3d_solve", "ow3d_set_surface", "ow3d_vertical_velocity",
           "ow3d_rhs_free_surface", "ow3d_add_relax_zone", "ow3d_set_state", "ow3d_get_state",
           "ow3d_rk4_steps", "ow3d_stats", "ow3d_laplace_apply_device", "ow3d_stream",
           "ow3d_time_matvec", "ow3d_batch_rk4_steps", "ow3d_rk4_steps_async", "ow3d_query", "ow3d_wait",
           "ow3d_device_bytes", "ow3d_estimate_bytes", "ow3d_step_stats", "ow3d_tune", "ow3d_mg_plan", "ow3d_kinematics", "ow3d_steps_status",
           "ow3d_solve_noalloc", "ow3d_kinematics_gradients", "ow3d_surface_slopes"]


class OW3DError(RuntimeError):
    pass


def lib():
    """Loads libow3d.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OW3DError(f"{LIB_PATH} is missing: build it with oceanwave3d-fortran90_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.c_int
        IP = ctypes.POINTER(ctypes.c_int)
        L.ow3d_create.restype = I
        L.ow3d_create.argtypes = [ctypes.POINTER(Config), D, D, D, D, D, ctypes.POINTER(P)]
        L.ow3d_destroy.argtypes = [P]
        L.ow3d_dims.argtypes = [P, IP, IP, IP, IP]
        L.ow3d_last_error.restype = ctypes.c_char_p
        L.ow3d_laplace_apply.argtypes = [P, D, D]
        L.ow3d_precond_apply.argtypes = [P, D, D]
        L.ow3d_solve.argtypes = [P, D, D, IP, D]
        L.ow3d_solve_noalloc.argtypes = [P, D, D, IP, D]
        L.ow3d_set_surface.argtypes = [P, D, D, D, D, D, D, D, D]
        L.ow3d_vertical_velocity.argtypes = [P, D, D]
        L.ow3d_rhs_free_surface.argtypes = [P, D, D, D, D]
        L.ow3d_add_relax_zone.argtypes = [P, I, I, I, I, I, D]
        L.ow3d_set_state.argtypes = [P, D, D, D, D]
        L.ow3d_get_state.argtypes = [P, D, D, D, D]
        L.ow3d_rk4_steps.argtypes = [P, I, ctypes.c_double, ctypes.c_double]
        L.ow3d_stats.argtypes = [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), IP, IP]
        L.ow3d_laplace_apply_device.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.ow3d_stream.restype = ctypes.c_void_p
        L.ow3d_stream.argtypes = [P]
        L.ow3d_time_matvec.argtypes = [P, I, D]
        L.ow3d_batch_rk4_steps.argtypes = [ctypes.POINTER(P), I, I, D]
        L.ow3d_rk4_steps_async.argtypes = [P, I, ctypes.c_double]
        L.ow3d_query.argtypes = [P]
        L.ow3d_wait.argtypes = [P]
        L.ow3d_steps_status.argtypes = [P]
        L.ow3d_device_bytes.restype = ctypes.c_longlong
        L.ow3d_device_bytes.argtypes = [P]
        L.ow3d_tune.argtypes = [P, I, I]
        L.ow3d_kinematics.argtypes = [P, D, D, D, D]
        L.ow3d_kinematics_gradients.argtypes = [P, D, D]
        L.ow3d_surface_slopes.argtypes = [P, D, D]
        L.ow3d_mg_plan.argtypes = [ctypes.POINTER(Config), IP, IP]
        L.ow3d_step_stats.argtypes = [P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong), D]
        L.ow3d_estimate_bytes.restype = ctypes.c_longlong
        L.ow3d_estimate_bytes.argtypes = [ctypes.POINTER(Config)]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("arrays must be C-contiguous float64")
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(rc, what):
    if rc < 0:
        raise OW3DError(f"{what}: {lib().ow3d_last_error().decode()}")
    return rc


class Tank:
    """One wave tank resident on the GPU."""

    def __init__(self, cfg, h, hx=None, hy=None, hxx=None, hyy=None):
        c = dict(DEFAULTS)
        c.update(cfg)
        self.cfg = c
        self._c = Config(**c)
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (h, hx, hy, hxx, hyy)]
        h_ = ctypes.c_void_p()
        _check(lib().ow3d_create(ctypes.byref(self._c), *[_p(a) for a in arrs], ctypes.byref(h_)), "ow3d_create")
        self.h = h_
        d = [ctypes.c_int() for _ in range(4)]
        lib().ow3d_dims(self.h, *[ctypes.byref(v) for v in d])
        self.nxg, self.nyg, self.nzg, self.mg_levels = [v.value for v in d]
        self.shape3 = (self.nyg, self.nxg, self.nzg)
        self.shape2 = (self.nyg, self.nxg)
        self.n = self.nyg * self.nxg * self.nzg

    def close(self):
        if getattr(self, "h", None):
            lib().ow3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def laplace_apply(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(self.shape3)
        y = np.empty(self.shape3)
        _check(lib().ow3d_laplace_apply(self.h, _p(phi), _p(y)), "ow3d_laplace_apply")
        return y

    def laplace_apply_device(self, d_phi, d_y, stream=None):
        """y = A phi on device buffers given by address (e.g. a torch tensor's data_ptr(); d_phi 16-byte
        aligned), enqueued on `stream` (a HIP stream handle) or the handle's own stream."""
        _check(lib().ow3d_laplace_apply_device(self.h, ctypes.c_void_p(int(d_phi)), ctypes.c_void_p(int(d_y)),
                                               ctypes.c_void_p(int(stream)) if stream else None),
               "ow3d_laplace_apply_device")

    def precond_apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64).reshape(self.shape3)
        u = np.empty(self.shape3)
        _check(lib().ow3d_precond_apply(self.h, _p(r), _p(u)), "ow3d_precond_apply")
        return u

    def solve(self, rhs, sol):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(self.shape3)
        x = np.array(sol, dtype=np.float64, copy=True).reshape(self.shape3)
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        st = _check(lib().ow3d_solve(self.h, _p(rhs), _p(x), ctypes.byref(it), ctypes.pointer(res)), "ow3d_solve")
        return x, st, it.value, res.value

    def solve_noalloc(self, rhs, sol):
        """solve() with its workspace in the library's static device array (no allocation)."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(self.shape3)
        x = np.array(sol, dtype=np.float64, copy=True).reshape(self.shape3)
        it = ctypes.c_int(0)
        res = ctypes.c_double(0.0)
        st = _check(lib().ow3d_solve_noalloc(self.h, _p(rhs), _p(x), ctypes.byref(it), ctypes.pointer(res)),
                    "ow3d_solve_noalloc")
        return x, st, it.value, res.value

    def set_surface(self, E, P):
        E = np.array(E, dtype=np.float64, copy=True).reshape(self.shape2)
        P = np.array(P, dtype=np.float64, copy=True).reshape(self.shape2)
        out = {k: np.empty(self.shape2) for k in ("Ex", "Exx", "Ey", "Eyy", "Px", "Py")}
        _check(lib().ow3d_set_surface(self.h, _p(E), _p(P), *[_p(out[k]) for k in ("Ex", "Exx", "Ey", "Eyy", "Px", "Py")]),
               "ow3d_set_surface")
        return E, P, out

    def vertical_velocity(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64).reshape(self.shape3)
        W = np.empty(self.shape2)
        _check(lib().ow3d_vertical_velocity(self.h, _p(phi), _p(W)), "ow3d_vertical_velocity")
        return W

    def rhs_free_surface(self, E, W):
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(self.shape2)
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(self.shape2)
        rE = np.empty(self.shape2)
        rP = np.empty(self.shape2)
        _check(lib().ow3d_rhs_free_surface(self.h, _p(E), _p(W), _p(rE), _p(rP)), "ow3d_rhs_free_surface")
        return rE, rP

    def add_relax_zone(self, axis, i0, i1, j0, j1, gam):
        gam = np.ascontiguousarray(gam, dtype=np.float64)
        _check(lib().ow3d_add_relax_zone(self.h, axis, i0, i1, j0, j1, _p(gam)), "ow3d_add_relax_zone")

    def set_state(self, E, P, W=None, phi=None):
        E = np.ascontiguousarray(E, dtype=np.float64).reshape(self.shape2)
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(self.shape2)
        W = None if W is None else np.ascontiguousarray(W, dtype=np.float64).reshape(self.shape2)
        phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64).reshape(self.shape3)
        _check(lib().ow3d_set_state(self.h, _p(E), _p(P), _p(W), _p(phi)), "ow3d_set_state")

    def get_state(self):
        E = np.empty(self.shape2); P = np.empty(self.shape2); W = np.empty(self.shape2); phi = np.empty(self.shape3)
        _check(lib().ow3d_get_state(self.h, _p(E), _p(P), _p(W), _p(phi)), "ow3d_get_state")
        return E, P, W, phi

    def rk4_steps(self, nsteps, t0, dt):
        return _check(lib().ow3d_rk4_steps(self.h, int(nsteps), float(t0), float(dt)), "ow3d_rk4_steps")

    def stats(self):
        tot = ctypes.c_longlong(); sol = ctypes.c_longlong(); mn = ctypes.c_int(); mx = ctypes.c_int()
        lib().ow3d_stats(self.h, ctypes.byref(tot), ctypes.byref(sol), ctypes.byref(mn), ctypes.byref(mx))
        return tot.value, sol.value, mn.value, mx.value

    def time_matvec(self, n_launches):
        ms = ctypes.c_double(0.0)
        _check(lib().ow3d_time_matvec(self.h, int(n_launches), ctypes.pointer(ms)), "ow3d_time_matvec")
        return ms.value

    def rk4_steps_async(self, nsteps, dt):
        """Enqueues nsteps RK4 steps on the tank's stream and returns at once."""
        _check(lib().ow3d_rk4_steps_async(self.h, int(nsteps), float(dt)), "ow3d_rk4_steps_async")

    def query(self):
        """True once the steps enqueued by rk4_steps_async are done."""
        return _check(lib().ow3d_query(self.h), "ow3d_query") == 1

    def wait(self):
        """Blocks until the enqueued steps are done; 1 if a solve missed its tolerance."""
        return _check(lib().ow3d_wait(self.h), "ow3d_wait")

    def steps_status(self):
        """1 if a solve of the steps enqueued by the last rk4_steps / rk4_steps_async / batch call missed
        its tolerance, else 0 (the steps must have completed)."""
        return _check(lib().ow3d_steps_status(self.h), "ow3d_steps_status")

    def device_bytes(self):
        return int(lib().ow3d_device_bytes(self.h))

    def step_stats(self):
        """(replays, captures, capture_ms): speculative steps rolled back and replayed with the whole
        iteration budget, stage graphs captured, host milliseconds spent capturing."""
        r = ctypes.c_longlong(); c = ctypes.c_longlong(); ms = ctypes.c_double()
        _check(lib().ow3d_step_stats(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.pointer(ms)), "ow3d_step_stats")
        return r.value, c.value, ms.value

    def replays(self):
        return self.step_stats()[0]

    def kinematics(self, phi=None):
        """(u, v, w) at every grid point (StoreKinematicData.f90:266-326) from phi (default: the tank's
        current solution) and the current surface."""
        p = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64).reshape(self.shape3)
        U, V, W = (np.empty(self.shape3) for _ in range(3))
        _check(lib().ow3d_kinematics(self.h, _p(p), _p(U), _p(V), _p(W)), "ow3d_kinematics")
        return U, V, W

    def kinematics_gradients(self, phi=None):
        """The 9 fields StoreKinematicData.f90:327-478 writes after u, v, w, shape (9,) + shape3: d w_z,
        (d w)_x, (d w)_y, u_z, u_x, u_y, v_z, v_x, v_y."""
        p = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64).reshape(self.shape3)
        G = np.empty((9,) + self.shape3)
        _check(lib().ow3d_kinematics_gradients(self.h, _p(p), _p(G)), "ow3d_kinematics_gradients")
        return G

    def store_kinematics(self, directory, io, it, win, dt, grid=None):
        """Writes Kinematics_<io>_<it>.bin (formattype 21, StoreKinematicData.f90:90-478) for the tank's
        current state: with it == 0 the grid file (grid = dict(x, y, h, hx, hy, z), the tank's set-up
        fields), else eta and its slopes, phi, u, v, w and the velocity gradients of the window."""
        from ow3d import io as owio
        if it == 0:
            return owio.write_kinematics(directory, io, 0, win, dt, grid=grid)
        E, P, W, phi = self.get_state()
        Ex, Ey = self.surface_slopes()
        U, V, Wv = self.kinematics()
        G = self.kinematics_gradients()
        return owio.write_kinematics(directory, io, it, win, dt,
                                     fields=dict(E=E, Ex=Ex, Ey=Ey, phi=phi, U=U, V=V, W=Wv, G=G))

    def surface_slopes(self):
        """(eta_x, eta_y) of the tank's state, as the last time step or set_state computed them."""
        Ex, Ey = np.empty(self.shape2), np.empty(self.shape2)
        _check(lib().ow3d_surface_slopes(self.h, _p(Ex), _p(Ey)), "ow3d_surface_slopes")
        return Ex, Ey

    def tune(self, matvec_variant=0, sweep_threads=0):
        """Launch choices (see ow3d_tune in include/ow3d.h); results do not depend on them."""
        _check(lib().ow3d_tune(self.h, int(matvec_variant), int(sweep_threads)), "ow3d_tune")


def estimate_bytes(cfg):
    """Device memory Tank(cfg, ...) would allocate (before the allocator's rounding)."""
    c = dict(DEFAULTS)
    c.update(cfg)
    v = lib().ow3d_estimate_bytes(ctypes.byref(Config(**c)))
    if v < 0:
        raise OW3DError(f"ow3d_estimate_bytes: {lib().ow3d_last_error().decode()}")
    return int(v)


def mg_plan(cfg):
    """(levels, coarsest unknowns) of the multigrid hierarchy Tank(cfg, ...) would build (host only)."""
    c = dict(DEFAULTS)
    c.update(cfg)
    nl = ctypes.c_int(); nc = ctypes.c_int()
    if lib().ow3d_mg_plan(ctypes.byref(Config(**c)), ctypes.byref(nl), ctypes.byref(nc)) < 0:
        raise OW3DError(f"ow3d_mg_plan: {lib().ow3d_last_error().decode()}")
    return nl.value, nc.value


def batch_rk4_steps(tanks, nsteps, dts):
    """Advances every tank by nsteps RK4 steps (tank i with time step dts[i]) concurrently on the device;
    returns 1 if any solve missed its tolerance."""
    n = len(tanks)
    hs = (ctypes.c_void_p * n)(*[t.h for t in tanks])
    d = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (n,)))
    return _check(lib().ow3d_batch_rk4_steps(hs, n, int(nsteps), _p(d)), "ow3d_batch_rk4_steps")
