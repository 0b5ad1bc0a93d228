"""ctypes wrapper of the CPU oracle (oracle/ow3d_oracle.c).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker / the timed CPU
baseline.  The product path (oceanwave3d-fortran90_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libow3d_oracle.so")
_lib = None


class Params(ctypes.Structure):
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
        ("mg_mode", ctypes.c_int), ("cd_quirk", ctypes.c_int), ("gmres_side", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "libow3d_oracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.c_int
        L.owo_create.restype = P
        L.owo_create.argtypes = [ctypes.POINTER(Params), D, D, D, D, D]
        L.owo_destroy.argtypes = [P]
        L.owo_dims.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.owo_grid.argtypes = [P, D, D, D]
        L.owo_stencils.argtypes = [P, D, D, D]
        L.owo_set_state.argtypes = [P, D, D, D, D]
        L.owo_get_state.argtypes = [P, D, D, D, D]
        L.owo_surface.argtypes = [P, D, D, D, D, D, D, D, D]
        L.owo_get_dsigma.argtypes = [P, D]
        L.owo_matvec.argtypes = [P, D, D]
        L.owo_lowmatvec.argtypes = [P, I, D, D]
        L.owo_mg_level_dims.argtypes = [P, I, ctypes.POINTER(ctypes.c_int)]
        L.owo_precond.argtypes = [P, D, D]
        L.owo_solve.restype = I
        L.owo_solve.argtypes = [P, D, D, ctypes.POINTER(ctypes.c_int), D]
        L.owo_vertical_velocity.argtypes = [P, D, D]
        L.owo_kinematics.argtypes = [P, D, D, D, D, D]
        L.owo_kinematics_gradients.argtypes = [P, D, D, D]
        L.owo_rhs.argtypes = [P, D, D, D, D, D]
        L.owo_add_relax_zone.argtypes = [P, I, I, I, I, I, D]
        L.owo_rk4_step.restype = I
        L.owo_rk4_step.argtypes = [P, ctypes.c_double, ctypes.c_double]
        L.owo_stats.argtypes = [P, ctypes.POINTER(ctypes.c_long)]
        L.owo_restrict.argtypes = [P, I, D, D]
        L.owo_prolong.argtypes = [P, I, D, D]
        L.owo_smooth.argtypes = [P, I, D, D, I]
        L.owo_csr.restype = I
        L.owo_csr.argtypes = [P, I, ctypes.POINTER(I), ctypes.POINTER(I),
                              ctypes.POINTER(ctypes.POINTER(I)), ctypes.POINTER(ctypes.POINTER(I)),
                              ctypes.POINTER(D)]
        L.owo_fornberg.argtypes = [ctypes.c_double, D, I, I, D]
        L.owo_level_grid.restype = I
        L.owo_level_grid.argtypes = [P, I, ctypes.POINTER(I), D, D, D, D, D, D, D, D]
        L.owo_level_ghosts.restype = I
        L.owo_level_ghosts.argtypes = [P, I, ctypes.POINTER(I)]
        L.owo_gs_once.argtypes = [P, I, D, D]
        L.owo_zebra.argtypes = [P, I, D, D, I]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "arrays must be contiguous float64"
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def fornberg(z, x, m):
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = len(x) - 1
    c = np.zeros((n + 1) * (m + 1))
    lib().owo_fornberg(z, _p(x), n, m, _p(c))
    return c.reshape(n + 1, m + 1)


class OracleTank:
    """One wave tank on the CPU oracle.  `p` is a dict of Params fields; bathymetry
    arrays are plane arrays of shape (nyg, nxg) (x fastest)."""

    def __init__(self, p, h, hx=None, hy=None, hxx=None, hyy=None):
        L = lib()
        self.p = Params(**p)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                      for a in (h, hx, hy, hxx, hyy)]
        self.h = L.owo_create(ctypes.byref(self.p), *[_p(a) for a in self._keep])
        if not self.h:
            raise ValueError("oracle rejected the tank parameters")
        d = (ctypes.c_int * 6)()
        L.owo_dims(self.h, d)
        self.nxg, self.nyg, self.nzg, self.nlev, self.n, self.nxy = list(d)
        self.shape3 = (self.nyg, self.nxg, self.nzg)
        self.shape2 = (self.nyg, self.nxg)

    def close(self):
        if self.h:
            lib().owo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def grid(self):
        x = np.zeros(self.nxg); y = np.zeros(self.nyg); z = np.zeros(self.nzg)
        lib().owo_grid(self.h, _p(x), _p(y), _p(z))
        return x, y, z

    def stencils(self):
        ra = 2 * (self.p.alpha if self.p.nx > 1 else 0) + 1
        rb = 2 * (self.p.beta if self.p.ny > 1 else 0) + 1
        rc = 2 * self.p.gamma + 1
        sx = np.zeros((self.nxg, 2, ra)); sy = np.zeros((self.nyg, 2, rb)); sz = np.zeros((self.nzg, 2, rc))
        lib().owo_stencils(self.h, _p(sx), _p(sy), _p(sz))
        return sx, sy, sz

    def set_state(self, E, P, W=None, phi=None):
        E = np.ascontiguousarray(E, dtype=np.float64); P = np.ascontiguousarray(P, dtype=np.float64)
        W = None if W is None else np.ascontiguousarray(W, dtype=np.float64)
        phi = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
        lib().owo_set_state(self.h, _p(E), _p(P), _p(W), _p(phi))

    def get_state(self):
        E = np.zeros(self.shape2); P = np.zeros(self.shape2); W = np.zeros(self.shape2); phi = np.zeros(self.shape3)
        lib().owo_get_state(self.h, _p(E), _p(P), _p(W), _p(phi))
        return E, P, W, phi

    def surface(self, E, P):
        """Ghost update + plane derivatives + sigma coefficients; returns (E, P, derivs dict)."""
        E = np.array(E, dtype=np.float64, copy=True); P = np.array(P, dtype=np.float64, copy=True)
        out = {k: np.zeros(self.shape2) for k in ("Ex", "Exx", "Ey", "Eyy", "Px", "Py")}
        lib().owo_surface(self.h, _p(E), _p(P), *[_p(out[k]) for k in ("Ex", "Exx", "Ey", "Eyy", "Px", "Py")])
        return E, P, out

    def dsigma(self):
        out = np.zeros((5,) + self.shape3)
        lib().owo_get_dsigma(self.h, _p(out))
        return out

    def matvec(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros_like(phi)
        lib().owo_matvec(self.h, _p(phi), _p(out))
        return out

    def level_dims(self, level):
        d = (ctypes.c_int * 3)()
        if lib().owo_mg_level_dims(self.h, level, d) != 0:
            raise IndexError(level)
        return tuple(d)

    def lowmatvec(self, level, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros_like(phi)
        lib().owo_lowmatvec(self.h, level, _p(phi), _p(out))
        return out

    def precond(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        u = np.zeros_like(r)
        lib().owo_precond(self.h, _p(r), _p(u))
        return u

    def solve(self, rhs, sol):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        sol = np.array(sol, dtype=np.float64, copy=True)
        it = ctypes.c_int(0); res = ctypes.c_double(0)
        st = lib().owo_solve(self.h, _p(rhs), _p(sol), ctypes.byref(it), ctypes.pointer(res))
        return sol, st, it.value, res.value

    def vertical_velocity(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        W = np.zeros(self.shape2)
        lib().owo_vertical_velocity(self.h, _p(phi), _p(W))
        return W

    def kinematics(self, phi, E):
        """(u, v, w) at every grid point (StoreKinematicData.f90:266-326) for the surface E last given to
        surface() / set_state()."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        U, V, W = (np.zeros(self.shape3) for _ in range(3))
        lib().owo_kinematics(self.h, _p(phi), _p(E), _p(U), _p(V), _p(W))
        return U, V, W

    def kinematics_gradients(self, phi, E):
        """The 9 fields StoreKinematicData.f90:327-478 writes after u, v, w, shape (9,) + shape3: d w_z,
        (d w)_x, (d w)_y, u_z, u_x, u_y, v_z, v_x, v_y (the first three from W = dphi/dsigma)."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        E = np.ascontiguousarray(E, dtype=np.float64)
        G = np.zeros((9,) + self.shape3)
        lib().owo_kinematics_gradients(self.h, _p(phi), _p(E), _p(G))
        return G

    def rhs(self, E, P, W):
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (E, P, W)]
        rE = np.zeros(self.shape2); rP = np.zeros(self.shape2)
        lib().owo_rhs(self.h, *[_p(v) for v in a], _p(rE), _p(rP))
        return rE, rP

    def add_relax_zone(self, axis, i0, i1, j0, j1, gam):
        gam = np.ascontiguousarray(gam, dtype=np.float64)
        lib().owo_add_relax_zone(self.h, axis, i0, i1, j0, j1, _p(gam))

    def rk4_step(self, t, dt):
        return lib().owo_rk4_step(self.h, t, dt)

    def stats(self):
        s = (ctypes.c_long * 4)()
        lib().owo_stats(self.h, s)
        return tuple(s)

    def restrict(self, fine_level, rf):
        cd = self.level_dims(fine_level - 1)
        rf = np.ascontiguousarray(rf, dtype=np.float64)
        rc = np.zeros((cd[1], cd[0], cd[2]))
        lib().owo_restrict(self.h, fine_level, _p(rf), _p(rc))
        return rc

    def prolong(self, fine_level, ec):
        fd = self.level_dims(fine_level)
        ec = np.ascontiguousarray(ec, dtype=np.float64)
        ef = np.zeros((fd[1], fd[0], fd[2]))
        lib().owo_prolong(self.h, fine_level, _p(ec), _p(ef))
        return ef

    def smooth(self, level, b, x, sweeps=1):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        lib().owo_smooth(self.h, level, _p(b), _p(x), sweeps)
        return x

    def level_grid(self, level):
        d = (ctypes.c_int * 6)()
        if lib().owo_level_grid(self.h, level, d, *([None] * 8)) != 0:
            raise IndexError(level)
        nx, ny, nz, nxg, nyg, nzg = list(d)
        out = dict(nx=nx, ny=ny, nz=nz, nxg=nxg, nyg=nyg, nzg=nzg, n=nxg * nyg * nzg,
                   x=np.zeros(nxg), y=np.zeros(nyg), z=np.zeros(nzg))
        for k in ("h", "hx", "hy", "hxx", "hyy"):
            out[k] = np.zeros((nyg, nxg))
        lib().owo_level_grid(self.h, level, d, *[_p(out[k]) for k in ("x", "y", "z", "h", "hx", "hy", "hxx", "hyy")])
        return out

    def level_ghosts(self, level):
        n = lib().owo_level_ghosts(self.h, level, None)
        out = np.zeros(n, dtype=np.int32)
        lib().owo_level_ghosts(self.h, level, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out

    def smooth_gs_once(self, level, b, x):
        b = np.ascontiguousarray(b, dtype=np.float64).ravel()
        x = np.array(x, dtype=np.float64, copy=True).ravel()
        lib().owo_gs_once(self.h, level, _p(b), _p(x))
        return x

    def zebra(self, level, b, x, colour):
        """One zebra pass on a level: colour 0 / 1 relaxes the interior columns with (i + j) even / odd,
        colour 2 is the wall-ghost pass."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.array(x, dtype=np.float64, copy=True)
        lib().owo_zebra(self.h, level, _p(b), _p(x), colour)
        return x

    def csr(self, level):
        nrow = ctypes.c_int(); nnz = ctypes.c_int()
        rp = ctypes.POINTER(ctypes.c_int)(); ci = ctypes.POINTER(ctypes.c_int)(); va = ctypes.POINTER(ctypes.c_double)()
        if lib().owo_csr(self.h, level, ctypes.byref(nrow), ctypes.byref(nnz), ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va)) != 0:
            return None
        n, z = nrow.value, nnz.value
        return (np.ctypeslib.as_array(rp, (n + 1,)).copy(), np.ctypeslib.as_array(ci, (z,)).copy(),
                np.ctypeslib.as_array(va, (z,)).copy())


# ---------------------------------------------------------------------------------------------
# The reference's own Fortran (compiled from /root/reference by oracle/ref.mk into oracle/_ref).
# ---------------------------------------------------------------------------------------------
_REF = os.path.join(_HERE, "_ref", "libow3d_ref.so")
_ref = None


def ref_available():
    return os.path.exists(_REF)


def ref_lib():
    """The reference checker (lazy binding: the curvilinear branch of
    DifferentiationsFreeSurfacePlane is compiled but never called)."""
    global _ref
    if _ref is None:
        if not ref_available():
            if os.path.isdir("/root/reference/src"):
                subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)
            else:
                raise FileNotFoundError("oracle/_ref is not built and /root/reference is absent")
        R = ctypes.CDLL(_REF, mode=os.RTLD_LAZY)
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.c_int
        IP = ctypes.POINTER(ctypes.c_int)
        R.ref_setup.argtypes = [I, I, I, ctypes.c_double, ctypes.c_double, ctypes.c_double, I, I, D, D, D, D, D]
        R.ref_grid.argtypes = [D, D, D]
        R.ref_stencils.argtypes = [D, D, D]
        R.ref_surface.argtypes = [D] * 8
        R.ref_dsigma.argtypes = [D]
        R.ref_matvec.argtypes = [D, D]
        R.ref_vertical_velocity.argtypes = [D, D]
        R.ref_rhs.argtypes = [ctypes.c_double, D, D, D]
        R.ref_lowmatrix.argtypes = [I, I, I, D, D, D, D, D, D, D, D, IP]
        R.ref_lowmatrix_get.argtypes = [D, IP, IP]
        R.ref_gs.argtypes = [I, I, IP, IP, D, D, D, I, IP]
        R.ref_transfer.argtypes = [I, I, I, I, D, D, D, I, I, I, D, D, D, D, D]
        R.ref_store_ep.argtypes = [D, D, I]
        R.ref_jonswap.argtypes = [I, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D]
        R.ref_store_kinematics.argtypes = [D, IP, ctypes.c_double, I, I]
        _ref = R
    return _ref


def _ip(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


class RefTank:
    """The reference routines on one grid (module-global state: one RefTank at a time)."""

    def __init__(self, nx, ny, nz, lx, ly, alpha, gridz=1, lz=1.0, bath=None):
        R = ref_lib()
        self.nx, self.ny, self.nz, self.alpha = nx, ny, nz, alpha
        self.nxg, self.nyg, self.nzg = nx + 2, (ny + 2 if ny > 1 else 1), nz + 1
        self.shape2 = (self.nyg, self.nxg)
        self.shape3 = (self.nyg, self.nxg, self.nzg)
        b = [np.ascontiguousarray(bath[k], dtype=np.float64) for k in ("h", "hx", "hy", "hxx", "hyy")]
        self._keep = b
        R.ref_setup(nx, ny, nz, lx, ly, lz, gridz, alpha, *[_p(a) for a in b])

    def grid(self):
        x = np.zeros(self.nxg); y = np.zeros(self.nyg); z = np.zeros(self.nzg)
        ref_lib().ref_grid(_p(x), _p(y), _p(z))
        return x, y, z

    def stencils(self):
        r = 2 * self.alpha + 1
        sx = np.zeros((self.nxg, 2, r)); sy = np.zeros((self.nyg, 2, r) if self.ny > 1 else (1, 2, 1)); sz = np.zeros((self.nzg, 2, r))
        ref_lib().ref_stencils(_p(sx), _p(sy), _p(sz))
        return sx, sy, sz

    def surface(self, E, P):
        E = np.array(E, dtype=np.float64, copy=True); P = np.array(P, dtype=np.float64, copy=True)
        out = {k: np.zeros(self.shape2) for k in ("Ex", "Exx", "Ey", "Eyy", "Px", "Py")}
        ref_lib().ref_surface(_p(E), _p(P), *[_p(out[k]) for k in ("Ex", "Exx", "Ey", "Eyy", "Px", "Py")])
        return E, P, out

    def dsigma(self):
        out = np.zeros((5,) + self.shape3)
        ref_lib().ref_dsigma(_p(out))
        return out

    def matvec(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        out = np.zeros_like(phi)
        ref_lib().ref_matvec(_p(phi), _p(out))
        return out

    def vertical_velocity(self, phi):
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        W = np.zeros(self.shape2)
        ref_lib().ref_vertical_velocity(_p(phi), _p(W))
        return W

    def rhs(self, g, W):
        W = np.ascontiguousarray(W, dtype=np.float64)
        rE = np.zeros(self.shape2); rP = np.zeros(self.shape2)
        ref_lib().ref_rhs(g, _p(W), _p(rE), _p(rP))
        return rE, rP

    def store_kinematics(self, directory, phi, win, dt, it, nonlinear=1):
        """StoreKinematicData's Kinematics_01_<it>.bin (formattype 21) for phi and the surface last given to
        surface(), written into `directory`; win = (xbeg, xend, xstride, ybeg, yend, ystride, tbeg, tend,
        tstride).  Returns the file path."""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        w = np.ascontiguousarray(win, dtype=np.int32)
        cwd = os.getcwd()
        try:
            os.chdir(directory)
            ref_lib().ref_store_kinematics(_p(phi), _ip(w), float(dt), int(nonlinear), int(it))
        finally:
            os.chdir(cwd)
        return os.path.join(directory, "Kinematics_01_%05d.bin" % it)

    def store_ep(self, directory, E, P, nr=0):
        """StoreData's EP_<nr>.bin for E, P, written into `directory`; returns the file path."""
        E = np.ascontiguousarray(E, dtype=np.float64); P = np.ascontiguousarray(P, dtype=np.float64)
        cwd = os.getcwd()
        try:
            os.chdir(directory)
            ref_lib().ref_store_ep(_p(E), _p(P), int(nr))
        finally:
            os.chdir(cwd)
        return os.path.join(directory, "EP_%05d.bin" % nr)


def ref_lowmatrix(nx, ny, nz, x, y, z, bath):
    """Raw COO (value, 0-based row, 0-based col) of BuildLinearSystemMatrix on a level grid."""
    R = ref_lib()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, bath["h"], bath["hx"], bath["hy"], bath["hxx"], bath["hyy"])]
    nnz = ctypes.c_int(0)
    R.ref_lowmatrix(nx, ny, nz, *[_p(a) for a in arrs], ctypes.byref(nnz))
    val = np.zeros(nnz.value); row = np.zeros(nnz.value, dtype=np.int32); col = np.zeros(nnz.value, dtype=np.int32)
    R.ref_lowmatrix_get(_p(val), _ip(row), _ip(col))
    return val, row - 1, col - 1


def ref_gs(rowptr, col, val, x, b, ighost=None):
    """GaussSeidel (ighost None) or GaussSeidelSpecialPoints on a 0-based CSR matrix."""
    rp = np.ascontiguousarray(rowptr + 1, dtype=np.int32)
    ci = np.ascontiguousarray(col + 1, dtype=np.int32)
    va = np.ascontiguousarray(val, dtype=np.float64)
    x = np.array(x, dtype=np.float64, copy=True).ravel()
    b = np.ascontiguousarray(b, dtype=np.float64).ravel()
    gh = np.ascontiguousarray((ighost + 1) if ighost is not None else np.zeros(1), dtype=np.int32)
    ref_lib().ref_gs(len(rp) - 1, len(va), _ip(rp), _ip(ci), _p(va), _p(x), _p(b), 0 if ighost is None else len(gh), _ip(gh))
    return x


def ref_transfer(direction, fine_dims, fine_xyz, coarse_dims, coarse_xyz, fine, coarse):
    """direction 0: restriction fine -> coarse (returns coarse); 1: prolongation coarse -> fine."""
    f = [np.ascontiguousarray(a, dtype=np.float64) for a in fine_xyz]
    c = [np.ascontiguousarray(a, dtype=np.float64) for a in coarse_xyz]
    fine = np.array(fine, dtype=np.float64, copy=True)
    coarse = np.array(coarse, dtype=np.float64, copy=True)
    ref_lib().ref_transfer(direction, *fine_dims, *[_p(a) for a in f], *coarse_dims, *[_p(a) for a in c], _p(fine), _p(coarse))
    return coarse if direction == 0 else fine


def ref_jonswap(freq_hz, hs, tp, gamma):
    """JONSWAP_spectrum (utilities/JONSWAP_spectrum.f90) at the given frequencies (Hz)."""
    f = np.ascontiguousarray(freq_hz, dtype=np.float64)
    out = np.zeros_like(f)
    ref_lib().ref_jonswap(len(f), _p(f), float(hs), float(tp), float(gamma), _p(out))
    return out
