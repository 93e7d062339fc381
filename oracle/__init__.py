"""ctypes wrapper of the CPU oracle (oracle/libhd_oracle.so).

TEST INFRASTRUCTURE ONLY: the oracle restates hyper.deal's phase-space DG
advection operator on the CPU.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py may import this module, to check results or as
the timed CPU baseline.  The product (hyperdeal_amd) never imports it.

Every restated function cites its hyper.deal source in oracle/hd_oracle.cpp.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FUNC_ZERO = 0
FUNC_HYPERRECTANGLE = 1
FUNC_CONSTANT_ONE = 2
FUNC_SIN_COS = 3
FUNC_PRODUCT_SINES = 4
FUNC_LANDAU = 5

BC_HOMOGENEOUS = 0
BC_INHOMOGENEOUS = 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib_path():
    return os.path.join(_HERE, "libhd_oracle.so")


def build(native_out=None):
    """Compile the oracle (make -C oracle).  native_out: also build an
    -march=native copy at that path and return it."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    if native_out:
        subprocess.run(["make", "-s", "-C", _HERE, "native",
                        "NATIVE_OUT=" + native_out], check=True)
        return native_out
    return lib_path()


def load(path=None):
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or lib_path()
    if not os.path.exists(p):
        build()
    lib = ctypes.CDLL(p)
    lib.hdo_create.restype = ctypes.c_void_p
    lib.hdo_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int, _ip, _dp, _dp,
                               ctypes.c_int, ctypes.c_int, ctypes.c_double,
                               _dp, ctypes.c_int, ctypes.c_int, _dp]
    lib.hdo_destroy.argtypes = [ctypes.c_void_p]
    lib.hdo_n_dofs.restype = ctypes.c_longlong
    lib.hdo_n_dofs.argtypes = [ctypes.c_void_p]
    lib.hdo_n_cells.restype = ctypes.c_longlong
    lib.hdo_n_cells.argtypes = [ctypes.c_void_p]
    lib.hdo_apply.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double,
                              ctypes.c_int]
    lib.hdo_apply_cells.argtypes = [ctypes.c_void_p, _dp, _dp,
                                    ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_longlong),
                                    ctypes.c_int]
    lib.hdo_interpolate.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
    lib.hdo_norm_and_error.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double,
                                       _dp]
    lib.hdo_critical_time_step.restype = ctypes.c_double
    lib.hdo_critical_time_step.argtypes = [ctypes.c_void_p]
    lib.hdo_set_deformation.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_double,
                                        ctypes.c_int]
    lib.hdo_geometry.argtypes = [ctypes.c_void_p,
                                 ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp,
                                 _dp]
    lib.hdo_lsrk_step.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp, _dp,
                                  _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int]
    lib.hdo_lsrk_ode.restype = ctypes.c_double
    lib.hdo_lsrk_ode.argtypes = [ctypes.c_char_p, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_double]
    lib.hdo_lsrk_n_stages.argtypes = [ctypes.c_char_p]
    lib.hdo_lsrk_coefficients.argtypes = [ctypes.c_char_p, _dp, _dp]
    lib.hdo_run_advection.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                      ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int,
                                      ctypes.c_int, _dp, ctypes.c_int, _dp]
    lib.hdo_tables.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               _dp, _dp, _dp, _dp, _dp, _dp, _dp]
    lib.hdo_splitmix_uniform.restype = ctypes.c_double
    lib.hdo_splitmix_uniform.argtypes = [ctypes.c_ulonglong,
                                         ctypes.c_ulonglong]
    lib.hdo_add_noise.argtypes = [ctypes.c_void_p, _dp, ctypes.c_ulonglong,
                                  ctypes.c_double]
    lib.hdo_set_vlasov_poisson.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.hdo_x_n_dofs.restype = ctypes.c_longlong
    lib.hdo_x_n_dofs.argtypes = [ctypes.c_void_p]
    for f in ("hdo_vp_density", "hdo_vp_rhs", "hdo_poisson_apply",
              "hdo_vp_diagnostics"):
        getattr(lib, f).argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.hdo_poisson_solve.argtypes = [ctypes.c_void_p, _dp, _dp,
                                      ctypes.c_double, ctypes.c_int]
    lib.hdo_vp_set_potential.argtypes = [ctypes.c_void_p, _dp]
    lib.hdo_vp_efield.argtypes = [ctypes.c_void_p, _dp]
    lib.hdo_vp_energy.argtypes = [ctypes.c_void_p, _dp]
    lib.hdo_run_vlasov_poisson.argtypes = [
        ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double, ctypes.c_double,
        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
        ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_int)]
    lib.hdo_num_threads.restype = ctypes.c_int
    if path is None:
        _LIB = lib
    return lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


class Problem:
    """A Cartesian phase-space advection problem on the oracle."""

    def __init__(self, dim_x, dim_v, degree, n_cells, lo, hi, a,
                 n_points=None, collocation=False, periodic_x=True,
                 periodic_v=True, skew=0.0, bc=BC_INHOMOGENEOUS,
                 func=FUNC_ZERO, func_params=None, lib=None):
        self.lib = lib or load()
        D = dim_x + dim_v
        self.dim_x, self.dim_v, self.D, self.degree = dim_x, dim_v, D, degree
        self.n_points = n_points or degree + 1
        self.n_cells = list(n_cells)
        nc = (ctypes.c_int * D)(*n_cells)
        lo_ = (ctypes.c_double * D)(*lo)
        hi_ = (ctypes.c_double * D)(*hi)
        a_ = (ctypes.c_double * D)(*a)
        fp = None
        if func_params is not None:
            fp = (ctypes.c_double * len(func_params))(*func_params)
        self.h = self.lib.hdo_create(dim_x, dim_v, degree, self.n_points,
                                     int(collocation), nc, lo_, hi_,
                                     int(periodic_x), int(periodic_v),
                                     float(skew), a_, bc, func, fp)
        if not self.h:
            raise ValueError("oracle: invalid problem")
        self.n_dofs = self.lib.hdo_n_dofs(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.hdo_destroy(self.h)
            self.h = None

    def vector(self):
        return np.zeros(self.n_dofs, dtype=np.float64)

    def apply(self, src, t=0.0, fcl=False, dst=None):
        dst = self.vector() if dst is None else dst
        rc = self.lib.hdo_apply(self.h, _ptr(src), _ptr(dst), t, int(fcl))
        assert rc == 0
        return dst

    def apply_cells(self, src, cell_ids, t=0.0, dst=None):
        """ECL apply on the given cells only; returns dst (other cells
        untouched)."""
        dst = self.vector() if dst is None else dst
        ids = np.ascontiguousarray(cell_ids, dtype=np.int64)
        rc = self.lib.hdo_apply_cells(
            self.h, _ptr(src), _ptr(dst), t,
            ids.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(ids))
        assert rc == 0
        return dst

    def interpolate(self, t=0.0):
        dst = self.vector()
        self.lib.hdo_interpolate(self.h, t, _ptr(dst))
        return dst

    def norm_and_error(self, src, t=0.0):
        out = np.zeros(2)
        self.lib.hdo_norm_and_error(self.h, _ptr(src), t, _ptr(out))
        return out

    def critical_time_step(self):
        return self.lib.hdo_critical_time_step(self.h)

    def set_deformation(self, deformation_x, frequency_x=2,
                        deformation_v=None, frequency_v=None):
        """The deformed mesh of grid_generator.cc:110-230 (exact map)."""
        dv = deformation_x if deformation_v is None else deformation_v
        fv = frequency_x if frequency_v is None else frequency_v
        assert self.lib.hdo_set_deformation(self.h, float(deformation_x),
                                            int(frequency_x), float(dv),
                                            int(fv)) == 0

    def geometry(self, cell_index, xi):
        """(physical point, inverse Jacobian d xi / d x, det d x / d xi) at
        reference point xi of the cell with the given multi-index."""
        D = self.D
        ci = (ctypes.c_int * D)(*cell_index)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        x = np.zeros(D)
        ji = np.zeros((D, D))
        det = np.zeros(1)
        assert self.lib.hdo_geometry(self.h, ci, _ptr(xi), _ptr(x), _ptr(ji),
                                     _ptr(det)) == 0
        return x, ji, float(det[0])

    def lsrk_step(self, rk, sol, ki, ti, t, dt, ecl=True):
        rc = self.lib.hdo_lsrk_step(self.h, rk.encode(), _ptr(sol), _ptr(ki),
                                    _ptr(ti), t, dt, int(ecl))
        assert rc == 0

    def add_noise(self, dst, seed, amplitude):
        self.lib.hdo_add_noise(self.h, _ptr(dst), seed, amplitude)

    def run_advection(self, rk="rk45", time_step=0.1, start_time=0.0,
                      final_time=1.0, cfl=0.3, diag_tick=0.1, ecl=True,
                      max_steps=100000000, max_rows=100000):
        rows = np.zeros((max_rows, 3))
        dt = np.zeros(1)
        n = self.lib.hdo_run_advection(self.h, rk.encode(), time_step,
                                       start_time, final_time, cfl, diag_tick,
                                       int(ecl), max_steps, _ptr(rows),
                                       max_rows, _ptr(dt))
        assert n >= 0
        return rows[:n].copy(), float(dt[0])


class VlasovPoisson(Problem):
    """The Vlasov-Poisson problem of examples/vlasov_poisson on the oracle:
    x-velocity v, v-velocity grad phi from the periodic SIPG Poisson solve.
    `transport` is the case's direction used only for the CFL time step."""

    def __init__(self, dim, degree, n_cells, lo, hi, transport, eps=0.01,
                 kx=0.5, n_perturbed=1, **kw):
        super().__init__(dim, dim, degree, n_cells, lo, hi, transport,
                         func=FUNC_LANDAU, func_params=[eps, kx, n_perturbed],
                         **kw)
        assert self.lib.hdo_set_vlasov_poisson(self.h, 1) == 0
        self.n_x_dofs = self.lib.hdo_x_n_dofs(self.h)

    def xvector(self):
        return np.zeros(self.n_x_dofs)

    def density(self, f):
        rho = self.xvector()
        self.lib.hdo_vp_density(self.h, _ptr(f), _ptr(rho))
        return rho

    def rhs(self, rho):
        rho = rho.copy()
        b = self.xvector()
        self.lib.hdo_vp_rhs(self.h, _ptr(rho), _ptr(b))
        return b

    def poisson_apply(self, x):
        y = self.xvector()
        self.lib.hdo_poisson_apply(self.h, _ptr(x), _ptr(y))
        return y

    def poisson_solve(self, b, phi=None, rel_tol=1e-12, max_it=100000):
        phi = self.xvector() if phi is None else phi
        it = self.lib.hdo_poisson_solve(self.h, _ptr(b), _ptr(phi), rel_tol,
                                        max_it)
        return phi, it

    def set_potential(self, phi):
        assert self.lib.hdo_vp_set_potential(self.h, _ptr(phi)) == 0

    def efield(self):
        nqx = self.n_points ** self.dim_x
        ncx = int(np.prod(self.n_cells[:self.dim_x]))
        out = np.zeros(ncx * nqx * self.dim_x)
        self.lib.hdo_vp_efield(self.h, _ptr(out))
        return out

    def energy(self):
        out = np.zeros(self.dim_x)
        self.lib.hdo_vp_energy(self.h, _ptr(out))
        return out

    def diagnostics(self, f):
        out = np.zeros(6)
        self.lib.hdo_vp_diagnostics(self.h, _ptr(f), _ptr(out))
        return out

    def run(self, rk="rk45", time_step=0.1, start_time=0.0, final_time=0.5,
            cfl=0.05, diag_tick=0.1, poisson_tol=1e-12,
            max_steps=100000000, max_rows=10000):
        width = 1 + self.dim_x + 6
        rows = np.zeros((max_rows, width))
        dt = np.zeros(1)
        it = ctypes.c_int(0)
        n = self.lib.hdo_run_vlasov_poisson(
            self.h, rk.encode(), time_step, start_time, final_time, cfl,
            diag_tick, poisson_tol, max_steps, _ptr(rows), max_rows,
            _ptr(dt), ctypes.byref(it))
        assert n >= 0
        return rows[:n].copy(), float(dt[0]), it.value


def tables(degree, n_points=None, collocation=False):
    lib = load()
    nq = n_points or degree + 1
    nd = degree + 1
    out = dict(nodes=np.zeros(nd), qpoints=np.zeros(nq), qweights=np.zeros(nq),
               S=np.zeros((nq, nd)), D=np.zeros((nq, nq)), F=np.zeros((2, nq)),
               Sinv=np.zeros((nd, nq)))
    lib.hdo_tables(degree, nq, int(collocation), *[_ptr(out[k]) for k in
                   ("nodes", "qpoints", "qweights", "S", "D", "F", "Sinv")])
    return out


def lsrk_ode(rk, dt, n_steps, y0=1.0):
    return load().hdo_lsrk_ode(rk.encode(), dt, n_steps, y0)


def lsrk_coefficients(rk):
    lib = load()
    s = lib.hdo_lsrk_n_stages(rk.encode())
    a, b = np.zeros(max(s - 1, 1)), np.zeros(s)
    lib.hdo_lsrk_coefficients(rk.encode(), _ptr(a), _ptr(b))
    return a[: s - 1], b


def splitmix_uniform(seed, i):
    return load().hdo_splitmix_uniform(seed, i)


def num_threads():
    return load().hdo_num_threads()
