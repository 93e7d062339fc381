"""hyperdeal_amd -- MI355X (gfx950) phase-space DG advection operator.

Python mirror of the hyper.deal interface for the hot path, over the C ABI
in include/hyperdeal_amd.h (libhyperdeal_amd.so, built in-tree):

    hyper.deal                                      here
    ----------------------------------------------  -------------------------
    hyperdeal::MatrixFree (reinit, n_dofs)           MatrixFree
    advection::AdvectionOperation (reinit, apply)    AdvectionOperation
    LowStorageRungeKuttaIntegrator                   LowStorageRungeKuttaIntegrator
      (perform_time_step, n_stages)
    VectorTools::interpolate / norm_and_error        VectorTools
    advection::compute_critical_time_step            compute_critical_time_step

Vectors are torch tensors on the GPU (torch is only the allocator here);
every compute call goes to the HIP library and raises HyperdealError on a
non-zero status.  There is no CPU fallback: if the library is missing the
import of the binding fails loudly.
"""

import ctypes
import os

__all__ = ["MatrixFree", "AdvectionOperation", "VlasovPoissonOperation",
           "vp_moments", "LowStorageRungeKuttaIntegrator",
           "VectorTools", "HyperdealError", "FP64", "FP32", "library_path",
           "load_library", "compute_critical_time_step"]

_HERE = os.path.dirname(os.path.abspath(__file__))
FP64, FP32 = 0, 1
ALGO_AUTO, ALGO_QUADRATURE, ALGO_COLLAPSED, ALGO_FCL = 0, 1, 2, 3
BC_HOMOGENEOUS, BC_INHOMOGENEOUS = 0, 1
FUNC_ZERO, FUNC_HYPERRECTANGLE, FUNC_CONSTANT_ONE, FUNC_SIN_COS, \
    FUNC_PRODUCT_SINES, FUNC_LANDAU = range(6)

# every symbol include/hyperdeal_amd.h declares
EXPORTED = [
    "hd_last_error", "hd_version", "hd_matrix_free_create",
    "hd_matrix_free_destroy", "hd_matrix_free_n_dofs",
    "hd_matrix_free_n_cells", "hd_matrix_free_dofs_per_cell",
    "hd_advection_create", "hd_advection_destroy",
    "hd_advection_set_algorithm", "hd_advection_algorithm",
    "hd_advection_apply", "hd_lsrk_stage", "hd_lsrk_coefficients",
    "hd_lsrk_step", "hd_interpolate", "hd_add_noise", "hd_norm_and_error",
    "hd_vector_convert", "hd_critical_time_step", "hd_device_malloc",
    "hd_device_free", "hd_memcpy_htod", "hd_memcpy_dtoh",
    "hd_device_synchronize", "hd_vlasov_poisson_create",
    "hd_vp_update_field", "hd_vp_set_auto_update",
    "hd_vp_set_poisson_tolerance", "hd_vp_solver_stats", "hd_vp_x_n_dofs",
    "hd_vp_potential", "hd_vp_field_energy", "hd_vp_moments",
    "hd_batch_create", "hd_batch_destroy", "hd_batch_size", "hd_batch_apply",
    "hd_batch_lsrk_step", "hd_batch_autotune", "hd_batch_cells_per_task",
    "hd_batch_set_cells_per_task", "hd_batch_graph_create",
    "hd_graph_launch", "hd_graph_destroy", "hd_vp_solver_start",
    "hd_vp_learned_guess_enable", "hd_vp_learned_guess_disable",
    "hd_vp_learned_guess_stats", "hd_vp_learned_guess_net",
    "hd_vp_learned_guess_schedule",
    "hd_cnn_n_params", "hd_cnn_create", "hd_cnn_destroy", "hd_cnn_params",
    "hd_cnn_forward", "hd_cnn_loss_grad", "hd_cnn_train_step",
    "hd_cnn_last_loss", "hd_vp_exact_start", "hd_cnn_dofs_to_image",
    "hd_cnn_image_to_dofs", "hd_norm_and_error_noalloc",
    "hd_matrix_free_set_deformation",
]
BATCH_MAX = 64


class HyperdealError(RuntimeError):
    pass


class _MeshDesc(ctypes.Structure):
    _fields_ = [("dim_x", ctypes.c_int), ("dim_v", ctypes.c_int),
                ("n_cells", ctypes.c_int * 6), ("lo", ctypes.c_double * 6),
                ("hi", ctypes.c_double * 6), ("periodic_x", ctypes.c_int),
                ("periodic_v", ctypes.c_int)]


_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


def library_path():
    return os.path.join(_HERE, "libhyperdeal_amd.so")


def load_library(path=None):
    """Load libhyperdeal_amd.so (built by `make -C hyperdeal_amd`)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or library_path()
    if not os.path.exists(p):
        raise HyperdealError(
            "libhyperdeal_amd.so not built: run __graft_entry__.build() or "
            "`make -C hyperdeal_amd`")
    lib = ctypes.CDLL(p)
    lib.hd_last_error.restype = ctypes.c_char_p
    lib.hd_matrix_free_create.argtypes = [ctypes.POINTER(_MeshDesc),
                                          ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(_vp)]
    lib.hd_matrix_free_destroy.argtypes = [_vp]
    lib.hd_matrix_free_set_deformation.argtypes = [_vp, ctypes.c_double,
                                                   ctypes.c_int,
                                                   ctypes.c_double,
                                                   ctypes.c_int]
    for f in ("hd_matrix_free_n_dofs", "hd_matrix_free_n_cells"):
        getattr(lib, f).restype = ctypes.c_longlong
        getattr(lib, f).argtypes = [_vp]
    lib.hd_matrix_free_dofs_per_cell.argtypes = [_vp]
    lib.hd_advection_create.argtypes = [_vp, _dp, ctypes.c_double,
                                        ctypes.c_int, ctypes.c_int, _dp,
                                        ctypes.POINTER(_vp)]
    lib.hd_advection_destroy.argtypes = [_vp]
    lib.hd_advection_set_algorithm.argtypes = [_vp, ctypes.c_int]
    lib.hd_advection_algorithm.argtypes = [_vp, ctypes.c_int]
    lib.hd_advection_apply.argtypes = [_vp, _vp, _vp, ctypes.c_int,
                                       ctypes.c_double, _vp]
    lib.hd_lsrk_stage.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp,
                                  ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, _vp]
    lib.hd_lsrk_coefficients.argtypes = [ctypes.c_char_p, _dp, _dp]
    lib.hd_lsrk_step.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int, _vp, _vp,
                                 _vp, ctypes.c_double, ctypes.c_double, _vp]
    lib.hd_interpolate.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_double,
                                   _vp, _vp]
    lib.hd_add_noise.argtypes = [_vp, _vp, ctypes.c_uint64, ctypes.c_double,
                                 _vp]
    lib.hd_norm_and_error.argtypes = [_vp, _vp, ctypes.c_int, _dp,
                                      ctypes.c_double, _dp, _vp]
    lib.hd_vector_convert.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
    lib.hd_norm_and_error_noalloc.argtypes = [_vp, _vp, ctypes.c_int, _dp,
                                              ctypes.c_double, _vp, _vp]
    lib.hd_critical_time_step.restype = ctypes.c_double
    lib.hd_critical_time_step.argtypes = [_vp, _dp]
    lib.hd_device_synchronize.argtypes = []
    lib.hd_vlasov_poisson_create.argtypes = [_vp, ctypes.c_double,
                                             ctypes.c_double,
                                             ctypes.POINTER(_vp)]
    lib.hd_vp_update_field.argtypes = [_vp, _vp, _vp]
    lib.hd_vp_set_auto_update.argtypes = [_vp, ctypes.c_int]
    lib.hd_vp_set_poisson_tolerance.argtypes = [_vp, ctypes.c_double,
                                                ctypes.c_int]
    lib.hd_vp_solver_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_longlong),
                                       ctypes.POINTER(ctypes.c_longlong), _dp]
    _llp = ctypes.POINTER(ctypes.c_longlong)
    lib.hd_vp_solver_start.argtypes = [_vp, _dp, _dp, _dp, _llp]
    lib.hd_vp_exact_start.argtypes = [_vp, ctypes.c_int]
    lib.hd_vp_learned_guess_enable.argtypes = [
        _vp, ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_float, ctypes.c_int]
    lib.hd_vp_learned_guess_disable.argtypes = [_vp]
    lib.hd_vp_learned_guess_schedule.argtypes = [_vp, ctypes.c_longlong,
                                                 ctypes.c_int]
    lib.hd_vp_learned_guess_stats.argtypes = [_vp, _llp, _llp, _llp, _dp]
    lib.hd_vp_learned_guess_net.restype = _vp
    lib.hd_vp_learned_guess_net.argtypes = [_vp]
    _fp = ctypes.POINTER(ctypes.c_float)
    lib.hd_cnn_n_params.argtypes = []
    lib.hd_cnn_create.argtypes = [ctypes.c_int, ctypes.c_int,
                                  ctypes.c_ulonglong, ctypes.POINTER(_vp)]
    lib.hd_cnn_destroy.argtypes = [_vp]
    lib.hd_cnn_params.argtypes = [_vp, _fp, ctypes.c_int]
    lib.hd_cnn_forward.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
    lib.hd_cnn_loss_grad.argtypes = [_vp, _vp, _vp, ctypes.c_int, _dp, _fp,
                                     _vp]
    lib.hd_cnn_train_step.argtypes = [_vp, _vp, _vp, ctypes.c_int,
                                      ctypes.c_float, _vp]
    lib.hd_cnn_last_loss.argtypes = [_vp, _dp]
    lib.hd_cnn_dofs_to_image.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _vp, _vp, _vp]
    lib.hd_cnn_image_to_dofs.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, _vp, _vp]
    lib.hd_vp_x_n_dofs.restype = ctypes.c_longlong
    lib.hd_vp_x_n_dofs.argtypes = [_vp]
    lib.hd_vp_potential.argtypes = [_vp, _dp, ctypes.c_int]
    lib.hd_vp_field_energy.argtypes = [_vp, _dp]
    lib.hd_vp_moments.argtypes = [_vp, _vp, _dp, _vp]
    _pa = ctypes.POINTER(_vp)
    lib.hd_batch_create.argtypes = [_pa, ctypes.c_int, ctypes.POINTER(_vp)]
    lib.hd_batch_destroy.argtypes = [_vp]
    lib.hd_batch_size.argtypes = [_vp]
    lib.hd_batch_apply.argtypes = [_vp, _pa, _pa, ctypes.c_int, _vp]
    lib.hd_batch_lsrk_step.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int,
                                       _pa, _pa, _pa, ctypes.c_double, _vp]
    lib.hd_batch_autotune.argtypes = [_vp, ctypes.c_int, _pa, _pa, _vp]
    lib.hd_batch_cells_per_task.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
    lib.hd_batch_set_cells_per_task.argtypes = [_vp, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int]
    lib.hd_batch_graph_create.argtypes = [_vp, ctypes.c_char_p, ctypes.c_int,
                                          _pa, _pa, _pa, ctypes.c_double,
                                          ctypes.POINTER(_vp)]
    lib.hd_graph_launch.argtypes = [_vp, _vp]
    lib.hd_graph_destroy.argtypes = [_vp]
    if path is None:
        _lib = lib
    return lib


def _check(rc, what):
    if rc < 0:
        msg = load_library().hd_last_error().decode()
        raise HyperdealError("%s failed (%d): %s" % (what, rc, msg))
    return rc


def _darr(vals):
    vals = list(vals)
    return (ctypes.c_double * max(len(vals), 1))(*vals)


def _ptr(t, dtype, n):
    """Device pointer of a contiguous torch tensor, with checks."""
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise HyperdealError("vectors must be torch tensors on the GPU")
    if t.dtype != dtype:
        raise HyperdealError("expected %s, got %s" % (dtype, t.dtype))
    if not t.is_contiguous() or t.numel() != n:
        raise HyperdealError("expected a contiguous vector of %d values" % n)
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream")
                           else int(stream))


def _torch_dtype(precision):
    import torch
    return torch.float64 if precision == FP64 else torch.float32


class MatrixFree:
    """Phase-space matrix-free data of a Cartesian box (hyperdeal::MatrixFree
    built from two dealii::MatrixFree objects, matrix_free.templates.h:860)."""

    def __init__(self, dim_x, dim_v, degree, n_cells, lo, hi, periodic_x=True,
                 periodic_v=True, n_points=None, collocation=False):
        self._lib = load_library()
        D = dim_x + dim_v
        if len(n_cells) != D or len(lo) != D or len(hi) != D:
            raise HyperdealError("n_cells/lo/hi need dim_x + dim_v entries")
        desc = _MeshDesc()
        desc.dim_x, desc.dim_v = dim_x, dim_v
        for d in range(D):
            desc.n_cells[d] = int(n_cells[d])
            desc.lo[d], desc.hi[d] = float(lo[d]), float(hi[d])
        desc.periodic_x, desc.periodic_v = int(periodic_x), int(periodic_v)
        self.dim_x, self.dim_v, self.dim, self.degree = dim_x, dim_v, D, degree
        self.n_points = n_points or degree + 1
        self.n_cells = list(n_cells)
        self.lo, self.hi = list(lo), list(hi)
        h = _vp()
        _check(self._lib.hd_matrix_free_create(ctypes.byref(desc), degree,
                                               self.n_points, int(collocation),
                                               ctypes.byref(h)),
               "hd_matrix_free_create")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.hd_matrix_free_destroy(self.handle)
            self.handle = None

    def set_deformation(self, deformation_x=0.1, frequency_x=2,
                        deformation_v=None, frequency_v=None):
        """Deform both spaces with the reference's DeformedCubeManifold
        (grid_generator.cc:110-230; defaults 0.1 and 2) before operators
        are created on this object."""
        dv = deformation_x if deformation_v is None else deformation_v
        fv = frequency_x if frequency_v is None else frequency_v
        _check(self._lib.hd_matrix_free_set_deformation(
            self.handle, float(deformation_x), int(frequency_x), float(dv),
            int(fv)), "hd_matrix_free_set_deformation")

    def n_dofs(self):
        return self._lib.hd_matrix_free_n_dofs(self.handle)

    def n_cells_total(self):
        return self._lib.hd_matrix_free_n_cells(self.handle)

    def dofs_per_cell(self):
        return self._lib.hd_matrix_free_dofs_per_cell(self.handle)

    def initialize_dof_vector(self, precision=FP64, device="cuda"):
        import torch
        return torch.zeros(self.n_dofs(), dtype=_torch_dtype(precision),
                           device=device)


def compute_critical_time_step(mf, velocity):
    """dt_crit = min over x and v of 1 / max |J^{-T} a| (cfl.h:94-124)."""
    return mf._lib.hd_critical_time_step(mf.handle, _darr(velocity))


class AdvectionOperation:
    """advection::AdvectionOperation (advection_operation.h:56-209) with a
    constant velocity field."""

    def __init__(self, matrix_free, velocity, skew_factor=0.0,
                 boundary=BC_INHOMOGENEOUS, boundary_function=FUNC_ZERO,
                 boundary_params=None):
        self.mf = matrix_free
        self._lib = matrix_free._lib
        self.velocity = [float(v) for v in velocity]
        if len(self.velocity) != matrix_free.dim:
            raise HyperdealError("velocity needs dim_x + dim_v entries")
        h = _vp()
        bp = _darr(boundary_params) if boundary_params is not None else None
        _check(self._lib.hd_advection_create(matrix_free.handle,
                                             _darr(self.velocity),
                                             float(skew_factor), boundary,
                                             boundary_function, bp,
                                             ctypes.byref(h)),
               "hd_advection_create")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.hd_advection_destroy(self.handle)
            self.handle = None

    def set_algorithm(self, algorithm):
        _check(self._lib.hd_advection_set_algorithm(self.handle, algorithm),
               "hd_advection_set_algorithm")

    def algorithm(self, precision=FP64):
        return _check(self._lib.hd_advection_algorithm(self.handle, precision),
                      "hd_advection_algorithm")

    def apply(self, dst, src, time=0.0, precision=FP64, stream=None):
        """dst = M^{-1}(cell + face integrals)(src) (apply, :134-209)."""
        n = self.mf.n_dofs()
        dt = _torch_dtype(precision)
        _check(self._lib.hd_advection_apply(self.handle, _ptr(src, dt, n),
                                            _ptr(dst, dt, n), precision,
                                            float(time), _stream(stream)),
               "hd_advection_apply")
        return dst

    def stage(self, ti_in, ti_out, sol, time, a_dt, b_dt, precision=FP64,
              stream=None):
        """One fused LSRK stage (time_integrators.templates.h:103-138)."""
        import torch
        n = self.mf.n_dofs()
        dt = _torch_dtype(precision)
        out = _ptr(ti_out, dt, n) if ti_out is not None else None
        _check(self._lib.hd_lsrk_stage(self.handle, precision,
                                       _ptr(ti_in, dt, n), out,
                                       _ptr(sol, torch.float64, n),
                                       float(time), float(a_dt), float(b_dt),
                                       _stream(stream)),
               "hd_lsrk_stage")


class VlasovPoissonOperation(AdvectionOperation):
    """The Vlasov-Poisson right-hand side (examples/vlasov_poisson/include/
    application.h:516-602): advection with velocity (v, grad phi), phi from
    the periodic SIPG Poisson problem for rho = int f dv, solved with FP64
    CG from the previous phi.  Every apply/stage recomputes the field from
    its input unless auto update is switched off."""

    def __init__(self, matrix_free, skew_factor=0.0, poisson_tol=1e-10):
        self.mf = matrix_free
        self._lib = matrix_free._lib
        self.velocity = None
        h = _vp()
        _check(self._lib.hd_vlasov_poisson_create(matrix_free.handle,
                                                  float(skew_factor),
                                                  float(poisson_tol),
                                                  ctypes.byref(h)),
               "hd_vlasov_poisson_create")
        self.handle = h

    def update_field(self, f, stream=None):
        import torch
        _check(self._lib.hd_vp_update_field(
            self.handle, _ptr(f, torch.float64, self.mf.n_dofs()),
            _stream(stream)), "hd_vp_update_field")

    def set_auto_update(self, on):
        _check(self._lib.hd_vp_set_auto_update(self.handle, int(on)),
               "hd_vp_set_auto_update")

    def set_poisson_tolerance(self, rel_tol, max_it=100000):
        _check(self._lib.hd_vp_set_poisson_tolerance(self.handle,
                                                     float(rel_tol), max_it),
               "hd_vp_set_poisson_tolerance")

    def solver_stats(self):
        it, n = ctypes.c_longlong(), ctypes.c_longlong()
        res = ctypes.c_double()
        _check(self._lib.hd_vp_solver_stats(self.handle, ctypes.byref(it),
                                            ctypes.byref(n),
                                            ctypes.byref(res)),
               "hd_vp_solver_stats")
        return it.value, n.value, res.value

    def solver_start(self):
        """(relative residual of the last solve's start, that of the previous
        phi alone, weights (previous phi, guess), solves from a guess)."""
        r0, rp = ctypes.c_double(), ctypes.c_double()
        w = (ctypes.c_double * 2)()
        g = ctypes.c_longlong()
        _check(self._lib.hd_vp_solver_start(self.handle, ctypes.byref(r0),
                                            ctypes.byref(rp), w,
                                            ctypes.byref(g)),
               "hd_vp_solver_start")
        return r0.value, rp.value, (w[0], w[1]), g.value

    def exact_start(self, enable=True):
        """Start the potential solves from the exact solution of the
        periodic 2D SIPG system (fast diagonalisation); CG still runs to the
        tolerance from there.  Disables a learned guess."""
        if enable:
            self._release_learned_net()
        _check(self._lib.hd_vp_exact_start(self.handle, 1 if enable else 0),
               "hd_vp_exact_start")

    def _release_learned_net(self):
        # a network borrowed from the learned guess (learned_guess_net) dies
        # with it: a new generation invalidates the borrowed handles
        self._learned_gen = getattr(self, "_learned_gen", 0) + 1

    def enable_learned_guess(self, seed=20261205, capacity=64, train_every=1,
                             train_batch=4, train_steps=1,
                             learning_rate=1e-3, min_pairs=4):
        """Start the potential solves from the previous phi combined with
        the CNN's correction for its residual; the network trains on this
        operator's own solves (residual -> correction pairs)."""
        self._release_learned_net()
        _check(self._lib.hd_vp_learned_guess_enable(
            self.handle, int(seed), int(capacity), int(train_every),
            int(train_batch), int(train_steps), float(learning_rate),
            int(min_pairs)), "hd_vp_learned_guess_enable")

    def learned_guess_schedule(self, dense_solves, train_every):
        """Train on every one of the first dense_solves solves (counted
        from enabling), then every train_every solves (0: keep the network
        as trained)."""
        _check(self._lib.hd_vp_learned_guess_schedule(
            self.handle, int(dense_solves), int(train_every)),
            "hd_vp_learned_guess_schedule")

    def disable_learned_guess(self):
        self._release_learned_net()
        _check(self._lib.hd_vp_learned_guess_disable(self.handle),
               "hd_vp_learned_guess_disable")

    def learned_guess_stats(self):
        """(Adam steps, solves that used the guess, pairs stored, last
        training loss)."""
        st, gs, pr = (ctypes.c_longlong() for _ in range(3))
        loss = ctypes.c_double()
        _check(self._lib.hd_vp_learned_guess_stats(
            self.handle, ctypes.byref(st), ctypes.byref(gs), ctypes.byref(pr),
            ctypes.byref(loss)), "hd_vp_learned_guess_stats")
        return st.value, gs.value, pr.value, loss.value

    def learned_guess_net(self):
        """The network of the learned guess, borrowed from this operator:
        usable until the learned guess is disabled, enabled again or
        replaced by the exact start (calls raise HyperdealError after
        that)."""
        h = self._lib.hd_vp_learned_guess_net(self.handle)
        if not h:
            raise HyperdealError("no learned guess on this operator")
        return Cnn(_handle=h, _owner=self)

    def x_n_dofs(self):
        return self._lib.hd_vp_x_n_dofs(self.handle)

    def potential(self):
        import numpy as np
        out = np.zeros(self.x_n_dofs())
        _check(self._lib.hd_vp_potential(
            self.handle, out.ctypes.data_as(_dp), 1), "hd_vp_potential")
        return out

    def set_potential(self, phi):
        import numpy as np
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        if phi.ndim != 1 or phi.size != self.x_n_dofs():
            raise HyperdealError("set_potential: %d values, expected %d"
                                 % (phi.size, self.x_n_dofs()))
        _check(self._lib.hd_vp_potential(
            self.handle, phi.ctypes.data_as(_dp), 0), "hd_vp_potential")

    def field_energy(self):
        out = (ctypes.c_double * 3)()
        _check(self._lib.hd_vp_field_energy(self.handle, out),
               "hd_vp_field_energy")
        return [out[i] for i in range(self.mf.dim_x)]


class Cnn:
    """The Poisson-guess network on its own: four 3x3 periodic convolutions
    1 -> 32 -> 32 -> 32 -> 1, leaky ReLU 1/8, BF16 weights, FP32 accumulation,
    Adam on FP32 master parameters.  Images are float32 CUDA tensors of shape
    (batch, height, width), height and width multiples of 16."""

    def __init__(self, height=None, width=None, seed=0, lib=None,
                 _handle=None, _owner=None):
        self._lib = lib or _lib or load_library()
        self._owner = _owner
        if _handle is not None:
            self._handle = _vp(_handle)
            self._owned = False
            self._gen = getattr(_owner, "_learned_gen", 0)
            return
        h = _vp()
        _check(self._lib.hd_cnn_create(int(height), int(width), int(seed),
                                       ctypes.byref(h)), "hd_cnn_create")
        self._handle = h
        self._owned = True

    @property
    def handle(self):
        """the C handle; for a borrowed network, only while its operator's
        learned guess lives"""
        if not self._owned and \
                getattr(self._owner, "_learned_gen", 0) != self._gen:
            raise HyperdealError("the learned guess this network belonged "
                                 "to was released")
        return self._handle

    @staticmethod
    def n_params(lib=None):
        return (lib or _lib or load_library()).hd_cnn_n_params()

    def params(self):
        import numpy as np
        out = np.zeros(self._lib.hd_cnn_n_params(), dtype=np.float32)
        _check(self._lib.hd_cnn_params(
            self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
            1), "hd_cnn_params")
        return out

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self._lib.hd_cnn_n_params():
            raise HyperdealError("wrong number of parameters")
        _check(self._lib.hd_cnn_params(
            self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 0),
            "hd_cnn_params")

    @staticmethod
    def _images(x):
        import torch
        if x.dtype != torch.float32 or not x.is_cuda or x.dim() != 3 or \
                not x.is_contiguous():
            raise HyperdealError("images must be contiguous float32 CUDA "
                                 "tensors (batch, height, width)")
        return ctypes.c_void_p(x.data_ptr())

    def forward(self, x, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty_like(x)
        _check(self._lib.hd_cnn_forward(self.handle, self._images(x),
                                        self._images(out), x.shape[0],
                                        _stream(stream)), "hd_cnn_forward")
        return out

    def loss_grad(self, x, target, stream=None):
        import numpy as np
        g = np.zeros(self._lib.hd_cnn_n_params(), dtype=np.float32)
        loss = ctypes.c_double()
        _check(self._lib.hd_cnn_loss_grad(
            self.handle, self._images(x), self._images(target), x.shape[0],
            ctypes.byref(loss), g.ctypes.data_as(
                ctypes.POINTER(ctypes.c_float)), _stream(stream)),
            "hd_cnn_loss_grad")
        return loss.value, g

    def train_step(self, x, target, learning_rate=1e-3, stream=None):
        _check(self._lib.hd_cnn_train_step(
            self.handle, self._images(x), self._images(target), x.shape[0],
            float(learning_rate), _stream(stream)), "hd_cnn_train_step")

    def last_loss(self):
        loss = ctypes.c_double()
        _check(self._lib.hd_cnn_last_loss(self.handle, ctypes.byref(loss)),
               "hd_cnn_last_loss")
        return loss.value

    def __del__(self):
        if getattr(self, "_owned", False) and self._handle:
            self._lib.hd_cnn_destroy(self._handle)
            self._handle = None


def cnn_dofs_to_image(u, n0, n1, N, sub=None, stream=None):
    """The learned guess's x-space DoFs -> network image: ((u - sub) - mean)
    / rms as a float32 (n1 N, n0 N) CUDA tensor, and rms (a float64 CUDA
    tensor of one element)."""
    import torch
    n = n0 * n1 * N * N
    lib = load_library()
    img = torch.empty((n1 * N, n0 * N), dtype=torch.float32, device=u.device)
    scale = torch.empty(1, dtype=torch.float64, device=u.device)
    _check(lib.hd_cnn_dofs_to_image(
        _ptr(u, torch.float64, n),
        _ptr(sub, torch.float64, n) if sub is not None else None, n0, n1, N,
        _ptr(img, torch.float32, n), _ptr(scale, torch.float64, 1),
        _stream(stream)), "hd_cnn_dofs_to_image")
    return img, scale


def cnn_image_to_dofs(img, scale, n0, n1, N, stream=None):
    """Network image -> x-space DoFs, scale * image with shared nodes
    averaged over their pixels (continuous across cell faces)."""
    import torch
    n = n0 * n1 * N * N
    lib = load_library()
    u = torch.empty(n, dtype=torch.float64, device=img.device)
    _check(lib.hd_cnn_image_to_dofs(
        _ptr(img, torch.float32, n), _ptr(scale, torch.float64, 1), n0, n1, N,
        _ptr(u, torch.float64, n), _stream(stream)), "hd_cnn_image_to_dofs")
    return u


def vp_moments(mf, f, stream=None):
    """mass, ||f||_L2, kinetic energy, momentum (3) of f
    (vlasov_poisson/include/diagnostics.h:28-87)."""
    import torch
    out = (ctypes.c_double * 6)()
    _check(mf._lib.hd_vp_moments(mf.handle, _ptr(f, torch.float64,
                                                   mf.n_dofs()), out,
                                 _stream(stream)), "hd_vp_moments")
    return [out[i] for i in range(6)]


class Batch:
    """Independent advection operators (2+2, collapsed, any degree <= 5)
    applied / time-stepped together, one launch per apply or stage (no
    counterpart in the reference, which steps one operator per rank)."""

    def __init__(self, ops):
        ops = list(ops)
        if not 1 <= len(ops) <= BATCH_MAX:
            raise HyperdealError("a batch holds 1 to %d operators" % BATCH_MAX)
        self.ops = ops
        self._lib = ops[0]._lib
        arr = (_vp * len(ops))(*[o.handle for o in ops])
        h = _vp()
        _check(self._lib.hd_batch_create(arr, len(ops), ctypes.byref(h)),
               "hd_batch_create")
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.hd_batch_destroy(self.handle)
            self.handle = None

    def __len__(self):
        return len(self.ops)

    def _arr(self, tensors, dtype):
        if len(tensors) != len(self.ops):
            raise HyperdealError("one vector per operator")
        return (_vp * len(self.ops))(*[
            _ptr(t, dtype, o.mf.n_dofs()) for t, o in zip(tensors, self.ops)])

    def apply(self, dst, src, precision=FP64, stream=None):
        dt = _torch_dtype(precision)
        _check(self._lib.hd_batch_apply(self.handle, self._arr(src, dt),
                                        self._arr(dst, dt), precision,
                                        _stream(stream)), "hd_batch_apply")
        return dst

    def lsrk_step(self, rk_type, sol, ti_a, ti_b, dt, precision=FP64,
                  stream=None):
        import torch
        t = _torch_dtype(precision)
        _check(self._lib.hd_batch_lsrk_step(
            self.handle, rk_type.encode(), precision,
            self._arr(sol, torch.float64), self._arr(ti_a, t),
            self._arr(ti_b, t), float(dt), _stream(stream)),
            "hd_batch_lsrk_step")

    def autotune(self, src, dst, precision=FP64, stream=None):
        dt = _torch_dtype(precision)
        _check(self._lib.hd_batch_autotune(self.handle, precision,
                                           self._arr(src, dt),
                                           self._arr(dst, dt),
                                           _stream(stream)),
               "hd_batch_autotune")

    def cells_per_task(self, degree, precision=FP64):
        return _check(self._lib.hd_batch_cells_per_task(self.handle, degree,
                                                        precision),
                      "hd_batch_cells_per_task")

    def set_cells_per_task(self, degree, cells, precision=FP64):
        _check(self._lib.hd_batch_set_cells_per_task(self.handle, degree,
                                                     precision, cells),
               "hd_batch_set_cells_per_task")

    def graph(self, rk_type, sol, ti_a, ti_b, dt, precision=FP64):
        """A hipGraph of one LSRK step of every operator (fixed vectors and
        dt); Graph.launch() replays it."""
        import torch
        t = _torch_dtype(precision)
        h = _vp()
        _check(self._lib.hd_batch_graph_create(
            self.handle, rk_type.encode(), precision,
            self._arr(sol, torch.float64), self._arr(ti_a, t),
            self._arr(ti_b, t), float(dt), ctypes.byref(h)),
            "hd_batch_graph_create")
        return Graph(self, h)


class Graph:
    def __init__(self, batch, handle):
        self.batch = batch          # the batch outlives its graphs
        self._lib = batch._lib
        self.handle = handle

    def launch(self, stream=None):
        _check(self._lib.hd_graph_launch(self.handle, _stream(stream)),
               "hd_graph_launch")

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.hd_graph_destroy(self.handle)
            self.handle = None


class LowStorageRungeKuttaIntegrator:
    """Two-register LSRK (time_integrators.h:47-115): rk33, rk45, rk47,
    rk59.  The registers ti_a, ti_b hold `precision` values; the solution is
    FP64."""

    def __init__(self, rk_type="rk45"):
        self._lib = load_library()
        self.type = rk_type
        a = (ctypes.c_double * 16)()
        b = (ctypes.c_double * 16)()
        n = _check(self._lib.hd_lsrk_coefficients(rk_type.encode(), a, b),
                   "hd_lsrk_coefficients")
        self.ai = [a[i] for i in range(n - 1)]
        self.bi = [b[i] for i in range(n)]

    def n_stages(self):
        return len(self.bi)

    def perform_time_step(self, op, solution, ti_a, ti_b, current_time,
                          time_step, precision=FP64, stream=None):
        import torch
        n = op.mf.n_dofs()
        dt = _torch_dtype(precision)
        _check(self._lib.hd_lsrk_step(op.handle, self.type.encode(), precision,
                                      _ptr(solution, torch.float64, n),
                                      _ptr(ti_a, dt, n), _ptr(ti_b, dt, n),
                                      float(current_time), float(time_step),
                                      _stream(stream)),
               "hd_lsrk_step")


class VectorTools:
    """numerics/vector_tools.h: interpolate and norm_and_error."""

    @staticmethod
    def interpolate(mf, dst, function, params=None, time=0.0, stream=None):
        import torch
        p = _darr(params) if params is not None else None
        _check(mf._lib.hd_interpolate(mf.handle, function, p, float(time),
                                      _ptr(dst, torch.float64, mf.n_dofs()),
                                      _stream(stream)), "hd_interpolate")
        return dst

    @staticmethod
    def add_noise(mf, dst, seed, amplitude, stream=None):
        import torch
        _check(mf._lib.hd_add_noise(mf.handle,
                                    _ptr(dst, torch.float64, mf.n_dofs()),
                                    int(seed), float(amplitude),
                                    _stream(stream)), "hd_add_noise")
        return dst

    @staticmethod
    def norm_and_error(mf, src, function, params=None, time=0.0, stream=None):
        import torch
        out = (ctypes.c_double * 2)()
        p = _darr(params) if params is not None else None
        _check(mf._lib.hd_norm_and_error(mf.handle,
                                         _ptr(src, torch.float64, mf.n_dofs()),
                                         function, p, float(time), out,
                                         _stream(stream)), "hd_norm_and_error")
        return out[0], out[1]

    @staticmethod
    def norm_and_error_noalloc(mf, src, function, out, params=None, time=0.0,
                               stream=None):
        """norm_and_error into `out` (a float64 CUDA tensor of two values),
        stream ordered, with no allocation (hd_norm_and_error_noalloc)."""
        import torch
        p = _darr(params) if params is not None else None
        _check(mf._lib.hd_norm_and_error_noalloc(
            mf.handle, _ptr(src, torch.float64, mf.n_dofs()), function, p,
            float(time), _ptr(out, torch.float64, 2), _stream(stream)),
            "hd_norm_and_error_noalloc")
        return out

    @staticmethod
    def convert(mf, src, dst, precision, stream=None):
        import torch
        n = mf.n_dofs()
        _check(mf._lib.hd_vector_convert(mf.handle,
                                         _ptr(src, torch.float64, n),
                                         _ptr(dst, _torch_dtype(precision), n),
                                         precision, _stream(stream)),
               "hd_vector_convert")
        return dst
