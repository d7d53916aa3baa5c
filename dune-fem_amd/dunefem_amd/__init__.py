"""Python mirror of the dune-fem operator / inverse-operator interface over
the MI355X C ABI (libdunefem_amd.so).

The objects follow the reference's Python-facing contract:

* ``DofVector``         -- DoF storage of a discrete function (``dofVector``,
  ``clear``, ``assign``, ``axpy``, ``scalarProductDofs``; buffer access as in
  dune/fempy/py/discretefunction.hh:55-167);
* ``GalerkinOperator``  -- ``op(u, w)`` computes w = A u
  (dune/fempy/py/operator.hh:49,215 -> galerkin.hh:1430-1433);
* ``InverseOperator``   -- ``bind(op[, pre])``, ``inv(b, x)``, ``iterations()``,
  ``info()`` (dune/fem/solver/inverseoperatorinterface.hh:66-166);
* ``solve()`` returns the info dict of ``scheme.solve``
  (dune/fempy/py/scheme.hh:156-187).

There is no fallback: every call goes through the HIP library, and the module
raises if the library is missing or no GPU is present.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libdunefem_amd.so")

F64, F32, BF16MFMA = 0, 1, 2
OP_DG_SIPG_3D, OP_CG_Q2_2D = 1, 2
KAPPA_CONST, KAPPA_QP, KAPPA_CELL = 0, 1, 2
PRE_NONE, PRE_JACOBI, PRE_BLOCK_JACOBI, PRE_LEARNED, PRE_MULTIGRID = 0, 1, 2, 3, 4
LEARNED_OFF, LEARNED_NET, LEARNED_COARSE_ONLY = 0, 1, 2
CG, GMRES, FGMRES, BICGSTAB = 0, 1, 2, 3
ABSOLUTE, RELATIVE, REDUCTION = 0, 1, 2
SPACE_LAGRANGE, SPACE_LEGENDRE_DG = 6, 8
LINESEARCH_NONE, LINESEARCH_SIMPLE = 0, 1
FORCING_NONE, FORCING_EISENSTATWALKER = 0, 1
# NewtonFailure (dune/fem/solver/newtoninverseoperator.hh:389-400)
NEWTON_FAILURE = {0: "Success", 1: "InvalidResidual", 2: "IterationsExceeded", 3: "LinearIterationsExceeded",
                  4: "LineSearchFailed", 5: "TooManyIterations", 6: "TooManyLinearIterations",
                  7: "LinearSolverFailed"}

# every symbol declared in include/dunefem_amd.h
EXPORTS = [
    "fem_last_error", "fem_version", "fem_ctx_create", "fem_ctx_destroy", "fem_ctx_synchronize",
    "fem_ctx_mark", "fem_ctx_elapsed_ms", "fem_vec_create", "fem_vec_wrap", "fem_vec_destroy", "fem_vec_size",
    "fem_vec_precision", "fem_vec_device_ptr", "fem_vec_upload", "fem_vec_download",
    "fem_vec_upload_f32", "fem_vec_fill", "fem_vec_copy", "fem_vec_axpy", "fem_vec_scal",
    "fem_vec_dot", "fem_vec_dot_noalloc", "fem_vec_norm2sq", "fem_op_create", "fem_op_destroy", "fem_op_ndofs",
    "fem_op_set_kappa_qp", "fem_op_set_kappa_cell", "fem_op_apply", "fem_op_assemble_rhs",
    "fem_op_diagonal", "fem_op_block_diagonal", "fem_op_set_boundary_data", "fem_op_constrain", "fem_precond_create", "fem_precond_destroy", "fem_precond_apply",
    "fem_solver_default_params", "fem_solver_params_set", "fem_solver_create",
    "fem_solver_destroy", "fem_solver_bind", "fem_solver_unbind", "fem_solver_solve",
    "fem_gauss_points", "fem_batch_create", "fem_batch_destroy", "fem_batch_ndofs",
    "fem_batch_dof_offsets", "fem_batch_set_kappa_qp", "fem_batch_apply", "fem_batch_assemble_rhs",
    "fem_batch_autotune", "fem_batch_solve", "fem_batch_instance_results",
    "fem_cnn_create", "fem_cnn_destroy", "fem_cnn_nparams", "fem_cnn_set_weights", "fem_cnn_get_weights",
    "fem_cnn_forward", "fem_cnn_loss_grad", "fem_cnn_set_coefficient", "fem_cnn_collect_pairs",
    "fem_cnn_train", "fem_precond_create_learned", "fem_precond_create_learned_ex", "fem_precond_set_learned_mode", "fem_cnn_coarse_apply",
    "fem_cnn_coarse_solve", "fem_op_assemble_csr", "fem_csr_destroy", "fem_csr_rows", "fem_csr_nnz",
    "fem_csr_apply", "fem_csr_diagonal", "fem_csr_download",
    "fem_dofs_encode", "fem_dofs_decode", "fem_vec_write", "fem_vec_read",
    "fem_op_set_reaction_qp", "fem_op_set_reaction_law", "fem_op_nonlinear", "fem_op_eval_qp",
    "fem_op_jacobian_create", "fem_op_jacobian", "fem_op_set_reaction_plugin",
    "fem_newton_default_params", "fem_newton_params_set", "fem_newton_create", "fem_newton_destroy",
    "fem_newton_bind", "fem_newton_unbind", "fem_newton_solve",
]


class FemError(RuntimeError):
    pass


class OpDesc(C.Structure):
    _fields_ = [("kind", C.c_int), ("n", C.c_int * 3), ("order", C.c_int), ("beta", C.c_double),
                ("b", C.c_double * 3), ("c", C.c_double), ("kappa_mode", C.c_int),
                ("kappa_const", C.c_double), ("precision", C.c_int)]


class SolverParams(C.Structure):
    _fields_ = [("method", C.c_int), ("tolerance", C.c_double), ("errormeasure", C.c_int),
                ("maxiterations", C.c_int), ("restart", C.c_int), ("verbose", C.c_int),
                ("refinement", C.c_int), ("inner_tolerance", C.c_double),
                ("inner_maxiterations", C.c_int), ("residual_check", C.c_int),
                ("preconditioning", C.c_int), ("relaxation", C.c_double), ("precond_iterations", C.c_int),
                ("graph", C.c_int)]


class BatchInfo(C.Structure):
    _fields_ = [("count", C.c_int), ("converged", C.c_int), ("outer_iterations", C.c_int),
                ("max_inner_iterations", C.c_int), ("total_inner_iterations", C.c_int),
                ("max_true_rel_residual", C.c_double), ("time_s", C.c_double),
                ("tile_x", C.c_int), ("tile_y", C.c_int)]


class SolverInfo(C.Structure):
    _fields_ = [("converged", C.c_int), ("iterations", C.c_int), ("outer_iterations", C.c_int),
                ("true_rel_residual", C.c_double), ("time_s", C.c_double), ("fallback", C.c_int)]


class NewtonParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("maxiterations", C.c_int), ("maxlineariterations", C.c_int),
                ("linesearch", C.c_int), ("maxlinesearchiterations", C.c_int), ("forcing", C.c_int),
                ("eisenstatwalker_etamax", C.c_double), ("eisenstatwalker_gamma", C.c_double),
                ("simplified", C.c_int), ("verbose", C.c_int)]


class NewtonInfo(C.Structure):
    _fields_ = [("converged", C.c_int), ("failure", C.c_int), ("iterations", C.c_int),
                ("linear_iterations", C.c_int), ("residual", C.c_double), ("timing", C.c_double * 3)]


_lib = None
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


def load_library(path=LIBPATH):
    """Load the C ABI library (no compute calls are made here)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FemError("libdunefem_amd.so not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    sig = {
        "fem_last_error": ([], C.c_char_p),
        "fem_version": ([], C.c_int),
        "fem_ctx_create": ([C.c_int, C.POINTER(_vp)], C.c_int),
        "fem_ctx_destroy": ([_vp], C.c_int),
        "fem_ctx_synchronize": ([_vp], C.c_int),
        "fem_ctx_mark": ([_vp, C.c_int], C.c_int),
        "fem_ctx_elapsed_ms": ([_vp, C.POINTER(C.c_float)], C.c_int),
        "fem_vec_create": ([_vp, C.c_int64, C.c_int, C.POINTER(_vp)], C.c_int),
        "fem_vec_wrap": ([_vp, C.c_int64, C.c_int, _vp, C.POINTER(_vp)], C.c_int),
        "fem_vec_destroy": ([_vp], C.c_int),
        "fem_vec_size": ([_vp], C.c_int64),
        "fem_vec_precision": ([_vp], C.c_int),
        "fem_vec_device_ptr": ([_vp], _vp),
        "fem_vec_upload": ([_vp, _dp], C.c_int),
        "fem_vec_download": ([_vp, _dp], C.c_int),
        "fem_vec_upload_f32": ([_vp, C.POINTER(C.c_float)], C.c_int),
        "fem_dofs_encode": ([C.POINTER(C.c_double), C.c_int64, C.c_char_p, C.c_int, C.c_void_p, C.c_int64,
                             C.POINTER(C.c_int64)], C.c_int),
        "fem_dofs_decode": ([C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_double), C.c_int64, C.c_char_p,
                             C.c_int64], C.c_int),
        "fem_vec_write": ([_vp, C.c_char_p, C.c_char_p, C.c_int], C.c_int),
        "fem_vec_read": ([_vp, C.c_char_p, C.c_int, C.c_char_p, C.c_int64], C.c_int),
        "fem_vec_fill": ([_vp, C.c_double], C.c_int),
        "fem_vec_copy": ([_vp, _vp], C.c_int),
        "fem_vec_axpy": ([_vp, C.c_double, _vp], C.c_int),
        "fem_vec_scal": ([_vp, C.c_double], C.c_int),
        "fem_vec_dot": ([_vp, _vp, _dp], C.c_int),
        "fem_vec_dot_noalloc": ([_vp, _vp, _dp], C.c_int),
        "fem_vec_norm2sq": ([_vp, _dp], C.c_int),
        "fem_op_create": ([_vp, C.POINTER(OpDesc), C.POINTER(_vp)], C.c_int),
        "fem_op_destroy": ([_vp], C.c_int),
        "fem_op_ndofs": ([_vp], C.c_int64),
        "fem_op_set_kappa_qp": ([_vp, _dp, _dp, _dp, _dp], C.c_int),
        "fem_op_set_kappa_cell": ([_vp, _dp], C.c_int),
        "fem_op_apply": ([_vp, _vp, _vp], C.c_int),
        "fem_op_assemble_rhs": ([_vp, _dp, _vp], C.c_int),
        "fem_op_diagonal": ([_vp, _vp], C.c_int),
        "fem_op_block_diagonal": ([_vp, _vp], C.c_int),
        "fem_op_set_boundary_data": ([_vp, _dp, _dp, _dp], C.c_int),
        "fem_op_constrain": ([_vp, _vp, _vp], C.c_int),
        "fem_op_assemble_csr": ([_vp, C.POINTER(_vp)], C.c_int),
        "fem_csr_destroy": ([_vp], C.c_int),
        "fem_csr_rows": ([_vp], C.c_int64),
        "fem_csr_nnz": ([_vp], C.c_int64),
        "fem_csr_apply": ([_vp, _vp, _vp], C.c_int),
        "fem_csr_diagonal": ([_vp, _vp], C.c_int),
        "fem_csr_download": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int), _dp], C.c_int),
        "fem_precond_create": ([_vp, C.c_int, C.POINTER(_vp)], C.c_int),
        "fem_precond_destroy": ([_vp], C.c_int),
        "fem_precond_apply": ([_vp, _vp, _vp], C.c_int),
        "fem_solver_default_params": ([C.POINTER(SolverParams)], None),
        "fem_solver_params_set": ([C.POINTER(SolverParams), C.c_char_p, C.c_char_p], C.c_int),
        "fem_solver_create": ([_vp, C.POINTER(SolverParams), C.POINTER(_vp)], C.c_int),
        "fem_solver_destroy": ([_vp], C.c_int),
        "fem_solver_bind": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
        "fem_solver_unbind": ([_vp], C.c_int),
        "fem_solver_solve": ([_vp, _vp, _vp, C.POINTER(SolverInfo)], C.c_int),
        "fem_gauss_points": ([C.c_int, _dp, _dp], C.c_int),
        "fem_batch_create": ([_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_double,
                              C.POINTER(_vp)], C.c_int),
        "fem_batch_destroy": ([_vp], C.c_int),
        "fem_batch_ndofs": ([_vp], C.c_int64),
        "fem_batch_dof_offsets": ([_vp, C.POINTER(C.c_int64)], C.c_int),
        "fem_batch_set_kappa_qp": ([_vp, _dp], C.c_int),
        "fem_batch_apply": ([_vp, _vp, _vp], C.c_int),
        "fem_batch_assemble_rhs": ([_vp, _dp, _vp], C.c_int),
        "fem_batch_autotune": ([_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)], C.c_int),
        "fem_batch_solve": ([_vp, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                             C.POINTER(BatchInfo)], C.c_int),
        "fem_batch_instance_results": ([_vp, C.POINTER(C.c_int), _dp], C.c_int),
        "fem_cnn_create": ([_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(_vp)], C.c_int),
        "fem_cnn_destroy": ([_vp], C.c_int),
        "fem_cnn_nparams": ([_vp], C.c_int64),
        "fem_cnn_set_weights": ([_vp, C.POINTER(C.c_float)], C.c_int),
        "fem_cnn_get_weights": ([_vp, C.POINTER(C.c_float)], C.c_int),
        "fem_cnn_forward": ([_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
        "fem_cnn_loss_grad": ([_vp, C.POINTER(C.c_float), C.POINTER(C.c_float), _dp, C.POINTER(C.c_float)], C.c_int),
        "fem_cnn_set_coefficient": ([_vp, _dp, C.c_double], C.c_int),
        "fem_cnn_coarse_apply": ([_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
        "fem_cnn_coarse_solve": ([_vp, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float), _dp], C.c_int),
        "fem_cnn_collect_pairs": ([_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "fem_cnn_train": ([_vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_uint, _dp, _dp], C.c_int),
        "fem_precond_create_learned": ([_vp, _vp, C.c_int, C.POINTER(_vp)], C.c_int),
        "fem_precond_create_learned_ex": ([_vp, _vp, C.c_int, C.c_int, C.POINTER(_vp)], C.c_int),
        "fem_precond_set_learned_mode": ([_vp, C.c_int], C.c_int),
        "fem_op_set_reaction_qp": ([_vp, _dp], C.c_int),
        "fem_op_set_reaction_law": ([_vp, C.c_double, C.c_int], C.c_int),
        "fem_op_nonlinear": ([_vp], C.c_int),
        "fem_op_set_reaction_plugin": ([_vp, C.c_char_p], C.c_int),
        "fem_op_eval_qp": ([_vp, _vp, _vp], C.c_int),
        "fem_op_jacobian_create": ([_vp, C.POINTER(_vp)], C.c_int),
        "fem_op_jacobian": ([_vp, _vp, _vp], C.c_int),
        "fem_newton_default_params": ([C.POINTER(NewtonParams)], None),
        "fem_newton_params_set": ([C.POINTER(NewtonParams), C.c_char_p, C.c_char_p], C.c_int),
        "fem_newton_create": ([_vp, C.POINTER(NewtonParams), C.POINTER(SolverParams), C.POINTER(_vp)], C.c_int),
        "fem_newton_destroy": ([_vp], C.c_int),
        "fem_newton_bind": ([_vp, _vp, _vp], C.c_int),
        "fem_newton_unbind": ([_vp], C.c_int),
        "fem_newton_solve": ([_vp, _vp, _vp, C.POINTER(NewtonInfo)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise FemError("dunefem_amd error %d: %s" % (rc, _lib.fem_last_error().decode()))


def encode_dofs(dofs, name="", space=SPACE_LEGENDRE_DG):
    """The bytes DiscreteFunction::write puts on a binary stream for these
    DoFs (host only, no device work)."""
    L = load_library()
    d = np.ascontiguousarray(dofs, dtype=np.float64)
    n = C.c_int64()
    L.fem_dofs_encode(_dptr(d), d.size, name.encode(), int(space), None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value)
    _check(L.fem_dofs_encode(_dptr(d), d.size, name.encode(), int(space), buf, n.value, C.byref(n)))
    return buf.raw[:n.value]


def decode_dofs(data, n, space=SPACE_LEGENDRE_DG):
    """DiscreteFunction::read of n DoFs from the bytes of a binary stream:
    (dofs, name)."""
    L = load_library()
    raw = C.create_string_buffer(bytes(data), len(data))
    out = np.empty(int(n))
    name = C.create_string_buffer(4096)
    _check(L.fem_dofs_decode(raw, len(data), int(space), _dptr(out), int(n), name, len(name)))
    return out, name.value.decode()


def _dptr(a):
    return a.ctypes.data_as(_dp)


def gauss_points(m):
    """m-point Gauss-Legendre rule on [0,1] from the library (host-only call)."""
    L = load_library()
    x = np.zeros(m)
    w = np.zeros(m)
    _check(L.fem_gauss_points(m, _dptr(x), _dptr(w)))
    return x, w


class Context:
    """Device context: one GPU, one stream."""

    def __init__(self, device=0):
        L = load_library()
        h = _vp()
        _check(L.fem_ctx_create(device, C.byref(h)))
        self.h = h

    def synchronize(self):
        _check(_lib.fem_ctx_synchronize(self.h))

    def mark(self, which):
        _check(_lib.fem_ctx_mark(self.h, which))

    def elapsed_ms(self):
        ms = C.c_float()
        _check(_lib.fem_ctx_elapsed_ms(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            _check(_lib.fem_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DofVector:
    """Device DoF vector (the dofVector of a discrete function)."""

    def __init__(self, ctx, n, precision=F64, data=None, device_ptr=None):
        """A new device vector, or (device_ptr given) a view of n DoFs the
        caller already holds in device memory: no copy, the memory is not
        released with the view (AdaptiveDiscreteFunction over existing DoF
        storage)."""
        self.ctx = ctx
        h = _vp()
        if device_ptr is None:
            _check(_lib.fem_vec_create(ctx.h, int(n), precision, C.byref(h)))
        else:
            _check(_lib.fem_vec_wrap(ctx.h, int(n), precision, _vp(int(device_ptr)), C.byref(h)))
        self.h = h
        self.n = int(n)
        self.precision = precision
        if data is not None:
            self.upload(data)

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        assert a.size == self.n
        _check(_lib.fem_vec_upload(self.h, _dptr(a)))

    def download(self):
        out = np.empty(self.n)
        _check(_lib.fem_vec_download(self.h, _dptr(out)))
        return out

    def numpy(self):
        return self.download()

    def write(self, path, name="", space=SPACE_LEGENDRE_DG):
        """DiscreteFunction::write through a binary file stream
        (function/common/discretefunction_inline.hh:270-305)."""
        _check(_lib.fem_vec_write(self.h, os.fsencode(path), name.encode(), int(space)))

    def read(self, path, space=SPACE_LEGENDRE_DG):
        """DiscreteFunction::read (discretefunction_inline.hh:213-267); returns
        the stored name."""
        buf = C.create_string_buffer(4096)
        _check(_lib.fem_vec_read(self.h, os.fsencode(path), int(space), buf, len(buf)))
        return buf.value.decode()

    def clear(self):
        _check(_lib.fem_vec_fill(self.h, 0.0))

    def fill(self, v):
        _check(_lib.fem_vec_fill(self.h, float(v)))

    def assign(self, other):
        _check(_lib.fem_vec_copy(self.h, other.h))

    def axpy(self, a, x):
        _check(_lib.fem_vec_axpy(self.h, float(a), x.h))

    def scale(self, a):
        _check(_lib.fem_vec_scal(self.h, float(a)))

    def scalarProductDofs(self, other):
        out = C.c_double()
        _check(_lib.fem_vec_dot(self.h, other.h, C.byref(out)))
        return out.value

    def normSquaredDofs(self):
        out = C.c_double()
        _check(_lib.fem_vec_norm2sq(self.h, C.byref(out)))
        return out.value

    def device_ptr(self):
        return _lib.fem_vec_device_ptr(self.h)

    def __del__(self):
        try:
            if self.h:
                _lib.fem_vec_destroy(self.h)
                self.h = None
        except Exception:
            pass


class GalerkinOperator:
    """Matrix-free Galerkin operator; ``op(u, w)`` overwrites w with A u."""

    def __init__(self, ctx, kind, n, order, beta=0.0, b=(0.0, 0.0, 0.0), c=0.0,
                 kappa_mode=KAPPA_CONST, kappa_const=1.0, precision=F64):
        self.ctx = ctx
        d = OpDesc()
        d.kind = kind
        nn = list(n) + [1] * (3 - len(n))
        d.n[:] = nn
        d.order = order
        d.beta = beta
        d.b[:] = list(b)
        d.c = c
        d.kappa_mode = kappa_mode
        d.kappa_const = kappa_const
        d.precision = precision
        h = _vp()
        _check(_lib.fem_op_create(ctx.h, C.byref(d), C.byref(h)))
        self.h = h
        self.desc = d
        self.precision = precision
        self.vec_precision = F64 if precision == F64 else F32
        self.N = _lib.fem_op_ndofs(h)

    @property
    def nqp(self):
        """number of volume quadrature points: E (p+1)^3 (DG), E 9 (Q2 in 2D)"""
        n = self.desc.n
        if self.desc.kind == OP_CG_Q2_2D:
            return n[0] * n[1] * 9
        return n[0] * n[1] * n[2] * (self.desc.order + 1) ** 3

    def set_kappa_qp(self, kvol, kfaces=None):
        kvol = np.ascontiguousarray(kvol, dtype=np.float64)
        if kfaces is None:
            _check(_lib.fem_op_set_kappa_qp(self.h, _dptr(kvol), None, None, None))
        else:
            kf = [np.ascontiguousarray(k, dtype=np.float64) for k in kfaces]
            _check(_lib.fem_op_set_kappa_qp(self.h, _dptr(kvol), _dptr(kf[0]), _dptr(kf[1]), _dptr(kf[2])))

    def set_kappa_cell(self, kcell):
        kcell = np.ascontiguousarray(kcell, dtype=np.float64)
        _check(_lib.fem_op_set_kappa_cell(self.h, _dptr(kcell)))

    def vector(self, data=None):
        return DofVector(self.ctx, self.N, self.vec_precision, data)

    def __call__(self, u, w):
        _check(_lib.fem_op_apply(self.h, u.h, w.h))

    def rhs(self, fq, F=None):
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        F = F if F is not None else self.vector()
        _check(_lib.fem_op_assemble_rhs(self.h, _dptr(fq), F.h))
        return F

    def diagonal(self, d=None):
        d = d if d is not None else self.vector()
        _check(_lib.fem_op_diagonal(self.h, d.h))
        return d

    def set_boundary_data(self, g):
        """u = g on the boundary.  DG: three arrays, g at the face quadrature
        points [faces_d][m^2] (kface layout); CG Q2: the nodal values [N]."""
        if isinstance(g, (list, tuple)):
            gs = [np.ascontiguousarray(x, dtype=np.float64) for x in g]
            _check(_lib.fem_op_set_boundary_data(self.h, _dptr(gs[0]), _dptr(gs[1]), _dptr(gs[2])))
        else:
            g = np.ascontiguousarray(g, dtype=np.float64)
            _check(_lib.fem_op_set_boundary_data(self.h, _dptr(g), None, None))

    def constrain(self, x, add=None):
        """x_D = g_D (+ add_D) on the Dirichlet DoFs (CG Q2; no effect for DG)."""
        _check(_lib.fem_op_constrain(self.h, x.h, add.h if add is not None else None))

    def block_diagonal(self):
        """Element blocks [E][nb][nb] (FP64, host array) of a DG operator."""
        nb = (self.desc.order + 1) ** 3
        E = self.N // nb
        B = DofVector(self.ctx, E * nb * nb, F64)
        _check(_lib.fem_op_block_diagonal(self.h, B.h))
        return B.download().reshape(E, nb, nb)

    def set_reaction_qp(self, cvol):
        """pointwise reaction coefficient at the volume quadrature points
        [E][m^3] (DG) or [E][9] (Q2), x fastest, added to c; None clears"""
        if cvol is None:
            _check(_lib.fem_op_set_reaction_qp(self.h, None))
        else:
            cvol = np.ascontiguousarray(cvol, dtype=np.float64)
            assert cvol.size == self.nqp, "one value per volume quadrature point"
            _check(_lib.fem_op_set_reaction_qp(self.h, _dptr(cvol)))

    def set_reaction_law(self, coef, pw):
        """the nonlinear reaction coef u^pw u (pw in 1..4; coef = 0 clears):
        the operator becomes S[u] = A u + int coef u^(pw+1) v"""
        _check(_lib.fem_op_set_reaction_law(self.h, float(coef), int(pw)))

    def set_reaction(self, r, build=True):
        """a pointwise reaction r(u, x): the volume term int r(u_h, x) v with
        r a sympy expression (or string) of u, x, y, z, compiled to HIP by
        dunefem_amd.forms (cached; build=False requires the compiled library
        to exist); None removes it.  FP64 DG and Q2 Lagrange operators (z = 0
        in 2D)."""
        if r is None:
            _check(_lib.fem_op_set_reaction_plugin(self.h, None))
            self.reaction = None
            return
        from . import forms
        path = forms.compile_reaction(r, build=build)
        _check(_lib.fem_op_set_reaction_plugin(self.h, path.encode()))
        self.reaction = forms.as_expr(r)

    def qp_coordinates(self):
        """x, y, z of the volume quadrature points [E][m^3] (DG) or [E][9]
        (Q2 in 2D, z = 0)"""
        from . import forms
        m = self.desc.order + 1
        if self.desc.kind == OP_CG_Q2_2D:
            return forms.qp_coordinates_2d(tuple(self.desc.n[:2]), m, gauss_points(m)[0])
        return forms.qp_coordinates(tuple(self.desc.n), m, gauss_points(m)[0])

    def l2_error(self, u, g=0):
        """||u_h - g||_L2 by the operator's volume quadrature (u_h at the
        points from eval_qp, g an expression of x, y, z or a number): the
        discrete counterpart of the reference's integrate(grid, (u_h - g)^2)
        used in convergence studies"""
        from . import forms
        m = self.desc.order + 1
        _, w = gauss_points(m)
        X, Y, Z = self.qp_coordinates()
        gq, _ = forms.evaluate(g, np.zeros_like(X), X, Y, Z)
        uq = self.eval_qp(u).download().reshape(X.shape)
        n = self.desc.n
        if self.desc.kind == OP_CG_Q2_2D:
            W = (w[None, :] * w[:, None]).ravel() / (n[0] * n[1])                       # [qy][qx]
        else:
            W = (w[None, None, :] * w[None, :, None] * w[:, None, None]).ravel() / (n[0] * n[1] * n[2])
        return float(np.sqrt(np.sum(W[None, :] * (uq - gq) ** 2)))

    def rhs_expr(self, f, F=None):
        """load vector of a source f(x, y, z) (a sympy expression or string
        of x, y, z), evaluated at the volume quadrature points"""
        from . import forms
        X, Y, Z = self.qp_coordinates()
        fq, _ = forms.evaluate(f, np.zeros_like(X), X, Y, Z)
        return self.rhs(fq, F)

    def nonlinear(self):
        """Operator::nonlinear() (dune/fem/operator/common/operator.hh:64)"""
        r = _lib.fem_op_nonlinear(self.h)
        if r < 0:
            _check(r)
        return bool(r)

    def eval_qp(self, u, out=None):
        """u at the volume quadrature points [E][m^dim] (x fastest)"""
        out = out if out is not None else DofVector(self.ctx, self.nqp, self.vec_precision)
        _check(_lib.fem_op_eval_qp(self.h, u.h, out.h))
        return out

    def jacobian(self, u, jop=None):
        """DifferentiableOperator::jacobian(u, jOp): the linear operator
        S'(u), matrix-free (a GalerkinOperator; pass jop to refill one)"""
        if jop is None:
            h = _vp()
            _check(_lib.fem_op_jacobian_create(self.h, C.byref(h)))
            jop = GalerkinOperator.__new__(GalerkinOperator)
            jop.ctx, jop.h, jop.desc = self.ctx, h, self.desc
            jop.precision, jop.vec_precision, jop.N = self.precision, self.vec_precision, self.N
        _check(_lib.fem_op_jacobian(self.h, u.h, jop.h))
        return jop

    def __del__(self):
        try:
            if self.h:
                _lib.fem_op_destroy(self.h)
                self.h = None
        except Exception:
            pass


class SparseRowMatrix:
    """The operator's Jacobian assembled on the device in CSR form
    (DifferentiableGalerkinOperator::jacobian into SparseRowMatrix):
    ``A(x, y)`` is y = A x, ``diagonal()`` extractDiagonal, ``to_scipy()`` a
    host copy."""

    def __init__(self, op):
        self.op = op
        h = _vp()
        _check(_lib.fem_op_assemble_csr(op.h, C.byref(h)))
        self.h = h
        self.rows = _lib.fem_csr_rows(h)
        self.nnz = _lib.fem_csr_nnz(h)

    def __call__(self, x, y):
        _check(_lib.fem_csr_apply(self.h, x.h, y.h))

    def diagonal(self):
        d = DofVector(self.op.ctx, self.rows, F64)
        _check(_lib.fem_csr_diagonal(self.h, d.h))
        return d

    def to_scipy(self):
        import scipy.sparse as sp
        rp = np.zeros(self.rows + 1, dtype=np.int64)
        col = np.zeros(self.nnz, dtype=np.int32)
        val = np.zeros(self.nnz)
        _check(_lib.fem_csr_download(self.h, rp.ctypes.data_as(C.POINTER(C.c_int64)),
                                     col.ctypes.data_as(C.POINTER(C.c_int)), _dptr(val)))
        return sp.csr_matrix((val, col, rp), shape=(self.rows, self.rows))

    def __del__(self):
        try:
            if self.h:
                _lib.fem_csr_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Preconditioner:
    def __init__(self, op, kind):
        self.op = op
        h = _vp()
        _check(_lib.fem_precond_create(op.h, kind, C.byref(h)))
        self.h = h

    def __call__(self, x, y):
        _check(_lib.fem_precond_apply(self.h, x.h, y.h))

    def __del__(self):
        try:
            if self.h:
                _lib.fem_precond_destroy(self.h)
                self.h = None
        except Exception:
            pass


def solver_parameters(**kw):
    p = SolverParams()
    load_library().fem_solver_default_params(C.byref(p))
    for k, v in kw.items():
        if k in ("method", "errormeasure", "maxiterations", "restart", "verbose", "refinement",
                 "inner_maxiterations", "residual_check", "preconditioning", "precond_iterations", "graph"):
            setattr(p, k, int(v))
        elif k in ("tolerance", "inner_tolerance", "relaxation"):
            setattr(p, k, float(v))
        else:
            _check(_lib.fem_solver_params_set(C.byref(p), k.encode(), str(v).encode()))
    return p


class InverseOperator:
    """Krylov inverse operator (CgInverseOperator / GmresInverseOperator / FGMRES)."""

    def __init__(self, ctx, parameters=None, **kw):
        self.ctx = ctx
        p = parameters if parameters is not None else solver_parameters(**kw)
        self.params = p
        h = _vp()
        _check(_lib.fem_solver_create(ctx.h, C.byref(p), C.byref(h)))
        self.h = h
        self._info = SolverInfo()
        self._bound = None

    def bind(self, op, pre=None, inner_op=None, inner_pre=None):
        _check(_lib.fem_solver_bind(self.h, op.h, pre.h if pre else None,
                                    inner_op.h if inner_op else None,
                                    inner_pre.h if inner_pre else None))
        self._bound = (op, pre, inner_op, inner_pre)

    def unbind(self):
        _check(_lib.fem_solver_unbind(self.h))
        self._bound = None

    def __call__(self, b, x):
        _check(_lib.fem_solver_solve(self.h, b.h, x.h, C.byref(self._info)))

    def iterations(self):
        return self._info.iterations

    def info(self):
        i = self._info
        return {"converged": bool(i.converged), "iterations": i.iterations,
                "linear_iterations": i.iterations, "outer_iterations": i.outer_iterations,
                "true_rel_residual": i.true_rel_residual, "timing": [i.time_s],
                "fallback": bool(i.fallback)}

    def __del__(self):
        try:
            if self.h:
                _lib.fem_solver_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Batch:
    """Many independent continuous-Q2 problems solved together (one launch per step)."""

    def __init__(self, ctx, sides, c=0.0):
        self.ctx = ctx
        nx = np.array([s[0] for s in sides], dtype=np.int32)
        ny = np.array([s[1] for s in sides], dtype=np.int32)
        h = _vp()
        _check(_lib.fem_batch_create(ctx.h, len(sides), nx.ctypes.data_as(C.POINTER(C.c_int)),
                                     ny.ctypes.data_as(C.POINTER(C.c_int)), c, C.byref(h)))
        self.h = h
        self.count = len(sides)
        self.N = _lib.fem_batch_ndofs(h)
        off = np.zeros(self.count + 1, dtype=np.int64)
        _check(_lib.fem_batch_dof_offsets(h, off.ctypes.data_as(C.POINTER(C.c_int64))))
        self.offsets = off

    def set_kappa_qp(self, kvol):
        kvol = np.ascontiguousarray(kvol, dtype=np.float64)
        _check(_lib.fem_batch_set_kappa_qp(self.h, _dptr(kvol)))

    def vector(self, data=None, precision=F64):
        return DofVector(self.ctx, self.N, precision, data)

    def __call__(self, u, w):
        _check(_lib.fem_batch_apply(self.h, u.h, w.h))

    def rhs(self, fq):
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        F = self.vector()
        _check(_lib.fem_batch_assemble_rhs(self.h, _dptr(fq), F.h))
        return F

    def autotune(self):
        tx, ty = C.c_int(), C.c_int()
        ms = (C.c_float * 5)()
        _check(_lib.fem_batch_autotune(self.h, C.byref(tx), C.byref(ty), ms))
        return (tx.value, ty.value), list(ms)

    def solve(self, b, x, tol=1e-8, inner_tol=1e-3, inner_maxit=5000, max_outer=20, use_graph=True):
        info = BatchInfo()
        _check(_lib.fem_batch_solve(self.h, b.h, x.h, tol, inner_tol, inner_maxit, max_outer,
                                    1 if use_graph else 0, C.byref(info)))
        its = np.zeros(self.count, dtype=np.int32)
        rr = np.zeros(self.count)
        _check(_lib.fem_batch_instance_results(self.h, its.ctypes.data_as(C.POINTER(C.c_int)), _dptr(rr)))
        return {f: getattr(info, f) for f, _ in BatchInfo._fields_} | {"iterations": its, "rel_residual": rr}

    def __del__(self):
        try:
            if self.h:
                _lib.fem_batch_destroy(self.h)
                self.h = None
        except Exception:
            pass


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class CNN:
    """Coarse-correction network on an element grid (BF16 MFMA kernels)."""

    LAYERS = ((2, 32), (32, 32), (32, 32), (32, 1))

    def __init__(self, ctx, n, dilations=(1, 2, 4, 1)):
        self.ctx = ctx
        self.n = tuple(int(v) for v in n)
        d = (C.c_int * 4)(*dilations)
        h = _vp()
        _check(_lib.fem_cnn_create(ctx.h, self.n[0], self.n[1], self.n[2], d, C.byref(h)))
        self.h = h
        self.dilations = tuple(dilations)
        self.nparams = _lib.fem_cnn_nparams(h)
        self.V = self.n[0] * self.n[1] * self.n[2]

    def split(self, flat):
        out, off = [], 0
        for ci, co in self.LAYERS:
            out.append(flat[off:off + 27 * ci * co].reshape(27, ci, co))
            off += 27 * ci * co
        return out

    def get_weights(self):
        w = np.zeros(self.nparams, dtype=np.float32)
        _check(_lib.fem_cnn_get_weights(self.h, _fptr(w)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32).ravel()
        assert w.size == self.nparams
        _check(_lib.fem_cnn_set_weights(self.h, _fptr(w)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.V, 2)
        y = np.zeros(self.V, dtype=np.float32)
        _check(_lib.fem_cnn_forward(self.h, _fptr(x), _fptr(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(self.V, 2)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(self.V)
        loss = C.c_double()
        g = np.zeros(self.nparams, dtype=np.float32)
        _check(_lib.fem_cnn_loss_grad(self.h, _fptr(x), _fptr(t), C.byref(loss), _fptr(g)))
        return loss.value, g

    def set_coefficient(self, kcell, beta):
        kcell = np.ascontiguousarray(kcell, dtype=np.float64)
        _check(_lib.fem_cnn_set_coefficient(self.h, _dptr(kcell), float(beta)))

    def coarse_apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32).ravel()
        y = np.zeros(self.V, dtype=np.float32)
        _check(_lib.fem_cnn_coarse_apply(self.h, _fptr(x), _fptr(y)))
        return y

    def coarse_solve(self, b, steps):
        """`steps` Jacobi-CG steps on A_c x = b from x = 0; returns x and the
        final and initial (r, D^-1 r)."""
        b = np.ascontiguousarray(b, dtype=np.float32).ravel()
        x = np.zeros(self.V, dtype=np.float32)
        rz = np.zeros(2)
        _check(_lib.fem_cnn_coarse_solve(self.h, _fptr(b), int(steps), _fptr(x), rz.ctypes.data_as(_dp)))
        return x, rz[0], rz[1]

    def collect_pairs(self, op64, bj64, b, max_pairs, every=10):
        n = C.c_int()
        _check(_lib.fem_cnn_collect_pairs(self.h, op64.h, bj64.h, b.h, max_pairs, every, C.byref(n)))
        return n.value

    def train(self, epochs=10, batch=8, lr=1e-3, val_frac=0.1, seed=0):
        tl, vr = C.c_double(), C.c_double()
        _check(_lib.fem_cnn_train(self.h, epochs, batch, lr, val_frac, seed, C.byref(tl), C.byref(vr)))
        return tl.value, vr.value

    def __del__(self):
        try:
            if self.h:
                _lib.fem_cnn_destroy(self.h)
                self.h = None
        except Exception:
            pass


class LearnedPreconditioner(Preconditioner):
    """Two-level preconditioner (FGMRES only): coarse solve on element means
    started from the network's output and finished by ``coarse_steps``
    Jacobi-CG steps, energy-optimal coarse step, and a smoother: the exact
    block Jacobi on what the coarse step leaves (``smoother=PRE_BLOCK_JACOBI``)
    or a multigrid V-cycle before the coarse step (``PRE_MULTIGRID``).
    ``mode``: LEARNED_NET, LEARNED_COARSE_ONLY (coarse CG from zero, i.e.
    without the network) or LEARNED_OFF (the smoother alone)."""

    def __init__(self, op, cnn, coarse_steps=5, mode=LEARNED_NET, smoother=PRE_BLOCK_JACOBI):
        self.op = op
        self.cnn = cnn
        h = _vp()
        _check(_lib.fem_precond_create_learned_ex(op.h, cnn.h, int(coarse_steps), int(smoother), C.byref(h)))
        self.h = h
        self.set_mode(mode)

    def set_mode(self, mode):
        _check(_lib.fem_precond_set_learned_mode(self.h, int(mode)))


def _is_newton_key(k):
    """a fem.solver.nonlinear.* (or newton.*) key that is not a linear-solver key"""
    k = k[len("fem.solver."):] if k.startswith("fem.solver.") else k
    for pre in ("nonlinear.", "newton."):
        if k.startswith(pre):
            rest = k[len(pre):]
            return not rest.startswith("linear.") or rest == "linear.tolerance.strategy"
    return False


def newton_parameters(parameters=None, **kw):
    """NewtonParameter (dune/fem/solver/newtoninverseoperator.hh:101-380) from
    fem.solver.nonlinear.* keys or the fields of NewtonParams"""
    p = NewtonParams()
    load_library().fem_newton_default_params(C.byref(p))
    fields = {f for f, _ in NewtonParams._fields_}
    for k, v in dict(parameters or {}, **kw).items():
        if k in fields:
            setattr(p, k, type(getattr(p, k))(v))
        else:
            _check(_lib.fem_newton_params_set(C.byref(p), k.encode(), str(v).encode()))
    return p


class NewtonInverseOperator:
    """NewtonInverseOperator (dune/fem/solver/newtoninverseoperator.hh:423-802):
    ``bind(op[, pre])``, ``inv(u, w)`` solves S[w] = u (u None: zero) from the
    initial guess w, ``iterations()``, ``linearIterations()``, ``info()``.
    ``parameters`` holds reference keys: fem.solver.nonlinear.* for Newton,
    the others (fem.solver.*, fem.solver.nonlinear.linear.*) for its linear
    solver; ``linear`` (a dict of solver_parameters fields) adds to those."""

    def __init__(self, ctx, parameters=None, linear=None):
        self.ctx = ctx
        parameters = dict(parameters or {})
        nl = {k: v for k, v in parameters.items() if _is_newton_key(k)}
        lin = {k: v for k, v in parameters.items() if not _is_newton_key(k)}
        lin.update(linear or {})
        self.params = newton_parameters(nl)
        if self.params.forcing == FORCING_EISENSTATWALKER and not any(
                k in ("errormeasure",) or k.endswith(".errormeasure") for k in lin):
            # the reference's default linear error measure under Eisenstat-Walker
            # forcing (newtoninverseoperator.hh:175-177)
            lin["errormeasure"] = REDUCTION
        self.linear_params = solver_parameters(**lin)
        h = _vp()
        _check(_lib.fem_newton_create(ctx.h, C.byref(self.params), C.byref(self.linear_params), C.byref(h)))
        self.h = h
        self._info = NewtonInfo()
        self._bound = None

    def bind(self, op, pre=None):
        _check(_lib.fem_newton_bind(self.h, op.h, pre.h if pre is not None else None))
        self._bound = (op, pre)

    def unbind(self):
        _check(_lib.fem_newton_unbind(self.h))
        self._bound = None

    def __call__(self, u, w):
        _check(_lib.fem_newton_solve(self.h, u.h if u is not None else None, w.h, C.byref(self._info)))

    def iterations(self):
        return self._info.iterations

    def linearIterations(self):
        return self._info.linear_iterations

    def info(self):
        i = self._info
        return {"converged": bool(i.converged), "iterations": i.iterations,
                "linear_iterations": i.linear_iterations, "timing": list(i.timing),
                "failure": NEWTON_FAILURE.get(i.failure, str(i.failure)), "residual": i.residual}

    def __del__(self):
        try:
            if self.h:
                _lib.fem_newton_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Scheme:
    """The scheme-level solve of the reference, FemScheme::solve/_solve
    (dune/fem/schemes/femscheme.hh:195-253), for the operators of this
    library: S[u] = A u + N(u) with the affine part F, R(u) = S[u] - F, the
    boundary data inside F (and, for the continuous Q2 operator, the
    Dirichlet rows u_D - g_D).

    ``solve(target, rhs=None)``: setConstraints (u_D = g_D) and
    addConstraints (u_D += rhs_D) on the target, then
    NewtonInverseOperator::operator() (newtoninverseoperator.hh:688-802) on
    S[u] = F + rhs: one step for a linear operator (no nonlinear iteration
    counted, :786-790), Newton iterations with matrix-free Jacobians for a
    nonlinear one.  Returns the info dict of scheme.solve
    (dune/fempy/py/scheme.hh:156-187): converged, iterations (nonlinear),
    linear_iterations, timing [total, jacobian, solve].

    ``parameters``: reference keys (fem.solver.*, fem.solver.nonlinear.*);
    keyword arguments: fields of solver_parameters for the linear solver."""

    def __init__(self, op, F, preconditioner=None, parameters=None, **solver_parameters):
        self.op, self.F = op, F
        self.pre = preconditioner
        self.inv = NewtonInverseOperator(op.ctx, parameters, linear=solver_parameters)

    def __call__(self, u, w):
        """w = R(u) = S[u] - F"""
        self.op(u, w)
        w.axpy(-1.0, self.F)

    def solve(self, target, rhs=None):
        op = self.op
        op.constrain(target, rhs)                 # u_D = g_D + rhs_D
        u = self.F
        if rhs is not None:
            u = op.vector()
            u.assign(self.F)
            u.axpy(1.0, rhs)                      # S[w] = F + rhs
        self.inv.bind(op, self.pre)
        try:
            self.inv(u, target)
        finally:
            self.inv.unbind()
        i = self.inv.info()
        return {"converged": i["converged"], "iterations": i["iterations"],
                "linear_iterations": i["linear_iterations"], "timing": i["timing"]}


def galerkin(form, space, solver="cg", parameters=None, build=True):
    """dune.fem.scheme.galerkin for the integrands of dunefem_amd.forms.Form
    on a dunefem_amd.forms.Space (see forms.galerkin)"""
    from .forms import galerkin as _galerkin
    return _galerkin(form, space, solver, parameters, build)
