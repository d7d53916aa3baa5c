"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker,
never as the measured or shipped path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KAPPA_CONST, KAPPA_QP, KAPPA_CELL = 0, 1, 2

_dp = C.POINTER(C.c_double)
_APPLY = C.CFUNCTYPE(None, C.c_void_p, _dp, _dp)


class DG(C.Structure):
    _fields_ = [("n", C.c_int * 3), ("p", C.c_int), ("beta", C.c_double),
                ("b", C.c_double * 3), ("c", C.c_double), ("kappa_mode", C.c_int),
                ("kappa_const", C.c_double), ("kvol", _dp), ("kface", _dp * 3),
                ("kcell", _dp), ("cvol", _dp)]


class CG(C.Structure):
    _fields_ = [("n", C.c_int * 2), ("c", C.c_double), ("kappa_mode", C.c_int),
                ("kappa_const", C.c_double), ("kvol", _dp), ("cvol", _dp)]


class Diag(C.Structure):
    _fields_ = [("n", C.c_int64), ("d", _dp)]


class Block(C.Structure):
    _fields_ = [("nblocks", C.c_int64), ("nb", C.c_int), ("inv", _dp)]


_LIBNAME = "liboracle.so"


def select_native():
    """Build (make native) and use the -march=native oracle of this host
    instead of the portable one; returns the library name in use.  Must run
    before the first oracle call."""
    global _LIBNAME
    import subprocess
    if _LIB is not None:
        return _LIBNAME
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True, capture_output=True, timeout=120)
        _LIBNAME = "liboracle_native.so"
    except Exception:
        pass
    return _LIBNAME


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _LIBNAME)
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(path)
        L.ora_gauss.argtypes = [C.c_int, _dp, _dp]
        L.ora_gauss_m_for_order.argtypes = [C.c_int]
        for f in ("ora_legendre", "ora_legendre_d", "ora_lagrange2", "ora_lagrange2_d"):
            getattr(L, f).argtypes = [C.c_int, C.c_double]
            getattr(L, f).restype = C.c_double
        L.ora_dg_ndofs.argtypes = [C.POINTER(DG)]
        L.ora_dg_ndofs.restype = C.c_int64
        L.ora_dg_apply.argtypes = [C.POINTER(DG), _dp, _dp, C.c_int]
        L.ora_dg_apply_sf.argtypes = [C.POINTER(DG), _dp, _dp, C.c_int]
        L.ora_dg_rhs.argtypes = [C.POINTER(DG), _dp, _dp]
        L.ora_dg_block_diag.argtypes = [C.POINTER(DG), _dp]
        L.ora_dg_rhs_bnd.argtypes = [C.POINTER(DG), C.POINTER(_dp), _dp]
        L.ora_cg_ndofs.argtypes = [C.POINTER(CG)]
        L.ora_cg_ndofs.restype = C.c_int64
        L.ora_cg_apply.argtypes = [C.POINTER(CG), _dp, _dp]
        L.ora_cg_rhs.argtypes = [C.POINTER(CG), _dp, _dp]
        L.ora_cg_eval_qp.argtypes = [C.POINTER(CG), _dp, _dp]
        L.ora_cg_diag.argtypes = [C.POINTER(CG), _dp]
        for f in ("ora_cg", "ora_bicgstab"):
            getattr(L, f).argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      _dp, _dp, C.c_double, C.c_int, C.c_int]
        for f in ("ora_gmres", "ora_fgmres"):
            getattr(L, f).argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]
        _LIB = L
    return _LIB


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def gauss(m):
    x = np.zeros(m); w = np.zeros(m)
    assert lib().ora_gauss(m, _ptr(x), _ptr(w)) == 0
    return x, w


def gauss_m_for_order(order):
    return lib().ora_gauss_m_for_order(order)


def legendre(n, x):
    return lib().ora_legendre(n, float(x))


def legendre_d(n, x):
    return lib().ora_legendre_d(n, float(x))


def lagrange2(i, x):
    return lib().ora_lagrange2(i, float(x))


def lagrange2_d(i, x):
    return lib().ora_lagrange2_d(i, float(x))


class DGOperator:
    """SIPG DG operator of the oracle.  kappa: scalar, ('qp', kvol, (kx, ky, kz))
    or ('cell', kcell)."""

    def __init__(self, n, p, beta, b=(0.0, 0.0, 0.0), c=0.0, kappa=1.0):
        self.n, self.p = tuple(int(v) for v in n), int(p)
        s = DG()
        s.n[:] = list(self.n)
        s.p = self.p
        s.beta = beta
        s.b[:] = list(b)
        s.c = c
        self._keep = []
        if isinstance(kappa, tuple) and kappa[0] == "qp":
            s.kappa_mode = KAPPA_QP
            kv = np.ascontiguousarray(kappa[1], dtype=np.float64)
            kf = [np.ascontiguousarray(k, dtype=np.float64) for k in kappa[2]]
            self._keep += [kv] + kf
            s.kvol = _ptr(kv)
            for d in range(3):
                s.kface[d] = _ptr(kf[d])
        elif isinstance(kappa, tuple) and kappa[0] == "cell":
            s.kappa_mode = KAPPA_CELL
            kc = np.ascontiguousarray(kappa[1], dtype=np.float64)
            self._keep.append(kc)
            s.kcell = _ptr(kc)
        else:
            s.kappa_mode = KAPPA_CONST
            s.kappa_const = float(kappa)
        self.s = s
        self.N = lib().ora_dg_ndofs(C.byref(s))

    def apply(self, u, nthreads=1):
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.empty(self.N)
        lib().ora_dg_apply(C.byref(self.s), _ptr(u), _ptr(w), nthreads)
        return w

    def set_reaction_qp(self, cvol):
        """pointwise reaction coefficient at the volume quadrature points
        [E][m^3] (x fastest), added to c; None to clear"""
        if cvol is None:
            self.s.cvol = None
            return
        cv = np.ascontiguousarray(cvol, dtype=np.float64).ravel()
        m = self.p + 1
        assert cv.size == self.N // m ** 3 * m ** 3
        self._cvol = cv
        self.s.cvol = _ptr(cv)

    def eval_qp(self, u):
        """u at the volume quadrature points [E][m^3] (x fastest), by the
        dense basis tables (numpy: phi_a(x) phi_b(y) phi_c(z))"""
        m = self.p + 1
        x, _ = gauss(m)
        B = np.array([[legendre(a, xq) for xq in x] for a in range(m)])   # [mode][point]
        E = self.N // m ** 3
        U = np.asarray(u, dtype=np.float64).reshape(E, m, m, m)   # [e][a][b][c]
        # uq[e][qz][qy][qx] = sum_abc U[a][b][c] B[a][qx] B[b][qy] B[c][qz]
        return np.einsum("eabc,ax,by,cz->ezyx", U, B, B, B).reshape(E, m ** 3)

    def apply_sf(self, u, nthreads=1):
        """the same apply evaluated by sum factorisation (timing baseline)"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.empty(self.N)
        lib().ora_dg_apply_sf(C.byref(self.s), _ptr(u), _ptr(w), nthreads)
        return w

    def rhs(self, fq):
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        F = np.empty(self.N)
        lib().ora_dg_rhs(C.byref(self.s), _ptr(fq), _ptr(F))
        return F

    def rhs_bnd(self, gfaces, F=None):
        """F (+)= boundary-data part of the load vector for u = g on the
        boundary; gfaces: three arrays [faces_d][m^2] (kface layout)."""
        F = np.zeros(self.N) if F is None else np.array(F, dtype=np.float64)
        gs = [np.ascontiguousarray(g, dtype=np.float64) for g in gfaces]
        arr = (_dp * 3)(*[_ptr(g) for g in gs])
        lib().ora_dg_rhs_bnd(C.byref(self.s), arr, _ptr(F))
        return F

    def block_diag(self):
        k = self.p + 1
        nb = k ** 3
        E = self.N // nb
        B = np.empty((E, nb, nb))
        lib().ora_dg_block_diag(C.byref(self.s), _ptr(B))
        return B

    def callback(self, nthreads=1):
        f = _APPLY(lambda ctx, x, y: lib().ora_dg_apply(C.byref(self.s), x, y, nthreads))
        self._keep.append(f)
        return f


class CGOperator:
    """Continuous Q2 operator (2D) of the oracle with strong zero Dirichlet."""

    def __init__(self, n, c=0.0, kappa=1.0):
        self.n = tuple(int(v) for v in n)
        s = CG()
        s.n[:] = list(self.n)
        s.c = c
        self._keep = []
        if isinstance(kappa, np.ndarray):
            s.kappa_mode = KAPPA_QP
            kv = np.ascontiguousarray(kappa, dtype=np.float64)
            self._keep.append(kv)
            s.kvol = _ptr(kv)
        else:
            s.kappa_mode = KAPPA_CONST
            s.kappa_const = float(kappa)
        self.s = s
        self.N = lib().ora_cg_ndofs(C.byref(s))

    def apply(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        w = np.empty(self.N)
        lib().ora_cg_apply(C.byref(self.s), _ptr(u), _ptr(w))
        return w

    def rhs(self, fq):
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        F = np.empty(self.N)
        lib().ora_cg_rhs(C.byref(self.s), _ptr(fq), _ptr(F))
        return F

    def set_reaction_qp(self, cvol):
        """pointwise reaction coefficient at the volume quadrature points
        [E][9] (qx + 3 qy), added to c; None to clear"""
        if cvol is None:
            self.s.cvol = None
            return
        cv = np.ascontiguousarray(cvol, dtype=np.float64).ravel()
        assert cv.size == self.n[0] * self.n[1] * 9
        self._keep.append(cv)
        self.s.cvol = _ptr(cv)

    def eval_qp(self, u):
        """u at the volume quadrature points [E][9]"""
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.empty((self.n[0] * self.n[1], 9))
        lib().ora_cg_eval_qp(C.byref(self.s), _ptr(u), _ptr(out))
        return out

    def diag(self):
        d = np.empty(self.N)
        lib().ora_cg_diag(C.byref(self.s), _ptr(d))
        return d

    def callback(self):
        f = _APPLY(lambda ctx, x, y: lib().ora_cg_apply(C.byref(self.s), x, y))
        self._keep.append(f)
        return f


class _Pre:
    def __init__(self, fn):
        self._keep = []
        self.cb = _APPLY(lambda ctx, x, y: fn(x, y))


def jacobi(dinv):
    """Pointwise Jacobi preconditioner y = dinv * x (as a callback)."""
    dinv = np.ascontiguousarray(dinv, dtype=np.float64)
    n = dinv.size

    def fn(x, y):
        xa = np.ctypeslib.as_array(x, shape=(n,))
        ya = np.ctypeslib.as_array(y, shape=(n,))
        ya[:] = dinv * xa
    p = _Pre(fn)
    p._keep.append(dinv)
    return p


def function_pre(fn, n):
    """Preconditioner y = fn(x) of any function on numpy arrays of n entries
    (as a callback)."""
    def cb(x, y):
        xa = np.ctypeslib.as_array(x, shape=(n,))
        ya = np.ctypeslib.as_array(y, shape=(n,))
        ya[:] = fn(xa.copy())
    return _Pre(cb)


def block_jacobi(inv_blocks):
    inv = np.ascontiguousarray(inv_blocks, dtype=np.float64)
    E, nb, _ = inv.shape

    def fn(x, y):
        xa = np.ctypeslib.as_array(x, shape=(E, nb))
        ya = np.ctypeslib.as_array(y, shape=(E, nb))
        ya[:] = np.einsum("eij,ej->ei", inv, xa)
    p = _Pre(fn)
    p._keep.append(inv)
    return p


def _solve(kind, op, b, x0=None, pre=None, tol=1e-10, maxit=1000, criteria=1, restart=20, nthreads=1):
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=np.float64)
    cb = op.callback(nthreads) if isinstance(op, DGOperator) else op.callback()
    pcb = pre.cb if pre is not None else None
    L = lib()
    if kind in ("cg", "bicgstab"):
        it = (L.ora_cg if kind == "cg" else L.ora_bicgstab)(b.size, C.cast(cb, C.c_void_p), None,
                      C.cast(pcb, C.c_void_p) if pcb else None, None,
                      _ptr(x), _ptr(b), tol, maxit, criteria)
    else:
        f = L.ora_gmres if kind == "gmres" else L.ora_fgmres
        it = f(b.size, C.cast(cb, C.c_void_p), None,
               C.cast(pcb, C.c_void_p) if pcb else None, None,
               _ptr(x), _ptr(b), restart, tol, maxit, criteria)
    return x, it


def cg(op, b, **kw):
    return _solve("cg", op, b, **kw)


def bicgstab(op, b, **kw):
    return _solve("bicgstab", op, b, **kw)


def gmres(op, b, **kw):
    return _solve("gmres", op, b, **kw)


def fgmres(op, b, **kw):
    return _solve("fgmres", op, b, **kw)
