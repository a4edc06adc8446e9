"""ctypes wrapper of the CPU oracle (oracle/_build/libikoracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker. The product package (ikarus_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libikoracle.so")

RAW, FULL, REDUCED = 0, 1, 2
NEOHOOKE, SVK = 0, 1
DIRECT, JACOBI_CG = 0, 1
PRED_CONSTANT, PRED_LINEAR = 0, 1

_lib = None


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
        L.ikor_last_error.restype = C.c_char_p
        L.ikor_create.restype = P
        L.ikor_create.argtypes = [I, I, I, P, P, I, D, D, I, I]
        L.ikor_destroy.argtypes = [P]
        L.ikor_set_fixed.argtypes = [P, P]
        L.ikor_set_ubar.argtypes = [P, P]
        L.ikor_set_traction.argtypes = [P, I, P, P, P, P]
        L.ikor_add_point_load.argtypes = [P, I64, D]
        L.ikor_set_faithful.argtypes = [P, I]
        L.ikor_set_threads.argtypes = [P, I]
        L.ikor_set_alpha.argtypes = [P, P]
        L.ikor_get_alpha.argtypes = [P, P]
        L.ikor_pattern_nnz.restype = I64
        L.ikor_pattern_nnz.argtypes = [P, I]
        L.ikor_pattern_n.restype = I64
        L.ikor_pattern_n.argtypes = [P, I]
        L.ikor_pattern.argtypes = [P, I, P, P]
        L.ikor_slots.argtypes = [P, I, P]
        L.ikor_matrix.argtypes = [P, P, D, I, P]
        L.ikor_vector.argtypes = [P, P, D, I, P]
        L.ikor_energy.argtypes = [P, P, D, P]
        L.ikor_stress_at.argtypes = [P, P, P, I, P]
        L.ikor_element.argtypes = [P, I, P, D, P, P, P, P, P]
        L.ikor_update_state.argtypes = [P, P, D, P]
        L.ikor_cg.argtypes = [P, P, D, P, P, D, I, I, P, P]
        L.ikor_direct_solve.argtypes = [P, P, D, P, P]
        L.ikor_load_control.argtypes = [P, I, D, D, D, I, I, D, I, I, I, P, P, P, P, P, P, P]
        L.ikor_newton.argtypes = [P, P, D, D, D, I, I, D, I, P, P, P]
        L.ikor_structured_mesh.argtypes = [I, P, P, D, C.c_uint64, P, P]
        L.ikor_lame.argtypes = [D, D, P, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().ikor_last_error().decode()}")


def lame(E, nu):
    lam, mu = C.c_double(), C.c_double()
    lib().ikor_lame(E, nu, C.byref(lam), C.byref(mu))
    return lam.value, mu.value


def structured_mesh(nel, L, jitter=0.0, seed=0):
    dim = len(nel)
    nv = np.prod([n + 1 for n in nel])
    ne = np.prod(nel)
    X = np.zeros(nv * dim)
    conn = np.zeros(ne * (1 << dim), dtype=np.int32)
    _check(lib().ikor_structured_mesh(dim, _p(np.asarray(nel, dtype=np.int32)), _p(np.asarray(L, dtype=np.float64)),
                                      float(jitter), C.c_uint64(seed), _p(X), _p(conn)))
    return X.reshape(nv, dim), conn.reshape(ne, 1 << dim)


class Oracle:
    """One problem: mesh + material + EAS + BCs + loads, with the reference's operator semantics."""

    def __init__(self, X, conn, mat_kind, lam, mu, neas=0, nq1d=2):
        X = np.ascontiguousarray(X, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.dim = X.shape[1]
        self.nn = X.shape[0]
        self.ne = conn.shape[0]
        self.neas = neas
        self.ndof = self.nn * self.dim
        self.nd = conn.shape[1] * self.dim
        self._X, self._conn = X, conn
        self.h = lib().ikor_create(self.dim, self.nn, self.ne, _p(X), _p(conn), mat_kind, lam, mu, neas, nq1d)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ikor_destroy(self.h)
            self.h = None

    def set_fixed(self, flags):
        self._fixed = np.ascontiguousarray(flags, dtype=np.uint8)
        lib().ikor_set_fixed(self.h, _p(self._fixed))

    def set_ubar(self, ubar):
        self._ubar = None if ubar is None else np.ascontiguousarray(ubar, dtype=np.float64)
        lib().ikor_set_ubar(self.h, _p(self._ubar))

    def set_traction(self, elems, axis, side, t0):
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        ax = np.full(len(elems), axis, dtype=np.int32) if np.isscalar(axis) else np.ascontiguousarray(axis, np.int32)
        sd = np.full(len(elems), side, dtype=np.int32) if np.isscalar(side) else np.ascontiguousarray(side, np.int32)
        t = np.zeros(3)
        t[:len(t0)] = t0
        lib().ikor_set_traction(self.h, len(elems), _p(elems), _p(ax), _p(sd), _p(t))

    def add_point_load(self, dof, coef):
        lib().ikor_add_point_load(self.h, int(dof), float(coef))

    def set_faithful(self, on):
        lib().ikor_set_faithful(self.h, 1 if on else 0)

    def set_threads(self, n):
        """Element loops of matrix()/vector() on n OpenMP threads over an element coloring (1: serial)."""
        lib().ikor_set_threads(self.h, int(n))

    def alpha(self):
        a = np.zeros(self.ne * self.neas)
        lib().ikor_get_alpha(self.h, _p(a))
        return a.reshape(self.ne, self.neas)

    def set_alpha(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        lib().ikor_set_alpha(self.h, _p(a))

    def pattern(self, reduced=False):
        n = lib().ikor_pattern_n(self.h, int(reduced))
        nnz = lib().ikor_pattern_nnz(self.h, int(reduced))
        if n < 0 or nnz < 0:
            _check(-1)
        outer = np.zeros(n + 1, dtype=np.int32)
        inner = np.zeros(nnz, dtype=np.int32)
        _check(lib().ikor_pattern(self.h, int(reduced), _p(outer), _p(inner)))
        return outer, inner

    def slots(self, e):
        s = np.zeros(self.nd * self.nd, dtype=np.int64)
        _check(lib().ikor_slots(self.h, int(e), _p(s)))
        return s

    def matrix(self, u, lam, dbc=FULL):
        u = np.ascontiguousarray(u, dtype=np.float64)
        nnz = lib().ikor_pattern_nnz(self.h, int(dbc == REDUCED))
        v = np.zeros(nnz)
        _check(lib().ikor_matrix(self.h, _p(u), float(lam), dbc, _p(v)))
        return v

    def matrix_dense(self, u, lam, dbc=FULL):
        outer, inner = self.pattern(dbc == REDUCED)
        v = self.matrix(u, lam, dbc)
        n = len(outer) - 1
        K = np.zeros((n, n))
        for c in range(n):
            K[inner[outer[c]:outer[c + 1]], c] = v[outer[c]:outer[c + 1]]
        return K

    def vector(self, u, lam, dbc=FULL):
        u = np.ascontiguousarray(u, dtype=np.float64)
        n = self.ndof if dbc != REDUCED else lib().ikor_pattern_n(self.h, 1)
        R = np.zeros(n)
        _check(lib().ikor_vector(self.h, _p(u), float(lam), dbc, _p(R)))
        return R

    def stress_at(self, u, xi, kind=0):
        """Per element, the PK2 (kind 0), Kirchhoff (1) or Cauchy (2) stress at local point xi, Voigt 6."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        out = np.zeros((self.ne, 6))
        _check(lib().ikor_stress_at(self.h, _p(u), _p(xi), int(kind), _p(out)))
        return out

    def energy(self, u, lam):
        u = np.ascontiguousarray(u, dtype=np.float64)
        e = C.c_double()
        _check(lib().ikor_energy(self.h, _p(u), float(lam), C.byref(e)))
        return e.value

    def element(self, e, u, lam):
        u = np.ascontiguousarray(u, dtype=np.float64)
        nd, ne = self.nd, self.neas
        K = np.zeros((nd, nd))
        R = np.zeros(nd)
        Dm = np.zeros((ne, ne)) if ne else None
        Lm = np.zeros((ne, nd)) if ne else None
        Rt = np.zeros(ne) if ne else None
        _check(lib().ikor_element(self.h, int(e), _p(u), float(lam), _p(K), _p(R), _p(Dm), _p(Lm), _p(Rt)))
        return (K, R) if not ne else (K, R, Dm, Lm, Rt)

    def update_state(self, u, lam, du):
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.ascontiguousarray(du, dtype=np.float64)
        _check(lib().ikor_update_state(self.h, _p(u), float(lam), _p(du)))

    def cg(self, u, lam, b, tol=2.220446049250313e-16, max_iter=-1, threads=1):
        u = np.ascontiguousarray(u, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        it, err = C.c_int(), C.c_double()
        _check(lib().ikor_cg(self.h, _p(u), float(lam), _p(b), _p(x), tol, max_iter, threads, C.byref(it),
                             C.byref(err)))
        return x, it.value, err.value

    def direct_solve(self, u, lam, b):
        u = np.ascontiguousarray(u, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        _check(lib().ikor_direct_solve(self.h, _p(u), float(lam), _p(b), _p(x)))
        return x

    def load_control(self, steps, tbegin=0.0, tend=1.0, tol=1e-8, max_iter=20, solver=DIRECT,
                     cg_tol=2.220446049250313e-16, cg_max_iter=-1, threads=1, predictor=PRED_CONSTANT, u=None,
                     lam=0.0):
        u = np.zeros(self.ndof) if u is None else np.array(u, dtype=np.float64)
        lamc = C.c_double(lam)
        iters = np.zeros(steps + 1, dtype=np.int32)
        res = np.zeros(steps + 1)
        total, cgit, ok = C.c_int(), C.c_int64(), C.c_int()
        _check(lib().ikor_load_control(self.h, steps, tbegin, tend, tol, max_iter, solver, cg_tol, cg_max_iter,
                                       threads, predictor, _p(u), C.byref(lamc), _p(iters), _p(res),
                                       C.byref(total), C.byref(cgit), C.byref(ok)))
        return dict(u=u, lam=lamc.value, iterations=iters, residuals=res, total=total.value,
                    cg_iterations=cgit.value, success=bool(ok.value))

    def newton(self, u, lam, step_size=0.0, tol=1e-8, max_iter=20, solver=DIRECT, cg_tol=2.220446049250313e-16,
               threads=1):
        u = np.array(u, dtype=np.float64)
        it, res, ok = C.c_int(), C.c_double(), C.c_int()
        _check(lib().ikor_newton(self.h, _p(u), float(lam), float(step_size), tol, max_iter, solver, cg_tol, threads,
                                 C.byref(it), C.byref(res), C.byref(ok)))
        return dict(u=u, iterations=it.value, residual=res.value, success=bool(ok.value))
