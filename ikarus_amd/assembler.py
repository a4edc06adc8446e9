"""Host-side mirror of Ikarus's assembler / solver / control operator API over the C ABI.

Names and argument meaning follow the reference:
  SparseFlatAssembler.matrix(req, affordance, dbc) / vector(...)   (ikarus/assembler/interface.hh:355-462)
  size / reducedSize / isConstrained / constraintsBelow / createFullVector / createReducedVector
                                                                   (interface.hh:51-132, simpleassemblers.inl:27-57)
  FERequirement(u, lambda)                                         (ikarus/finiteelements/ferequirements.hh:223-407)
  NewtonRaphson(settings).solve(req, stepSize)                      (ikarus/solver/nonlinearsolver/newtonraphson.hh:196)
  LoadControl(nr, loadSteps, (tbegin, tEnd)).run(req)              (ikarus/controlroutines/loadcontrol.hh:59-113)
Errors raise IkError (the reference throws Dune exceptions or aborts).
All arithmetic runs in libikarus_amd.so; this module only moves arguments.
"""
import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import IkError, ptr

STIFFNESS = "stiffness"
FORCES = "forces"
ENERGY = "mechanicalPotentialEnergy"


class FERequirement:
    """Global solution vector (full size) and load factor."""

    def __init__(self, u, lam=0.0):
        self.u = np.ascontiguousarray(u, dtype=np.float64)
        self.lam = float(lam)

    def globalSolution(self):
        return self.u

    def parameter(self):
        return self.lam


class Device:
    """One problem resident on the GPU (ik_handle)."""

    def __init__(self, problem):
        L = capi.lib()
        p = problem
        if p.dim not in (2, 3):
            raise IkError(capi.IK_ENOTIMPL, "the device path supports 3D hex8 and 2D plane-strain Q1 meshes")
        self.problem = p
        self._keep = dict(
            X=np.ascontiguousarray(p.X, dtype=np.float64),
            conn=np.ascontiguousarray(p.conn, dtype=np.int32),
            fixed=np.ascontiguousarray(p.fixed if p.fixed is not None else np.zeros(p.ndof, np.uint8), np.uint8),
            fe=np.ascontiguousarray(p.face_elem, np.int32),
            fa=np.ascontiguousarray(p.face_axis, np.int8),
            fs=np.ascontiguousarray(p.face_side, np.int8),
            pd=np.ascontiguousarray(p.point_dof, np.int64),
            pc=np.ascontiguousarray(p.point_coef, np.float64),
        )
        k = self._keep
        d = capi.ProblemDesc()
        d.dim = p.dim
        d.num_nodes = p.X.shape[0]
        d.num_elements = p.conn.shape[0]
        d.X = ptr(k["X"])
        d.conn = ptr(k["conn"])
        d.material = int(p.material)
        d.lam = float(p.lam)
        d.mu = float(p.mu)
        d.eas_modes = int(p.eas)
        d.dirichlet = ptr(k["fixed"])
        d.num_faces = len(k["fe"])
        d.face_elem = ptr(k["fe"]) if len(k["fe"]) else None
        d.face_axis = ptr(k["fa"]) if len(k["fe"]) else None
        d.face_side = ptr(k["fs"]) if len(k["fe"]) else None
        for i in range(3):
            d.traction[i] = float(p.traction[i]) if i < len(p.traction) else 0.0
        d.num_point_loads = len(k["pd"])
        d.point_dof = ptr(k["pd"]) if len(k["pd"]) else None
        d.point_coef = ptr(k["pc"]) if len(k["pd"]) else None
        h = C.c_void_p()
        rc = L.ik_create(C.byref(h), C.byref(d))
        self.h = h
        if rc != 0:
            msg = L.ik_last_error(h).decode() if h else "create failed"
            L.ik_destroy(h)
            self.h = None
            raise IkError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            capi.lib().ik_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            L = capi.lib()
            raise IkError(rc, L.ik_last_error(self.h).decode(), L.ik_bad_element(self.h))

    # --- raw C-ABI conveniences -----------------------------------------------------------------
    @property
    def ndof(self):
        return int(capi.lib().ik_size(self.h))

    def set_requirement(self, u, lam):
        u = None if u is None else np.ascontiguousarray(u, dtype=np.float64)
        self.check(capi.lib().ik_set_requirement(self.h, ptr(u), float(lam), capi.HOST))

    def set_dirichlet_values(self, ubar):
        """Inhomogeneous Dirichlet values u_D(lambda) = lambda * ubar on constrained DOFs
        (evaluateInhomogeneousBoundaryCondition, dirichletvalues.hh:214-257); None clears them."""
        self._ubar = None if ubar is None else np.ascontiguousarray(ubar, dtype=np.float64)
        self.check(capi.lib().ik_set_dirichlet_values(self.h, ptr(self._ubar), capi.HOST))

    def energy(self):
        """Potential energy at the bound requirement (plain element)."""
        e = C.c_double()
        self.check(capi.lib().ik_energy(self.h, C.byref(e)))
        return e.value

    def stress_at(self, xi, kind=0):
        """Per element, the PK2 (kind 0), Kirchhoff (1) or Cauchy (2) stress at local point xi, Voigt 6
        ([11, 22, 12] in plane strain)."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        out = np.zeros((self.problem.ne, 6 if self.problem.dim == 3 else 3))
        self.check(capi.lib().ik_stress_at(self.h, int(kind), ptr(xi), ptr(out), capi.HOST))
        return out

    def get_requirement(self):
        u = np.zeros(self.ndof)
        lam = C.c_double()
        self.check(capi.lib().ik_get_requirement(self.h, ptr(u), C.byref(lam), capi.HOST))
        return u, lam.value

    def vector(self, dbc=capi.DBC_FULL):
        n = self.ndof if dbc != capi.DBC_REDUCED else int(capi.lib().ik_reduced_size(self.h))
        R = np.zeros(n)
        self.check(capi.lib().ik_vector(self.h, dbc, ptr(R), capi.HOST))
        return R

    def matrix(self, dbc=capi.DBC_FULL, precision=capi.FP64):
        self.check(capi.lib().ik_matrix(self.h, dbc, precision))

    def assemble(self, dbc=capi.DBC_FULL, precision=capi.FP64):
        """Residual and tangent of the bound state in one call (ik_assemble); returns the residual."""
        n = self.ndof if dbc != capi.DBC_REDUCED else int(capi.lib().ik_reduced_size(self.h))
        R = np.zeros(n)
        self.check(capi.lib().ik_assemble(self.h, dbc, precision, ptr(R), capi.HOST))
        return R

    def csc(self, dbc=capi.DBC_FULL, values=True):
        L = capi.lib()
        n, nnz = C.c_int64(), C.c_int64()
        self.check(L.ik_matrix_csc(self.h, dbc, C.byref(n), C.byref(nnz), None, None, None))
        outer = np.zeros(n.value + 1, np.int32)
        inner = np.zeros(nnz.value, np.int32)
        vals = np.zeros(nnz.value) if values else None
        self.check(L.ik_matrix_csc(self.h, dbc, C.byref(n), C.byref(nnz), ptr(outer), ptr(inner), ptr(vals)))
        return outer, inner, vals

    def element(self, e):
        nd = self.problem.conn.shape[1] * self.problem.dim
        K = np.zeros((nd, nd))
        R = np.zeros(nd)
        self.check(capi.lib().ik_element(self.h, int(e), ptr(K), ptr(R)))
        return K, R

    def update_state(self, du):
        du = np.ascontiguousarray(du, dtype=np.float64)
        self.check(capi.lib().ik_update_state(self.h, ptr(du), capi.HOST))

    def alpha(self):
        a = np.zeros(self.problem.ne * self.problem.eas)
        self.check(capi.lib().ik_get_alpha(self.h, ptr(a), capi.HOST))
        return a.reshape(self.problem.ne, self.problem.eas)

    def set_alpha(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        self.check(capi.lib().ik_set_alpha(self.h, ptr(a), capi.HOST))

    def solve(self, rhs, opts):
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.zeros_like(rhs)
        info = capi.SolverInfo()
        self.check(capi.lib().ik_solve(self.h, ptr(rhs), ptr(x), capi.HOST, C.byref(opts), C.byref(info)))
        return x, info

    def newton(self, settings, step_size=0.0):
        info = capi.NRInfo()
        self.check(capi.lib().ik_newton_solve(self.h, C.byref(settings), float(step_size), C.byref(info)))
        return info

    def load_control(self, config, settings):
        info = capi.ControlInfo()
        iters = np.zeros(config.load_steps + 1, np.int32)
        res = np.zeros(config.load_steps + 1)
        self.check(capi.lib().ik_load_control(self.h, C.byref(config), C.byref(settings), C.byref(info), ptr(iters),
                                              ptr(res)))
        return info, iters, res

    def path_following(self, config, settings):
        """Arc-length PathFollowing::run: (info, iterations, residual norms, load factors) per solve."""
        info = capi.ControlInfo()
        iters = np.zeros(config.steps + 1, np.int32)
        res = np.zeros(config.steps + 1)
        lams = np.zeros(config.steps + 1)
        self.check(capi.lib().ik_path_following(self.h, C.byref(config), C.byref(settings), C.byref(info),
                                                ptr(iters), ptr(res), ptr(lams)))
        return info, iters, res, lams

    def timing(self, on=True):
        self.check(capi.lib().ik_timing_enable(self.h, 1 if on else 0))

    def timing_reset(self):
        self.check(capi.lib().ik_timing_reset(self.h))

    def timing_read(self, kernel):
        n, ms = C.c_int64(), C.c_double()
        self.check(capi.lib().ik_timing_read(self.h, kernel, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def synchronize(self):
        self.check(capi.lib().ik_synchronize(self.h))

    def assembly_layout(self):
        """(staged slots, gathered upper blocks, directly written upper blocks) of the tangent assembly."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self.check(capi.lib().ik_assembly_layout(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value


def solver_opts(kind=capi.PCG_JACOBI_FP64, rtol=1e-10, max_iter=-1, inner_rtol=1e-4, max_refine=20):
    o = capi.SolverOpts()
    o.kind, o.rtol, o.max_iter, o.inner_rtol, o.max_refine = kind, rtol, max_iter, inner_rtol, max_refine
    return o


def nr_settings(tol=1e-8, max_iter=20, min_iter=0, precision=capi.FP64, linear=None, forcing=0.0, eta_max=0.0):
    s = capi.NRSettings()
    s.tol, s.max_iter, s.min_iter, s.precision = tol, max_iter, min_iter, precision
    s.linear = linear if linear is not None else solver_opts()
    s.forcing, s.eta_max = forcing, eta_max
    return s


class SparseFlatAssembler:
    """Ikarus::SparseFlatAssembler mirror (ikarus/assembler/simpleassemblers.hh:109)."""

    def __init__(self, problem):
        self.dev = Device(problem)
        self.problem = problem
        fixed = problem.fixed if problem.fixed is not None else np.zeros(problem.ndof, np.uint8)
        self._fixed = fixed.astype(bool)
        self._cb = np.concatenate([[0], np.cumsum(self._fixed)[:-1]]).astype(np.int64)
        self._req = None
        self._dbc = capi.DBC_FULL

    def bind(self, req=None, affordances=None, dbc=None):
        if req is not None:
            self._req = req
        if dbc is not None:
            self._dbc = dbc

    def _sync(self, req):
        req = req if req is not None else self._req
        if req is None:
            raise IkError(capi.IK_ENOTBOUND, "The assembler is not bound to a requirement")
        self.dev.set_requirement(req.u, req.lam)

    def size(self):
        return self.problem.ndof

    def reducedSize(self):
        return int(self.problem.ndof - self._fixed.sum())

    def isConstrained(self, i):
        return bool(self._fixed[i])

    def constraintsBelow(self, i):
        return int(self._cb[i])

    def createFullVector(self, red):
        full = np.zeros(self.size())
        full[~self._fixed] = red
        return full

    def createReducedVector(self, full):
        return np.asarray(full)[~self._fixed].copy()

    def scalar(self, req=None, affordance=ENERGY):
        """ScalarAssembler::scalar (interface.hh:283-330): the potential energy at the requirement."""
        if affordance != ENERGY:
            raise IkError(capi.IK_ENOTIMPL, f"ScalarAffordance not implemented: {affordance}")
        self._sync(req)
        return self.dev.energy()

    def vector(self, req=None, affordance=FORCES, dbc=None):
        if affordance != FORCES:
            raise IkError(capi.IK_ENOTIMPL, f"VectorAffordance not implemented: {affordance}")
        self._sync(req)
        return self.dev.vector(self._dbc if dbc is None else dbc)

    def matrix(self, req=None, affordance=STIFFNESS, dbc=None, precision=capi.FP64):
        """Returns the assembled matrix in Eigen CSC layout as (outer, inner, values) arrays."""
        if affordance != STIFFNESS:
            raise IkError(capi.IK_ENOTIMPL, f"MatrixAffordance not implemented: {affordance}")
        dbc = self._dbc if dbc is None else dbc
        self._sync(req)
        self.dev.matrix(dbc, precision)
        return self.dev.csc(dbc)


class NewtonRaphson:
    """NewtonRaphson mirror: absolute ||R||_2 <= tol, maxIter, linear solver options."""

    def __init__(self, assembler, settings=None):
        self.assembler = assembler
        self.settings = settings if settings is not None else nr_settings()

    def solve(self, req, step_size=0.0):
        dev = self.assembler.dev
        dev.set_requirement(req.u, req.lam)
        info = dev.newton(self.settings, step_size)
        req.u, req.lam = dev.get_requirement()
        return info


class LoadControl:
    """LoadControl mirror: lambda += (tEnd - tbegin)/loadSteps per step, Newton per step."""

    def __init__(self, nonlinear_solver, load_steps, tbegin_end=(0.0, 1.0), predictor=capi.PRED_CONSTANT):
        if load_steps <= 0:
            raise IkError(capi.IK_EINVAL, "Number of load steps should be greater than zero.")
        self.nr = nonlinear_solver
        self.cfg = capi.LCConfig()
        self.cfg.load_steps, self.cfg.tbegin, self.cfg.tend = load_steps, tbegin_end[0], tbegin_end[1]
        self.cfg.predictor = predictor

    def run(self, req):
        dev = self.nr.assembler.dev
        dev.set_requirement(req.u, req.lam)
        info, iters, res = dev.load_control(self.cfg, self.nr.settings)
        req.u, req.lam = dev.get_requirement()
        return dict(success=bool(info.success), totalIterations=info.total_iterations,
                    linearIterations=info.linear_iterations, iterations=iters[:info.solves], residuals=res[:info.solves],
                    fallbacks=info.fallbacks)


class PathFollowing:
    """PathFollowing mirror with the ArcLength technique (pathfollowing.inl:28-77): `steps` steps of arc
    length `step_size` from an equilibrium start at lambda = 0; target_iterations > 0 selects
    AdaptiveStepSizing::IterationBased with that target (adaptivestepsizing.hh:48-90), 0 NoOp."""

    def __init__(self, nonlinear_solver, steps, step_size, target_iterations=0):
        if steps <= 0:
            raise IkError(capi.IK_EINVAL, "Number of load steps should be greater than zero.")
        self.nr = nonlinear_solver
        self.cfg = capi.PFConfig()
        self.cfg.steps, self.cfg.step_size, self.cfg.target_iterations = steps, step_size, target_iterations

    def run(self, req):
        dev = self.nr.assembler.dev
        dev.set_requirement(req.u, req.lam)
        info, iters, res, lams = dev.path_following(self.cfg, self.nr.settings)
        req.u, req.lam = dev.get_requirement()
        return dict(success=bool(info.success), totalIterations=info.total_iterations,
                    linearIterations=info.linear_iterations, iterations=iters[:info.solves],
                    residuals=res[:info.solves], loadFactors=lams[:info.solves])
