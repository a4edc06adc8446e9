"""Arc-length path following on the device against the oracle's restatement (oracle/pathfollowing.py,
pinned by the reference's two-DOF golden values in tests/test_oracle_golden.py), with the oracle's dense
FULL-DBC residual and tangent of the same mesh.

Bars: Newton iterations of every solve identical, load factor after every solve within 1e-10, final
displacements within 1e-8 relative (the device solves with FP64 PCG to 1e-13, the oracle densely).
"""
import numpy as np
import pytest

import ikarus_amd as ia
import pathfollowing as pf
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _oracle_for(p):
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    return o


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_arc_length_matches_oracle(material):
    p = pb.hex_block((4, 3, 3), material=material, jitter=0.05, seed=41, mag_range=(20.0, 30.0))
    print(f"case: arc length material={material}", flush=True)
    o = _oracle_for(p)
    res = lambda u, lam: o.vector(u, lam, po.FULL)
    jac = lambda u, lam: o.matrix_dense(u, lam, po.FULL)
    z = np.zeros(p.ndof)
    d1 = np.linalg.solve(jac(z, 1.0), -res(z, 1.0))
    s = 0.25 * np.sqrt(2.0) * np.linalg.norm(d1)  # about a quarter of the load per step
    r = pf.arc_length(res, jac, z, steps=4, step_size=s)
    assert r["success"]
    dev = ia.Device(p)
    dev.set_requirement(z, 0.0)
    cfg = capi.PFConfig()
    cfg.steps, cfg.step_size = 4, s
    info, iters, rn, lams = dev.path_following(cfg, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13)))
    ug, lam = dev.get_requirement()
    dev.close()
    assert info.success and info.solves == 5
    assert list(iters) == list(r["iterations"])
    assert np.all(np.abs(lams - np.array(r["lambdas"])) < 1e-10)
    assert abs(lam - r["lam"]) < 1e-10 and 0.8 < lam < 1.2
    assert np.linalg.norm(ug - r["u"]) / np.linalg.norm(r["u"]) < 1e-8
    assert np.all(rn[1:] <= 1e-8)


@pytest.mark.parametrize("kind", [capi.PCG_BJACOBI_FP32_IR, capi.PCG_TWOLEVEL_FP32_IR])
def test_arc_length_mixed_precision(kind):
    """The bordered solves by the mixed form (FP32 tangent and PCG with the block-Jacobi or the two-level
    preconditioner, FP64 refinement to 1e-10): same iterations as the oracle, the final state in FP64 equilibrium
    (oracle residual <= 2e-8), load factors and displacements within 1e-6 (the tolerance region's extent, see
    below)."""
    p = pb.hex_block((4, 3, 3), jitter=0.05, seed=41, mag_range=(20.0, 30.0))
    print(f"case: arc length, mixed precision, solver kind {kind}", flush=True)
    o = _oracle_for(p)
    res = lambda u, lam: o.vector(u, lam, po.FULL)
    jac = lambda u, lam: o.matrix_dense(u, lam, po.FULL)
    z = np.zeros(p.ndof)
    s = 0.25 * np.sqrt(2.0) * np.linalg.norm(np.linalg.solve(jac(z, 1.0), -res(z, 1.0)))
    r = pf.arc_length(res, jac, z, steps=4, step_size=s)
    dev = ia.Device(p)
    dev.set_requirement(z, 0.0)
    cfg = capi.PFConfig()
    cfg.steps, cfg.step_size = 4, s
    st = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                        linear=ia.solver_opts(kind=kind, rtol=1e-10, inner_rtol=1e-3))
    info, iters, rn, lams = dev.path_following(cfg, st)
    ug, lam = dev.get_requirement()
    dev.close()
    assert info.success and list(iters) == list(r["iterations"])
    # The FP32 tangent takes slightly different Newton steps, so the two runs stop at different points of the
    # region sqrt(|R|^2 + f^2) <= tol. There lambda is fixed only to about tol / (df/dDl) (df/dDl = psi^2 Dl
    # / root ~ 0.05 here): compare the equilibrium itself, then the states at that scale.
    assert np.linalg.norm(res(ug, lam)) <= 2e-8
    dl = np.abs(lams - np.array(r["lambdas"])).max()
    assert dl < 1e-6, dl
    assert np.linalg.norm(ug - r["u"]) / np.linalg.norm(r["u"]) < 1e-6  # the same shift along the path


def _arc_step(o, p, frac):
    z = np.zeros(p.ndof)
    d1 = np.linalg.solve(o.matrix_dense(z, 1.0, po.FULL), o.vector(z, 1.0, po.FULL))
    return frac * np.sqrt(2.0) * np.linalg.norm(d1)


def _device_path(p, steps, s, ubar=None, target=0):
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    if ubar is not None:
        dev.set_dirichlet_values(ubar)
    cfg = capi.PFConfig()
    cfg.steps, cfg.step_size, cfg.target_iterations = steps, s, target
    info, iters, rn, lams = dev.path_following(cfg, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13)))
    ug, lam = dev.get_requirement()
    dev.close()
    return info, iters, rn, lams, ug, lam


def _compare(r, info, iters, rn, lams, ug, lam, steps):
    assert r["success"] and info.success and info.solves == steps + 1
    assert list(iters) == list(r["iterations"])
    assert np.all(np.abs(lams - np.array(r["lambdas"])) < 1e-10)
    assert abs(lam - r["lam"]) < 1e-10
    assert np.linalg.norm(ug - r["u"]) / np.linalg.norm(r["u"]) < 1e-8
    assert np.all(rn[1:] <= 1e-8)


def test_arc_length_eas_matches_oracle():
    """EAS-21 elements: Fext0 = -R(0, 1) at the start of every solve keeps the current alpha (the reference does
    not reset it), and alpha is updated on every correction at the pre-update state."""
    p = pb.hex_block((3, 2, 2), eas=21, jitter=0.05, seed=44, mag_range=(20.0, 30.0))
    print("case: arc length EAS", flush=True)
    o = _oracle_for(p)
    res = lambda u, lam: o.vector(u, lam, po.FULL)
    jac = lambda u, lam: o.matrix_dense(u, lam, po.FULL)
    s = _arc_step(o, p, 0.3)
    r = pf.arc_length(res, jac, np.zeros(p.ndof), steps=3, step_size=s,
                      update_state=lambda u, lam, d: o.update_state(u, lam, d))
    _compare(r, *_device_path(p, 3, s), 3)


def test_arc_length_idbc_matches_oracle():
    """Inhomogeneous Dirichlet values on the loaded face (u_z = -0.02 lambda on top) beside the traction: the arc
    length includes their increment, the load direction subtracts K_raw(u) ubar, the DOFs follow lambda."""
    p = pb.hex_block((3, 3, 2), jitter=0.05, seed=45, mag_range=(10.0, 20.0))
    print("case: arc length IDBC", flush=True)
    top = np.nonzero(p.X[:, 2] > 1 - 1e-9)[0]
    p.fixed = p.fixed.copy()
    p.fixed[3 * top + 2] = 1
    ubar = np.zeros(p.ndof)
    ubar[3 * top + 2] = -0.02
    o = _oracle_for(p)
    fixed = p.fixed.astype(bool)
    res = lambda u, lam: o.vector(u, lam, po.FULL)
    jac = lambda u, lam: o.matrix_dense(u, lam, po.FULL)

    def idbc_force(u, lam):
        f = o.matrix_dense(u, lam, po.RAW) @ ubar
        f[fixed] = 0.0
        return f
    s = _arc_step(o, p, 0.25)
    r = pf.arc_length(res, jac, np.zeros(p.ndof), steps=4, step_size=s, ubar=ubar, idbc_force=idbc_force)
    out = _device_path(p, 4, s, ubar=ubar)
    _compare(r, *out, 4)
    ug, lam = out[4], out[5]
    assert np.allclose(ug[3 * top + 2], -0.02 * lam, rtol=1e-14)


def test_arc_length_adaptive_step_sizing_matches_oracle():
    """IterationBased step sizing with target 1: every step after a two-iteration solve is shortened by sqrt(1/2)."""
    p = pb.hex_block((3, 3, 2), jitter=0.05, seed=46, mag_range=(20.0, 30.0))
    print("case: arc length adaptive", flush=True)
    o = _oracle_for(p)
    res = lambda u, lam: o.vector(u, lam, po.FULL)
    jac = lambda u, lam: o.matrix_dense(u, lam, po.FULL)
    s = _arc_step(o, p, 0.3)
    r = pf.arc_length(res, jac, np.zeros(p.ndof), steps=4, step_size=s, target_iterations=1)
    plain = pf.arc_length(res, jac, np.zeros(p.ndof), steps=4, step_size=s)
    assert r["lam"] < plain["lam"]
    _compare(r, *_device_path(p, 4, s, target=1), 4)


def test_path_following_mirror():
    """The host mirror (SparseFlatAssembler, NewtonRaphson, PathFollowing, FERequirement) runs the same path."""
    p = pb.hex_block((3, 3, 2), jitter=0.05, seed=43, mag_range=(20.0, 30.0))
    print("case: path following mirror", flush=True)
    o = _oracle_for(p)
    z = np.zeros(p.ndof)
    d1 = np.linalg.solve(o.matrix_dense(z, 1.0, po.FULL), o.vector(z, 1.0, po.FULL))
    step = 0.2 * np.sqrt(2.0) * np.linalg.norm(d1)  # about a fifth of the load per step
    settings = ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13))
    asm = ia.SparseFlatAssembler(p)
    req = ia.FERequirement(np.zeros(p.ndof), 0.0)
    out = ia.PathFollowing(ia.NewtonRaphson(asm, settings), 3, step).run(req)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.PFConfig()
    cfg.steps, cfg.step_size = 3, step
    info, iters, _, lams = dev.path_following(cfg, settings)
    ug, lam = dev.get_requirement()
    dev.close()
    assert out["success"] and info.success
    assert list(out["iterations"]) == list(iters) and np.array_equal(out["loadFactors"], lams)
    assert np.array_equal(req.u, ug) and req.lam == lam
