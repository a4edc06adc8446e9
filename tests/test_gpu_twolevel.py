"""The two-level preconditioner of the FP32 inner PCG (IK_PCG_TWOLEVEL_FP32_IR, no reference counterpart):
block Jacobi plus an additive coarse correction over box aggregates spanned by the rigid-body modes.

Bars: linear solves equal the oracle's FP64 direct solution to the FP32-rounding level (1e-5 relative, as the
block-Jacobi form) with fewer inner iterations than block Jacobi; load paths (plain and EAS elements, plane
strain) reach ||R|| <= 1e-8 at every step with the oracle's FP64 displacements to 1e-8."""
import numpy as np
import pytest

import ikarus_amd as ia
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def _oracle_for(p):
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    if len(p.face_elem):
        o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    return o


@pytest.mark.parametrize("n", [(5, 4, 6), (12, 12, 12)])
def test_two_level_solve_matches_direct(n):
    p = pb.hex_block(n, jitter=0.05, seed=3)
    print(f"case: two-level solve {n}", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    u = 0.01 * np.random.default_rng(7).standard_normal(p.ndof) * (1 - p.fixed)
    dev.set_requirement(u, 0.5)
    dev.matrix(capi.DBC_FULL, capi.FP32)
    b = np.random.default_rng(8).standard_normal(p.ndof) * (1 - p.fixed)
    xd = o.direct_solve(u, 0.5, b)
    its = {}
    for kind in (capi.PCG_BJACOBI_FP32_IR, capi.PCG_TWOLEVEL_FP32_IR):
        x, info = dev.solve(b, ia.solver_opts(kind=kind, rtol=1e-10, inner_rtol=1e-3))
        assert info.converged == 1
        assert _rel(x, xd) < 1e-5
        its[kind] = info.iterations
    assert its[capi.PCG_TWOLEVEL_FP32_IR] < its[capi.PCG_BJACOBI_FP32_IR]
    if n == (12, 12, 12):
        assert 3 * its[capi.PCG_TWOLEVEL_FP32_IR] < its[capi.PCG_BJACOBI_FP32_IR]
    dev.close()


def test_two_level_solve_repeated_is_bitwise_identical():
    """Same matrix and right-hand side twice (the second solve reuses the coarse inverse): bitwise equal x."""
    p = pb.hex_block((10, 9, 8), jitter=0.05, seed=4)
    print("case: two-level repeat", flush=True)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    dev.matrix(capi.DBC_FULL, capi.FP32)
    b = np.random.default_rng(9).standard_normal(p.ndof) * (1 - p.fixed)
    opts = ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-8, inner_rtol=1e-3)
    x1, i1 = dev.solve(b, opts)
    x2, i2 = dev.solve(b, opts)
    assert i1.iterations == i2.iterations and np.array_equal(x1, x2)
    dev.close()


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_two_level_load_control_matches_fp64_oracle(material):
    p = pb.hex_block((8, 8, 6), material=material, jitter=0.05, seed=21)
    print(f"case: two-level load control material={material}", flush=True)
    o = _oracle_for(p)
    r = o.load_control(3, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 3, 0.0, 1.0, capi.PRED_CONSTANT
    s = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                       linear=ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    assert info.success and res[info.solves - 1] <= 1e-8
    assert list(iters[:info.solves]) == list(r["iterations"])
    assert _rel(ug, r["u"]) < 1e-8
    dev.close()


def test_two_level_inexact_secant_path_matches_fp64_oracle():
    """Inexact Newton (Eisenstat-Walker) with the secant predictor, 5 load steps on a 16^3 block: every step
    converges and the final displacement equals the oracle's FP64 load path to 1e-8."""
    p = pb.hex_block((16, 16, 16), jitter=0.05, seed=5)
    print("case: two-level inexact secant 16^3", flush=True)
    o = _oracle_for(p)
    r = o.load_control(5, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 5, 0.0, 1.0, capi.PRED_LINEAR
    s = ia.nr_settings(tol=1e-8, max_iter=30, precision=capi.FP32, forcing=0.9, eta_max=0.1,
                       linear=ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    assert info.success and max(res[:info.solves]) <= 1e-8
    assert _rel(ug, r["u"]) < 1e-8
    dev.close()


def test_two_level_eas_load_control_matches_oracle():
    p = pb.bending_block((6, 6, 3), seed=2)
    print("case: two-level EAS load control", flush=True)
    o = _oracle_for(p)
    r = o.load_control(2, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 2, 0.0, 1.0, capi.PRED_CONSTANT
    s = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                       linear=ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    assert info.success and res[info.solves - 1] <= 1e-8
    assert _rel(ug, r["u"]) < 1e-8
    dev.close()


def test_two_level_plane_strain_load_control_matches_oracle():
    p = pb.cantilever2d((16, 4), seed=1)
    print("case: two-level plane strain", flush=True)
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, 0, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction[:2])
    r = o.load_control(4, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 4, 0.0, 1.0, capi.PRED_CONSTANT
    s = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                       linear=ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    assert info.success and res[info.solves - 1] <= 1e-8
    assert _rel(ug, r["u"]) < 1e-8
    dev.close()


def test_two_level_solve_on_scrambled_mesh():
    """Box aggregation works from coordinates, so a mesh with holes, shuffled elements and relabelled nodes gives
    the same solution as the direct solve."""
    from test_gpu_parity import _scrambled_block
    p = _scrambled_block((9, 8, 7), seed=17)
    print(f"case: two-level scrambled mesh, {p.ne} elements", flush=True)
    o = _oracle_for(p)
    u = 0.01 * np.random.default_rng(5).standard_normal(p.ndof) * (1 - p.fixed)
    dev = ia.Device(p)
    dev.set_requirement(u, 0.9)
    dev.matrix(capi.DBC_FULL, capi.FP32)
    b = np.random.default_rng(6).standard_normal(p.ndof) * (1 - p.fixed)
    x, info = dev.solve(b, ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    xb, infob = dev.solve(b, ia.solver_opts(kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    assert info.converged == 1 and infob.converged == 1
    # both refine against the FP32-rounded matrix, so they agree to the refinement tolerance; against the FP64
    # direct solve they differ by that rounding, amplified by this mesh's conditioning (1.1e-5 measured)
    assert _rel(x, xb) < 1e-8
    assert _rel(x, o.direct_solve(u, 0.9, b)) < 1e-4
    assert info.iterations < infob.iterations
    dev.close()
