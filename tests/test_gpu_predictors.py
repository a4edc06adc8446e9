"""Load-step predictors and failure handling on the device, through the C ABI, against the FP64 CPU oracle.

* secant predictor (u += u_n - u_{n-1}, the ArcLength intermediate-prediction semantics of
  pathfollowingfunctions.hh:188-195) against the oracle's load_control(predictor=PRED_LINEAR);
* the learned predictor inside the mixed-precision load path with the committed network (the c5 form), and its
  fallback to the secant start when the Newton solve from the MLP prediction fails;
* the two error paths that end a solve without a GPU fault: a singular EAS matrix D (IK_ESINGULAR) and a residual
  norm that is not finite (Newton reports failure).
"""
import os

import numpy as np
import pytest

import ikarus_amd as ia
import mlp_oracle as mo
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu

MLP_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ikarus_amd", "data",
                        "predictor_mlp_32cubed.f32")


def _oracle_for(p):
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    if len(p.face_elem):
        o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    return o


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _run(p, steps, pred, settings, params=None):
    dev = ia.Device(p)
    if params is not None:
        dev.check(capi.lib().ik_mlp_set_params(dev.h, capi.ptr(params), capi.HOST))
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = steps, 0.0, 1.0, pred
    info, iters, res = dev.load_control(cfg, settings)
    u, lam = dev.get_requirement()
    dev.close()
    return info, iters, res, u, lam


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_secant_predictor_matches_oracle(material):
    """Device PRED_LINEAR against the oracle's PRED_LINEAR load path: identical Newton iterations per step,
    converged u within 1e-8."""
    p = pb.hex_block((4, 4, 3), material=material, jitter=0.05, seed=17)
    print(f"case: secant predictor material={material}", flush=True)
    r = _oracle_for(p).load_control(6, 0.0, 1.0, tol=1e-8, predictor=po.PRED_LINEAR)
    info, iters, res, u, lam = _run(p, 6, capi.PRED_LINEAR, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13)))
    assert r["success"] and info.success and abs(lam - 1.0) < 1e-14
    assert list(iters) == list(r["iterations"])
    # the secant start needs fewer iterations than the constant one on this path
    rc = _oracle_for(p).load_control(6, 0.0, 1.0, tol=1e-8)
    assert r["total"] < rc["total"]
    assert np.all(res[1:] <= 1e-8)
    assert _rel(u, r["u"]) < 1e-8


def _committed_params():
    return np.fromfile(MLP_PATH, dtype=np.float32)


def _c5_settings():
    """The c5 solver form: FP32 tangent, FP32 block-Jacobi PCG with FP64 refinement, inexact Newton (forcing 0.9,
    eta_max 1e-3), FP64 residual and ||R|| <= 1e-8."""
    return ia.nr_settings(tol=1e-8, precision=capi.FP32,
                          linear=ia.solver_opts(kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, inner_rtol=1e-3,
                                                max_refine=30),
                          forcing=0.9, eta_max=1e-3)


def test_learned_predictor_in_mixed_load_path_matches_fp64_oracle():
    """The committed 32^3 network inside the mixed-precision load path (40 load steps, 8^3 mesh, unseen seed):
    every step reaches ||R|| <= 1e-8 without a fallback, and u matches the oracle's FP64 load path."""
    params = _committed_params()
    p = pb.hex_block((8, 8, 8), seed=2077)
    print("case: c5 form 8^3 x 40 steps", flush=True)
    info, iters, res, u, _ = _run(p, 40, capi.PRED_MLP, _c5_settings(), params)
    assert info.success and info.solves == 41 and info.fallbacks == 0
    assert np.all(res <= 1e-8)
    r = _oracle_for(p).load_control(40, 0.0, 1.0, tol=1e-8)
    assert r["success"]
    assert _rel(u, r["u"]) < 1e-8
    # the learned start saves Newton iterations against the secant start on the same solver form
    info_s, _, _, _, _ = _run(p, 40, capi.PRED_LINEAR, _c5_settings())
    print("newton iterations mlp / secant:", info.total_iterations, info_s.total_iterations, flush=True)
    assert info.total_iterations <= info_s.total_iterations


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32])
def test_learned_predictor_falls_back_to_secant(precision):
    """A network with its output layer scaled by 1e3 predicts far-off states: the solves from those predictions
    fail (det C <= 0 or no convergence), each such step is solved again from the secant start, and the path still
    ends at the oracle's solution."""
    d = mo.unpack(_committed_params().astype(np.float64))
    d["W4"] *= 1e3
    d["b4"] *= 1e3
    params = mo.pack(d).astype(np.float32)
    p = pb.hex_block((4, 4, 4), seed=2003)
    print(f"case: mlp fallback precision={precision}", flush=True)
    if precision == capi.FP64:
        s = ia.nr_settings(tol=1e-8, max_iter=12, linear=ia.solver_opts(rtol=1e-13))
    else:
        s = _c5_settings()
        s.max_iter = 12
    info, iters, res, u, _ = _run(p, 8, capi.PRED_MLP, s, params)
    assert info.success and info.solves == 9
    assert info.fallbacks >= 1
    assert np.all(res <= 1e-8)
    r = _oracle_for(p).load_control(8, 0.0, 1.0, tol=1e-8, predictor=po.PRED_LINEAR)
    assert _rel(u, r["u"]) < 1e-8
    if precision == capi.FP64 and info.fallbacks == 8:
        # every step fell back: the per-step iterations are those of the secant path
        assert list(iters) == list(r["iterations"])


def test_singular_enhanced_strain_matrix_is_reported():
    """With lambda = mu = 0 the material tangent vanishes, so every EAS element's D is zero: the call returns
    IK_ESINGULAR naming element 0 instead of writing Inf/NaN into K and R, and the handle stays usable."""
    p = pb.hex_block((2, 2, 2), material=capi.SVK, jitter=0.0, seed=0, eas=21, traction=(0.0, 0.0, 1.0))
    print("case: singular EAS D", flush=True)
    p.lam, p.mu = 0.0, 0.0
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 1.0)
    with pytest.raises(capi.IkError) as ei:
        dev.vector()
    assert ei.value.code == capi.IK_ESINGULAR
    assert ei.value.bad_element == 0
    with pytest.raises(capi.IkError) as ei:
        dev.assemble()
    assert ei.value.code == capi.IK_ESINGULAR
    dev.close()


def test_nonfinite_residual_fails_newton():
    """A vanishing material (lambda = mu = 0, SVK) under traction has no equilibrium: the correction and then the
    residual become non-finite, and Newton reports failure instead of success (NaN > tol is false)."""
    p = pb.hex_block((2, 2, 2), material=capi.SVK, jitter=0.0, seed=0, traction=(0.0, 0.0, 1.0))
    print("case: non-finite residual", flush=True)
    p.lam, p.mu = 0.0, 0.0
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 1.0)
    info = dev.newton(ia.nr_settings(tol=1e-8, max_iter=3, linear=ia.solver_opts(rtol=1e-10, max_iter=20)))
    assert info.success == 0
    assert not np.isfinite(info.residual_norm)
    dev.close()
