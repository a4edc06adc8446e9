"""Inhomogeneous Dirichlet values on the device load path against the oracle (the oracle's treatment is
pinned by the reference's elastic-strip golden values, tests/test_oracle_golden.py).

The reference's semantics (newtonraphson.hh:215-218,236-237; nonlinearsolverfactory.hh:45-54;
functionhelper.hh:170-184): each load step adds step * K_raw(u) ubar (constrained entries zeroed) to its
first residual and resets the constrained DOFs to lambda * ubar after every correction. Bars: Newton
iterations per step identical, converged displacements within 1e-8 relative.
"""
import numpy as np
import pytest

import ikarus_amd as ia
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _stretch_problem(material, eas, n=(4, 3, 3)):
    """z=0 clamped, traction on z=1, and the x=1 face (above z=0) pulled to u_x = 0.05 lambda."""
    p = pb.hex_block(n, material=material, jitter=0.05, seed=31, eas=eas, mag_range=(5.0, 10.0))
    on = (np.abs(p.X[:, 0] - 1.0) < 1e-8) & (p.X[:, 2] > 1e-8)
    ubar = np.zeros(p.ndof)
    p.fixed[np.nonzero(on)[0] * 3] = 1
    ubar[np.nonzero(on)[0] * 3] = 0.05
    return p, ubar


def _oracle_for(p):
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    return o


@pytest.mark.parametrize("material,eas", [(capi.NEOHOOKE, 0), (capi.SVK, 0), (capi.NEOHOOKE, 21)])
def test_inhomogeneous_dirichlet_load_control(material, eas):
    p, ubar = _stretch_problem(material, eas)
    print(f"case: IDBC load control material={material} eas={eas}", flush=True)
    o = _oracle_for(p)
    o.set_ubar(ubar)
    r = o.load_control(4, 0.0, 1.0, tol=1e-8)
    assert r["success"]
    dev = ia.Device(p)
    dev.set_dirichlet_values(ubar)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 4, 0.0, 1.0, capi.PRED_CONSTANT
    info, iters, res = dev.load_control(cfg, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13)))
    ug, lam = dev.get_requirement()
    dev.close()
    assert info.success and abs(lam - 1.0) < 1e-14
    assert list(iters) == list(r["iterations"])
    on = ubar != 0
    assert np.array_equal(ug[on], ubar[on] * 1.0)  # the constrained values are set exactly
    assert np.linalg.norm(ug - r["u"]) / np.linalg.norm(r["u"]) < 1e-8


def test_inhomogeneous_dirichlet_mixed_precision():
    """The mixed form (FP32 tangent and PCG, FP64 refinement to 1e-10) on the same path: same Newton
    iterations as the oracle and displacements within 1e-8."""
    p, ubar = _stretch_problem(capi.NEOHOOKE, 0)
    print("case: IDBC load control, mixed precision", flush=True)
    o = _oracle_for(p)
    o.set_ubar(ubar)
    r = o.load_control(4, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_dirichlet_values(ubar)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 4, 0.0, 1.0, capi.PRED_CONSTANT
    s = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                       linear=ia.solver_opts(kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    dev.close()
    assert info.success and list(iters) == list(r["iterations"])
    assert np.linalg.norm(ug - r["u"]) / np.linalg.norm(r["u"]) < 1e-8


def test_inhomogeneous_dirichlet_rejects_free_dofs():
    p, ubar = _stretch_problem(capi.NEOHOOKE, 0, n=(2, 2, 2))
    print("case: IDBC on a free DOF", flush=True)
    dev = ia.Device(p)
    bad = ubar.copy()
    bad[np.nonzero(p.fixed == 0)[0][0]] = 1.0
    with pytest.raises(RuntimeError):
        dev.set_dirichlet_values(bad)
    dev.set_dirichlet_values(None)
    dev.close()
