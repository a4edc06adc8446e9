"""Edge cases of the device path against the FP64 CPU oracle, through the C ABI: a single element, meshes one
element thick in one or two directions (every node on the boundary, tiles mostly padding), a mesh smaller than one
16-element tile, EAS on such meshes, and a fully clamped mesh (no free DOF: empty REDUCED system, Newton done at
once)."""
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


@pytest.mark.parametrize("n,material,eas", [
    ((1, 1, 1), capi.NEOHOOKE, 0), ((1, 1, 5), capi.SVK, 0), ((7, 1, 1), capi.NEOHOOKE, 0),
    ((2, 3, 1), capi.NEOHOOKE, 0), ((1, 2, 1), capi.NEOHOOKE, 21), ((3, 1, 2), capi.SVK, 21)])
def test_thin_meshes_match_oracle(n, material, eas):
    p = pb.hex_block(n, L=(0.7, 1.1, 0.9), material=material, eas=eas, seed=31, mag_range=(5.0, 10.0))
    print(f"case: thin mesh {n} material={material} eas={eas}", flush=True)
    o = _oracle_for(p)
    dev = ia.Device(p)
    u = 0.01 * np.random.default_rng(2).standard_normal(p.ndof) * (1 - p.fixed)
    lam = 0.6
    dev.set_requirement(u, lam)
    for dbc in (capi.DBC_RAW, capi.DBC_FULL, capi.DBC_REDUCED):
        assert _rel(dev.vector(dbc), o.vector(u, lam, dbc)) < 1e-10, dbc
        dev.matrix(dbc)
        outer, inner, vals = dev.csc(dbc)
        oo, io = o.pattern(dbc == capi.DBC_REDUCED)
        assert np.array_equal(outer, oo) and np.array_equal(inner, io), dbc
        assert _rel(vals, o.matrix(u, lam, dbc)) < 1e-10, dbc
    # a short load path from rest
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    if eas:
        dev.set_alpha(np.zeros((p.ne, eas)))
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 2, 0.0, 1.0, capi.PRED_CONSTANT
    info, iters, res = dev.load_control(cfg, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-13)))
    ug, _ = dev.get_requirement()
    # the same path by the mixed form with the two-level preconditioner (aggregates of a few nodes, some modes
    # dropped on these thin meshes)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    if eas:
        dev.set_alpha(np.zeros((p.ne, eas)))
    s2 = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                        linear=ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info2, iters2, res2 = dev.load_control(cfg, s2)
    u2, _ = dev.get_requirement()
    dev.close()
    o2 = _oracle_for(p)
    r = o2.load_control(2, 0.0, 1.0, tol=1e-8)
    assert info.success and r["success"] and list(iters) == list(r["iterations"])
    assert _rel(ug, r["u"]) < 1e-8
    assert info2.success and res2[info2.solves - 1] <= 1e-8
    assert _rel(u2, r["u"]) < 1e-8


def test_fully_clamped_mesh():
    """Every DOF constrained: the REDUCED system is empty, FULL K is the identity, R vanishes and Newton needs no
    iteration at any load step."""
    p = pb.hex_block((2, 2, 1), seed=1)
    p.fixed = np.ones(p.ndof, np.uint8)
    print("case: fully clamped", flush=True)
    dev = ia.Device(p)
    assert int(capi.lib().ik_reduced_size(dev.h)) == 0
    dev.set_requirement(np.zeros(p.ndof), 1.0)
    assert np.all(dev.vector(capi.DBC_FULL) == 0.0)
    assert dev.vector(capi.DBC_REDUCED).size == 0
    dev.matrix(capi.DBC_FULL)
    outer, inner, vals = dev.csc(capi.DBC_FULL)
    K = np.zeros((p.ndof, p.ndof))
    for c in range(p.ndof):
        K[inner[outer[c]:outer[c + 1]], c] = vals[outer[c]:outer[c + 1]]
    assert np.array_equal(K, np.eye(p.ndof))
    dev.matrix(capi.DBC_REDUCED)
    outer, inner, vals = dev.csc(capi.DBC_REDUCED)
    assert len(outer) == 1 and len(inner) == 0
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 3, 0.0, 1.0, capi.PRED_LINEAR
    info, iters, res = dev.load_control(cfg, ia.nr_settings(tol=1e-8))
    assert info.success and np.all(iters == 0) and np.all(res == 0.0)
    # two-level solve with every rigid-body mode dropped (no free DOF): the coarse part vanishes and K = I
    dev.matrix(capi.DBC_FULL, capi.FP32)
    b = np.linspace(1.0, 2.0, p.ndof)
    x, sinfo = dev.solve(b, ia.solver_opts(kind=capi.PCG_TWOLEVEL_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    dev.close()
    assert sinfo.converged == 1 and _rel(x, b) < 1e-9
