"""EAS-21 (enhanced Green-Lagrange) hex8 on the GPU against the oracle and the reference's golden values."""
import numpy as np
import pytest

import ikarus_amd as ia
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _oracle_for(p):
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    if len(p.face_elem):
        o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    for d, c in zip(p.point_dof, p.point_coef):
        o.add_point_load(int(d), float(c))
    return o


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_eas_element_matrix_and_vector(material):
    """Condensed K_e, R_e of EnhancedAssumedStrains (enhancedassumedstrains.hh:258-348) at alpha != 0."""
    p = pb.hex_block((3, 3, 2), material=material, jitter=0.08, seed=13, eas=21)
    print(f"case: eas element material={material}", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    rng = np.random.default_rng(2)
    u = 0.02 * rng.standard_normal(p.ndof)
    # a large enhanced strain makes det C <= 0 in element 7: both sides must refuse it
    alpha = 1e-4 * rng.standard_normal((p.ne, 21))
    o.set_alpha(alpha)
    dev.set_alpha(alpha)
    dev.set_requirement(u, 0.0)
    if material == capi.NEOHOOKE:  # SVK has no det C check (svk.hh), Neo-Hooke aborts (neohooke.hh:90-93)
        with pytest.raises(RuntimeError):
            o.element(7, u, 0.0)
        with pytest.raises(capi.IkError) as ei:
            dev.element(7)
        assert ei.value.code == capi.IK_ENEGDET and ei.value.bad_element == 7
    alpha *= 1e-2
    o.set_alpha(alpha)
    dev.set_alpha(alpha)
    dev.set_requirement(u, 0.0)
    for e in (0, 7, 17):
        K, R = dev.element(e)
        Ko, Ro, *_ = o.element(e, u, 0.0)
        assert _rel(K, Ko) < 1e-10, e
        assert _rel(R, Ro) < 1e-10, e
        assert np.array_equal(K, K.T)  # mirrored from the upper triangle


def test_eas_assembly_and_alpha_update():
    p = pb.hex_block((4, 3, 3), jitter=0.05, seed=17, eas=21)
    print("case: eas assembly + update", flush=True)
    dev = ia.Device(p)
    o = _oracle_for(p)
    rng = np.random.default_rng(3)
    u = 0.01 * rng.standard_normal(p.ndof) * (1 - p.fixed)
    alpha = 1e-6 * rng.standard_normal((p.ne, 21))
    o.set_alpha(alpha)
    dev.set_alpha(alpha)
    dev.set_requirement(u, 0.8)
    dev.matrix(capi.DBC_FULL)
    outer, inner, vals = dev.csc()
    oo, io = o.pattern()
    assert np.array_equal(outer, oo) and np.array_equal(inner, io)
    assert _rel(vals, o.matrix(u, 0.8, po.FULL)) < 1e-10
    assert _rel(dev.vector(), o.vector(u, 0.8, po.FULL)) < 1e-10
    du = 1e-3 * rng.standard_normal(p.ndof) * (1 - p.fixed)
    dev.update_state(du)
    o.update_state(u, 0.8, du)
    assert _rel(dev.alpha(), o.alpha()) < 1e-10


def test_eas_fp32_matrix_is_the_rounded_fp64_matrix():
    """The EAS element computes and stages FP64 blocks; the FP32 matrix of the mixed solver is gathered from
    the same staging in the same order, so every entry is the FP64 entry rounded once."""
    p = pb.hex_block((4, 3, 3), jitter=0.05, seed=17, eas=21)
    print("case: eas fp32 matrix", flush=True)
    dev = ia.Device(p)
    rng = np.random.default_rng(5)
    dev.set_alpha(1e-6 * rng.standard_normal((p.ne, 21)))
    dev.set_requirement(0.01 * rng.standard_normal(p.ndof) * (1 - p.fixed), 0.8)
    dev.matrix(capi.DBC_FULL, precision=capi.FP64)
    _, _, v64 = dev.csc()
    dev.matrix(capi.DBC_FULL, precision=capi.FP32)
    _, _, v32 = dev.csc()
    assert np.array_equal(v32, v64.astype(np.float32).astype(np.float64))


def test_eas_mixed_precision_newton_matches_fp64():
    """Load control with the mixed solver (FP32 condensed tangent, FP32 block-Jacobi PCG, FP64 refinement)
    on an EAS-21 bending block reaches the FP64 residual tolerance and the plain FP64 path's solution. Both
    stop at ||R|| <= 1e-8 (FP64 residual), so the displacements differ by about K^-1 times that residual."""
    p = pb.bending_block((6, 4, 3), seed=3)
    print("case: eas mixed newton", flush=True)
    out = {}
    for prec in (capi.FP64, capi.FP32):
        dev = ia.Device(p)
        dev.set_requirement(np.zeros(p.ndof), 0.0)
        cfg = capi.LCConfig()
        cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 3, 0.0, 1.0, capi.PRED_LINEAR
        kind = capi.PCG_JACOBI_FP64 if prec == capi.FP64 else capi.PCG_BJACOBI_FP32_IR
        s = ia.nr_settings(tol=1e-8, precision=prec,
                           linear=ia.solver_opts(kind=kind, rtol=1e-12, inner_rtol=1e-3, max_iter=20000))
        info, iters, res = dev.load_control(cfg, s)
        assert info.success and res[info.solves - 1] <= 1e-8
        out[prec] = (dev.get_requirement()[0], dev.alpha())
        dev.close()
    print("rel u", _rel(out[capi.FP32][0], out[capi.FP64][0]), "rel alpha",
          _rel(out[capi.FP32][1], out[capi.FP64][1]), flush=True)
    assert _rel(out[capi.FP32][0], out[capi.FP64][0]) < 1e-8
    assert _rel(out[capi.FP32][1], out[capi.FP64][1]) < 1e-6


# tests/src/testcantileverbeamEAS.cpp:61-67 with tests/src/testcantileverbeam.hh:83-198
@pytest.mark.parametrize("material,maxu", [(capi.NEOHOOKE, 4.781820664768682), (capi.SVK, 4.7692391315649365)])
def test_cantilever_eas21_golden_on_gpu(material, maxu):
    p = pb.cantilever3d(material=material)
    print(f"case: cantilever golden material={material}", flush=True)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 20, 0.0, 1.0, capi.PRED_CONSTANT
    s = ia.nr_settings(tol=1e-10, linear=ia.solver_opts(rtol=1e-15, max_iter=4000))
    info, iters, res = dev.load_control(cfg, s)
    u, lam = dev.get_requirement()
    assert info.success
    assert info.total_iterations == 80
    assert abs(np.max(np.abs(u)) - maxu) <= 1e-10 * maxu
    assert abs(lam - 1.0) < 1e-12


@pytest.mark.parametrize("material", ["nh", "svk"])
def test_cpp_mirror_api_cantilever(material):
    """The same golden test through the header-only C++ mirror (include/ikarus_amd/ikarus.hh)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "cpp", "_build",
                       "cantilever_eas")
    assert os.path.exists(exe), "C++ test program not built (run __graft_entry__.build())"
    print(f"case: C++ cantilever {material}", flush=True)
    r = subprocess.run([exe, material], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
