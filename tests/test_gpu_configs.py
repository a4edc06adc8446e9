"""The benchmark's larger configurations on the GPU: EAS-21 element matrices on the full 96x96x48 mesh
against the oracle, and the precision variant for both materials against the oracle's FP64 path."""
import numpy as np
import pytest

import ikarus_amd as ia
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_eas_96x96x48_element_matrices():
    """Condensed EAS-21 K_e, R_e on the full bending mesh (442,368 elements) at a perturbed state with
    alpha != 0, sampled elements (first, last, clamped face, loaded face, interior) to 1e-10."""
    p = pb.bending_block((96, 96, 48), seed=0)
    print(f"case: eas 96x96x48 elements ({p.ne})", flush=True)
    dev = ia.Device(p)
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, p.eas, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    rng = np.random.default_rng(11)
    u = 2e-4 * rng.standard_normal(p.ndof) * (1 - p.fixed)  # |grad u| stays well below 1 at h = 1/48
    # the enhanced-strain operator T0^-1 M~ / detJ (T0 carrying detJ0) is of order 1e13 at h = 1/48, so
    # alpha ~ 1e-16 gives enhanced strains ~ 1e-3 (K_e changes by ~0.3 %, R_e by ~1 %)
    alpha = 1e-16 * rng.standard_normal((p.ne, 21))
    o.set_alpha(alpha)
    dev.set_alpha(alpha)
    dev.set_requirement(u, 1.0)
    nx, ny = 96, 96
    picks = [0, p.ne - 1, nx - 1, 95 * nx * ny // 2, p.ne // 2 + 17] + list(rng.integers(0, p.ne, 27))
    for e in picks:
        K, R = dev.element(int(e))
        Ko, Ro, *_ = o.element(int(e), u, 1.0)
        assert _rel(K, Ko) < 1e-10, e
        assert _rel(R, Ro) < 1e-10, e


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_precision_variant_matches_fp64_oracle(material):
    """FP32 tangent + FP32 block-Jacobi PCG + FP64 refinement (plain Newton, constant predictor) reaches
    ||R|| <= 1e-8 with the oracle's iteration counts and its FP64 displacement to 1e-8."""
    p = pb.hex_block((8, 8, 6), material=material, jitter=0.05, seed=21)
    print(f"case: precision variant material={material}", flush=True)
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, 0, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    r = o.load_control(3, 0.0, 1.0, tol=1e-8)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    cfg = capi.LCConfig()
    cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 3, 0.0, 1.0, capi.PRED_CONSTANT
    s = ia.nr_settings(tol=1e-8, precision=capi.FP32,
                       linear=ia.solver_opts(kind=capi.PCG_BJACOBI_FP32_IR, rtol=1e-10, inner_rtol=1e-3))
    info, iters, res = dev.load_control(cfg, s)
    ug, _ = dev.get_requirement()
    assert info.success and res[info.solves - 1] <= 1e-8
    assert list(iters[:info.solves]) == list(r["iterations"])
    assert _rel(ug, r["u"]) < 1e-8
