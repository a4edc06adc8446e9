"""The assembler's scalar (potential energy, interface.hh:283-330) on the device against the oracle, whose
energy is pinned by R = dPi/du (tests/test_oracle_golden.py), and the same derivative check on the device
itself. Bars: 1e-12 relative to the oracle; the finite-difference slope within 1e-6 of the residual."""
import numpy as np
import pytest

import ikarus_amd as ia
import pyoracle as po
from ikarus_amd import _capi as capi
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_energy_matches_oracle(material):
    p = pb.hex_block((5, 4, 3), material=material, jitter=0.08, seed=51, traction=(4.0, -3.0, 6.0))
    print(f"case: energy material={material}", flush=True)
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, 0, 2)
    o.set_fixed(p.fixed)
    o.set_traction(p.face_elem, p.face_axis, p.face_side, p.traction)
    u = 0.02 * np.random.default_rng(2).standard_normal(p.ndof)
    lam = 0.7
    asm = ia.SparseFlatAssembler(p)
    req = ia.FERequirement(u, lam)
    E = asm.scalar(req)
    Eo = o.energy(u, lam)
    assert abs(E - Eo) <= 1e-12 * abs(Eo)
    # R = dPi/du on the device (RAW residual), central differences at loaded and random DOFs
    dev = asm.dev
    R = asm.vector(req, dbc=capi.DBC_RAW)
    h = 1e-6
    rng = np.random.default_rng(3)
    top = np.nonzero(p.X[:, 2] > 1 - 1e-9)[0] * 3 + 2
    for j in np.concatenate([top[:3], rng.choice(p.ndof, 4, replace=False)]):
        e = np.zeros(p.ndof)
        e[j] = h
        dev.set_requirement(u + e, lam)
        ep = dev.energy()
        dev.set_requirement(u - e, lam)
        em = dev.energy()
        assert abs((ep - em) / (2 * h) - R[j]) <= 1e-6 * max(1.0, np.abs(R).max())


def test_energy_rejects_eas():
    p = pb.hex_block((2, 2, 2), eas=21, seed=3)
    print("case: energy EAS", flush=True)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    with pytest.raises(RuntimeError):
        dev.energy()
    dev.close()


@pytest.mark.parametrize("material", [capi.NEOHOOKE, capi.SVK])
def test_stress_at_points_matches_oracle(material):
    """PK2, Kirchhoff and Cauchy stresses at local points (nonlinearelastic.hh:237-270) against the oracle,
    whose stresses are pinned by the closed form of a homogeneous deformation. Bar 1e-12 relative."""
    p = pb.hex_block((4, 3, 3), material=material, jitter=0.08, seed=53)
    print(f"case: stresses material={material}", flush=True)
    o = po.Oracle(p.X, p.conn, p.material, p.lam, p.mu, 0, 2)
    u = 0.03 * np.random.default_rng(4).standard_normal(p.ndof)
    dev = ia.Device(p)
    dev.set_requirement(u, 0.5)
    for xi in ([0.5, 0.5, 0.5], [0.2113248654051871, 0.7886751345948129, 0.5], [0.0, 1.0, 0.3]):
        for kind in (capi.PK2, capi.KIRCHHOFF, capi.CAUCHY):
            s = dev.stress_at(xi, kind)
            so = o.stress_at(u, xi, kind)
            assert np.abs(s - so).max() <= 1e-12 * np.abs(so).max()
    dev.close()


def test_stress_at_rejects_eas():
    p = pb.hex_block((2, 2, 2), eas=21, seed=3)
    print("case: stresses EAS", flush=True)
    dev = ia.Device(p)
    dev.set_requirement(np.zeros(p.ndof), 0.0)
    with pytest.raises(RuntimeError):
        dev.stress_at([0.5, 0.5, 0.5])
    dev.close()
