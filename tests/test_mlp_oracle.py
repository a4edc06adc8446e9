"""The MLP oracle pinned by finite differences and by closed-form checks (CPU only)."""
import numpy as np

import mlp_oracle as mo


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(mo.N_PARAMS) * scale


def test_pack_roundtrip_and_count():
    p = _params()
    assert mo.N_PARAMS == 14 * 128 + 128 + 2 * (128 * 128 + 128) + 3 * 128 + 3
    np.testing.assert_array_equal(mo.pack(mo.unpack(p)), p)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(1, 0.1)
    x = rng.standard_normal((16, 14))
    t = rng.standard_normal((16, 3))
    loss, g = mo.loss_grad(p, x, t, bf16=False)
    h = 1e-6
    for i in rng.choice(mo.N_PARAMS, 40, replace=False):
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss_grad(p + e, x, t)[0] - mo.loss_grad(p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9


def test_silu_closed_form_and_zero_network():
    z = np.linspace(-5, 5, 11)
    np.testing.assert_allclose(mo.silu(z), z / (1 + np.exp(-z)))
    h = 1e-6
    np.testing.assert_allclose(mo.dsilu(z), (mo.silu(z + h) - mo.silu(z - h)) / (2 * h), rtol=1e-6, atol=1e-8)
    p = _params(2)
    d = mo.unpack(p)
    d["W4"][:] = 0
    d["b4"][:] = 0.5
    y, _, _ = mo.forward(mo.pack(d), np.ones((3, 14)))
    np.testing.assert_array_equal(y, 0.5)


def test_bf16_rounding_ties_to_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9], np.float32)
    r = mo.to_bf16(vals)
    np.testing.assert_array_equal(r, np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0], np.float32))


def test_adam_first_step_is_signed_lr():
    p = np.zeros(4)
    g = np.array([1.0, -2.0, 0.5, 0.0])
    p2, m, v = mo.adam(p, g, np.zeros(4), np.zeros(4), 1, lr=0.1)
    np.testing.assert_allclose(p2[:3], -0.1 * np.sign(g[:3]), rtol=1e-6)
    assert p2[3] == 0.0
