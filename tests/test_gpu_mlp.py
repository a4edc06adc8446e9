"""GPU MLP kernels (BF16 MFMA, FP32 accumulation) against the numpy oracle, and the predictor in LoadControl."""
import numpy as np
import pytest

import ikarus_amd as ia
import mlp_oracle as mo
from ikarus_amd import _capi as capi
from ikarus_amd import learned
from ikarus_amd import problems as pb

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 14)).astype(np.float32)
    t = (0.3 * np.sin(x[:, :3]) + 0.1 * x[:, 3:6] * x[:, 6:7]).astype(np.float32)
    return x, t


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [1, 37, 1000])
def test_forward_matches_oracle(n):
    print(f"case: mlp forward n={n}", flush=True)
    m = learned.MLP(seed=3)
    p = m.params()
    x, _ = _data(n, 1)
    y = m.forward(x)
    yo, _, _ = mo.forward(p, x, bf16=True)
    assert _rel(y, yo) < 2e-3
    y64, _, _ = mo.forward(p.astype(np.float64), x.astype(np.float64), bf16=False)
    assert _rel(y, y64) < 5e-2  # BF16 operands vs exact FP64


@pytest.mark.parametrize("n", [64, 3000])
def test_gradient_matches_oracle(n):
    print(f"case: mlp gradient n={n}", flush=True)
    m = learned.MLP(seed=4)
    p = m.params()
    x, t = _data(n, 2)
    loss, g = m.loss_grad(x, t)
    lo, go = mo.loss_grad(p, x, t, bf16=True)
    assert abs(loss - lo) <= 1e-3 * lo
    for name, sl in _slices():
        assert _rel(g[sl], go[sl]) < 1e-2, name
    _, g64 = mo.loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), bf16=False)
    assert _rel(g, g64) < 5e-2


def _slices():
    o = 0
    for name, shp in mo.SHAPES:
        k = int(np.prod(shp))
        yield name, slice(o, o + k)
        o += k


def test_adam_step_matches_oracle_and_training_reduces_loss():
    print("case: mlp training", flush=True)
    m = learned.MLP(seed=5)
    p0 = m.params().astype(np.float64)
    x, t = _data(4096, 3)
    lo, go = mo.loss_grad(p0.astype(np.float32), x, t, bf16=True)
    l1 = m.train_step(x, t, 1e-3)
    p1 = m.params()
    po1, _, _ = mo.adam(p0, go, np.zeros_like(p0), np.zeros_like(p0), 1, lr=1e-3)
    # the first Adam step moves every parameter by about lr in the direction of -sign(g)
    moved = np.abs(p1 - p0)
    big = np.abs(go) > 1e-3 * np.abs(go).max()
    assert np.allclose(moved[big], 1e-3, rtol=0.05)
    assert np.mean(np.sign(p1 - p0)[big] == np.sign(po1 - p0)[big]) > 0.98
    for _ in range(300):
        last = m.train_step(x, t, 2e-3)
    assert last < 0.2 * l1


def test_zero_network_predictor_equals_secant_predictor():
    """With W4 = b4 = 0 the MLP predicts v_{n+1} = v_n, i.e. u_n + du_n: the secant predictor."""
    print("case: mlp predictor vs secant", flush=True)
    p = pb.hex_block((4, 4, 4), seed=8)
    params = learned.MLP(seed=1).params()
    d = mo.unpack(params.astype(np.float64))
    d["W4"][:] = 0
    d["b4"][:] = 0
    params = mo.pack(d).astype(np.float32)
    res = {}
    for pred in (capi.PRED_LINEAR, capi.PRED_MLP):
        dev = ia.Device(p)
        if pred == capi.PRED_MLP:
            dev.check(capi.lib().ik_mlp_set_params(dev.h, capi.ptr(params), capi.HOST))
        dev.set_requirement(np.zeros(p.ndof), 0.0)
        cfg = capi.LCConfig()
        cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 6, 0.0, 1.0, pred
        info, iters, r = dev.load_control(cfg, ia.nr_settings(tol=1e-9, linear=ia.solver_opts(rtol=1e-12)))
        u, _ = dev.get_requirement()
        res[pred] = (list(iters), u)
    assert res[capi.PRED_LINEAR][0] == res[capi.PRED_MLP][0]
    assert _rel(res[capi.PRED_MLP][1], res[capi.PRED_LINEAR][1]) < 1e-12


def test_trained_predictor_reduces_newton_iterations():
    """Train on solver paths at 6^3 and evaluate on an unseen seed at 8^3: the learned predictor needs no
    more Newton iterations than the secant predictor and fewer than the constant one."""
    print("case: mlp train + eval small", flush=True)
    rng = np.random.default_rng(0)
    xs, ts = [], []
    for s in range(12):
        p = pb.hex_block((6, 6, 6), seed=300 + s)
        hist = learned.solve_path(p, 10)
        x, t = learned.samples_from_path(p, hist, 128, rng)
        xs.append(x)
        ts.append(t)
    m, losses = learned.train(np.concatenate(xs), np.concatenate(ts), epochs=60, batch=1024, lr=3e-3)
    assert losses[-1] < 0.5 * losses[0]
    params = m.params()
    p = pb.hex_block((8, 8, 8), seed=999)
    tot = {}
    for pred in (capi.PRED_CONSTANT, capi.PRED_LINEAR, capi.PRED_MLP):
        dev = ia.Device(p)
        if pred == capi.PRED_MLP:
            dev.check(capi.lib().ik_mlp_set_params(dev.h, capi.ptr(params), capi.HOST))
        dev.set_requirement(np.zeros(p.ndof), 0.0)
        cfg = capi.LCConfig()
        cfg.load_steps, cfg.tbegin, cfg.tend, cfg.predictor = 10, 0.0, 1.0, pred
        info, iters, res = dev.load_control(cfg, ia.nr_settings(tol=1e-8, linear=ia.solver_opts(rtol=1e-12)))
        assert info.success and res[-1] <= 1e-8
        tot[pred] = info.total_iterations
    print("newton iterations constant/secant/mlp:", tot, flush=True)
    assert tot[capi.PRED_MLP] <= tot[capi.PRED_LINEAR] < tot[capi.PRED_CONSTANT]


def test_gradients_are_reproducible():
    """The batch-split weight and bias gradients are summed in a fixed slice order: two evaluations on the same
    batch give bit-identical gradients (and so training is reproducible)."""
    print("case: mlp gradient reproducibility", flush=True)
    m = learned.MLP(seed=7)
    x, t = _data(20000, 4)
    l1, g1 = m.loss_grad(x, t)
    l2, g2 = m.loss_grad(x, t)
    assert l1 == l2 and np.array_equal(g1, g2)
    m2 = learned.MLP(seed=7)
    for _ in range(3):
        m.train_step(x, t, 1e-3)
        m2.train_step(x, t, 1e-3)
    assert np.array_equal(m.params(), m2.params())


@pytest.mark.parametrize("n", [(3, 4, 5), (33, 32, 32)])
def test_fused_predictor_matches_oracle(n):
    """ik_mlp_predict (features, four layers and the update of u in one kernel) against the oracle's BF16 forward
    on the host-built features, for no, one and two previous increments: the network output recovered from the
    update agrees to 2e-3 (as the layer-by-layer forward) and with ik_mlp_forward on the same features to 2e-4.
    Both round the same operands to BF16, but the fused kernel sums each 32-input step in another FP32 order,
    which now and then flips a BF16 rounding of an activation (3.7e-5 measured at 37026 nodes).
    (33, 32, 32) has 37026 nodes, more than one pass of the 512-workgroup grid."""
    p = pb.hex_block(n, jitter=0.05, seed=11)
    print(f"case: fused predictor {n}, {p.X.shape[0]} nodes", flush=True)
    m = learned.MLP(seed=6)
    params = m.params()
    rng = np.random.default_rng(12)
    free = p.fixed == 0
    u0 = 0.01 * rng.standard_normal(p.ndof) * free
    du_n = 0.002 * rng.standard_normal(p.ndof) * free
    du_n1 = 0.002 * rng.standard_normal(p.ndof) * free
    traction, char_len, dl = np.array([0.0, -3.0, 12.0]), 2.0, 0.05
    dev = ia.Device(p)
    dev.check(capi.lib().ik_mlp_set_params(dev.h, capi.ptr(params), capi.HOST))
    dev.check(capi.lib().ik_mlp_set_load(dev.h, capi.ptr(traction), char_len))
    for dprev, dprev2 in ((0.0, 0.0), (dl, 0.0), (dl, dl)):
        lam_next = 0.4 + dl
        dev.set_requirement(u0, 0.4)
        dev.check(capi.lib().ik_mlp_predict(dev.h, capi.ptr(du_n), capi.ptr(du_n1), lam_next, dl, dprev, dprev2,
                                            capi.HOST))
        ug, lam = dev.get_requirement()
        assert lam == 0.4
        assert np.array_equal(ug[~free], u0[~free])
        vn = du_n / dprev if dprev > 0 else np.zeros(p.ndof)
        y = ((ug - u0) / dl - vn) / 0.1
        f = learned.features(p.X, traction, lam_next, dl, du_n, du_n1, dprev, dprev2, char_len)
        yo, _, _ = mo.forward(params, f, bf16=True)
        yt = m.forward(f)
        yo, yt = yo.reshape(-1)[free], yt.reshape(-1)[free]
        assert _rel(y[free], yo) < 2e-3
        assert _rel(y[free], yt) < 2e-4
    dev.close()


def test_fused_predictor_errors():
    p = pb.hex_block((2, 2, 2), seed=1)
    print("case: fused predictor errors", flush=True)
    dev = ia.Device(p)
    z = np.zeros(p.ndof)
    L = capi.lib()
    assert L.ik_mlp_predict(dev.h, capi.ptr(z), capi.ptr(z), 0.1, 0.1, 0.0, 0.0, capi.HOST) == capi.IK_EINVAL
    dev.check(L.ik_mlp_set_params(dev.h, capi.ptr(learned.MLP(seed=1).params()), capi.HOST))
    assert L.ik_mlp_predict(dev.h, capi.ptr(z), capi.ptr(z), 0.1, 0.1, 0.0, 0.0, capi.HOST) == capi.IK_ENOTBOUND
    dev.close()
    # plane-strain handles: the network's features are 3D
    p2 = pb.cantilever2d((8, 2), seed=1)
    dev = ia.Device(p2)
    dev.check(L.ik_mlp_set_params(dev.h, capi.ptr(learned.MLP(seed=1).params()), capi.HOST))
    dev.set_requirement(np.zeros(p2.ndof), 0.0)
    z2 = np.zeros(3 * dev.ndof)
    assert L.ik_mlp_predict(dev.h, capi.ptr(z2), capi.ptr(z2), 0.1, 0.1, 0.0, 0.0, capi.HOST) == capi.IK_ENOTIMPL
    dev.close()
