"""Learned load-step predictor: data generation from the solver's own load paths, GPU training through
the C ABI (ik_mlp_*), and evaluation inside LoadControl (IK_PRED_MLP).

The network and the feature definition are fixed by libikarus_amd.so (include/ikarus_amd.h): per node,
  x = [X/L (3), t/40 (3), lambda_{n+1}, 40 dlambda, v_n/0.1 (3), v_{n-1}/0.1 (3)],  v = du/dlambda,
  y = (v_{n+1} - v_n) / 0.1,   u_pred = u_n + dlambda (v_n + 0.1 y).

Usage (on a GPU):  python -m ikarus_amd.learned train --paths 256 --n 32 --steps 40 --out <file.f32>
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

from . import _capi as capi
from . import problems as pb
from ._capi import IkError, ptr
from .assembler import Device, nr_settings, solver_opts

VEL_SCALE = 0.1
TRAC_SCALE = 1.0 / 40.0
DLAM_SCALE = 40.0
N_IN, N_OUT = 14, 3
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "predictor_mlp_32cubed.f32")


class MLP:
    """Training context (ik_mlp): FP32 master parameters, BF16-MFMA forward/backward, Adam."""

    def __init__(self, seed=0):
        L = capi.lib()
        h = C.c_void_p()
        rc = L.ik_mlp_create(C.byref(h), C.c_uint64(seed))
        self.h = h
        self._check(rc)

    def _check(self, rc):
        if rc != 0:
            raise IkError(rc, capi.lib().ik_mlp_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            capi.lib().ik_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        p = np.zeros(capi.lib().ik_mlp_num_params(), np.float32)
        self._check(capi.lib().ik_mlp_params(self.h, ptr(p), capi.HOST))
        return p

    def load(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(capi.lib().ik_mlp_load(self.h, ptr(p), capi.HOST))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], N_OUT), np.float32)
        self._check(capi.lib().ik_mlp_forward(self.h, ptr(x), x.shape[0], ptr(y), capi.HOST))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(capi.lib().ik_mlp_num_params(), np.float32)
        loss = C.c_double()
        self._check(capi.lib().ik_mlp_loss_grad(self.h, ptr(x), ptr(t), x.shape[0], C.byref(loss), ptr(g), capi.HOST))
        return loss.value, g

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_double()
        self._check(capi.lib().ik_mlp_train_step(self.h, ptr(x), ptr(t), x.shape[0], C.c_float(lr), C.byref(loss),
                                                 capi.HOST))
        return loss.value


def features(X, traction, lam_next, dlam, du_n, du_n1, dlam_prev, dlam_prev2, char_len=1.0):
    """Host restatement of the device feature builder (same scalings, FP32 output)."""
    nn = X.shape[0]
    f = np.zeros((nn, N_IN), np.float32)
    f[:, 0:3] = X / char_len
    f[:, 3:6] = np.asarray(traction) * TRAC_SCALE
    f[:, 6] = lam_next
    f[:, 7] = dlam * DLAM_SCALE
    if dlam_prev > 0:
        f[:, 8:11] = (du_n.reshape(nn, 3) / dlam_prev / VEL_SCALE)
    if dlam_prev2 > 0:
        f[:, 11:14] = (du_n1.reshape(nn, 3) / dlam_prev2 / VEL_SCALE)
    return f


def targets(du_next, du_n, dlam, dlam_prev):
    v_next = du_next / dlam
    v_n = du_n / dlam_prev if dlam_prev > 0 else np.zeros_like(du_next)
    return ((v_next - v_n) / VEL_SCALE).reshape(-1, 3).astype(np.float32)


def mixed_settings(tol=1e-8):
    return nr_settings(tol=tol, precision=capi.FP32,
                       linear=solver_opts(capi.PCG_BJACOBI_FP32_IR, rtol=1e-6, inner_rtol=1e-3, max_refine=20),
                       forcing=1e-2)


def solve_path(p, steps, settings=None):
    """Load path with the secant predictor, one Newton solve per step; returns the converged u_0..u_steps."""
    dev = Device(p)
    s = settings if settings is not None else mixed_settings()
    u = np.zeros(p.ndof)
    lam = 0.0
    dl = 1.0 / steps
    hist = [u.copy()]
    free = p.fixed == 0
    for n in range(steps):
        pred = u + (hist[-1] - hist[-2] if n > 0 else 0.0)
        pred = np.where(free, pred, 0.0)
        lam += dl
        dev.set_requirement(pred, lam)
        info = dev.newton(s, dl)
        if not info.success:
            raise RuntimeError(f"Newton failed on step {n}")
        u, _ = dev.get_requirement()
        hist.append(u.copy())
    dev.close()
    return hist


def samples_from_path(p, hist, nodes_per_step, rng):
    """(features, targets) for every step of one path on a random subset of nodes."""
    steps = len(hist) - 1
    dl = 1.0 / steps
    nn = p.X.shape[0]
    free_nodes = np.nonzero(p.fixed.reshape(nn, 3).sum(1) == 0)[0]
    xs, ts = [], []
    for n in range(steps):
        du_n = hist[n] - hist[n - 1] if n >= 1 else np.zeros(p.ndof)
        du_n1 = hist[n - 1] - hist[n - 2] if n >= 2 else np.zeros(p.ndof)
        du_next = hist[n + 1] - hist[n]
        sel = rng.choice(free_nodes, size=min(nodes_per_step, len(free_nodes)), replace=False)
        f = features(p.X, p.traction, (n + 1) * dl, dl, du_n, du_n1, dl if n >= 1 else 0.0, dl if n >= 2 else 0.0)
        t = targets(du_next, du_n, dl, dl if n >= 1 else 0.0)
        xs.append(f[sel])
        ts.append(t[sel])
    return np.concatenate(xs), np.concatenate(ts)


def train(x, t, epochs=20, batch=8192, lr=2e-3, seed=0, log=None):
    m = MLP(seed)
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    nb = max(1, n // batch)
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        lr_ep = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
        tot = 0.0
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            tot += m.train_step(x[idx], t[idx], lr_ep)
        losses.append(tot / nb)
        if log:
            log(f"epoch {ep + 1}/{epochs} lr {lr_ep:.2e} loss {losses[-1]:.4e}")
    return m, losses


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--paths", type=int, default=256)
    tr.add_argument("--n", type=int, default=32)
    tr.add_argument("--steps", type=int, default=40)
    tr.add_argument("--nodes-per-step", type=int, default=1024)
    tr.add_argument("--epochs", type=int, default=40)
    tr.add_argument("--seed0", type=int, default=1000)
    tr.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args(argv)

    def log(msg):
        print(f"[learned] {msg}", flush=True)

    rng = np.random.default_rng(7)
    xs, ts = [], []
    t0 = time.time()
    for i in range(args.paths):
        p = pb.hex_block((args.n,) * 3, seed=args.seed0 + i)
        hist = solve_path(p, args.steps)
        x, t = samples_from_path(p, hist, args.nodes_per_step, rng)
        xs.append(x)
        ts.append(t)
        if (i + 1) % 8 == 0 or i == 0:
            log(f"path {i + 1}/{args.paths} done ({time.time() - t0:.0f} s), samples so far {sum(len(a) for a in xs)}")
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    log(f"training on {len(x)} samples")
    m, losses = train(x, t, epochs=args.epochs, log=log)
    p = m.params()
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    p.tofile(args.out)
    log(f"wrote {args.out} ({p.nbytes} bytes), final loss {losses[-1]:.4e}, target variance {float(np.mean(t ** 2)):.4e}")


if __name__ == "__main__":
    main(sys.argv[1:])
