"""CPU oracle of arc-length path following (numpy, dense, small systems). TEST INFRASTRUCTURE ONLY: the
checker for the device path-following entry point; the product never imports it.

Restates, step by step:
  PathFollowing::run (ikarus/controlroutines/pathfollowing.inl:28-77): a solve at the start state, the
    initial prediction and its solve, then per further step the intermediate prediction and a solve;
  ArcLength (ikarus/controlroutines/pathfollowingfunctions.hh:88-190): f = sqrt(DD.DD + psi^2 Dl^2) - s,
    df/dDD = DD / root, df/dDl = psi^2 Dl / root once the initial prediction has run (zero before); the
    initial prediction solves K DD = -R at lambda = 1, sets psi = |DD|, scales DD and Dl = 1 to the arc
    length s, and rescales psi = |DD| / Dl; the intermediate prediction adds the last step's DD and Dl;
  NewtonRaphsonWithSubsidiaryFunction::solve (ikarus/solver/nonlinearsolver/
    newtonraphsonwithscalarsubsidiaryfunction.hh:190-307): Fext0 = -R(0, 1); while ||R|| (then
    sqrt(||R||^2 + f^2)) > tol: K [dR dL] = [-R Fext0], dl = (-f - df/dDD.dR) / (df/dDD.dL + df/dDl),
    u += dR + dl dL, lambda += dl, DD and Dl accumulate the same increments.
Inhomogeneous Dirichlet values and adaptive step sizing (both absent from the device path) are left out.
"""
import numpy as np


def arc_length(residual, jacobian, u0, steps, step_size, tol=1e-8, max_iter=20, solve=np.linalg.solve):
    """residual(u, lam) -> R, jacobian(u, lam) -> dense K. Returns u, lam, the Newton iterations of every
    solve (the start-state solve first), the load factor after each solve and success."""
    u = np.array(u0, dtype=np.float64)
    lam = 0.0
    DD = np.zeros_like(u)
    Dl = 0.0
    psi = 0.0
    predicted = False
    iterations, lambdas = [], []

    def subsidiary():
        root = np.sqrt(DD @ DD + psi * psi * Dl * Dl)
        f = root - step_size
        if not predicted:
            return f, np.zeros_like(DD), 0.0
        return f, DD / root, psi * psi * Dl / root

    def newton():
        nonlocal u, lam, DD, Dl
        fext0 = -residual(np.zeros_like(u), 1.0)
        rx = residual(u, lam)
        K = jacobian(u, lam)
        f, dfdDD, dfdDl = subsidiary()
        r_norm = np.sqrt(rx @ rx)
        it = 0
        while r_norm > tol and it < max_iter:
            dR = solve(K, -rx)
            dL = solve(K, fext0)
            dl = (-f - dfdDD @ dR) / (dfdDD @ dL + dfdDl)
            d = dR + dl * dL
            u = u + d
            DD = DD + d
            lam += dl
            Dl += dl
            rx = residual(u, lam)
            K = jacobian(u, lam)
            f, dfdDD, dfdDl = subsidiary()
            r_norm = np.sqrt(rx @ rx + f * f)
            it += 1
        iterations.append(it)
        lambdas.append(lam)
        return it != max_iter

    ok = newton()
    if ok:
        # initial prediction at lambda = 1 (the load factor is set, not incremented)
        lam = 1.0
        DD = solve(jacobian(u, lam), -residual(u, lam))
        dd2 = DD @ DD
        psi = np.sqrt(dd2)
        s = np.sqrt(dd2 + psi * psi)
        DD = DD * step_size / s
        Dl = step_size / s
        u = u + DD
        lam = Dl
        psi = np.sqrt(DD @ DD / (Dl * Dl))
        predicted = True
        ok = newton()
        for _ in range(1, steps):
            if not ok:
                break
            u = u + DD
            lam += Dl
            ok = newton()
    return dict(u=u, lam=lam, iterations=iterations, lambdas=lambdas, success=bool(ok))
