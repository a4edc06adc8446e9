"""CPU oracle of arc-length path following (numpy, dense, small systems). TEST INFRASTRUCTURE ONLY: the
checker for the device path-following entry point; the product never imports it.

Restates, step by step:
  PathFollowing::run (ikarus/controlroutines/pathfollowing.inl:28-77): a solve at the start state, the
    initial prediction and its solve, then per further step the adaptive step-size rule, the intermediate
    prediction and a solve;
  ArcLength (ikarus/controlroutines/pathfollowingfunctions.hh:88-190): with the inhomogeneous Dirichlet increment
    of the step idbcDelta = Dl ubar (idbcIncrement, ikarus/controlroutines/common.hh:63-76),
    f = sqrt(DD.DD + |idbcDelta|^2 + psi^2 Dl^2) - s, df/dDD = DD / root, df/dDl = (|ubar|^2 + psi^2) Dl / root
    once the initial prediction has run (zero before); the initial prediction solves K DD = -(R + F_idbc) at
    lambda = 1, sets psi = sqrt(|DD|^2 + |ubar|^2), scales DD and Dl = 1 to the arc length s, moves the
    inhomogeneous DOFs to Dl ubar and rescales psi = sqrt(|Dl ubar|^2 + |DD|^2) / Dl; the intermediate prediction
    adds the last step's DD and Dl and moves the inhomogeneous DOFs;
  NewtonRaphsonWithSubsidiaryFunction::solve (ikarus/solver/nonlinearsolver/
    newtonraphsonwithscalarsubsidiaryfunction.hh:190-307): Fext0 = -R(0, 1) (with the current internal state:
    for EAS elements alpha is not reset); while ||R|| (then sqrt(||R||^2 + f^2)) > tol:
    K [dR dL] = [-R, Fext0 - F_idbc], dl = (-f - df/dDD.dR) / (df/dDD.dL + df/dDl), the state update on the
    correction (CORRECTION_UPDATED: EAS alpha), u += dR + dl dL, lambda += dl, the inhomogeneous DOFs set to
    lambda ubar, DD and Dl accumulate the increments;
  AdaptiveStepSizing::IterationBased (ikarus/controlroutines/adaptivestepsizing.hh:48-90): from the second
    step on, s <- sqrt(target / its) s when the last solve took its > target iterations.
F_idbc = K_raw(u) ubar with the constrained entries zeroed (utils::obtainForcesDueToIDBC,
ikarus/utils/functionhelper.hh:170-184).
"""
import numpy as np


def arc_length(residual, jacobian, u0, steps, step_size, tol=1e-8, max_iter=20, solve=np.linalg.solve, ubar=None,
               idbc_force=None, update_state=None, target_iterations=0):
    """residual(u, lam) -> R, jacobian(u, lam) -> dense K (FULL Dirichlet treatment when some DOFs are fixed).
    ubar: inhomogeneous Dirichlet values per unit load factor (nonzero on constrained DOFs only) and
    idbc_force(u, lam) -> K_raw(u) ubar with the constrained entries zeroed; update_state(u, lam, du) is called with
    every correction before the update (EAS). Returns u, lam, the Newton iterations of every solve (the start-state
    solve first), the load factor and the arc length after each solve, and success."""
    u = np.array(u0, dtype=np.float64)
    lam = 0.0
    DD = np.zeros_like(u)
    Dl = 0.0
    psi = 0.0
    s_cur = step_size
    predicted = False
    iterations, lambdas, sizes = [], [], []
    has_idbc = ubar is not None and np.any(ubar != 0)
    ub = np.zeros_like(u) if not has_idbc else np.asarray(ubar, dtype=np.float64)
    ub2 = float(ub @ ub)
    on = ub != 0

    def sync():
        if has_idbc:
            u[on] = lam * ub[on]

    def f_idbc():
        return idbc_force(u, lam) if has_idbc else 0.0

    def subsidiary():
        root = np.sqrt(DD @ DD + Dl * Dl * ub2 + psi * psi * Dl * Dl)
        f = root - s_cur
        if not predicted:
            return f, np.zeros_like(DD), 0.0
        return f, DD / root, (ub2 + psi * psi) * Dl / root

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
            dL = solve(K, fext0 - f_idbc())
            dl = (-f - dfdDD @ dR) / (dfdDD @ dL + dfdDl)
            d = dR + dl * dL
            if update_state is not None:
                update_state(u, lam, d)
            u = u + d
            DD = DD + d
            lam += dl
            Dl += dl
            sync()
            rx = residual(u, lam)
            K = jacobian(u, lam)
            f, dfdDD, dfdDl = subsidiary()
            r_norm = np.sqrt(rx @ rx + f * f)
            it += 1
        iterations.append(it)
        lambdas.append(lam)
        sizes.append(s_cur)
        return it != max_iter

    ok = newton()
    if ok:
        # initial prediction at lambda = 1 (the load factor is set, not incremented; u is not synced to it)
        lam = 1.0
        DD = solve(jacobian(u, lam), -(residual(u, lam) + f_idbc()))
        dd2 = DD @ DD
        psi = np.sqrt(dd2 + ub2)
        s = np.sqrt(dd2 + ub2 + psi * psi)
        DD = DD * s_cur / s
        Dl = s_cur / s
        u = u + DD
        lam = Dl
        sync()
        psi = np.sqrt((Dl * Dl * ub2 + DD @ DD) / (Dl * Dl))
        predicted = True
        ok = newton()
        for _ in range(1, steps):
            if not ok:
                break
            if target_iterations > 0 and iterations[-1] > target_iterations:
                s_cur = np.sqrt(target_iterations / iterations[-1]) * s_cur
            u = u + DD
            lam += Dl
            sync()
            ok = newton()
    return dict(u=u, lam=lam, iterations=iterations, lambdas=lambdas, step_sizes=sizes, success=bool(ok))
