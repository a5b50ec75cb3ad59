"""Shared inputs and checks for the GPU parity tests (test infrastructure)."""
import numpy as np

import oracle as O

L2_VAR = (0, 2, 4, 6)


def spiked(m, n, k=20, strength=4.0, seed=0):
    """Noise N(0,1) plus one planted sparse loading (k nonzeros) of the given strength."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    v = np.zeros(n)
    idx = rng.choice(n, k, replace=False)
    v[idx] = rng.standard_normal(k)
    v /= np.linalg.norm(v)
    u = rng.standard_normal(m)
    A += strength * np.sqrt(n / k) * 0.2 * np.outer(u, v)
    return A.astype(np.float32), idx


def unit_gaussian_starts(n, B, seed=1):
    rng = np.random.default_rng(seed)
    X0 = rng.standard_normal((n, B))
    X0 /= np.linalg.norm(X0, axis=0)
    return X0.astype(np.float32)


def default_gamma(f, m):
    """Penalties that leave a few dozen survivors on spiked(m, 1000)."""
    return {4: 9.0, 5: 9.0 * m, 6: 3.0, 7: 3.0 * np.sqrt(m)}.get(f, 0.0)


def explained_variance(A, X, f):
    """FP64 EV of each normalised column: |A x|_2^2 (L2 variance) or |A x|_1."""
    A = np.asarray(A, dtype=np.float64)  # no copy when A is FP64 already
    Xn = X.astype(np.float64) / np.linalg.norm(X.astype(np.float64), axis=0)
    Z = A @ Xn
    return np.sum(Z * Z, axis=0) if f in L2_VAR else np.sum(np.abs(Z), axis=0)


def tie_margin(A, X_prev, f, s, gamma, ztol=0.0):
    """Relative gap between the threshold and the nearest competing value in the last
    step, per SP, computed from the oracle's previous iterate (FP64).  A support that
    differs only where this gap is within rounding is a documented magnitude tie."""
    A = np.asarray(A, dtype=np.float64)  # no copy when A is FP64 already
    Z = A @ X_prev
    # L1 variance maps z to sgn z: an entry of z within the product's rounding (ztol, relative
    # to the rms of z) of zero may flip sign, which moves every entry of A^T sgn(z); such an
    # SP counts as a tie (margin 0)
    # (an exact zero of z is no tie: its terms are all zero, so every product form gives 0 too)
    with np.errstate(invalid="ignore", divide="ignore"):
        zrel = (np.min(np.where(Z != 0, np.abs(Z), np.inf), axis=0) / np.sqrt(np.mean(Z * Z, axis=0))
                if f not in L2_VAR else None)
    if f in L2_VAR:
        if f >= 4:
            Z = Z / np.linalg.norm(Z, axis=0)
    else:
        Z = np.sign(Z)
    V = A.T @ Z
    out = np.empty(V.shape[1])
    for j in range(V.shape[1]):
        a = np.abs(V[:, j])
        if f in (0, 1):
            srt = np.sort(a)[::-1]
            out[j] = (srt[s - 1] - srt[s]) / srt[s - 1] if s < a.size else np.inf
        elif f in (2, 3):
            _, _, lam, _ = O.soft_threshold(V[:, j], s, O.LAMBDA_GPU)
            out[j] = np.min(np.abs(a - lam)) / lam
        elif f in (4, 5):
            out[j] = np.min(np.abs(a * a - gamma)) / gamma
        else:
            out[j] = np.min(np.abs(a - gamma)) / gamma
    return out if zrel is None else np.where(zrel < ztol, 0.0, out)


def trajectory_tie_margin(A, X0, f, s, gamma, steps, ztol=0.0, gemm_backend=0, threads=8):
    """Smallest one-step tie margin (tie_margin) over the FP64 oracle's trajectory from X0, per SP:
    steps 0 .. steps-1 each start from the oracle's own iterate, advanced one iteration at a time
    (the iteration depends on the iterate alone, so the chain is the full solve's trajectory).
    An SP whose solve ends on another support than the oracle's must have passed a step whose
    margin is within the product's rounding: a documented magnitude tie.  NaN margins (an SP
    zeroed out by a penalty) do not count as ties.  gemm_backend 1 advances the chain with host
    BLAS (faster on large matrices; its sums differ from the fixed-order solve's by rounding).
    Returns (margins, final FP64 iterates)."""
    A64 = np.asarray(A, dtype=np.float64)
    X = np.asarray(X0, dtype=np.float64)
    out = np.full(X.shape[1], np.inf)
    for _ in range(steps):
        with np.errstate(invalid="ignore", divide="ignore"):
            out = np.fmin(out, tie_margin(A64, X, f, s, gamma, ztol=ztol))
        o = O.solve(A64, f, s=s, gamma=gamma, tol=0.0, max_iter=1, total_sp=X.shape[1], X0=X, want_X=True,
                    lambda_rule=O.LAMBDA_GPU, gemm_backend=gemm_backend, threads=threads)
        X = o["X"]
    return out, X


SOFT = (2, 3, 6, 7)  # formulations whose step is a soft threshold x = sgn(v)(|v| - lam)_+


def soft_condition(A, X, f, s, gamma):
    """Per SP, how much the soft threshold amplifies a relative error of v into x: kappa =
    |v_S|_2 / |S_lam(v)|_2 over the kept set S, for the step from the (FP64) iterate X.  The
    step subtracts lam from every kept entry, so where the kept |v_i| sit barely above lam (data
    whose entries are all alike) an error of eps |v_i| in v becomes an error of about eps kappa in
    x.  1 for the hard-threshold formulations, whose kept entries are v's own."""
    out = np.ones(X.shape[1])
    if f not in SOFT:
        return out
    A = np.asarray(A, dtype=np.float64)  # no copy when A is FP64 already
    Z = A @ X
    if f in L2_VAR:
        if f >= 4:
            Z = Z / np.linalg.norm(Z, axis=0)
    else:
        Z = np.sign(Z)
    V = A.T @ Z
    for j in range(V.shape[1]):
        if f in (2, 3):
            _, _, lam, _ = O.soft_threshold(V[:, j], s, O.LAMBDA_GPU)
        else:
            lam = gamma
        a = np.abs(V[:, j])
        kept = a > lam
        d = np.linalg.norm(a[kept] - lam)
        out[j] = np.linalg.norm(a[kept]) / d if d > 0 else np.inf
    return out
