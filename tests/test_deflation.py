"""Multi-component solves by deflation (spca_solve_components), against the reference's
deflated iteration (c++/src/gpower/sparse_PCA_solver_for_CSC.h:56-106, applied before each
A*X at :207 and after each A^T*Z at :311): x <- x - (p . x) p for every earlier component p.

Pinned by the power method: with s = n the L0-constrained L2-variance step is x <- v/|v|, so
component c is the top eigenvector of (I - PP^T) A^T A (I - PP^T), the c-th right singular
vector of A (LAPACK SVD).  The sparse case is checked against an FP64 restatement of the
deflated iteration on the same data and starting points.  Needs an MI355X.
"""
import numpy as np
import pytest

import spca_mi355x as S
from spca_cases import spiked, unit_gaussian_starts

pytestmark = pytest.mark.gpu


def _low_rank(m, n, sig, seed):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, len(sig))))
    V, _ = np.linalg.qr(rng.standard_normal((n, len(sig))))
    A = U @ np.diag(sig) @ V.T + 1e-3 * rng.standard_normal((m, n))
    return A.astype(np.float32)


@pytest.mark.parametrize("prec,tol", [(S.PREC_FP32, 1e-5), (S.PREC_BF16X3, 1e-4)])
def test_power_method_components_are_singular_vectors(ctx, prec, tol):
    m, n, B, k = 64, 48, 8, 3
    A = _low_rank(m, n, (10.0, 7.0, 5.0, 3.5, 2.0), seed=11)
    _, sv, Vt = np.linalg.svd(A.astype(np.float64))
    X0 = unit_gaussian_starts(n, B, seed=12)
    st = S.OptimizationSettings(formulation=S.SPCA_Formulation.L0_constrained_L2_PCA, constraintParameter=n,
                                tolerance=0.0, maximumIterations=300, totalStartingPoints=B, batchSize=B,
                                precision=prec)
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    try:
        X, stats = M.solve_components(st, k, X0=X0)
    finally:
        M.release()
    for c in range(k):
        align = abs(float(X[:, c].astype(np.float64) @ Vt[c]))
        print(f"case prec={prec} component {c}: |<x, v_c>| = {align:.9f}, fval {stats[c].fval:.6f} "
              f"(sigma_c {sv[c]:.6f})", flush=True)
        assert align > 1 - tol
        assert stats[c].fval == pytest.approx(sv[c], rel=10 * tol)


def _deflated_am_fp64(A, X0, p, s, iters):
    """FP64 restatement of the deflated L0-constrained L2-variance iteration (GPU tie rule:
    |v| >= the s-th largest |v|); returns the best SP's unit x and its value."""
    A = A.astype(np.float64)
    X = X0.astype(np.float64).copy()
    f = np.zeros(X.shape[1])
    for _ in range(iters):
        X -= np.outer(p, p @ X)
        Z = A @ X
        f = np.linalg.norm(Z, axis=0)
        V = A.T @ Z
        V -= np.outer(p, p @ V)
        a = np.abs(V)
        tau = -np.partition(-a, s - 1, axis=0)[s - 1]
        V = np.where(a >= tau, V, 0.0)
        X = V / np.linalg.norm(V, axis=0)
    j = int(np.argmax(f))
    return X[:, j], f[j]


def test_sparse_second_component_matches_fp64(ctx):
    m, n, B, s, it = 64, 200, 16, 10, 30
    A, _ = spiked(m, n, k=20, strength=4.0, seed=21)
    X0 = unit_gaussian_starts(n, B, seed=22)
    st = S.OptimizationSettings(formulation=S.SPCA_Formulation.L0_constrained_L2_PCA, constraintParameter=s,
                                tolerance=0.0, maximumIterations=it, totalStartingPoints=B, batchSize=B)
    M = ctx.upload(A)
    try:
        X, stats = M.solve_components(st, 2, X0=X0)
        x1 = M.solve(st, X0=X0)
    finally:
        M.release()
    np.testing.assert_array_equal(X[:, 0], x1)  # component 0 is the plain solve
    p = X[:, 0].astype(np.float64)
    x2, f2 = _deflated_am_fp64(A, X0, p, s, it)
    g = X[:, 1].astype(np.float64)
    align = abs(float(g @ x2))
    print(f"case sparse: component 1 |<x_gpu, x_fp64>| = {align:.9f}, fval {stats[1].fval:.6f} vs {f2:.6f}, "
          f"|<x1, p>| = {abs(float(g @ p)):.3e}", flush=True)
    assert np.array_equal(g != 0, x2 != 0)
    assert align > 1 - 1e-6
    assert stats[1].fval == pytest.approx(f2, rel=1e-5)


def test_components_reject_bad_arguments(ctx):
    A = np.ones((8, 16), dtype=np.float32)
    st = S.OptimizationSettings(formulation=0, constraintParameter=2, totalStartingPoints=4, batchSize=4,
                                start=S.START_REFERENCE)
    M = ctx.upload(A)
    try:
        with pytest.raises(S.SpcaError):
            M.solve_components(st, 0)
        with pytest.raises(S.SpcaError):
            M.solve_components(st, 17)
    finally:
        M.release()
