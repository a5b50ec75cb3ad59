"""Sparse (CSC) input with implicit centring (spca_matrix_upload_csc), against the reference's
sparse solver semantics (c++/src/gpower/sparse_PCA_solver_for_CSC.h:112-495): the solve runs
on A_c = A - 1 mu^T - r 1^T without densifying A (doMean / doRowMean, :215-250, :287-316).

Pinned by the FP64 oracle run on the densified A_c with the same starting points, by the
dense upload of the same matrix (no centring), by the SVD of A_c (power method with
deflation), and by NumPy for the explained variance the sparse kernels report.  Needs an
MI355X.
"""
import numpy as np
import pytest

import oracle as O
import spca_mi355x as S
from spca_cases import explained_variance, unit_gaussian_starts

pytestmark = pytest.mark.gpu


def _sparse_problem(m=400, n=1500, density=0.03, seed=71):
    """Sparse nonnegative data (as the reference's bag-of-words inputs, with continuous
    values so that no two coordinates tie exactly) and a planted block of correlated columns;
    returns the dense FP32 matrix and its CSC arrays."""
    rng = np.random.default_rng(seed)
    mask = rng.random((m, n)) < density
    A = np.where(mask, rng.exponential(1.0, size=(m, n)), 0.0).astype(np.float32)
    rows = rng.choice(m, 60, replace=False)
    cols = rng.choice(n, 12, replace=False)
    A[np.ix_(rows, cols)] += (2.0 + 2.0 * rng.random((60, 12))).astype(np.float32)
    cp = np.zeros(n + 1, dtype=np.int32)
    ri, va = [], []
    for j in range(n):
        nz = np.nonzero(A[:, j])[0]
        ri.append(nz)
        va.append(A[nz, j])
        cp[j + 1] = cp[j] + nz.size
    return A, cp, np.concatenate(ri).astype(np.int32), np.concatenate(va).astype(np.float32)


def _gamma(f, Ac, X0):
    """A penalty that keeps ~20 coordinates after the first step from the first start."""
    z = Ac.astype(np.float64) @ X0[:, 0].astype(np.float64)
    z = z / np.linalg.norm(z) if f in (4, 6) else np.sign(z)
    a = np.sort(np.abs(Ac.astype(np.float64).T @ z))[::-1]
    t = 0.5 * (a[19] + a[20])
    return float(t * t) if f in (4, 5) else float(t)


@pytest.mark.parametrize("f", range(8))
def test_centred_csc_matches_fp64_oracle(ctx, f):
    A, cp, ri, va = _sparse_problem()
    m, n = A.shape
    mu = A.mean(axis=0).astype(np.float32)
    r = (0.1 * A.mean(axis=1)).astype(np.float32)
    Ac = (A.astype(np.float64) - mu[None, :].astype(np.float64) - r[:, None].astype(np.float64))
    B, it, s = 32, 40, 10
    X0 = unit_gaussian_starts(n, B, seed=72)
    gamma = _gamma(f, Ac, X0) if f >= 4 else 0.0
    o = O.solve(Ac, f, s=s, gamma=gamma, tol=0.0, max_iter=it, total_sp=B, batch=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    M = ctx.upload_csc(cp, ri, va, m, n, col_means=mu, row_means=r)
    try:
        st = S.OptimizationSettings(formulation=f, constraintParameter=s, penaltyParameter=gamma, tolerance=0.0,
                                    maximumIterations=it, totalStartingPoints=B, batchSize=B)
        stats = S.OptimizationStatistics()
        _, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
    finally:
        M.release()
    live = np.any(Xg != 0, axis=0) & np.any(o["X"] != 0, axis=0)  # a penalty can zero an SP out
    same = np.all((Xg != 0) == (o["X"] != 0), axis=0) & live
    ev_g = explained_variance(Ac, Xg[:, live], f)
    ev_o = explained_variance(Ac, o["X"][:, live], f)
    rel = np.full(B, np.inf)
    rel[live] = np.abs(ev_g - ev_o) / ev_o
    ev_kernel = np.abs(stats.sp_ev_fp64[live] - ev_g) / ev_g
    print(f"case f={f}: live SPs {live.sum()}, identical supports {same.sum()}/{live.sum()}, EV rel max "
          f"{rel[same].max():.2e}, sparse EV kernel vs NumPy {ev_kernel.max():.2e}", flush=True)
    assert live.sum() >= B // 2 and same.sum() >= 0.9 * live.sum()
    assert np.all(rel[same] <= 1e-6)
    assert np.all(ev_kernel <= 1e-9)


def test_uncentred_csc_matches_dense_upload(ctx):
    A, cp, ri, va = _sparse_problem(seed=73)
    m, n = A.shape
    B = 32
    X0 = unit_gaussian_starts(n, B, seed=74)
    st = S.OptimizationSettings(formulation=0, constraintParameter=10, tolerance=0.0, maximumIterations=30,
                                totalStartingPoints=B, batchSize=B)
    Ms = ctx.upload_csc(cp, ri, va, m, n)
    Md = ctx.upload(A)
    try:
        ss, sd = S.OptimizationStatistics(), S.OptimizationStatistics()
        _, Xs = Ms.solve(st, X0=X0, stats=ss, want_X=True)
        _, Xd = Md.solve(st, X0=X0, stats=sd, want_X=True)
    finally:
        Ms.release()
        Md.release()
    same = np.all((Xs != 0) == (Xd != 0), axis=0)
    rel = np.abs(ss.sp_ev_fp64 - sd.sp_ev_fp64) / sd.sp_ev_fp64
    print(f"case uncentred: identical supports {same.mean():.3f}, EV rel max {rel[same].max():.2e}", flush=True)
    assert same.mean() >= 0.9 and np.all(rel[same] <= 1e-6)


def test_csc_deflation_gives_singular_vectors(ctx):
    """Power method (s = n) with deflation on the centred sparse matrix: components are the
    leading right singular vectors of A_c (the reference's use of its sparse solver with a
    SparseDeflationCollection)."""
    rng = np.random.default_rng(75)
    m, n, k = 120, 80, 3
    A = np.where(rng.random((m, n)) < 0.2, rng.standard_normal((m, n)), 0.0)
    for c, strength in enumerate((6.0, 4.0, 2.0)):  # three planted directions, independent rows
        A[:, c] += strength * rng.standard_normal(m)
    A = A.astype(np.float32)
    mu = A.mean(axis=0).astype(np.float32)
    Ac = A.astype(np.float64) - mu[None, :]
    _, sv, Vt = np.linalg.svd(Ac)
    cp = np.zeros(n + 1, dtype=np.int32)
    ri, va = [], []
    for j in range(n):
        nz = np.nonzero(A[:, j])[0]
        ri.append(nz)
        va.append(A[nz, j])
        cp[j + 1] = cp[j] + nz.size
    M = ctx.upload_csc(cp, np.concatenate(ri), np.concatenate(va), m, n, col_means=mu)
    try:
        st = S.OptimizationSettings(formulation=0, constraintParameter=n, tolerance=0.0, maximumIterations=400,
                                    totalStartingPoints=8, batchSize=8)
        X, stats = M.solve_components(st, k, X0=unit_gaussian_starts(n, 8, seed=76))
    finally:
        M.release()
    for c in range(k):
        align = abs(float(X[:, c].astype(np.float64) @ Vt[c]))
        print(f"case deflation: component {c} |<x, v_c>| = {align:.9f} (sigma {sv[c]:.4f}, gap to next "
              f"{sv[c] / sv[c + 1]:.3f})", flush=True)
        assert align > 1 - 1e-5


def test_csc_rejects_what_it_does_not_support(ctx):
    A, cp, ri, va = _sparse_problem(m=64, n=100, seed=77)
    bad_cp = cp.copy()
    bad_cp[0] = 1
    with pytest.raises(S.SpcaError):
        ctx.upload_csc(bad_cp, ri, va, 64, 100)  # col_ptr not starting at 0
    bad = ri.copy()
    bad[0] = 64
    with pytest.raises(S.SpcaError):
        ctx.upload_csc(cp, bad, va, 64, 100)  # row index out of range
    M = ctx.upload_csc(cp, ri, va, 64, 100)
    try:
        st = S.OptimizationSettings(formulation=0, constraintParameter=5, totalStartingPoints=4, batchSize=4,
                                    precision=S.PREC_BF16X3, start=S.START_REFERENCE)
        with pytest.raises(S.SpcaError):
            M.solve(st)
        with pytest.raises(S.SpcaError):
            M.download()
    finally:
        M.release()
