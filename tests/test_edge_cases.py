"""Edge cases of the HIP solver against the FP64 oracle.  Needs an MI355X.

Tiny and ragged shapes (one row, one column, n < 32, m and n far from the 256 padding),
s >= n (the reference clamp), s = 1, a single starting point, total_sp not a multiple of the
batch (the reference rounds it up), a penalty that zeroes every coordinate (the NaN rule of
computeTheError), and starting points with exact zeros.  FP32 throughout; tolerances as
test_gpu_parity.py (objective 1e-5 relative; supports identical, these shapes being tie-free).
"""
import numpy as np
import pytest

import oracle as O
import spca_mi355x as S
from spca_cases import unit_gaussian_starts

pytestmark = pytest.mark.gpu


def _solve_both(ctx, A, f, s, gamma, B, it, X0, total=None):
    M = ctx.upload(A)
    try:
        total = total or B
        st = S.OptimizationSettings(formulation=f, constraintParameter=s, penaltyParameter=gamma, tolerance=0.0,
                                    maximumIterations=it, totalStartingPoints=total, batchSize=B)
        stats = S.OptimizationStatistics()
        x, X = M.solve(st, X0=X0, stats=stats, want_X=True)
        o = O.solve(A.astype(np.float64), f, s=s, gamma=gamma, tol=0.0, max_iter=it, total_sp=total, batch=B,
                    X0=X0.astype(np.float64), want_X=True, lambda_rule=O.LAMBDA_GPU)
        return x, X, stats, o
    finally:
        M.release()


def _l1_ball_steps(ctx, A, f, s, B, X0, steps=6):
    """The L1 ball with s close to n: an iterate's v can lie inside the ball's cone
    (|v|_1 <= sqrt(s) |v|_2), where the constraint is inactive and the lambda scan has no root
    (the oracle's restatement returns NaN there).  Checked one step at a time from the GPU's own
    iterate: an inactive step against its closed form x = v / |v|_2 (lambda = 0), an active one
    against the oracle's step."""
    X = X0
    inactive = 0
    for t in range(steps):
        _, Xn, _, o = _solve_both(ctx, A, f, s, 0.0, B, 1, X)
        Ad, Xd = A.astype(np.float64), X.astype(np.float64)
        Z = Ad @ Xd
        Z = np.sign(Z) if f == 3 else Z / np.linalg.norm(Z, axis=0)
        V = Ad.T @ Z
        ratio = np.abs(V).sum(0) / np.linalg.norm(V, axis=0)
        for j in range(B):
            if ratio[j] <= np.sqrt(s) * (1 + 1e-6):
                inactive += 1
                want = V[:, j] / np.linalg.norm(V[:, j])
            else:
                want = o["X"][:, j]
            np.testing.assert_allclose(Xn[:, j], want, rtol=1e-4, atol=1e-5, err_msg=f"step {t} SP {j}")
        X = Xn.astype(np.float32)
    print(f"  {inactive} of {steps * B} steps inside the ball's cone", flush=True)


@pytest.mark.parametrize("m,n", [(1, 7), (3, 1), (2, 5), (5, 31), (257, 33), (13, 300)])
@pytest.mark.parametrize("f", [0, 2, 4, 7])
def test_tiny_and_ragged_shapes(ctx, m, n, f):
    rng = np.random.default_rng(m * 1000 + n)
    A = rng.standard_normal((m, n)).astype(np.float32)
    B, s = 4, min(3, n)
    gamma = {4: 0.05, 7: 0.05 * np.sqrt(m)}.get(f, 0.0)
    X0 = unit_gaussian_starts(n, B, seed=m + n)
    print(f"case m={m} n={n} f={f}", flush=True)
    if f in (2, 3) and 1 < s and n < 31:
        _l1_ball_steps(ctx, A, f, s, B, X0)
        return
    x, X, stats, o = _solve_both(ctx, A, f, s, gamma, B, 6, X0)
    if f in (2, 3) and s <= 1:
        # the L1 ball with s = 1 holds only 1-sparse points (the oracle's lambda scan
        # degenerates to NaN there): each iterate is +-e_k
        assert np.all((X != 0).sum(axis=0) == 1) and np.allclose(np.abs(X).sum(axis=0), 1.0)
        return
    alive = np.all(np.isfinite(o["X"]), axis=0) & np.any(o["X"] != 0, axis=0)
    np.testing.assert_array_equal(np.all(np.isfinite(X), axis=0) & np.any(X != 0, axis=0), alive)
    np.testing.assert_array_equal((X != 0)[:, alive], (o["X"] != 0)[:, alive])
    np.testing.assert_allclose(stats.values[alive], o["sp_val"][alive], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(X[:, alive], o["X"][:, alive], rtol=1e-4, atol=1e-5)


def test_s_clamped_to_n_and_s_one(ctx):
    """L0 constraint against the oracle for s = 1, s = n and s > n (clamped to n).  The L1
    constraint with s = 1 or s >= n is the soft threshold's degenerate branch, where the
    reference GPU scan has no root (the oracle's restatement returns NaN; parity unpinned):
    here s >= n gives lambda = 0, i.e. the L0 iteration with s = n, and s = 1 keeps the
    largest coordinate alone, both checked against those closed forms."""
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 20)).astype(np.float32)
    X0 = unit_gaussian_starts(20, 8, seed=4)
    vals = {}
    for s in (1, 20, 500):  # 500 > n: the reference clamps s to n
        x, X, stats, o = _solve_both(ctx, A, 0, s, 0.0, 8, 10, X0)
        np.testing.assert_array_equal(X != 0, o["X"] != 0)
        np.testing.assert_allclose(stats.values, o["sp_val"], rtol=1e-5)
        assert np.all((X != 0).sum(axis=0) == min(s, 20))
        vals[s] = stats.values
    M = ctx.upload(A)
    try:
        for s in (1, 20, 500):
            st = S.OptimizationSettings(formulation=2, constraintParameter=s, tolerance=0.0, maximumIterations=10,
                                        totalStartingPoints=8, batchSize=8)
            stats = S.OptimizationStatistics()
            _, X = M.solve(st, X0=X0, stats=stats, want_X=True)
            print(f"L1 constraint s={s}: nnz {(X != 0).sum(axis=0)} values {stats.values}", flush=True)
            if s == 1:
                assert np.all((X != 0).sum(axis=0) == 1)
            else:
                np.testing.assert_allclose(stats.values, vals[20], rtol=1e-5)
    finally:
        M.release()


def test_single_starting_point_and_batch_rounding(ctx):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((60, 150)).astype(np.float32)
    # one SP
    X0 = unit_gaussian_starts(150, 1, seed=6)
    x, X, stats, o = _solve_both(ctx, A, 0, 10, 0.0, 1, 15, X0)
    np.testing.assert_allclose(stats.values, o["sp_val"], rtol=1e-5)
    np.testing.assert_array_equal(X != 0, o["X"] != 0)
    # 10 SPs in batches of 4: the reference rounds the total up to 12
    X0 = unit_gaussian_starts(150, 12, seed=7)
    x, X, stats, o = _solve_both(ctx, A, 0, 10, 0.0, 4, 15, X0, total=10)
    assert len(stats.values) == 12 and stats.it == 3 * 15
    np.testing.assert_allclose(stats.values, o["sp_val"], rtol=1e-5)
    assert stats.best_sp == o["best_sp"]


def test_penalty_zeroes_everything(ctx):
    """gamma above every |v|: the iterate becomes zero, its objective 0 and the next error
    0/0 = NaN, which the batch max skips (reference computeTheError and the '<' max)."""
    rng = np.random.default_rng(8)
    A = rng.standard_normal((30, 80)).astype(np.float32)
    X0 = unit_gaussian_starts(80, 4, seed=9)
    for f in (4, 6):
        x, X, stats, o = _solve_both(ctx, A, f, 0, 1e6, 4, 5, X0)
        assert np.all(stats.values == 0) and np.all(o["sp_val"] == 0)
        assert np.all(stats.cardinalities == 0)


def test_starting_points_with_zeros(ctx):
    """Sparse starting points (the reference's constrained SPs have s nonzeros): the first
    product already runs on a mostly-zero X."""
    rng = np.random.default_rng(10)
    A = rng.standard_normal((50, 400)).astype(np.float32)
    X0 = np.zeros((400, 16), dtype=np.float32)
    for j in range(16):
        idx = rng.choice(400, 5, replace=False)
        X0[idx, j] = rng.standard_normal(5)
    X0 /= np.linalg.norm(X0, axis=0)
    x, X, stats, o = _solve_both(ctx, A, 0, 8, 0.0, 16, 12, X0)
    np.testing.assert_array_equal(X != 0, o["X"] != 0)
    np.testing.assert_allclose(stats.values, o["sp_val"], rtol=1e-5)
