"""Pins the oracle where the reference publishes no result: closed forms,
invariants, special cases that reduce to textbook routines, and brute force
on tiny inputs.  Each check is chosen so a plausible slip (wrong sign, lost
term, transposed operand, wrong tie rule) fails it."""
import itertools

import numpy as np
import pytest

import oracle as O


def _lam_bisect(v, s):
    """lambda with |S(v)|_1 = sqrt(s) |S(v)|_2 by bisection on the monotone ratio."""
    a = np.abs(v)
    lo, hi = 0.0, a.max()
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        t = np.maximum(a - mid, 0)
        r = t.sum() / np.sqrt((t * t).sum())
        if r > np.sqrt(s):
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("n,s", [(50, 2), (200, 7), (1000, 30), (4096, 256)])
def test_soft_threshold_root(n, s):
    rng = np.random.default_rng(n + s)
    v = rng.standard_normal(n)
    x, nrm, lam, found = O.soft_threshold(v, s)
    assert found
    ref = _lam_bisect(v, s)
    assert abs(lam - ref) <= 1e-10 * abs(ref)
    np.testing.assert_allclose(x, np.sign(v) * np.maximum(np.abs(v) - lam, 0), rtol=0, atol=1e-12)
    assert abs(np.abs(x).sum() / np.linalg.norm(x) - np.sqrt(s)) < 1e-9
    assert abs(nrm - np.linalg.norm(x)) < 1e-12 * nrm
    # The float instantiation of the reference scan either lands on the same root, or
    # stops in the first segment: with one element the discriminant b^2-4ac is 0 in exact
    # arithmetic, float rounding makes it positive, and the 'w < linfty - 1e-6' guard
    # (c++/src/utils/thresh_functions.h:211) is below float resolution, so it accepts a
    # root just under max|v| that leaves a single nonzero.  The restatement must show
    # exactly that, not something in between.
    vf = v.astype(np.float32)
    xf, _, lamf, foundf = O.soft_threshold(vf, s)
    a = np.sort(np.abs(vf))[::-1]
    assert foundf
    if abs(lamf - ref) > 2e-5 * abs(ref):
        assert a[1] < lamf < a[0] and (xf != 0).sum() == 1
    # and for the double instantiation the first-segment discriminant is exact zero
    assert (x != 0).sum() > 1


@pytest.mark.parametrize("n,s", [(50, 2), (200, 7), (1000, 30), (4096, 256), (65536, 256)])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 2e-5)])
def test_soft_threshold_gpu_rule_root(n, s, dtype, tol):
    # The reference GPU scan (gpugpower/my_cublas_wrapper.h:335-375) lands on the exact
    # root in both precisions: no first-segment artifact.
    rng = np.random.default_rng(n * 7 + s)
    v = (rng.standard_normal(n) * 300.0).astype(dtype)
    ref = _lam_bisect(v.astype(np.float64), s)
    x, nrm, lam, found = O.soft_threshold(v, s, O.LAMBDA_GPU)
    assert found and abs(lam - ref) <= tol * abs(ref)
    assert abs(np.abs(x.astype(np.float64)).sum() / np.linalg.norm(x.astype(np.float64)) - np.sqrt(s)) < \
        (1e-9 if dtype == np.float64 else 1e-3)
    # c2-scale magnitudes make the CPU rule's first-segment artifact likely; the GPU rule
    # never leaves a single survivor unless s == 1
    assert (x != 0).sum() >= s


def test_soft_threshold_cpu_rule_artifact_at_scale():
    # The CPU scan's artifact is a property of max|v| alone (the one-element quadratic),
    # so at magnitudes like those of v = A^T z at 4096 x 65536 it fires for a sizable share
    # of columns in float; it never fires for the GPU rule.
    rng = np.random.default_rng(0)
    fired = 0
    for t in range(64):
        v = (rng.standard_normal(4096) * 1000.0).astype(np.float32)
        x, _, _, _ = O.soft_threshold(v, 256, O.LAMBDA_CPU)
        fired += (x != 0).sum() == 1
        xg, _, _, _ = O.soft_threshold(v, 256, O.LAMBDA_GPU)
        assert (xg != 0).sum() > 256
    assert fired > 0


def test_soft_threshold_degenerate_branch_is_flagged():
    # s >= (|v|_1/|v|_2)^2: no lambda > 0 solves the equation; the reference finds no root
    v = np.array([1.0, 1.0, 1.0, 1.0])
    _, _, _, found = O.soft_threshold(v, 4)
    assert not found


@pytest.mark.parametrize("n,s", [(10, 1), (300, 17), (5000, 256)])
def test_hard_threshold_top_s(n, s):
    rng = np.random.default_rng(s)
    v = rng.standard_normal(n)
    for rule in (O.HARD_SORT, O.HARD_SET):
        x, nrm, tau = O.hard_threshold(v, s, rule)
        keep = np.argsort(-np.abs(v), kind="stable")[:s]
        expect = np.zeros(n)
        expect[keep] = v[keep]
        np.testing.assert_array_equal(x, expect)
        assert tau == np.abs(v)[keep].min()
        assert abs(nrm - np.linalg.norm(expect)) < 1e-12 * nrm


def test_hard_threshold_ties():
    v = np.array([3.0, -2.0, 2.0, 1.0, -2.0, 0.5])
    # SORT rule: tau = 2nd largest |v| counting multiplicity = 2 -> all three 2's kept
    x, _, tau = O.hard_threshold(v, 2, O.HARD_SORT)
    assert tau == 2.0 and (x != 0).sum() == 4
    # SET rule de-duplicates magnitudes: distinct top-2 of {3,2,2,1,2,.5} is {3,2}
    x, _, tau = O.hard_threshold(v, 2, O.HARD_SET)
    assert tau == 2.0
    # The reference std::set is seeded with the first s magnitudes and de-duplicates them,
    # so it can hold fewer than s entries and never regrows: for s=3 it is seeded with
    # {3,2,2} -> {3,2}; no later magnitude beats 2, so tau = 2 (not the 3rd distinct, 1).
    _, _, tau_sort = O.hard_threshold(v, 3, O.HARD_SORT)
    _, _, tau_set = O.hard_threshold(v, 3, O.HARD_SET)
    assert tau_sort == 2.0 and tau_set == 2.0
    # Seeded with {2,2} -> {2}; 3 then evicts 2, so only the 3 survives (nnz 1 < s=2),
    # while the SORT rule keeps 3,2,2.
    w = np.array([2.0, -2.0, 3.0, 1.0, 0.5])
    xs, _, ts = O.hard_threshold(w, 2, O.HARD_SET)
    xo, _, to = O.hard_threshold(w, 2, O.HARD_SORT)
    assert ts == 3.0 and (xs != 0).sum() == 1
    assert to == 2.0 and (xo != 0).sum() == 3


def test_penalized_thresholds_closed_form():
    rng = np.random.default_rng(3)
    v = rng.standard_normal(1000)
    g = 0.7
    x, f = O.pen_threshold(v, g, l1=False)
    keep = v * v > g
    np.testing.assert_array_equal(x, np.where(keep, v, 0))
    assert abs(f - (v[keep] ** 2 - g).sum()) < 1e-10
    x, f = O.pen_threshold(v, g, l1=True)
    np.testing.assert_allclose(x, np.sign(v) * np.maximum(np.abs(v) - g, 0), atol=1e-15)
    assert abs(f - np.linalg.norm(x)) < 1e-12


def test_starting_points_match_rand_r():
    # constrained SP: s draws of (idx, value) from rand_r seeded with the SP index
    seq = O.rand_r_sequence(5, 6).astype(np.float64)
    n, s = 100, 3
    v = O.starting_point(n, s, True, 5)
    RM = 2147483647.0
    expect = np.zeros(n)
    for i in range(s):
        idx = int(n * seq[2 * i] / RM)
        expect[min(idx, n - 1)] = seq[2 * i + 1] / RM
    expect /= np.sqrt((seq[1::2] / RM) ** 2).sum() ** 0 * np.sqrt(((seq[1::2] / RM) ** 2).sum())
    np.testing.assert_allclose(v, expect, rtol=1e-15)
    w = O.starting_point(n, s, False, 7)
    np.testing.assert_allclose(w, -1 + 2 * O.rand_r_sequence(7, n) / RM, rtol=0, atol=1e-15)


def _spiked(m, n, seed=0, k=8, strength=6.0):
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(m)
    v = np.zeros(n)
    idx = rng.choice(n, k, replace=False)
    v[idx] = rng.standard_normal(k)
    v /= np.linalg.norm(v)
    return strength * np.outer(u, v) / np.sqrt(m) * np.sqrt(m) + rng.standard_normal((m, n)) * 0.3, idx


def test_power_method_special_cases():
    # s = n (L0 constrained, L2 variance) and gamma = 0 (penalised, L2 variance) are the power method
    rng = np.random.default_rng(11)
    A = rng.standard_normal((40, 25))
    U, S, Vt = np.linalg.svd(A)
    o = O.solve(A, 0, s=25, tol=0, max_iter=300, total_sp=4, threads=2)
    assert abs(o["fval"] - S[0]) < 1e-10 * S[0]
    assert abs(abs(o["x"] @ Vt[0]) - 1) < 1e-10
    o = O.solve(A, 4, gamma=0.0, tol=0, max_iter=300, total_sp=4, threads=2)
    assert abs(o["fval"] - S[0] ** 2) < 1e-9 * S[0] ** 2
    o = O.solve(A, 6, gamma=0.0, tol=0, max_iter=300, total_sp=4, threads=2)
    assert abs(o["fval"] - S[0]) < 1e-10 * S[0]
    assert abs(abs(o["x"] @ Vt[0]) - 1) < 1e-10


def test_l1_variance_brute_force():
    # max_{|x|=1} |A x|_1 = max_{u in {+-1}^m} |A^T u|_2 ; AM (s = n) reaches it from many SPs
    rng = np.random.default_rng(5)
    A = rng.standard_normal((9, 6))
    best = max(np.linalg.norm(A.T @ np.array(u)) for u in itertools.product((-1, 1), repeat=9))
    o = O.solve(A, 1, s=6, tol=0, max_iter=50, total_sp=64, threads=2)
    assert abs(o["fval"] - best) < 1e-10 * best
    # L0 penalised, L1 variance with gamma = 0: f = |A^T sgn(Ax)|^2 -> same optimum squared
    o = O.solve(A, 5, gamma=0.0, tol=0, max_iter=50, total_sp=64, threads=2)
    assert abs(o["fval"] - best ** 2) < 1e-9 * best ** 2


def test_l0_constrained_brute_force_support():
    # tiny problem: the best of many SPs equals the best support found by enumeration
    rng = np.random.default_rng(9)
    m, n, s = 12, 10, 3
    A = rng.standard_normal((m, n))
    best = max(np.linalg.svd(A[:, list(S)], compute_uv=False)[0] for S in itertools.combinations(range(n), s))
    o = O.solve(A, 0, s=s, tol=0, max_iter=100, total_sp=256, threads=2)
    assert o["fval"] <= best * (1 + 1e-12)
    assert abs(o["fval"] - best) < 1e-9 * best
    assert (o["x"] != 0).sum() == s


def test_l1_constrained_recovers_planted_support():
    A, idx = _spiked(60, 200, seed=4, k=8, strength=8.0)
    for form in (2, 3):
        o = O.solve(A, form, s=8, tol=1e-10, max_iter=300, total_sp=32, threads=4)
        x = o["x"]
        assert abs(np.abs(x).sum() / np.linalg.norm(x) - np.sqrt(8)) < 1e-8
        top = set(np.argsort(-np.abs(x))[:8])
        assert len(top & set(idx)) >= 6


@pytest.mark.parametrize("form", range(8))
def test_monotone_ascent_and_invariants(form):
    # The objective never decreases (the reference's DEBUG check, termination_criteria.h:31-33)
    # under the GPU lambda rule.  The CPU rule breaks ascent on this data (formulation 2, SP 2,
    # iteration 5 drops 9.49 -> 5.45) through its first-segment artifact; see
    # test_cpu_lambda_rule_breaks_ascent.
    rng = np.random.default_rng(form)
    A = rng.standard_normal((30, 80))
    gamma = {4: 4.0, 5: 40.0, 6: 1.0, 7: 6.0}.get(form, 0.0)
    prev = None
    for it in range(1, 12):
        o = O.solve(A, form, s=6, gamma=gamma, tol=0, max_iter=it, total_sp=8, threads=2, want_X=True,
                    lambda_rule=O.LAMBDA_GPU)
        if prev is not None:
            assert np.all(o["sp_val"] >= prev - 1e-9 * np.abs(prev))
        prev = o["sp_val"]
    if form < 4:
        np.testing.assert_allclose(np.linalg.norm(o["X"], axis=0), 1.0, rtol=1e-12)
        if form < 2:
            assert np.all((o["X"] != 0).sum(axis=0) == 6)


@pytest.mark.parametrize("form", [0, 3, 5, 6])
def test_per_sp_mode_equals_batch_of_one(form):
    # per-SP freezing is the reference batch rule applied to a batch of one SP
    rng = np.random.default_rng(20 + form)
    A = rng.standard_normal((25, 60))
    gamma = {5: 30.0, 6: 0.8}.get(form, 0.0)
    a = O.solve(A, form, s=5, gamma=gamma, tol=1e-4, max_iter=40, total_sp=16, termination=O.TERM_PER_SP,
                want_X=True)
    b = O.solve(A, form, s=5, gamma=gamma, tol=1e-4, max_iter=40, total_sp=16, batch=1, want_X=True)
    np.testing.assert_array_equal(a["sp_val"], b["sp_val"])
    np.testing.assert_array_equal(a["sp_iters"], b["sp_iters"])
    np.testing.assert_array_equal(a["X"], b["X"])


def test_scale_covariance_and_permutation():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((20, 50))
    o1 = O.solve(A, 0, s=5, tol=0, max_iter=30, total_sp=16)
    o2 = O.solve(3.0 * A, 0, s=5, tol=0, max_iter=30, total_sp=16)
    assert abs(o2["fval"] - 3 * o1["fval"]) < 1e-12 * o2["fval"]
    perm = rng.permutation(50)
    X0 = rng.standard_normal((50, 16))
    X0 /= np.linalg.norm(X0, axis=0)
    a = O.solve(A, 0, s=5, tol=0, max_iter=30, total_sp=16, X0=X0, want_X=True)
    b = O.solve(A[:, perm], 0, s=5, tol=0, max_iter=30, total_sp=16, X0=X0[perm], want_X=True)
    np.testing.assert_allclose(a["X"][perm], b["X"], atol=1e-12)


def test_explained_variance():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((30, 40))
    x = rng.standard_normal(40)
    assert abs(O.explained_variance(A, x, True) - np.sum((A @ x) ** 2)) < 1e-9
    assert abs(O.explained_variance(A, x, False) - np.sum(np.abs(A @ x))) < 1e-9


def test_cpu_lambda_rule_breaks_ascent():
    # The reference CPU scan accepts the spurious one-element root at this iterate (double
    # precision), leaving x = e_j; the objective then falls.  Pinned so that the oracle keeps
    # reproducing the reference's CPU rule rather than silently fixing it.
    rng = np.random.default_rng(2)
    A = rng.standard_normal((30, 80))
    a = O.solve(A, 2, s=6, tol=0, max_iter=5, total_sp=8, threads=2, want_X=True)
    b = O.solve(A, 2, s=6, tol=0, max_iter=6, total_sp=8, threads=2, want_X=True)
    assert (a["X"][:, 2] != 0).sum() == 1
    assert b["sp_val"][2] < 0.7 * a["sp_val"][2]


@pytest.mark.parametrize("form", [0, 3, 4, 7])
def test_blas_backend_with_per_sp_termination_matches_loops(form):
    """The host-BLAS backend packs the columns still iterating under per-SP termination; per SP it
    gives what the fixed-order loops give (iterations and nonzeros equal, values to rounding)."""
    rng = np.random.default_rng(5)
    m, n, B = 60, 300, 24
    A = rng.standard_normal((m, n))
    A[:, :5] += 2.0 * rng.standard_normal((m, 1))
    X0 = rng.standard_normal((n, B))
    X0 /= np.linalg.norm(X0, axis=0)
    g = {4: 4.0, 7: 2.0 * np.sqrt(m)}.get(form, 0.0)
    kw = dict(s=8, gamma=g, tol=1e-7, max_iter=60, total_sp=B, X0=X0, termination=O.TERM_PER_SP, want_X=True,
              lambda_rule=O.LAMBDA_GPU)
    a = O.solve(A, form, gemm_backend=0, **kw)
    b = O.solve(A, form, gemm_backend=1, **kw)
    assert O.cblas_name(), "no host BLAS"
    assert len(set(a["sp_iters"])) > 1  # the SPs stop at different iterations
    np.testing.assert_array_equal(a["sp_iters"], b["sp_iters"])
    np.testing.assert_array_equal(a["sp_nnz"], b["sp_nnz"])
    np.testing.assert_allclose(b["sp_val"], a["sp_val"], rtol=1e-12)
    np.testing.assert_allclose(b["X"], a["X"], atol=1e-12)
