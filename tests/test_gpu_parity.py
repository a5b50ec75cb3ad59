"""Parity of the HIP solver (through the C ABI) with the CPU oracle and with results the
reference published.  Needs an MI355X.

Tolerances, per starting point (SP), against the FP64 oracle on the same seeded inputs:
  FP32     explained variance (EV) within 1e-6 relative, objective within 1e-5
  bf16x3   EV within 1e-5 relative, objective within 1e-4
  fp16     EV within 1e-3 relative, objective within 5e-3 (fp16 keeps 11 bits; the
           penalised objectives subtract gamma, which loses more)
Supports after one step from the same iterate must be identical, except for an SP whose
step had a magnitude tie: the threshold and its nearest competitor within TIE of each
other (relative: 1e-5 FP32, 1e-4 bf16x3, 1e-3 fp16), where rounding of A*X decides the
order; such SPs are counted and bounded (<= 5%; fp16, whose A*X error is ~1e-3, <= 50%).
For L1-variance formulations (z -> sgn z) an entry of z = A x within the product's
rounding of zero (ZTOL relative to the rms of z) can flip sign and move every entry of
A^T sgn(z): such an SP also counts as a tie.
"""
import json
import os
import time

import numpy as np
import pytest

import oracle as O
import spca_mi355x as S
from spca_cases import (default_gamma, explained_variance, spiked, tie_margin, trajectory_tie_margin,
                        unit_gaussian_starts)

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
TOL = {  # precision: (EV, objective, tie, share of SPs allowed to sit on a tie in one step)
    S.PREC_FP32: (1e-6, 1e-5, 1e-5, 0.05),
    S.PREC_BF16X3: (1e-5, 1e-4, 1e-4, 0.05),
    S.PREC_FP16: (1e-3, 5e-3, 1e-3, 0.5),
}


# relative size (to the rms of z = A x) below which an entry of z may change sign through
# the product's rounding: about 3x the rms product error, which is ~sqrt(2) times the
# operand rounding (fp16 2^-12 rms ~1.4e-4, bf16x3 ~2^-17, FP32 ~2^-24 plus accumulation)
ZTOL = {S.PREC_FP32: 1e-6, S.PREC_BF16X3: 3e-5, S.PREC_FP16: 6e-4}


@pytest.fixture(scope="module")
def small():
    m, n, B = 200, 1000, 64
    A, planted = spiked(m, n)
    return A, unit_gaussian_starts(n, B), planted


@pytest.fixture(scope="module")
def small_mat(ctx, small):
    A = small[0]
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3 | S.STORE_FP16)
    yield M
    M.release()


def _settings(f, it, prec, B, s=10, m=200, tol=0.0):
    return S.OptimizationSettings(formulation=f, constraintParameter=s, penaltyParameter=default_gamma(f, m),
                                  tolerance=tol, maximumIterations=it, totalStartingPoints=B, batchSize=B,
                                  precision=prec)


@pytest.fixture(scope="module")
def wide():
    """Padded batch and rows multiples of 256: the 16-bit products run the 256x256-tile kernel."""
    m, n, B = 300, 1000, 256
    A, planted = spiked(m, n, seed=7)
    return A, unit_gaussian_starts(n, B, seed=8), planted


@pytest.fixture(scope="module")
def wide_mat(ctx, wide):
    M = ctx.upload(wide[0], storage=S.STORE_FP32 | S.STORE_BF16X3 | S.STORE_FP16)
    yield M
    M.release()


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3, S.PREC_FP16])
@pytest.mark.parametrize("f", range(8))
def test_step_parity(small, small_mat, f, prec):
    """One AM iteration from identical iterates (the GPU's own iterate after t steps),
    GPU vs FP64 oracle: supports identical except at this step's magnitude ties."""
    _step_parity(small, small_mat, f, prec)


@pytest.mark.parametrize("prec", [S.PREC_BF16X3, S.PREC_FP16])
@pytest.mark.parametrize("f", range(8))
def test_step_parity_large_tiles(wide, wide_mat, f, prec):
    """As test_step_parity with B = 256 and 300 x 1000 (padded 512 x 1024), where both
    16-bit products run the pipelined 256x256-tile kernel."""
    _step_parity(wide, wide_mat, f, prec)


def _step_parity(data, mat, f, prec):
    small_mat = mat
    A, X0, _ = data
    m, n = A.shape
    B, s, g = X0.shape[1], 10, default_gamma(f, m)
    ev_tol, val_tol, tie_tol, tie_share = TOL[prec]
    for t in (0, 5, 25):
        print(f"case f={f} prec={prec} t={t} m={m} n={n} B={B}", flush=True)
        Xt = X0 if t == 0 else small_mat.solve(_settings(f, t, prec, B, m=m), X0=X0, want_X=True)[1]
        # a penalty can zero an SP out; its next iterate is then NaN (0/0) in both
        alive = np.all(np.isfinite(Xt), axis=0) & np.any(Xt != 0, axis=0)
        stats = S.OptimizationStatistics()
        _, Xg = small_mat.solve(_settings(f, 1, prec, B, m=m), X0=Xt, stats=stats, want_X=True)
        o = O.solve(A.astype(np.float64), f, s=s, gamma=g, tol=0.0, max_iter=1, total_sp=B,
                    X0=Xt.astype(np.float64), want_X=True, lambda_rule=O.LAMBDA_GPU)
        tie = tie_margin(A, Xt.astype(np.float64), f, s, g, ztol=ZTOL[prec]) < tie_tol
        supp = np.all((Xg != 0) == (o["X"] != 0), axis=0)
        chk = alive & ~tie
        scale = np.abs(o["X"]).max(axis=0)
        xerr = np.abs(Xg - o["X"]).max(axis=0) / np.where(scale > 0, scale, 1)
        with np.errstate(invalid="ignore", divide="ignore"):
            vrel = np.where(stats.values == o["sp_val"], 0.0,
                            np.abs(stats.values - o["sp_val"]) / np.abs(o["sp_val"]))
        print(f"  t={t}: alive {alive.sum()} ties {tie[alive].sum()} supports equal {supp[alive].mean():.3f} "
              f"x err {np.nanmax(xerr[chk]) if chk.any() else 0:.2e} value rel {np.nanmax(vrel[chk]) if chk.any() else 0:.2e}",
              flush=True)
        assert np.all(supp[chk]), np.nonzero(chk & ~supp)[0]
        assert tie[alive].mean() <= tie_share
        assert np.all(xerr[chk] <= 10 * val_tol), np.nanmax(xerr[chk])
        assert np.all(vrel[chk] <= val_tol), np.nanmax(vrel[chk])
        np.testing.assert_array_equal(stats.cardinalities[alive], (Xg != 0).sum(axis=0)[alive])


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3, S.PREC_FP16])
@pytest.mark.parametrize("f", range(8))
def test_trajectory_parity(small, small_mat, f, prec):
    """60 iterations from the same seeded starts against the FP64 oracle.  Every SP whose FP64
    trajectory passes no magnitude tie (a step whose tie margin is within TIE, see
    trajectory_tie_margin) must end on the oracle's support, with the explained variance of its
    final iterate within the precision's tolerance.  SPs whose trajectory passed a tie may branch
    to another fixed point; they are counted and printed, not bounded by a share."""
    A, X0, _ = small
    m, n = A.shape
    B, it, s, g = X0.shape[1], 60, 10, default_gamma(f, m)
    ev_tol, _, tie_tol, _ = TOL[prec]
    print(f"case f={f} prec={prec} m={m} n={n} B={B} it={it}", flush=True)
    stats = S.OptimizationStatistics()
    x, Xg = small_mat.solve(_settings(f, it, prec, B), X0=X0, stats=stats, want_X=True)
    o = O.solve(A.astype(np.float64), f, s=s, gamma=g, tol=0.0, max_iter=it, total_sp=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    margin, Xchain = trajectory_tie_margin(A, X0, f, s, g, it, ztol=ZTOL[prec])
    alive_o = np.all(np.isfinite(o["X"]), axis=0) & np.any(o["X"] != 0, axis=0)
    alive_g = np.all(np.isfinite(Xg), axis=0) & np.any(Xg != 0, axis=0)
    np.testing.assert_array_equal(alive_g, alive_o)
    np.testing.assert_array_equal(Xchain[:, alive_o], o["X"][:, alive_o])  # the chain is the solve's trajectory
    supp = np.all((Xg != 0) == (o["X"] != 0), axis=0)
    ev_o = np.full(B, np.nan)
    ev_o[alive_o] = explained_variance(A, o["X"][:, alive_o], f)
    with np.errstate(invalid="ignore"):
        ev_rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
    tied = margin < tie_tol
    check = alive_o & ~tied
    # the explained variance is compared once an SP has settled: its last relative change below a
    # tenth of the tolerance (an SP still moving after 60 iterations is compared mid-way, where
    # the FP32 reference method itself differs from FP64 by ~1e-5)
    settled = check & (o["sp_err"] <= ev_tol / 10)
    print(f"  alive {alive_o.sum()} tied {tied[alive_o].sum()} settled {settled.sum()} supports equal "
          f"{supp[alive_o].mean():.3f} (untied {supp[check].mean() if check.any() else 1:.3f}) EV rel settled max "
          f"{np.nanmax(ev_rel[settled]) if settled.any() else 0:.2e}; off-support SPs "
          f"{np.nonzero(alive_o & ~supp)[0].tolist()} with margins {np.round(margin[alive_o & ~supp], 7).tolist()}; "
          f"still moving {np.nonzero(check & ~settled)[0].tolist()}", flush=True)
    assert stats.it == it
    assert np.all(supp[check]), np.nonzero(check & ~supp)[0]
    assert np.all(ev_rel[settled] <= ev_tol), np.sort(ev_rel[settled])[-5:]
    if alive_o.any():
        assert abs(np.linalg.norm(x.astype(np.float64)) - 1) < 1e-6
        assert stats.best_sp == int(np.nanargmax(np.where(alive_g, stats.values, -np.inf)))


def test_constrained_invariants(small, small_mat):
    # ||x||_2 = 1, nnz = s for L0 (tie-free data), ||x||_1/||x||_2 = sqrt(s) for L1
    A, X0, _ = small
    for f in range(4):
        st = S.OptimizationSettings(formulation=f, constraintParameter=10, tolerance=0.0, maximumIterations=15,
                                    totalStartingPoints=64, batchSize=64)
        _, X = small_mat.solve(st, X0=X0, want_X=True)
        Xd = X.astype(np.float64)
        np.testing.assert_allclose(np.linalg.norm(Xd, axis=0), 1.0, atol=2e-6)
        if f < 2:
            assert np.all((X != 0).sum(axis=0) == 10)
        else:
            np.testing.assert_allclose(np.abs(Xd).sum(0) / np.linalg.norm(Xd, axis=0), np.sqrt(10), rtol=2e-6)


@pytest.mark.parametrize("f", [0, 2])
def test_long_candidate_lists(ctx, f):
    """A 128 x 4096 spiked matrix with s = 10, where the list steps from the second iteration on
    work from the bound a column step left: the constrained invariants hold after every step, and
    the trajectory follows the FP64 oracle's supports and explained variance, as in
    test_trajectory_parity."""
    A, _ = spiked(128, 4096, seed=3)
    m, n = A.shape
    B, s, it = 128, 10, 6
    X0 = unit_gaussian_starts(n, B, seed=2)
    M = ctx.upload(A, storage=S.STORE_FP32)
    try:
        stats = S.OptimizationStatistics()
        x, Xg = M.solve(_settings(f, it, S.PREC_FP32, B, s=s, m=m), X0=X0, stats=stats, want_X=True)
    finally:
        M.release()
    print(f"case f={f} list steps {stats.list_steps}/{stats.list_tries}", flush=True)
    assert stats.list_steps > 0
    Xd = Xg.astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(Xd, axis=0), 1.0, atol=2e-6)
    if f == 0:
        assert np.all((Xg != 0).sum(axis=0) == s) and np.all(stats.cardinalities == s)
    else:
        np.testing.assert_allclose(np.abs(Xd).sum(0) / np.linalg.norm(Xd, axis=0), np.sqrt(s), rtol=2e-6)
    o = O.solve(A.astype(np.float64), f, s=s, tol=0.0, max_iter=it, total_sp=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    margin, _ = trajectory_tie_margin(A, X0, f, s, 0.0, it, ztol=ZTOL[S.PREC_FP32])
    check = ~(margin < TOL[S.PREC_FP32][2])
    supp = np.all((Xg != 0) == (o["X"] != 0), axis=0)
    assert np.all(supp[check]), np.nonzero(check & ~supp)[0]
    ev_o = explained_variance(A, o["X"], f)
    ev_rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
    assert np.all(ev_rel[check] <= 1e-5), np.sort(ev_rel[check])[-5:]


def _clustered_matrix():
    """128 x 4096: 20 strong columns, 1300 near-copies of one column aligned with them (|v| within
    0.2% of each other) and weak noise elsewhere.  A column step puts the list bound at the
    cluster's bin, so the next L0 list of many SPs holds the whole cluster: more entries than
    the 4096-element chunk's 1024 stage slots."""
    rng = np.random.default_rng(7)
    m, n = 128, 4096
    A = 0.01 * rng.standard_normal((m, n))
    strong = np.arange(0, 40, 2)
    A[:, strong] += rng.standard_normal((m, 20))
    c = A[:, strong].sum(1)
    c /= np.linalg.norm(c)
    cl = np.arange(100, 100 + 3 * 1300, 3)
    A[:, cl] = np.outer(c, 0.8 * (1 + 0.002 * np.arange(1300) / 1300)) + 0.001 * rng.standard_normal((m, 1300))
    return A.astype(np.float32)


def test_overflowing_l0_lists(ctx):
    """L0 lists that overflow a chunk's stage slots: the step histograms the staged keys, rescans
    its column in its own workgroup with a higher bound and resolves on that list (or runs the
    column passes).  Invariants and the FP64 oracle as in test_long_candidate_lists."""
    A = _clustered_matrix()
    m, n = A.shape
    B, s, it = 64, 10, 6
    X0 = unit_gaussian_starts(n, B, seed=5)
    M = ctx.upload(A, storage=S.STORE_FP32)
    try:
        stats = S.OptimizationStatistics()
        x, Xg = M.solve(_settings(0, it, S.PREC_FP32, B, s=s, m=m), X0=X0, stats=stats, want_X=True)
    finally:
        M.release()
    print(f"case overflowing lists: list steps {stats.list_steps}/{stats.list_tries}", flush=True)
    assert stats.list_tries > 0
    np.testing.assert_allclose(np.linalg.norm(Xg.astype(np.float64), axis=0), 1.0, atol=2e-6)
    assert np.all((Xg != 0).sum(axis=0) == s) and np.all(stats.cardinalities == s)
    o = O.solve(A.astype(np.float64), 0, s=s, tol=0.0, max_iter=it, total_sp=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    margin, _ = trajectory_tie_margin(A, X0, 0, s, 0.0, it, ztol=ZTOL[S.PREC_FP32])
    check = ~(margin < TOL[S.PREC_FP32][2])
    supp = np.all((Xg != 0) == (o["X"] != 0), axis=0)
    assert np.all(supp[check]), np.nonzero(check & ~supp)[0]
    ev_o = explained_variance(A, o["X"], 0)
    ev_rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
    assert np.all(ev_rel[check] <= 1e-5), np.sort(ev_rel[check])[-5:]


def test_planted_support_recovered(small, small_mat):
    A, X0, planted = small
    st = S.OptimizationSettings(formulation=0, constraintParameter=20, tolerance=1e-6, maximumIterations=100,
                                totalStartingPoints=64, batchSize=64)
    x = small_mat.solve(st, X0=X0)
    o = O.solve(A.astype(np.float64), 0, s=20, tol=1e-6, max_iter=100, total_sp=64, X0=X0.astype(np.float64),
                lambda_rule=O.LAMBDA_GPU)
    assert set(np.nonzero(x)[0]) == set(np.nonzero(o["x"])[0])
    assert len(set(np.nonzero(x)[0]) & set(planted)) >= 15


def test_deterministic_reruns(small, small_mat):
    # every constrained formulation resolves most steps on the previous threshold's list, whose
    # FP64 sums must not depend on the order waves reach LDS
    A, X0, _ = small
    for f in (0, 1, 2, 3, 5):
        st = S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=default_gamma(f, 200),
                                    tolerance=0.0, maximumIterations=20, totalStartingPoints=64, batchSize=64,
                                    precision=S.PREC_BF16X3)
        a = small_mat.solve(st, X0=X0, want_X=True)[1]
        b = small_mat.solve(st, X0=X0, want_X=True)[1]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("term", [S.TERM_BATCH_MAX, S.TERM_PER_SP])
def test_batches_and_termination(small, small_mat, term):
    # 100 SPs in batches of 32 -> 4 batches, 128 SPs (reference clamp); tol > 0
    A, _, _ = small
    n = A.shape[1]
    X0 = unit_gaussian_starts(n, 128, seed=7)
    st = S.OptimizationSettings(formulation=0, constraintParameter=10, tolerance=1e-4, maximumIterations=50,
                                totalStartingPoints=100, batchSize=32, termination=term)
    stats = S.OptimizationStatistics()
    x = small_mat.solve(st, X0=X0, stats=stats)
    o = O.solve(A.astype(np.float64), 0, s=10, tol=1e-4, max_iter=50, total_sp=100, batch=32,
                X0=X0.astype(np.float64), termination=term, lambda_rule=O.LAMBDA_GPU)
    assert len(stats.values) == 128 and o["total_sp"] == 128
    same_it = np.mean(stats.iters == o["sp_iters"])
    rel = np.abs(stats.values - o["sp_val"]) / o["sp_val"]
    print(f"  iterations equal {same_it:.3f}, total it {stats.it}/{o['it']}, max rel {rel.max():.2e}")
    assert same_it >= 0.95
    assert np.median(rel) < 1e-6
    assert abs(stats.fval - o["fval"]) <= 1e-5 * o["fval"]


@pytest.mark.parametrize("f", [0, 3, 4, 7])
def test_otf_refill_matches_oracle(small, small_mat, f):
    # on-the-fly refill (reference OTF driver): 200 SPs through 32 slots, refilled on the device
    A, _, _ = small
    n = A.shape[1]
    X0 = unit_gaussian_starts(n, 224, seed=11)  # 200 rounded up to a multiple of 32
    g = default_gamma(f, A.shape[0])
    print(f"case otf f={f}", flush=True)
    st = S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=g, tolerance=1e-4,
                                maximumIterations=30, totalStartingPoints=200, batchSize=32, useOTF=True)
    stats = S.OptimizationStatistics()
    x = small_mat.solve(st, X0=X0, stats=stats)
    o = O.solve(A.astype(np.float64), f, s=10, gamma=g, tol=1e-4, max_iter=30, total_sp=200, batch=32, otf=True,
                X0=X0.astype(np.float64), lambda_rule=O.LAMBDA_GPU)
    assert len(stats.values) == 224
    same_it = np.mean(stats.iters == o["sp_iters"])
    with np.errstate(invalid="ignore"):
        rel = np.abs(stats.values - o["sp_val"]) / np.abs(o["sp_val"])
    print(f"  total it {stats.it}/{o['it']} iterations equal {same_it:.3f} median rel {np.nanmedian(rel):.2e} "
          f"best {stats.fval:.7g}/{o['fval']:.7g} (SP {stats.best_sp}/{o['best_sp']})")
    assert same_it >= 0.95
    assert abs(int(stats.it) - int(o["it"])) <= 2
    assert np.nanmedian(rel) < 1e-6
    assert abs(stats.fval - o["fval"]) <= 1e-5 * abs(o["fval"])
    # the best SP may differ only as an argmax tie: its FP64 value equals the best to rounding
    assert stats.best_sp == o["best_sp"] or o["sp_val"][stats.best_sp] >= o["fval"] * (1 - 1e-6)
    np.testing.assert_allclose(x, o["x"], atol=2e-5)
    assert abs(stats.ev_fp64 - explained_variance(A, o["x"][:, None], f)[0]) <= 1e-5 * stats.ev_fp64


@pytest.mark.parametrize("f", [0, 4])
def test_otf_many_slots_run_every_sp_once(small, small_mat, f):
    """OTF with 256 slots (four waves of the refill-ranking workgroup): slots that finish in the
    same iteration in different waves take consecutive SP indices in slot order, so every one of
    the 2048 SPs runs exactly once.  A skipped SP would keep 0 iterations (its record is never
    written); a repeated one would displace another.  Checked against the oracle's OTF driver
    (sparse_PCA_solver.h:120-163)."""
    A, _, _ = small
    n = A.shape[1]
    X0 = unit_gaussian_starts(n, 2048, seed=12)  # 2000 SPs rounded up to whole batches of 256
    g = default_gamma(f, A.shape[0])
    print(f"case otf256 f={f}", flush=True)
    st = S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=g, tolerance=1e-4,
                                maximumIterations=30, totalStartingPoints=2000, batchSize=256, useOTF=True)
    stats = S.OptimizationStatistics()
    small_mat.solve(st, X0=X0, stats=stats)
    o = O.solve(A.astype(np.float64), f, s=10, gamma=g, tol=1e-4, max_iter=30, total_sp=2000, batch=256, otf=True,
                X0=X0.astype(np.float64), lambda_rule=O.LAMBDA_GPU)
    assert len(stats.iters) == 2048
    never = np.flatnonzero(stats.iters == 0)
    assert never.size == 0, f"SPs that never ran: {never[:20]}"
    same_it = np.mean(stats.iters == o["sp_iters"])
    with np.errstate(invalid="ignore"):
        rel = np.abs(stats.values - o["sp_val"]) / np.abs(o["sp_val"])
    print(f"  total it {stats.it}/{o['it']} iterations equal {same_it:.4f} median rel {np.nanmedian(rel):.2e}")
    assert same_it >= 0.99
    assert stats.sp_iterations == int(np.sum(stats.iters))
    assert abs(int(stats.it) - int(o["it"])) <= 2
    assert np.nanmedian(rel) < 1e-6


def test_resolve_l2_is_top_eigenvector(small, small_mat):
    # FP64 re-solve on the support of the returned x: for L2 variance it is the top
    # eigenvector of A_S^T A_S (LAPACK eigh in FP64 pins it)
    A, X0, _ = small
    st = S.OptimizationSettings(formulation=2, constraintParameter=10, tolerance=1e-6, maximumIterations=100,
                                totalStartingPoints=64, batchSize=64, resolve_iters=2000, resolve_tol=1e-15)
    stats = S.OptimizationStatistics()
    x = small_mat.solve(st, X0=X0, stats=stats)
    Sx = np.nonzero(x)[0]
    A64 = A.astype(np.float64)
    w, V = np.linalg.eigh(A64[:, Sx].T @ A64[:, Sx])
    print(f"  support {len(Sx)}: ev(x) {stats.ev_fp64:.12g} re-solve {stats.ev_resolve_fp64:.12g} eigh {w[-1]:.12g}")
    assert abs(stats.ev_resolve_fp64 - w[-1]) <= 1e-10 * w[-1]
    assert stats.ev_resolve_fp64 >= stats.ev_fp64 * (1 - 1e-12)
    y, ev, _ = small_mat.resolve_support(x, 2, max_iter=2000, tol=1e-15)
    assert abs(ev - w[-1]) <= 1e-10 * w[-1]
    assert np.all(y[np.setdiff1d(np.arange(A.shape[1]), Sx)] == 0)
    assert abs(abs(y[Sx] @ V[:, -1]) - 1) < 1e-8


def test_resolve_l1_brute_force(ctx):
    # L1 variance on a 10 x 4 support: max |A_S y|_1 over unit y = max over sign vectors u of
    # |A_S^T u|_2; the best of a few starts reaches it and no start exceeds it
    import itertools
    rng = np.random.default_rng(3)
    A = rng.standard_normal((10, 50)).astype(np.float32)
    Sx = np.array([3, 17, 29, 41])
    A64 = A.astype(np.float64)[:, Sx]
    best = max(np.linalg.norm(A64.T @ np.array(u)) for u in itertools.product((-1.0, 1.0), repeat=10))
    M = ctx.upload(A)
    try:
        evs = []
        for t in range(8):
            x = np.zeros(50, dtype=np.float32)
            x[Sx] = rng.standard_normal(4)
            y, ev, _ = M.resolve_support(x, 1, max_iter=500, tol=1e-15)
            evs.append(ev)
            xs = x[Sx].astype(np.float64) / np.linalg.norm(x[Sx])
            assert ev >= np.abs(A64 @ xs).sum() * (1 - 1e-12)
            assert ev <= best * (1 + 1e-12)
            assert abs(np.abs(A64 @ y[Sx]).sum() - ev) <= 1e-12 * ev
        assert abs(max(evs) - best) <= 1e-12 * best
    finally:
        M.release()


def _boxplot_rows():
    rows = []
    with open(os.path.join(GOLD, "ref_paper_experiment_boxplots.txt")) as fh:
        for line in fh:
            parts = line.strip().rstrip(",").split(",")
            rows.append((int(parts[0]), np.array([float(v) for v in parts[1:]])))
    return rows


@pytest.fixture(scope="module")
def boxplot_mat(ctx):
    A = O.generate_problem(512, 2048, uniform_scale=True).astype(np.float32)
    M = ctx.upload(A)
    yield M
    M.release()


@pytest.mark.parametrize("row", range(12))
def test_reference_boxplot_values(boxplot_mat, row):
    # c++/results/paper_experiment_boxplots.txt: per-SP objective of 1000 reference SPs,
    # L0_constrained_L2_PCA, 20 iterations, tol 1e-6, printed with 6 significant digits.
    s, ref = _boxplot_rows()[row]
    print(f"case boxplot s={s}", flush=True)
    st = S.OptimizationSettings(formulation=0, constraintParameter=s, tolerance=1e-6, maximumIterations=20,
                                totalStartingPoints=1000, batchSize=1000, start=S.START_REFERENCE_F64)
    stats = S.OptimizationStatistics()
    boxplot_mat.solve(st, stats=stats)
    rel = np.abs(stats.values.astype(np.float64) - ref) / ref
    exact = np.mean(["%g" % v == "%g" % r for v, r in zip(stats.values, ref)])
    print(f"  s={s}: 6-digit matches {exact:.3f}, rel max {rel.max():.2e}, share <= 1e-5: {np.mean(rel <= 1e-5):.4f}")
    assert np.mean(rel <= 1e-5) >= 0.99
    assert abs(stats.fval - ref.max()) <= 1e-5 * ref.max()


def _integer_rows():
    return json.load(open(os.path.join(GOLD, "ref_cpu_integer_results.json")))["rows"]


@pytest.mark.parametrize("idx", range(len(_integer_rows())))
def test_reference_integer_results(ctx, idx):
    # nnz of the returned x and the total iteration count of the reference's 1-thread runs
    r = _integer_rows()[idx]
    m, n = r["m"], r["n"]
    print(f"case {r['source']}:{r['line']}", flush=True)
    A = O.generate_problem(m, n, uniform_scale=False)
    if r["ld_quirk"]:
        ld = m if m % 16 == 0 else (m // 16 + 1) * 16
        flat = np.zeros(ld * n)
        flat[: m * n] = np.asfortranarray(A).reshape(-1, order="F")
        A = flat.reshape((ld, n), order="F")[:m]
    M = ctx.upload(np.asarray(A, dtype=np.float32))
    try:
        st = S.OptimizationSettings(formulation=r["formulation"], constraintParameter=r["s"],
                                    penaltyParameter=r["gamma"], tolerance=r["tol"], maximumIterations=r["max_iter"],
                                    totalStartingPoints=r["total_sp"], batchSize=r["batch"],
                                    useOTF=r["otf"], start=S.START_REFERENCE_F64)
        stats = S.OptimizationStatistics()
        x = M.solve(st, stats=stats)
    finally:
        M.release()
    assert stats.it == r["it"]
    nnz = int((x != 0).sum())
    if nnz != r["nnz"]:
        # Allowed only as an argmax tie: the GPU's best SP is another SP whose FP64 value is
        # within FP32 rounding of the reference best, and its own FP64 iterate has the nnz
        # the GPU returned.  Rerun all SPs in FP64 to check both.
        o = O.solve(np.asarray(A, dtype=np.float64), r["formulation"], s=r["s"], gamma=r["gamma"], tol=r["tol"],
                    max_iter=r["max_iter"], total_sp=r["total_sp"], batch=r["batch"], otf=r["otf"], want_X=True,
                    lambda_rule=O.LAMBDA_GPU)
        j = stats.best_sp
        print(f"  nnz {nnz} vs {r['nnz']}: GPU best SP {j}, FP64 value {o['sp_val'][j]:.9g} vs best "
              f"{o['fval']:.9g} (SP {o['best_sp']})")
        assert int((o["x"] != 0).sum()) == r["nnz"]
        assert j != o["best_sp"]
        assert o["sp_val"][j] >= o["fval"] * (1 - 1e-6)
        assert int((o["X"][:, j] != 0).sum()) == nnz


def _hip():
    import ctypes
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipFree.argtypes = [ctypes.c_void_p]
    return h


def test_drop_in_device_entry(ctx, small):
    # GPUSolver::denseDataSolver-shaped entry: caller-owned device matrix with LD_M padding
    import ctypes
    A = small[0]
    m, n = A.shape
    ldm = 256
    Ap = np.zeros((ldm, n), dtype=np.float32, order="F")
    Ap[:m] = A
    hip = _hip()
    d = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d), Ap.nbytes) == 0
    try:
        assert hip.hipMemcpy(d, Ap.ctypes.data_as(ctypes.c_void_p), Ap.nbytes, 1) == 0  # host to device
        st = S.OptimizationSettings(formulation=3, constraintParameter=10, tolerance=0.0, maximumIterations=10,
                                    totalStartingPoints=64, batchSize=64, start=S.START_REFERENCE)
        x1 = np.zeros(n, dtype=np.float32)
        rc = S.lib().spca_dense_solver_device(ctx.ptr, m, n, d, x1.ctypes.data_as(ctypes.c_void_p),
                                              ctypes.byref(st._c()), None, ldm, n)
        assert rc == S.SPCA_OK, S.lib().spca_last_error(ctx.ptr)
    finally:
        hip.hipFree(d)
    M = ctx.upload(A)
    x2 = M.solve(st)
    M.release()
    assert np.array_equal(x1, x2)


def test_errors_are_reported(ctx, small_mat, small):
    X0 = small[1]
    bad = S.OptimizationSettings(formulation=9, totalStartingPoints=64, batchSize=64)
    with pytest.raises(S.SpcaError) as e:
        small_mat.solve(bad, X0=X0)
    assert e.value.code == S.SPCA_EINVAL
    bad_start = S.OptimizationSettings(start=7, totalStartingPoints=64, batchSize=64)
    with pytest.raises(S.SpcaError):
        small_mat.solve(bad_start, X0=X0)
    # budget: a context with a 1 MiB budget cannot hold a 200 x 1000 matrix in two layouts
    with S.Context(hbm_budget_bytes=1 << 20) as c2:
        with pytest.raises(S.SpcaError) as e:
            c2.upload(small[0])
        assert e.value.code == S.SPCA_EBUDGET


def test_serving_alone_equals_mix(ctx):
    """Continuous batching: jobs of different shapes, formulations and precisions submitted
    while others run return bitwise the results they give when solved alone (own buffers,
    deterministic kernels), and a queued job can be cancelled."""
    rng = np.random.default_rng(5)
    mats = []
    for (m, n, seed) in ((200, 1000, 1), (300, 2500, 2), (128, 4096, 3)):
        A, _ = spiked(m, n, seed=seed)
        mats.append(ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3, key=100 + seed))
    specs = []
    for k in range(12):
        M = mats[k % 3]
        f = int(rng.integers(0, 8))
        B = [32, 64, 128][k % 3]
        prec = [S.PREC_FP32, S.PREC_BF16X3][k % 2]
        st = S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=default_gamma(f, M.m),
                                    tolerance=1e-5, maximumIterations=40, totalStartingPoints=B, batchSize=B,
                                    precision=prec, useOTF=(k % 4 == 3))
        specs.append((M, st, unit_gaussian_starts(M.n, B, seed=k)))
    alone = []
    for M, st, X0 in specs:
        stats = S.OptimizationStatistics()
        x = M.solve(st, X0=X0, stats=stats)
        alone.append((x, stats.values.copy(), stats.fval))
    ctx.start_server(4)
    try:
        jobs = [M.submit(st, X0=X0) for M, st, X0 in specs]
        # one extra job, cancelled while it may still be queued
        extra = specs[0][0].submit(specs[0][1], X0=specs[0][2])
        extra.cancel()
        for j, (xa, va, fa) in zip(jobs, alone):
            stats = S.OptimizationStatistics()
            x, times = j.wait(stats=stats)
            assert times["latency_ms"] >= times["run_ms"] >= 0
            assert np.array_equal(x.view(np.uint32), xa.view(np.uint32))
            assert np.array_equal(stats.values.view(np.uint32), va.view(np.uint32)) and stats.fval == fa
        with pytest.raises(S.SpcaError) as e:
            extra.wait()
        assert e.value.code == S.SPCA_ECANCELLED
    finally:
        ctx.stop_server()
        for M in mats:
            M.release()


def test_serving_jobs_join_a_running_group(ctx):
    """Slot-level continuous batching: with one worker, jobs submitted while a fused group
    iterates join it at a chunk boundary (they start at iteration 0 beside the running job),
    a running member can be cancelled, and every result is bitwise the job's solo result."""
    mats = []
    for (m, n, seed) in ((256, 3000, 31), (200, 1000, 32)):
        A, _ = spiked(m, n, seed=seed)
        mats.append(ctx.upload(A, storage=S.STORE_FP32, key=200 + seed))
    long_st = S.OptimizationSettings(formulation=0, constraintParameter=12, tolerance=0.0, maximumIterations=400,
                                     totalStartingPoints=256, batchSize=256)
    specs = [(mats[0], long_st, unit_gaussian_starts(mats[0].n, 256, seed=40))]
    for k, f in enumerate((1, 2, 4, 7)):
        M = mats[k % 2]
        B = (32, 64)[k % 2]
        st = S.OptimizationSettings(formulation=f, constraintParameter=8, penaltyParameter=default_gamma(f, M.m),
                                    tolerance=1e-5, maximumIterations=60, totalStartingPoints=B, batchSize=B)
        specs.append((M, st, unit_gaussian_starts(M.n, B, seed=41 + k)))
    alone = []
    for M, st, X0 in specs:
        stats = S.OptimizationStatistics()
        x = M.solve(st, X0=X0, stats=stats)
        alone.append((x, stats.values.copy(), stats.fval, stats.it))
    ctx.start_server(1)
    try:
        first = specs[0][0].submit(specs[0][1], X0=specs[0][2])
        t0 = time.time()
        while first.poll() != S.JOB_RUNNING and time.time() - t0 < 30:
            time.sleep(0.001)
        assert first.poll() == S.JOB_RUNNING
        later = [M.submit(st, X0=X0) for M, st, X0 in specs[1:]]
        victim = specs[0][0].submit(long_st, X0=specs[0][2])  # taken by the group, then cancelled
        t0 = time.time()
        while victim.poll() != S.JOB_RUNNING and time.time() - t0 < 30:
            time.sleep(0.001)
        victim.cancel()
        for j, (xa, va, fa, ita) in zip([first] + later, alone):
            stats = S.OptimizationStatistics()
            x, _ = j.wait(stats=stats)
            assert stats.it == ita
            assert np.array_equal(x.view(np.uint32), xa.view(np.uint32))
            assert np.array_equal(stats.values.view(np.uint32), va.view(np.uint32)) and stats.fval == fa
        with pytest.raises(S.SpcaError) as e:
            victim.wait()
        assert e.value.code == S.SPCA_ECANCELLED
        info = ctx.server_info()
        print(f"  server: {info}", flush=True)
        assert info["groups"] == 1 and info["solo_jobs"] == 0
        # the four later jobs joined (the victim may be cancelled before it starts iterating)
        assert info["joined_running"] >= 4 and info["max_members"] >= 3
    finally:
        ctx.stop_server()
        for M in mats:
            M.release()


def test_memory_cache_reuses_blocks():
    """Freed device blocks are kept by size class and handed to the next request of the same
    class; held memory (live + cached) stays within the budget, which evicts cached blocks."""
    A, _ = spiked(200, 1000, seed=21)
    with S.Context(0, hbm_budget_bytes=64 << 20) as c:
        M = c.upload(A)
        used1, cached1, budget = c.memory()  # the host-to-device staging block is cached already
        assert budget == 64 << 20 and 0 < cached1 < used1 <= budget
        M.release()
        used2, cached2, _ = c.memory()
        assert used2 == used1 and cached2 == used1  # every block kept, still counted
        M = c.upload(A)
        assert c.memory() == (used1, cached1, budget)  # the same blocks came back
        M.release()
        # a larger matrix needs fresh blocks: cached ones are released to stay in budget
        B, _ = spiked(1000, 4000, seed=22)
        M = c.upload(B)
        used3, cached3, _ = c.memory()
        assert used3 <= budget
        x = M.solve(S.OptimizationSettings(formulation=0, constraintParameter=10, tolerance=0.0, maximumIterations=3,
                                           totalStartingPoints=32, batchSize=32),
                    X0=unit_gaussian_starts(4000, 32, seed=23))
        assert np.isfinite(x).all()
        M.release()


@pytest.mark.parametrize("f", [0, 1, 3, 4])
def test_sparse_products_match_dense(small, small_mat, f):
    """Once the iterates are sparse, A*X is gathered over their nonzeros in FP64; the solve
    agrees with the dense-product solve (and the sparse form actually ran)."""
    A, X0, _ = small
    out = {}
    for sparse in (0, 1):
        st = _settings(f, 30, S.PREC_FP32, X0.shape[1])
        st.sparse_products = sparse
        stats = S.OptimizationStatistics()
        x = small_mat.solve(st, X0=X0, stats=stats)
        out[sparse] = (x, stats)
    (xd, sd), (xs, ss) = out[0], out[1]
    assert sd.sparse_iterations == 0
    if f in (0, 1):  # L0 constraint: s = 10 nonzeros <= n/16, so every iteration after the first
        assert ss.sparse_iterations == ss.it - 1
    rel = np.abs(ss.values - sd.values) / np.abs(sd.values)
    assert np.mean(rel <= 1e-5) >= 0.95, np.sort(rel)[-4:]
    assert abs(ss.fval - sd.fval) <= 1e-5 * abs(sd.fval)


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
def test_sparse_flags_per_group_of_256(ctx, prec):
    """A*X switches between the dense product and the sparse gather per group of 256 SPs.
    Batch of 512, L1 variance with the L1 constraint, per-SP termination.  The L1 bound
    sqrt(150) keeps at least 150 nonzeros per iterate, more than the gather caps at n = 1000
    (62 in FP32, 7 in bf16x3), so iterates are never gathered.  The first 256 SPs start at
    fixed points and stop after one check, which leaves their group without an active SP, so
    its dense tiles stand down while the second group keeps the dense product.  Results are
    bitwise those of the all-dense solve."""
    A, _ = spiked(300, 1000, seed=11)
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    try:
        def run(X0, sparse, it):
            st = S.OptimizationSettings(formulation=3, constraintParameter=150, tolerance=1e-5, maximumIterations=it,
                                        totalStartingPoints=X0.shape[1], batchSize=X0.shape[1], precision=prec)
            st.termination = S.TERM_PER_SP
            st.sparse_products = sparse
            stats = S.OptimizationStatistics()
            x, X = M.solve(st, X0=X0, stats=stats, want_X=True)
            return X, stats
        Xc, _ = run(unit_gaussian_starts(1000, 256, seed=12), 0, 200)
        Xc = np.asfortranarray(Xc / np.linalg.norm(Xc, axis=0))
        X0 = np.asfortranarray(np.concatenate([Xc, unit_gaussian_starts(1000, 256, seed=13)], axis=1))
        Xd, sd = run(X0, 0, 60)
        Xs, ss = run(X0, 1, 60)
        nnz = np.count_nonzero(Xd, axis=0)
        print(f"prec {prec}: iterations {ss.it}, gathered {ss.sparse_iterations}, nonzeros per iterate "
              f"{nnz.min()}..{nnz.max()}, first group's iterations {np.max(np.asarray(sd.iters)[:256])}", flush=True)
        assert nnz.min() > 62
        assert sd.sparse_iterations == 0 and ss.sparse_iterations > 0
        np.testing.assert_array_equal(Xs, Xd)
        np.testing.assert_array_equal(ss.values, sd.values)
        assert ss.best_sp == sd.best_sp
    finally:
        M.release()


@pytest.mark.parametrize("B", [64, 256])
@pytest.mark.parametrize("f", [1, 3, 5, 7])
def test_fixed_point_sign_products_bitwise(ctx, f, B):
    """A^T sgn(z) from the 25-bit fixed-point form of A is exact when A is: on a matrix whose
    entries are k * 2^-19 with |k| < 2^19 (so each column's scale is 1 and every entry is a
    whole number of 2^-25 units) and whose FP32 sums are all exact (32 rows: every partial sum
    is a multiple of 2^-19 below 2^5), the fixed-point product and the FP32 product agree
    bitwise, and so do whole L1-variance trajectories.  The entries carry up to 20 significant
    bits, so all four pieces are used.  Penalties keep about 5% of the coordinates
    ((A^T sgn z)_i has standard deviation sqrt(32/3))."""
    rng = np.random.default_rng(30 + f)
    m, n = 32, 1024
    A = (rng.integers(-2**19, 2**19, size=(m, n)) * 2.0**-19).astype(np.float32)
    X0 = unit_gaussian_starts(n, B, seed=31)
    gamma = {5: 43.0, 7: 6.5}.get(f, 0.0)
    M = ctx.upload(A)
    try:
        out, nbytes = {}, {}
        for exact in (0, 1):
            st = S.OptimizationSettings(formulation=f, constraintParameter=24, penaltyParameter=gamma,
                                        tolerance=0.0, maximumIterations=6, totalStartingPoints=B, batchSize=B,
                                        exact_sign_products=exact)
            stats = S.OptimizationStatistics()
            _, X = M.solve(st, X0=X0, stats=stats, want_X=True)
            out[exact] = (X, stats)
            nbytes[exact] = M.nbytes()
        nnz = np.count_nonzero(out[1][0], axis=0)
        print(f"case f={f} B={B}: nnz per SP min {nnz.min()} max {nnz.max()}", flush=True)
        assert nnz.min() > 0 and np.all(np.isfinite(out[1][0]))
        np.testing.assert_array_equal(out[1][0], out[0][0])
        np.testing.assert_array_equal(out[1][1].values, out[0][1].values)
        assert nbytes[1] > nbytes[0]  # the fixed-point pieces were added on first use
    finally:
        M.release()


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
def test_fixed_point_sign_products_against_fp64(small, small_mat, prec):
    """L1 variance with the L1 constraint, 40 iterations: the fixed-point product and the
    precision's own A^T*Z product both stay within the precision's EV bound of the FP64
    oracle, and the fixed-point product's error (seen through the final iterates) is no
    larger than that of the product it replaces."""
    A, X0, _ = small
    f, B, it = 3, X0.shape[1], 40
    o = O.solve(A.astype(np.float64), f, s=10, gamma=0.0, tol=0.0, max_iter=it, total_sp=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    ev_o = explained_variance(A, o["X"], f)
    errs = {}
    for exact in (0, 1):
        st = _settings(f, it, prec, B)
        st.exact_sign_products = exact
        stats = S.OptimizationStatistics()
        _, Xg = small_mat.solve(st, X0=X0, stats=stats, want_X=True)
        same = np.all((Xg != 0) == (o["X"] != 0), axis=0)
        rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
        errs[exact] = np.median(rel[same])
        print(f"prec {prec} exact {exact}: supports equal {same.mean():.3f} EV rel median {errs[exact]:.2e} "
              f"max {rel[same].max():.2e}", flush=True)
        assert same.mean() >= 0.9
        assert np.mean(rel[same] <= TOL[prec][0]) >= 0.95
    assert errs[1] <= 2 * errs[0] + 1e-9


def _planted_near_zero(m=512, n=4096, B=128, planted_sps=8, rows_per_sp=8, delta=1e-8, seed=41):
    """A (FP32) and starting points where, for SPs 0..planted_sps-1, rows_per_sp rows of
    z = A x_j sit at +-delta (FP64 of the FP32 data, alternating signs): far outside FP64's
    rounding, far inside that of the FP32 and bf16x3 products (~1e-7 to 1e-5 here).  Per
    planted row, A[i, k] moves at the coordinate k where |x_j| is largest, then A[i, k2] at the
    one where it is smallest takes up what FP32 rounding of the first left."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)).astype(np.float32)
    X0 = unit_gaussian_starts(n, B, seed=seed + 1)
    planted = []
    for j in range(planted_sps):
        x = X0[:, j].astype(np.float64)
        k, k2 = int(np.argmax(np.abs(x))), int(np.argmin(np.abs(x)))
        for r in range(rows_per_sp):
            i = j * rows_per_sp + r
            target = delta if r % 2 == 0 else -delta
            for kk in (k, k2):
                z = A[i].astype(np.float64) @ x
                A[i, kk] = np.float32(A[i, kk] - (z - target) / x[kk])
            planted.append((i, j))
    Z = A.astype(np.float64) @ X0.astype(np.float64)
    for i, j in planted:
        assert 0.5 * delta < abs(Z[i, j]) < 2 * delta, Z[i, j]
    return A, X0, planted


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
def test_exact_signs_of_near_zero_z(ctx, prec):
    """L1 variance, L0 constraint, one step from dense starting points (the dense A*X
    product): entries of z planted at +-1e-8, inside the product's rounding, get the sign of
    exact arithmetic (the FP64 recompute of flagged entries), so the step equals the FP64
    oracle's: identical supports and iterates within 1e-5.  With exact_sign_products = 0
    (the precision's own products, no recompute) some planted signs come out wrong, which
    shows up as iterates off by far more than rounding."""
    A, X0, planted = _planted_near_zero()
    B, f, s = X0.shape[1], 1, 64
    o = O.solve(A.astype(np.float64), f, s=s, tol=0.0, max_iter=1, total_sp=B, batch=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    Xo = o["X"]
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    try:
        out = {}
        for exact in (1, 0):
            st = S.OptimizationSettings(formulation=f, constraintParameter=s, tolerance=0.0, maximumIterations=1,
                                        totalStartingPoints=B, batchSize=B, precision=prec, exact_sign_products=exact)
            _, X = M.solve(st, X0=X0, want_X=True)
            out[exact] = X
        dev = {e: np.abs(out[e] - Xo).max(axis=0) for e in out}
        pj = sorted({j for _, j in planted})
        print(f"case prec={prec}: max |x - x_fp64| exact {dev[1].max():.2e}, plain on planted SPs "
              f"{np.round(dev[0][pj], 6).tolist()}", flush=True)
        assert np.all((out[1] != 0) == (Xo != 0))
        assert dev[1].max() <= 1e-5
        assert dev[0][pj].max() > 1e-4  # the plain products mis-sign planted entries
    finally:
        M.release()


def _planted_top_s_ties(m=2048, n=4096, B=32, s=32, gap=1.5e-7, seed=51):
    """FP32 A and one-coordinate starting points e_c (so z = A[:, c] exactly), with, for half
    the SPs, the s-th and (s+1)-th largest |v| of v = A^T z moved to within `gap` (relative) of
    each other: above FP32's resolution of v, below the FP32 product's rounding (~1e-6 at this
    m).  The (s+1)-th column is moved at the row where |z| is largest.  Returns A, X0 and the
    SPs whose planted gap holds in FP64 after all the moves."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)).astype(np.float32)
    cols = rng.choice(n, B, replace=False)
    X0 = np.zeros((n, B), dtype=np.float32)
    X0[cols, np.arange(B)] = 1.0
    for j in range(B // 2):
        c = cols[j]
        z = A[:, c].astype(np.float64)
        v = A.T.astype(np.float64) @ z
        order = np.argsort(-np.abs(v), kind="stable")
        i_s, i_t = order[s - 1], order[s]
        if c in (i_s, i_t):
            continue
        r = int(np.argmax(np.abs(z)))
        target = np.sign(v[i_t]) * abs(v[i_s]) * (1 - gap)
        A[r, i_t] = np.float32(A[r, i_t] + (target - v[i_t]) / z[r])
    V = A.T.astype(np.float64) @ A[:, cols].astype(np.float64)
    planted = []
    for j in range(B // 2):
        a = np.sort(np.abs(V[:, j]))[::-1]
        rel = (a[s - 1] - a[s]) / a[s - 1]
        if 0.5 * gap < rel < 2 * gap:
            planted.append(j)
    return A, X0, planted


def test_exact_ties_at_the_s_th_place(ctx):
    """L0 constraint, L2 variance, FP32, one step: candidates for the s-th place within the
    product's rounding of each other are re-evaluated in FP64 (exact_ties), so the step equals
    the FP64 oracle's on every SP, planted near-ties included; with exact_ties = 0 the
    product's own order decides and some planted SPs keep the wrong coordinate."""
    A, X0, planted = _planted_top_s_ties()
    B, s = X0.shape[1], 32
    assert len(planted) >= 12, planted
    o = O.solve(A.astype(np.float64), 0, s=s, tol=0.0, max_iter=1, total_sp=B, batch=B, X0=X0.astype(np.float64),
                want_X=True, lambda_rule=O.LAMBDA_GPU)
    Xo = o["X"]
    M = ctx.upload(A)
    try:
        out = {}
        for ties in (1, 0):
            st = S.OptimizationSettings(formulation=0, constraintParameter=s, tolerance=0.0, maximumIterations=1,
                                        totalStartingPoints=B, batchSize=B, exact_ties=ties)
            _, X = M.solve(st, X0=X0, want_X=True)
            out[ties] = X
    finally:
        M.release()
    off = {t: np.nonzero(~np.all((out[t] != 0) == (Xo != 0), axis=0))[0] for t in out}
    print(f"case planted {len(planted)}: SPs off the FP64 support, exact ties {off[1].tolist()}, "
          f"plain {off[0].tolist()}", flush=True)
    assert off[1].size == 0
    assert np.abs(out[1] - Xo).max() <= 1e-5
    assert np.intersect1d(off[0], planted).size >= 1  # the plain product misorders some planted ties


@pytest.mark.parametrize("f", [4, 6])
def test_exact_ties_at_the_penalty_threshold(ctx, f):
    """Penalised L2 variance, FP32, one step from e_c: gamma (an FP32 value) is set so that one
    coordinate's |v| sits 1.5e-7 (relative, 1.2e-7 to 2.4e-7 after rounding gamma) above or
    below the threshold (sqrt(gamma) for the L0 penalty, gamma for the L1 penalty): outside the
    rounding that FP32 z leaves in v (at most 9e-8 here), inside the product's (median 1.7e-7).
    With exact_ties the band entry is recomputed in FP64 and the step equals the FP64 oracle's;
    with exact_ties = 0 some of the 24 planted cases come out the other way."""
    rng = np.random.default_rng(61)
    m, n, g_rel = 2048, 4096, 1.5e-7
    A = rng.standard_normal((m, n)).astype(np.float32)
    A64 = A.astype(np.float64)
    M = ctx.upload(A)
    wrong = {0: 0, 1: 0}
    try:
        for case, c in enumerate(rng.choice(n, 24, replace=False)):
            z = A64[:, c] / np.linalg.norm(A64[:, c])
            v = A64.T @ z
            sgn = 1 if case % 2 == 0 else -1  # coordinate just above / just below the threshold
            for rank in range(40, 80):
                a = abs(v[np.argsort(-np.abs(v))[rank]])
                t = a / (1 + sgn * g_rel)
                gamma = float(np.float32(t * t if f == 4 else t))
                thr = np.sqrt(gamma) if f == 4 else gamma
                if 0.8 * g_rel <= sgn * (a - thr) / thr <= 1.6 * g_rel:
                    break
            else:
                pytest.fail("no coordinate to plant")
            X0 = np.zeros((n, 1), dtype=np.float32)
            X0[c, 0] = 1.0
            o = O.solve(A64, f, gamma=gamma, tol=0.0, max_iter=1, total_sp=1, batch=1, X0=X0.astype(np.float64),
                        want_X=True, lambda_rule=O.LAMBDA_GPU)
            for ties in (1, 0):
                st = S.OptimizationSettings(formulation=f, penaltyParameter=gamma, tolerance=0.0, maximumIterations=1,
                                            totalStartingPoints=1, batchSize=1, exact_ties=ties)
                _, X = M.solve(st, X0=X0, want_X=True)
                same = np.array_equal(X[:, 0] != 0, o["X"][:, 0] != 0)
                if ties:
                    assert same, (case, c)
                    np.testing.assert_allclose(X[:, 0], o["X"][:, 0], rtol=1e-5, atol=1e-6 * np.abs(o["X"]).max())
                wrong[ties] += 0 if same else 1
    finally:
        M.release()
    print(f"case f={f}: planted threshold ties decided off FP64: exact ties {wrong[1]}, plain {wrong[0]} of 24",
          flush=True)
    assert wrong[0] >= 1


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
@pytest.mark.parametrize("f", range(8))
def test_union_products_match_dense(ctx, f, prec):
    """sparse_products = 2: once the union U of the iterates' supports is small, A*X is the FP32
    product over U (the per-SP gather keeps the groups it covers).  The solve agrees with the
    all-dense solve of the same precision; with a strong planted loading the L1-constrained
    iterates (too many nonzeros for the gather) share a small U, so the union product runs.
    The gather takes iterates of up to n/16 nonzeros in FP32 and n/128 in bf16x3, the union
    product unions of up to n/4 and n/12.  The FP32 L1-constrained cases run at n = 400 with a
    30-coordinate loading and the L1 bound sqrt(20): from the sixth iteration on the FP64
    oracle's iterates hold 49-58 nonzeros (above 25) in unions of 52-75 (below 100).  The
    others run at n = 1000 with a 20-coordinate loading and the bound sqrt(10)."""
    n, k, c = (400, 30, 20) if f in (2, 3) and prec == S.PREC_FP32 else (1000, 20, 10)
    A, _ = spiked(200, n, k=k, strength=8.0, seed=41)
    X0 = unit_gaussian_starts(n, 64, seed=42)
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    try:
        out = {}
        for sparse in (0, 2):
            st = _settings(f, 30, prec, 64, s=c)
            st.sparse_products = sparse
            stats = S.OptimizationStatistics()
            x = M.solve(st, X0=X0, stats=stats)
            out[sparse] = (x, stats)
        (xd, sd), (xu, su) = out[0], out[2]
        print(f"f={f} prec={prec}: iterations {su.it}, union {su.union_iterations}, gathered "
              f"{su.sparse_iterations}", flush=True)
        assert sd.union_iterations == 0 and sd.sparse_iterations == 0
        assert su.union_iterations + su.sparse_iterations <= su.it - 1  # the first A*X is dense
        if f in (2, 3):
            assert su.union_iterations > 0
        rel = np.abs(su.values - sd.values) / np.abs(sd.values)
        tol = 1e-5 if prec == S.PREC_FP32 else 1e-4
        assert np.mean(rel <= tol) >= 0.95, np.sort(rel)[-4:]
        assert abs(su.fval - sd.fval) <= tol * abs(sd.fval)
        assert set(np.nonzero(xu)[0]) == set(np.nonzero(xd)[0])
    finally:
        M.release()


@pytest.mark.parametrize("f", range(8))
def test_narrow_fp32_tiles_bitwise(ctx, f):
    """FP32 batches of at most 64 SPs run the products in 128 x 32 / 128 x 64 tiles instead of
    padding the batch to 128 columns.  Every output element sees the same MFMA sequence over
    the same K split, so with dense products, tol = 0 and a fixed iteration count (no coupling
    between SPs) a batch of 32 (narrow, 32 wide), 64 (narrow, 64 wide) and 96 (the 128 x 128
    tile) give bitwise the same iterates and values for the starting points they share."""
    A, _ = spiked(300, 2000, seed=51)
    M = ctx.upload(A)
    X0 = unit_gaussian_starts(2000, 96, seed=52)
    try:
        res = {}
        for B in (32, 64, 96):
            st = _settings(f, 12, S.PREC_FP32, B, m=300)
            st.sparse_products = 0
            stats = S.OptimizationStatistics()
            _, X = M.solve(st, X0=np.asfortranarray(X0[:, :B]), stats=stats, want_X=True)
            res[B] = (X, stats.values)
        for B in (32, 64):
            np.testing.assert_array_equal(res[B][0], res[96][0][:, :B])
            np.testing.assert_array_equal(res[B][1], res[96][1][:B])
    finally:
        M.release()


@pytest.mark.parametrize("f", [0, 1, 4])
def test_otf_sparse_products_match_dense(small, small_mat, f):
    """OTF solves gather A*X over the nonzeros once every active slot of a group (refilled
    slots included) is sparse.  With the reference's constrained starting points (s nonzeros
    each) the gather runs from the first step on; the per-SP results agree with the all-dense
    OTF solve, and so do the iteration counts."""
    A, _, _ = small
    g = default_gamma(f, A.shape[0])
    out = {}
    for sparse in (0, 1):
        st = S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=g, tolerance=1e-4,
                                    maximumIterations=30, totalStartingPoints=200, batchSize=32, useOTF=True,
                                    start=S.START_REFERENCE_F64)
        st.sparse_products = sparse
        stats = S.OptimizationStatistics()
        x = small_mat.solve(st, stats=stats)
        out[sparse] = (x, stats)
    (xd, sd), (xs, ss) = out[0], out[1]
    print(f"f={f}: iterations {ss.it}/{sd.it}, gathered {ss.sparse_iterations}", flush=True)
    assert sd.sparse_iterations == 0
    if f in (0, 1):
        assert ss.sparse_iterations >= ss.it // 2
    with np.errstate(invalid="ignore"):
        rel = np.abs(ss.values - sd.values) / np.abs(sd.values)
    assert np.nanmean(rel <= 1e-5) >= 0.95, np.sort(rel)[-4:]
    assert np.mean(ss.iters == sd.iters) >= 0.95
    assert abs(ss.fval - sd.fval) <= 1e-5 * abs(sd.fval)
