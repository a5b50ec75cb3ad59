"""FP32 and bf16x3 against FP64 at the headline shape, all 8 formulations.  Needs an MI355X;
slow (the FP64 oracle solves on the host with host BLAS).

4096 x 65536 spiked model (bench.make_problem), s = 256, gamma per penalised formulation as
the bench calibrates it, tol = 1e-6, at most 200 iterations, on the first 512 of the bench's
unit Gaussian starting points.  Per starting point: the explained variance of the GPU's final
iterate (FP64, on the device) against that of the FP64 oracle's final iterate, and whether the
supports are identical.  The explained variance is compared over the starting points both
sides converged (their last relative change below tol); the L1 penalty with L2 variance
leaves some points still moving after 200 iterations, where the two trajectories are compared
mid-way and differ by up to ~1e-4 without a change of support.

The solve is nonconvex: a starting point whose trajectory passes near a tie can settle on
another fixed point.  Two kinds of tie occur:
  * two candidates for the s-th place (or for the gamma / lambda threshold) within the
    rounding of the product that computed them;
  * for L1 variance, an entry of z = A x within rounding of zero, whose sign flips.  One flip
    moves v = A^T sgn(z) by 2 |A_i| (about 1.5% of |v| here), and the trajectory can leave
    for another fixed point.  Such entries arise at about 1e-6 of the entries' scale, so
    A^T sgn(z) runs in exact fixed point: an FP32 accumulation of it put ~4e-7 of error into
    x, enough to flip such an entry at the next iteration.
and the entries of z within the dense product's rounding of zero are recomputed in FP64.

Sample: 512 of the bench's 1024 starting points, with per-SP termination on both sides (each SP
stops once its own relative change is below tol, as the reference OTF driver terminates), so
the FP64 oracle's host work is the SPs' own iterations rather than the slowest SP's.
Bars (the parity contract): on every SP that converged, the support is the oracle's except at a
documented magnitude tie -- each coordinate in which the supports differ sat within 1e-5 (FP32)
/ 1e-4 (bf16x3) of the last step's threshold, recomputed in FP64, or (an SP that settled on
another fixed point) a step of its FP64 trajectory had a tie margin within that -- and the FP64 explained
variance is within 1e-6 (FP32) / 1e-5 (bf16x3) on every SP that converged on the oracle's
support, unless the reference method itself computed in FP32 lands farther from FP64 on that SP
(a trajectory too ill-conditioned for single precision; measured on slowly converging SPs of
the L1 penalty, where the reference's FP32 path is off by 6e-5 to 1e-3 and this solve by ~1e-5):
there the solve must be at least as close as the reference's FP32 path.  An SP whose stop lands one iteration apart (its last change within rounding of tol)
is compared with the oracle run to the GPU's own iteration count.  The share of identical
supports is printed for the >= 99% target; per-SP termination stops an SP as soon as its
objective settles, which can leave a coordinate of the L1 thresholds sitting on lambda.
"""
import os
import time

import numpy as np
import pytest

import oracle as O
import spca_mi355x as S
from spca_cases import trajectory_tie_margin

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

NSP = 512
EV_TOL = {S.PREC_FP32: 1e-6, S.PREC_BF16X3: 1e-5}


@pytest.fixture(scope="module")
def headline(ctx):
    import bench
    A, X0, loadings = bench.make_problem()
    gam = bench.calibrate_gamma(A, loadings[0])
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3 | S.STORE_FP16)
    yield A, A.astype(np.float64), np.asfortranarray(X0[:, :NSP]), gam, M
    M.release()


TIE_TOL = {S.PREC_FP32: 1e-5, S.PREC_BF16X3: 1e-4}
# fp16's stated bound: median relative FP64 explained-variance error over the converged SPs
# (measured 5e-10 .. 5e-4 over the 8 formulations; identical supports on 6-77% of the SPs, the
# others on other fixed points: fp16 carries no parity claim)
FP16_BOUND_MEDIAN = 1e-3
ZTOL = {S.PREC_FP32: 1e-6, S.PREC_BF16X3: 3e-5}  # an entry of z = A x within rounding of zero (L1 variance)


def _ev(A64, X, l2):
    nrm = np.linalg.norm(X, axis=0)
    Z = A64 @ (X / np.where(nrm > 0, nrm, 1.0))
    return np.sum(Z * Z, axis=0) if l2 else np.sum(np.abs(Z), axis=0)


def _last_step_v(A64, f, x_prev):
    """v = A^T map(A x) of the step from x_prev (FP64): the values the threshold was applied to."""
    z = A64 @ x_prev
    if f in (0, 2):
        pass
    elif f in (4, 6):
        z = z / np.linalg.norm(z)
    else:
        z = np.sign(z)
    return A64.T @ z


def _threshold(v, f, s, gamma):
    a = np.abs(v)
    if f in (0, 1):
        return np.sort(a)[::-1][s - 1]
    if f in (2, 3):
        return O.soft_threshold(v, s, O.LAMBDA_GPU)[2]
    return np.sqrt(gamma) if f in (4, 5) else gamma


@pytest.mark.parametrize("f", range(8))
def test_headline_precision_against_fp64(headline, f):
    """Every converged SP ends on the oracle's support or differs from it only at a documented
    magnitude tie (coordinates on the last step's threshold, or a tie at an earlier step of its FP64
    trajectory), and the explained variance is within the precision's tolerance on every SP that
    converged on the oracle's support."""
    A, A64, X0, gam, M = headline
    threads = min(os.cpu_count() or 1, 16)
    gamma = gam.get(f, 0.0)
    l2 = f in (0, 2, 4, 6)
    common = dict(s=256, gamma=gamma, threads=threads, gemm_backend=1, want_X=True, lambda_rule=O.LAMBDA_GPU)
    t0 = time.time()
    o = O.solve(A64, f, tol=1e-6, max_iter=200, total_sp=NSP, batch=NSP, X0=X0.astype(np.float64),
                termination=O.TERM_PER_SP, **common)
    t1 = time.time()
    ev_o = _ev(A64, o["X"], l2)
    print(f"\nf={f}: oracle {o['loop_s']:.1f} s loop, {t1 - t0:.1f} s wall, EV {time.time() - t1:.1f} s; iterations "
          f"per SP median {np.median(o['sp_iters']):.0f} max {o['sp_iters'].max()}, converged "
          f"{(o['sp_iters'] < 200).sum()}/{NSP}", flush=True)

    def oracle_snapshots(need):
        """need: {SP: set of iteration counts}.  The oracle's iterates of those SPs after each of
        their counts, from one chain of single iterations over all of them (one host-BLAS product
        pair per step, the solve's own trajectory)."""
        cols = np.array(sorted(need), dtype=int)
        out = {}
        if not cols.size:
            return out
        X = X0[:, cols].astype(np.float64)
        top = max(max(need[int(j)]) for j in cols)
        for it in range(0, top + 1):
            if it > 0:
                X = O.solve(A64, f, tol=0.0, max_iter=1, total_sp=cols.size, batch=cols.size, X0=X, **common)["X"]
            for c, j in enumerate(cols):
                if it in need[int(j)]:
                    out[(int(j), it)] = X[:, c].copy()
            if it and it % 25 == 0:
                print(f"  chain step {it}/{top} over {cols.size} SPs", flush=True)
        return out

    for prec in (S.PREC_FP32, S.PREC_BF16X3):
        t0 = time.time()
        st = S.OptimizationSettings(formulation=f, constraintParameter=256, penaltyParameter=gamma, tolerance=1e-6,
                                    maximumIterations=200, totalStartingPoints=NSP, batchSize=NSP, precision=prec,
                                    termination=S.TERM_PER_SP)
        stats = S.OptimizationStatistics()
        _, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
        t1 = time.time()
        conv = (o["sp_iters"] < 200) & (stats.iters < 200)
        # The reference iterate of each SP is the oracle's at the GPU's own iteration count (an SP
        # whose last change sits within rounding of tol can stop one iteration apart).  SPs off the
        # oracle's support also need the iterate before their last step, for the tie check.
        late = np.nonzero(conv & (stats.iters != o["sp_iters"]))[0]
        off0 = np.nonzero(conv & ~np.all((Xg != 0) == (o["X"] != 0), axis=0))[0]
        need = {}
        for j in late:
            need.setdefault(int(j), set()).update({int(stats.iters[j]), int(stats.iters[j]) - 1})
        for j in off0:
            need.setdefault(int(j), set()).add(int(stats.iters[j]) - 1)
        print(f"f={f} prec={prec}: GPU {t1 - t0:.1f} s; {late.size} stops apart, chain over {len(need)} SPs",
              flush=True)
        snap = oracle_snapshots(need)
        Xref, ev_ref = o["X"].copy(), ev_o.copy()
        for j in late:
            Xref[:, j] = snap[(int(j), int(stats.iters[j]))]
        if late.size:
            ev_ref[late] = _ev(A64, Xref[:, late], l2)
        same = np.all((Xg != 0) == (Xref != 0), axis=0)
        # a converged SP off the oracle's support: every coordinate in which the supports differ
        # must have sat within the product's rounding of the threshold at the last step (FP64)
        untied = []
        for j in np.nonzero(~same & conv)[0]:
            key = (int(j), int(stats.iters[j]) - 1)
            if key not in snap:  # the support differs only at the GPU's stop count (a late SP)
                snap.update(oracle_snapshots({int(j): {key[1]}}))
            v = _last_step_v(A64, f, snap[key])
            thr = _threshold(v, f, 256, gamma)
            diff = np.nonzero((Xg[:, j] != 0) != (Xref[:, j] != 0))[0]
            gap = np.abs(np.abs(v[diff]) - thr) / thr
            if gap.max() >= TIE_TOL[prec]:
                untied.append((int(j), int(diff.size), float(gap.max())))
        # not a tie at the last step: the SP branched earlier; its own FP64 trajectory must pass a
        # tie (spca_cases.trajectory_tie_margin, one SP at a time)
        branched = []
        for (j, nd, gap) in untied:
            margin, _ = trajectory_tie_margin(A64, X0[:, [j]], f, 256, gamma, int(stats.iters[j]), ztol=ZTOL[prec],
                                              gemm_backend=1, threads=threads)
            print(f"  SP {j}: {nd} coordinates off, last-step gap {gap:.2e}; smallest tie margin on its trajectory "
                  f"{margin[0]:.2e}", flush=True)
            if not margin[0] < TIE_TOL[prec]:
                branched.append((j, float(margin[0])))
        rel = np.abs(stats.sp_ev_fp64 - ev_ref) / ev_ref
        chk = conv & same
        print(f"f={f} prec={prec}: checks {time.time() - t1:.1f} s; identical supports {same.sum()}/{NSP} "
              f"({same[conv].mean():.4f} of the {conv.sum()} converged; off: {np.nonzero(~same)[0].tolist()}, "
              f"converged and not at a tie: {untied}); EV rel over {chk.sum()} converged on-support SPs max "
              f"{rel[chk].max():.2e} median {np.median(rel[chk]):.2e}; unconverged GPU/oracle "
              f"{(stats.iters >= 200).sum()}/{(o['sp_iters'] >= 200).sum()}", flush=True)
        over = np.nonzero(chk & (rel > EV_TOL[prec]))[0]
        rel32 = np.zeros(0)
        if over.size:
            # how the reference method computed in FP32 does on the same SPs (its own rounding)
            r32 = O.solve(A, f, tol=1e-6, max_iter=200, total_sp=over.size, batch=over.size,
                          X0=X0[:, over], termination=O.TERM_PER_SP, s=256, gamma=gamma, threads=threads,
                          gemm_backend=1, want_X=True, lambda_rule=O.LAMBDA_GPU)
            rel32 = np.abs(_ev(A64, r32["X"].astype(np.float64), l2) - ev_o[over]) / ev_o[over]
            for c, j in enumerate(over):
                print(f"  over the bar: SP {j} EV rel {rel[j]:.2e}, iterations GPU {stats.iters[j]} oracle "
                      f"{o['sp_iters'][j]} (last change {o['sp_err'][j]:.1e}); the FP32 oracle: iterations "
                      f"{r32['sp_iters'][c]}, EV rel {rel32[c]:.2e}", flush=True)
        assert not branched, (prec, branched)
        # over the bar only where the trajectory is so ill-conditioned that the reference method in
        # FP32 lands farther from FP64 than this solve does (slowly converging SPs of the L1 penalty)
        assert np.all(rel[over] <= rel32), (prec, over, rel[over], rel32)

    # fp16 (11-bit significands): no parity claim (a starting point can settle on another fixed
    # point); its stated error bound against FP64
    st = S.OptimizationSettings(formulation=f, constraintParameter=256, penaltyParameter=gamma, tolerance=1e-6,
                                maximumIterations=200, totalStartingPoints=NSP, batchSize=NSP,
                                precision=S.PREC_FP16, termination=S.TERM_PER_SP)
    stats = S.OptimizationStatistics()
    _, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
    conv = (o["sp_iters"] < 200) & (stats.iters < 200)
    same = np.all((Xg != 0) == (o["X"] != 0), axis=0)
    rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
    chk = conv & same
    q = lambda x: (f"median {np.median(x):.2e} p99 {np.percentile(x, 99):.2e} max {x.max():.2e}" if x.size else "-")
    print(f"f={f} fp16: identical supports {same.sum()}/{NSP} ({same[conv].mean():.4f} of the {conv.sum()} "
          f"converged); EV rel, converged on-support: {q(rel[chk])}; all converged: {q(rel[conv])}", flush=True)
    assert FP16_BOUND_MEDIAN is None or np.median(rel[conv]) <= FP16_BOUND_MEDIAN, (f, np.median(rel[conv]))
