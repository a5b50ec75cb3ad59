"""The workload configurations the solver is specified on, run through the C ABI against the
FP64 oracle.  The generators live in bench.py (one definition for the bench and the tests).

c1  the reference CPU path's configuration: 400 x 10304 clipped, centred low-rank + noise
    matrix, L2 variance / L0 constraint, s = 16, 64 unit Gaussian starting points, tol 1e-6,
    at most 200 iterations (c++/src/gpower/sparse_PCA_solver.h:36-243 is the method).
c4  64 uneven jobs (bench.uneven_specs) in one fused solve against 64 solo solves, and a sample
    of jobs per formulation against the oracle.
c5  a burst of requests of the serving stream (bench.serving_draws) through the request server,
    each result against the same job solved alone, and the smallest jobs against the oracle.

Parity bar (per starting point, against the FP64 oracle on the same FP32 data and starts):
every SP whose FP64 trajectory passes no magnitude tie ends on the oracle's support with its
explained variance within 1e-6 (FP32) or 1e-5 (bf16x3) relative; a tie is a step whose
threshold and nearest competitor lie within 1e-5 (FP32) / 1e-4 (bf16x3) of each other
(spca_cases.trajectory_tie_margin).
"""
import hashlib
import os

import numpy as np
import pytest

import bench
import oracle as O
import spca_mi355x as S
from spca_cases import explained_variance, trajectory_tie_margin

# SHA-256 of the c1 matrix (FP64, column major) and starting points from bench.make_c1_problem(1)
C1_SHA = {"A": "f78994672cb9a363c89c0ffc8e34bdd22779f502cb5c9b7cafd6516da97c3841",
          "X0": "15532fccb880cb4b7c80919287e1440cf3ba0e7560537a599cb06c20fabb638c"}
EV_TOL = {S.PREC_FP32: 1e-6, S.PREC_BF16X3: 1e-5}
TIE_TOL = {S.PREC_FP32: 1e-5, S.PREC_BF16X3: 1e-4}


def test_c1_generator_is_pinned():
    """The c1 data are what the stated construction gives: the recorded SHA-256, centred columns,
    and the clip at 0 visible as each column's repeated minimum (about half the entries)."""
    A, X0 = bench.make_c1_problem()
    assert A.shape == bench.C1_SHAPE and X0.shape == (bench.C1_SHAPE[1], bench.C1_B)
    assert hashlib.sha256(A.tobytes(order="F")).hexdigest() == C1_SHA["A"]
    assert hashlib.sha256(X0.tobytes(order="F")).hexdigest() == C1_SHA["X0"]
    np.testing.assert_allclose(A.mean(axis=0), 0.0, atol=1e-15)
    np.testing.assert_allclose(np.linalg.norm(X0, axis=0), 1.0)
    lo = A.min(axis=0)
    clipped = np.mean(np.abs(A - lo) <= 1e-15, axis=0)  # entries that were clipped to 0
    assert 0.3 < np.median(clipped) < 0.7
    assert np.all(A.max(axis=0) - lo <= 1.0)


@pytest.fixture(scope="module")
def c1(ctx):
    A64, X0 = bench.make_c1_problem()
    A = A64.astype(np.float32)
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    yield A, np.asfortranarray(X0.astype(np.float32)), M
    M.release()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
def test_c1_against_fp64_oracle(c1, prec):
    A, X0, M = c1
    B, s = bench.C1_B, bench.C1_S
    print(f"case c1 prec={prec}", flush=True)
    st = S.OptimizationSettings(formulation=0, constraintParameter=s, tolerance=1e-6, maximumIterations=200,
                                totalStartingPoints=B, batchSize=B, precision=prec)
    stats = S.OptimizationStatistics()
    x, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
    A64 = A.astype(np.float64)
    o = O.solve(A64, 0, s=s, tol=1e-6, max_iter=200, total_sp=B, X0=X0.astype(np.float64), want_X=True,
                lambda_rule=O.LAMBDA_GPU)
    margin, _ = trajectory_tie_margin(A, X0, 0, s, 0.0, int(o["it"]), gemm_backend=1,
                                      threads=min(os.cpu_count() or 1, 16))
    tied = margin < TIE_TOL[prec]
    ev_o = explained_variance(A, o["X"], 0)
    ev_rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
    supp = np.all((Xg != 0) == (o["X"] != 0), axis=0)
    print(f"  iterations {stats.it}/{o['it']} fval {stats.fval:.9g}/{o['fval']:.9g} tied {tied.sum()} supports "
          f"{supp.mean():.3f} EV rel max {ev_rel[~tied].max() if (~tied).any() else 0:.2e}", flush=True)
    assert abs(int(stats.it) - int(o["it"])) <= 1
    assert np.all(supp[~tied]), np.nonzero(~tied & ~supp)[0]
    assert np.all(ev_rel[~tied] <= EV_TOL[prec]), np.sort(ev_rel[~tied])[-4:]
    assert abs(stats.fval - o["fval"]) <= EV_TOL[prec] * 10 * o["fval"]
    assert stats.ev_fp64 == pytest.approx(explained_variance(A, x[:, None], 0)[0], rel=1e-9)


# ---------------------------------------------------------------- c4: 64 uneven jobs
C4_ITERS = 5
C4_ORACLE_SPS = 32


def _gaussian_starts(n, count, seed):
    """The device's SPCA_START_GAUSSIAN starting points restated in NumPy (SP k from Philox
    counters (i, k, 4)); they agree with the device's to float rounding."""
    from test_generate import normal
    X = np.stack([normal(np.arange(n), k, 4, seed) for k in range(count)], axis=1)
    return (X / np.linalg.norm(X, axis=0)).astype(np.float32)


@pytest.mark.gpu
def test_c4_uneven_jobs_fused_equal_solo_and_oracle(ctx):
    specs = bench.uneven_specs(64)
    mats, jobs = [], []
    print(f"case c4: {len(specs)} jobs, {C4_ITERS} iterations", flush=True)
    try:
        for sp in specs:
            M = ctx.generate(sp["m"], sp["n"], sp["seed"], theta=bench.THETA, spike_nnz=bench.SPARSITY,
                             key=sp["seed"])
            mats.append(M)
            st = S.OptimizationSettings(formulation=sp["f"], constraintParameter=sp["s"],
                                        penaltyParameter=bench.noise_gamma(sp["f"], sp["s"], sp["m"], sp["n"]),
                                        tolerance=0.0, maximumIterations=C4_ITERS, totalStartingPoints=sp["B"],
                                        batchSize=sp["B"], precision=S.PREC_FP32, start=S.START_GAUSSIAN,
                                        seed=sp["seed"], check_every=C4_ITERS)
            jobs.append((M, st, None))
        fused = ctx.solve_fused(jobs, graph=True)
        print("  fused done", flush=True)
        for k, ((M, st, _), (xf, sf)) in enumerate(zip(jobs, fused)):
            ss = S.OptimizationStatistics()
            xs = M.solve(st, stats=ss)
            assert np.array_equal(xf.view(np.uint32), xs.view(np.uint32)), k
            assert np.array_equal(sf.values.view(np.uint32), ss.values.view(np.uint32)), k
            assert np.array_equal(sf.cardinalities, ss.cardinalities) and sf.it == ss.it == C4_ITERS, k
        print("  fused == solo bitwise for all 64 jobs", flush=True)
        # the smallest job of each formulation against the FP64 oracle, on its first 32 SPs (with
        # tol = 0 every SP runs its own trajectory, independent of the rest of the batch)
        by_f = {}
        for k, sp in enumerate(specs):
            if sp["f"] not in by_f or sp["m"] * sp["n"] < specs[by_f[sp["f"]]]["m"] * specs[by_f[sp["f"]]]["n"]:
                by_f[sp["f"]] = k
        threads = min(os.cpu_count() or 1, 16)
        for f, k in sorted(by_f.items()):
            sp, (M, st, _), (_, sf) = specs[k], jobs[k], fused[k]
            A = M.download()
            nsp = min(C4_ORACLE_SPS, sp["B"])
            X0 = _gaussian_starts(sp["n"], nsp, sp["seed"])
            g = st.penaltyParameter
            o = O.solve(A.astype(np.float64), f, s=sp["s"], gamma=g, tol=0.0, max_iter=C4_ITERS, total_sp=nsp,
                        X0=X0.astype(np.float64), lambda_rule=O.LAMBDA_GPU, threads=threads, gemm_backend=1)
            margin, _ = trajectory_tie_margin(A, X0, f, sp["s"], g, C4_ITERS, gemm_backend=1, threads=threads)
            tied = margin < TIE_TOL[S.PREC_FP32]
            vals = sf.values[:nsp]
            with np.errstate(invalid="ignore", divide="ignore"):
                rel = np.abs(vals - o["sp_val"]) / np.abs(o["sp_val"])
            ok = ~tied & np.isfinite(o["sp_val"])
            print(f"  job {k} f={f} {sp['m']}x{sp['n']} s={sp['s']} B={sp['B']}: tied {tied.sum()}/{nsp}, "
                  f"value rel max {np.nanmax(rel[ok]) if ok.any() else 0:.2e}, nnz equal "
                  f"{np.mean(sf.cardinalities[:nsp] == o['sp_nnz']):.3f}", flush=True)
            assert np.all(rel[ok] <= 1e-5), (k, np.sort(rel[ok])[-4:])
            assert np.array_equal(sf.cardinalities[:nsp][ok], o["sp_nnz"][ok]), k
    finally:
        for M in mats:
            M.release()


# ---------------------------------------------------------------- c5: serving burst
C5_BURST = 24
C5_WORKERS = 4


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
def test_c5_serving_burst_alone_equals_mix(prec):
    """The first requests of the c5 stream (200 jobs/s) arrive together: the server runs them as
    fused groups under the 200 GB budget, and every result is bitwise the result of the same job
    submitted alone and of its solo solve.  The two smallest jobs' answers are then
    checked against the FP64 oracle."""
    pool_size = 24
    specs = bench.uneven_specs(pool_size, seed=bench.SERVE_POOL_SEED)
    _, draws = bench.serving_draws(200.0, 1.0, pool_size)
    draws = draws[:C5_BURST]
    store = S.STORE_FP32 if prec == S.PREC_FP32 else S.STORE_BF16X3
    print(f"case c5 prec={prec}: burst of {len(draws)} requests", flush=True)
    with S.Context(0, hbm_budget_bytes=int(200e9)) as ctx:
        pool = {}
        for pi, _, _, _ in draws:
            if pi not in pool:
                sp = specs[pi]
                pool[pi] = ctx.generate(sp["m"], sp["n"], sp["seed"], theta=bench.THETA, spike_nnz=bench.SPARSITY,
                                        storage=store, key=sp["seed"])
        sts = [bench.serving_settings(S, prec, specs[pi], f, sv, B, k) for k, (pi, f, sv, B) in enumerate(draws)]
        ctx.start_server(C5_WORKERS)
        handles = [pool[pi].submit(st, per_sp=False) for (pi, _, _, _), st in zip(draws, sts)]
        mix = []
        for h in handles:
            stats = S.OptimizationStatistics()
            x, times = h.wait(stats=stats, per_sp=False)
            mix.append((x, stats))
        info = ctx.server_info()
        print(f"  mix done: {info}", flush=True)
        assert info["max_members"] >= 2  # the burst ran as fused groups
        # alone: each request submitted to the idle server by itself
        for k, ((pi, f, sv, B), st, (xm, sm)) in enumerate(zip(draws, sts, mix)):
            sa = S.OptimizationStatistics()
            xa, _ = pool[pi].submit(st, per_sp=False).wait(stats=sa, per_sp=False)
            assert np.array_equal(xa.view(np.uint32), xm.view(np.uint32)), k
            assert sa.it == sm.it and sa.fval == sm.fval, k
        ctx.stop_server()
        print("  alone == mix bitwise", flush=True)
        # fused jobs run their solo tiles, plan and summation order (shapes whose solo plan takes the
        # large 16-bit tiles are solved alone): the server's result is the solo solve's
        for k, ((pi, f, sv, B), st, (xm, sm)) in enumerate(zip(draws, sts, mix)):
            sa = S.OptimizationStatistics()
            xa = pool[pi].solve(st, stats=sa, per_sp=False)
            assert np.array_equal(xa.view(np.uint32), xm.view(np.uint32)), k
        print("  server == solo bitwise", flush=True)
        cost = [specs[pi]["m"] * specs[pi]["n"] * B * sm.it for (pi, _, _, B), (_, sm) in zip(draws, mix)]
        threads = min(os.cpu_count() or 1, 16)
        for k in np.argsort(cost)[:2]:
            pi, f, sv, B = draws[k]
            st, (xm, sm) = sts[k], mix[k]
            A = pool[pi].download()
            X0 = _gaussian_starts(A.shape[1], B, st.seed)
            o = O.solve(A.astype(np.float64), f, s=st.constraintParameter, gamma=st.penaltyParameter, tol=1e-5,
                        max_iter=200, total_sp=B, X0=X0.astype(np.float64), lambda_rule=O.LAMBDA_GPU, threads=threads,
                        gemm_backend=1)
            ev_g = explained_variance(A, xm[:, None], f)[0]
            ev_o = explained_variance(A, o["x"][:, None], f)[0]
            print(f"  request {k} f={f} {A.shape} s={st.constraintParameter} B={B}: iterations {sm.it}/{o['it']}, "
                  f"fval rel {abs(sm.fval - o['fval']) / abs(o['fval']):.2e}, EV rel {abs(ev_g - ev_o) / ev_o:.2e}",
                  flush=True)
            assert abs(int(sm.it) - int(o["it"])) <= 2
            assert abs(sm.fval - o["fval"]) <= 1e-5 * abs(o["fval"])
            assert abs(ev_g - ev_o) <= 1e-5 * ev_o
        for M in pool.values():
            M.release()
