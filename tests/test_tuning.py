"""Product-plan autotuner (spca_set_tuning / spca_get_plan).  Needs an MI355X.

With tuning on, the first solve of a shape times every valid plan (K split of A*X, tile
kernel) and keeps the fastest; the plan is cached in memory and in a JSON file that a new
context loads.  A plan fixes the summation order, so a solve is bitwise reproducible for a
given plan, and FP32 solo and fused solves of a shape share the plan.
"""
import json

import numpy as np
import pytest

import spca_mi355x as S
from spca_cases import default_gamma, spiked, unit_gaussian_starts

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def test_tuned_plan_cached_shared_and_reloaded(tmp_path):
    path = str(tmp_path / "plans.json")
    A, _ = spiked(300, 2500, seed=11)
    X0 = unit_gaussian_starts(2500, 96, seed=12)
    st = S.OptimizationSettings(formulation=2, constraintParameter=20, tolerance=1e-6, maximumIterations=30,
                                totalStartingPoints=96, batchSize=96, precision=S.PREC_FP32)
    with S.Context(0) as c1:
        c1.set_tuning(1, path)
        assert c1.plan(S.PREC_FP32, 300, 2500, 96)[2] == 0.0  # heuristic until measured
        M = c1.upload(A)
        stats = S.OptimizationStatistics()
        x1 = M.solve(st, X0=X0, stats=stats)
        splits, big, ms = c1.plan(S.PREC_FP32, 300, 2500, 96)
        assert ms > 0 and big == 0 and splits >= 1
        assert c1.plan(S.PREC_FP32, 300, 2500, 96, fused=True) == (splits, big, ms)  # one plan for both paths
        (xf, sf), = c1.solve_fused([(M, st, X0)])
        assert np.array_equal(_bits(xf), _bits(x1)) and sf.fval == stats.fval
        M.release()
    plans = json.load(open(path))["plans"]
    assert any(p["np"] == 2560 and p["mp"] == 512 and p["bp"] == 128 and p["splits"] == splits for p in plans)
    with S.Context(0) as c2:
        c2.set_tuning(0, path)  # fixed plans, loaded from the file
        assert c2.plan(S.PREC_FP32, 300, 2500, 96) == (splits, big, pytest.approx(ms, rel=1e-4))
        M = c2.upload(A)
        x2 = M.solve(st, X0=X0)
        assert np.array_equal(_bits(x2), _bits(x1))
        M.release()


def test_tuning_16bit_tile_choice(tmp_path):
    """bf16x3 with a 256-multiple batch: both tile kernels are candidates; the tuned solve
    agrees with the FP32 solve to bf16x3 accuracy."""
    A, _ = spiked(300, 2500, seed=13)
    X0 = unit_gaussian_starts(2500, 256, seed=14)
    with S.Context(0) as c:
        c.set_tuning(1, str(tmp_path / "p.json"))
        M = c.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
        out = {}
        for prec in (S.PREC_FP32, S.PREC_BF16X3):
            st = S.OptimizationSettings(formulation=0, constraintParameter=20, tolerance=1e-6, maximumIterations=40,
                                        totalStartingPoints=256, batchSize=256, precision=prec)
            stats = S.OptimizationStatistics()
            M.solve(st, X0=X0, stats=stats)
            out[prec] = stats
        splits, big, ms = c.plan(S.PREC_BF16X3, 300, 2500, 256)
        assert ms > 0 and big in (0, 1)
        a, b = out[S.PREC_FP32], out[S.PREC_BF16X3]
        assert abs(b.fval - a.fval) <= 1e-4 * a.fval
        M.release()
