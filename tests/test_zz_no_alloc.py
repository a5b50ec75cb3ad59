"""spca_solve_noalloc (the allocation-free solve) against spca_solve, through the C ABI.  Needs an
MI355X.

The allocation-free solve runs the same kernels with the same product plan on buffers carved
from the library's fixed device array, so its results are bitwise those of spca_solve (the FP64
explained variance may sum its partials in another grouping: 1e-12).  Two calls in a row, the
second on a smaller matrix and batch, check that nothing of the first call's contents leaks into
the second; the context's device memory is unchanged by the calls; shapes beyond the array and
modes that would allocate are rejected with SPCA_EINVAL.
"""
import numpy as np
import pytest

import spca_mi355x as S
from spca_cases import default_gamma, spiked, unit_gaussian_starts

pytestmark = pytest.mark.gpu

STORE_ALL = S.STORE_FP32 | S.STORE_BF16X3 | S.STORE_FP16 | S.STORE_FIX8


def _bits(a):
    return np.asarray(a, dtype=np.float32).view(np.uint32)


def _settings(f, prec, B, m, it=25, tol=0.0):
    return S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=default_gamma(f, m),
                                  tolerance=tol, maximumIterations=it, totalStartingPoints=B, batchSize=B,
                                  precision=prec)


def _same(a, b):
    """(x, X, stats) of spca_solve_noalloc against those of spca_solve."""
    (xa, Xa, sa), (xb, Xb, sb) = a, b
    np.testing.assert_array_equal(_bits(xb), _bits(xa))
    np.testing.assert_array_equal(_bits(Xb), _bits(Xa))
    np.testing.assert_array_equal(_bits(sb.values), _bits(sa.values))
    np.testing.assert_array_equal(sb.iters, sa.iters)
    np.testing.assert_array_equal(sb.cardinalities, sa.cardinalities)
    assert sb.it == sa.it and sb.fval == sa.fval and sb.best_sp == sa.best_sp
    np.testing.assert_allclose(sb.sp_ev_fp64, sa.sp_ev_fp64, rtol=1e-12)
    assert sb.ev_fp64 == pytest.approx(sa.ev_fp64, rel=1e-12)


def _run(M, st, X0, noalloc):
    stats = S.OptimizationStatistics()
    fn = M.solve_noalloc if noalloc else M.solve
    x, X = fn(st, X0=X0, stats=stats, want_X=True)
    return x, X, stats


@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3, S.PREC_FP16])
def test_noalloc_equals_solve_twice_second_smaller(ctx, prec):
    """Two allocation-free solves in a row, the second on a smaller matrix and batch (and with a
    tolerance, so it stops early), each bitwise equal to spca_solve of the same request; the
    context's device memory (live and cached blocks) does not change across the two calls."""
    A1, _ = spiked(300, 2000, seed=21)
    A2, _ = spiked(150, 700, seed=22)
    M1 = ctx.upload(A1, storage=STORE_ALL)
    M2 = ctx.upload(A2, storage=STORE_ALL)
    try:
        for f in range(8):
            print(f"case noalloc prec={prec} f={f}", flush=True)
            X1 = unit_gaussian_starts(2000, 200, seed=30 + f)
            X2 = unit_gaussian_starts(700, 40, seed=40 + f)
            st1, st2 = _settings(f, prec, 200, 300), _settings(f, prec, 40, 150, tol=1e-6, it=60)
            ref1, ref2 = _run(M1, st1, X1, False), _run(M2, st2, X2, False)
            mem0 = ctx.memory()[:2]
            got1 = _run(M1, st1, X1, True)
            got2 = _run(M2, st2, X2, True)
            assert ctx.memory()[:2] == mem0  # neither call allocated (nor freed) device memory
            _same(ref1, got1)
            _same(ref2, got2)
    finally:
        M1.release()
        M2.release()


def test_noalloc_rejects_what_would_allocate(ctx):
    A, _ = spiked(200, 1000, seed=23)
    X0 = unit_gaussian_starts(1000, 64)
    M = ctx.upload(A, storage=S.STORE_FP32)
    try:
        with pytest.raises(S.SpcaError) as e:  # OTF needs the pool of starting points
            M.solve_noalloc(S.OptimizationSettings(formulation=0, totalStartingPoints=64, batchSize=32, useOTF=True),
                            X0=X0)
        assert e.value.code == S.SPCA_EINVAL
        with pytest.raises(S.SpcaError) as e:  # bf16x3 storage is not there and is not added
            M.solve_noalloc(_settings(0, S.PREC_BF16X3, 64, 200), X0=X0)
        assert e.value.code == S.SPCA_EINVAL and "storage" in str(e.value)
        with pytest.raises(S.SpcaError) as e:  # the exact sign product's fixed-point form is missing
            M.solve_noalloc(_settings(1, S.PREC_FP32, 64, 200), X0=X0)
        assert e.value.code == S.SPCA_EINVAL
        with pytest.raises(S.SpcaError) as e:
            M.solve_noalloc(S.OptimizationSettings(formulation=0, totalStartingPoints=64, batchSize=64,
                                                   resolve_iters=10), X0=X0)
        assert e.value.code == S.SPCA_EINVAL
    finally:
        M.release()
    # a shape beyond the fixed array: 4096 x 65536 with 1024 SPs
    big = ctx.generate(4096, 65536, 5)
    try:
        st = S.OptimizationSettings(formulation=0, constraintParameter=16, totalStartingPoints=1024, batchSize=1024,
                                    start=S.START_GAUSSIAN, seed=1, maximumIterations=2)
        with pytest.raises(S.SpcaError) as e:
            big.solve_noalloc(st, per_sp=False)
        assert e.value.code == S.SPCA_EINVAL and "bytes of scratch" in str(e.value)
    finally:
        big.release()
