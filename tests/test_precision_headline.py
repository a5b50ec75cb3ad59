"""bf16x3 against FP64 at the headline shape.  Needs an MI355X; slow (the FP64 oracle solves
on the host, about half a minute).

4096 x 65536 spiked model (bench.make_problem), s = 256, tol = 1e-6, at most 200
iterations, the two formulations of the precision workload (L2 variance / L0 constraint and
L1 variance / L1 constraint), on the first 128 of the bench's unit Gaussian starting points.
Per starting point: the explained variance of the GPU's final iterate (FP64, on the device)
against that of the FP64 oracle's final iterate, and whether the supports are identical.
The bar: EV within 1e-5 relative and identical supports on >= 99% of starting points for
bf16x3, and the same for FP32 with EV within 1e-6.
"""
import os

import numpy as np
import pytest

import oracle as O
import spca_mi355x as S

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

NSP = 128


@pytest.fixture(scope="module")
def headline(ctx):
    import bench
    A, X0, _ = bench.make_problem()
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    yield A, np.asfortranarray(X0[:, :NSP]), M
    M.release()


@pytest.mark.parametrize("f", [0, 3])
def test_headline_precision_against_fp64(headline, f):
    A, X0, M = headline
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    A64 = A.astype(np.float64)
    o = O.solve(A64, f, s=256, tol=1e-6, max_iter=200, total_sp=NSP, batch=NSP, X0=X0.astype(np.float64),
                threads=threads, gemm_backend=1, want_X=True, lambda_rule=O.LAMBDA_GPU)
    Xo = o["X"]
    Zo = A64 @ (Xo / np.linalg.norm(Xo, axis=0))
    ev_o = np.sum(Zo * Zo, axis=0) if f == 0 else np.sum(np.abs(Zo), axis=0)
    for prec, ev_tol in ((S.PREC_BF16X3, 1e-5), (S.PREC_FP32, 1e-6)):
        st = S.OptimizationSettings(formulation=f, constraintParameter=256, tolerance=1e-6, maximumIterations=200,
                                    totalStartingPoints=NSP, batchSize=NSP, precision=prec)
        stats = S.OptimizationStatistics()
        _, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
        rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
        same = np.all((Xg != 0) == (Xo != 0), axis=0)
        print(f"f={f} prec={prec}: iterations gpu {stats.it} oracle {o['it']}; EV rel max {rel.max():.2e} "
              f"median {np.median(rel):.2e}; within tol {np.mean(rel <= ev_tol):.3f}; identical supports "
              f"{same.mean():.3f}", flush=True)
        assert np.mean(rel <= ev_tol) >= 0.99, np.sort(rel)[-4:]
        assert same.mean() >= 0.99, np.nonzero(~same)[0]
