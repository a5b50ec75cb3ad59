"""bf16x3 against FP64 at the headline shape.  Needs an MI355X; slow (the FP64 oracle solves
on the host, about half a minute).

4096 x 65536 spiked model (bench.make_problem), s = 256, tol = 1e-6, at most 200
iterations, the two formulations of the precision workload (L2 variance / L0 constraint and
L1 variance / L1 constraint), on the first 128 of the bench's unit Gaussian starting points.
Per starting point: the explained variance of the GPU's final iterate (FP64, on the device)
against that of the FP64 oracle's final iterate, and whether the supports are identical.

The solve is nonconvex: a starting point whose trajectory passes near a tie (two candidates
for the s-th place within the products' rounding, or for L1 variance an entry of z = A x
within rounding of zero) can settle on another fixed point.  FP32 does so for 1 of 128
points here, bf16x3 (~2^-17 relative products) for 2-4 of 128.  Bars: FP32 EV within 1e-6 and
identical supports on >= 99% of points; bf16x3 EV within 1e-5 on >= 97% and identical
supports on >= 96% (measured 99.2% / 98.4% for L2 variance with the L0 constraint, 97.7% /
96.9% for L1 variance with the L1 constraint).
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
    res = {}
    for prec, ev_tol in ((S.PREC_BF16X3, 1e-5), (S.PREC_FP32, 1e-6)):
        st = S.OptimizationSettings(formulation=f, constraintParameter=256, tolerance=1e-6, maximumIterations=200,
                                    totalStartingPoints=NSP, batchSize=NSP, precision=prec)
        stats = S.OptimizationStatistics()
        _, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
        rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
        same = np.all((Xg != 0) == (Xo != 0), axis=0)
        res[prec] = (rel, same, ev_tol)
        print(f"f={f} prec={prec}: iterations gpu {stats.it} oracle {o['it']}; EV rel max {rel.max():.2e} "
              f"median {np.median(rel):.2e}; within tol {np.mean(rel <= ev_tol):.3f}; identical supports "
              f"{same.mean():.3f}; off-support SPs {np.nonzero(~same)[0].tolist()}", flush=True)
    bars = {S.PREC_FP32: (0.99, 0.99), S.PREC_BF16X3: (0.97, 0.96)}
    for prec, (rel, same, ev_tol) in res.items():
        ev_bar, supp_bar = bars[prec]
        assert np.mean(rel <= ev_tol) >= ev_bar, (prec, np.sort(rel)[-4:])
        assert same.mean() >= supp_bar, (prec, np.nonzero(~same)[0])
