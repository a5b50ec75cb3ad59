"""The reference's own test inputs, solved on the MI355X and checked against the FP64 oracle.

* cluster.dat.all (800 x 300): the reference's gpu_unit_test (c++/src/test/gpu_unit_test.cu:115-133)
  runs all 8 formulations on it with 1024 starting points in one batch, s = 20, tol 1e-4, at
  most 100 iterations and the default penalty 0, and compares the GPU solver with the
  multicore one.  The reference prints both objectives and asserts nothing; here every
  starting point is compared.  Starting points are the reference CPU method's rand_r ones
  drawn in double (test_solver<double>), rounded to float.
* small.csv (2 x 3): the reference's test_multicore target (c++/multicore.mk:18-21) runs
  multicore_console three times: -f 1 -s 3; -l 1000 -r 64 -f 1 -s 2; and the same with -u 1
  (OTF).  Here spca_console runs the same options; the reference adds -d double, which this
  solver rejects (it has no double products), so the FP32 solve of the same CSV is checked
  against the FP64 oracle with the same (float rand_r) starting points.

Fixtures: tests/golden/cluster.dat.all.gz and small.csv, copied by
tests/golden/make_reference_inputs.py with their SHA-256 in reference_inputs.json.
"""
import gzip
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import spca_mi355x as S
from spca_cases import explained_variance, soft_condition, trajectory_tie_margin

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONSOLE = os.path.join(os.path.dirname(S.LIB_PATH), "spca_console")


def _read_csv_bytes(data):
    """The reference CSV reader's shape rule (c++/src/utils/file_reader.h:31-86): m = lines,
    n = cells of the first line; values row by row."""
    lines = data.decode().splitlines()
    n = len(lines[0].split(","))
    A = np.array([[float(v) for v in l.split(",")] for l in lines], dtype=np.float64)
    assert A.shape == (len(lines), n)
    return A


def _fixture(name):
    raw = open(os.path.join(GOLD, name), "rb").read()
    return gzip.decompress(raw) if name.endswith(".gz") else raw


def test_fixtures_are_the_reference_files():
    meta = json.load(open(os.path.join(GOLD, "reference_inputs.json")))
    for name, m in meta.items():
        data = _fixture(name)
        assert len(data) == m["bytes"] and hashlib.sha256(data).hexdigest() == m["sha256"], name
    assert _read_csv_bytes(_fixture("cluster.dat.all.gz")).shape == (800, 300)
    assert _read_csv_bytes(_fixture("small.csv")).shape == (2, 3)


# ---------------------------------------------------------------- gpu_unit_test on cluster.dat.all
UT_SP, UT_S, UT_TOL, UT_IT = 1024, 20, 1e-4, 100
EV_TOL = {S.PREC_FP32: 1e-6, S.PREC_BF16X3: 1e-5}
VAL_TOL = {S.PREC_FP32: 1e-5, S.PREC_BF16X3: 1e-4}
# relative rounding of v = A^T z in each product form: FP32 accumulation, bf16x3 split
V_EPS = {S.PREC_FP32: 1e-7, S.PREC_BF16X3: 1e-6}
# Tie tolerance: a step whose threshold and nearest competitor are within the product's
# rounding.  The L1-variance formulations' A^T sgn(z) is the exact fixed-point product (error
# ~2^-25 relative), so their ties are those within 1e-6; the others use the product's own.
TIE_TOL = {S.PREC_FP32: 1e-5, S.PREC_BF16X3: 1e-4}
TIE_TOL_EXACT = 1e-6
ZTOL = {S.PREC_FP32: 1e-6, S.PREC_BF16X3: 3e-5}


@pytest.fixture(scope="module")
def cluster(ctx):
    A = _read_csv_bytes(_fixture("cluster.dat.all.gz")).astype(np.float32)
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    yield A, M
    M.release()


def _reference_starts(n, s, constrained, count, dtype):
    return np.stack([O.starting_point(n, s, constrained, k, dtype=dtype) for k in range(count)], axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [S.PREC_FP32, S.PREC_BF16X3])
@pytest.mark.parametrize("f", range(8))
def test_gpu_unit_test_cluster(cluster, f, prec):
    A, M = cluster
    n = A.shape[1]
    X0 = _reference_starts(n, UT_S, f < 4, UT_SP, np.float64).astype(np.float32)
    print(f"case cluster f={f} prec={prec}", flush=True)
    st = S.OptimizationSettings(formulation=f, constraintParameter=UT_S, penaltyParameter=0.0, tolerance=UT_TOL,
                                maximumIterations=UT_IT, totalStartingPoints=UT_SP, batchSize=UT_SP, precision=prec)
    stats = S.OptimizationStatistics()
    x, Xg = M.solve(st, X0=X0, stats=stats, want_X=True)
    threads = min(os.cpu_count() or 1, 16)
    o = O.solve(A.astype(np.float64), f, s=UT_S, gamma=0.0, tol=UT_TOL, max_iter=UT_IT, total_sp=UT_SP,
                X0=X0.astype(np.float64), want_X=True, lambda_rule=O.LAMBDA_GPU, threads=threads, gemm_backend=1)
    margin, _ = trajectory_tie_margin(A, X0, f, UT_S, 0.0, int(o["it"]), ztol=ZTOL[prec], gemm_backend=1,
                                      threads=threads)
    tied = margin < (TIE_TOL_EXACT if f in (1, 3, 5, 7) else TIE_TOL[prec])
    alive = np.all(np.isfinite(o["X"]), axis=0) & np.any(o["X"] != 0, axis=0)
    ok = alive & ~tied
    # explained-variance bar per SP: 1e-6 (FP32) / 1e-5 (bf16x3), or the product's rounding of v
    # amplified by the soft threshold's condition where that is larger (this data's columns
    # are all alike, so lambda sits just below the kept |v_i|: kappa ~40 for the L1 constraint)
    kappa = soft_condition(A, o["X"], f, UT_S, 0.0)
    ev_bar = np.maximum(EV_TOL[prec], kappa * V_EPS[prec])
    supp = np.all((Xg != 0) == (o["X"] != 0), axis=0)
    ev_o = np.full(UT_SP, np.nan)
    ev_o[alive] = explained_variance(A, o["X"][:, alive], f)
    with np.errstate(invalid="ignore", divide="ignore"):
        ev_rel = np.abs(stats.sp_ev_fp64 - ev_o) / ev_o
        vrel = np.abs(stats.values - o["sp_val"]) / np.abs(o["sp_val"])
    mx = lambda a: float(np.nanmax(a)) if a.size else 0.0
    print(f"  iterations {stats.it}/{o['it']} fval {stats.fval:.9g}/{o['fval']:.9g} alive {alive.sum()} tied "
          f"{tied.sum()} supports {supp[alive].mean():.4f} EV rel max {mx(ev_rel[ok]):.2e} (bar max "
          f"{mx(ev_bar[ok]):.1e}, kappa max {mx(kappa):.1f}) value rel max {mx(vrel[ok]):.2e}", flush=True)
    assert abs(int(stats.it) - int(o["it"])) <= 1
    assert np.all(supp[ok]), np.nonzero(ok & ~supp)[0]
    assert np.all(ev_rel[ok] <= ev_bar[ok]), np.sort(ev_rel[ok] / ev_bar[ok])[-4:]
    assert np.all(vrel[ok] <= np.maximum(VAL_TOL[prec], kappa[ok] * V_EPS[prec])), np.sort(vrel[ok])[-4:]
    assert abs(stats.fval - o["fval"]) <= VAL_TOL[prec] * abs(o["fval"])


# ---------------------------------------------------------------- test_multicore on small.csv
@pytest.mark.gpu
@pytest.mark.parametrize("opts", [["-f", 1, "-s", 3], ["-l", 1000, "-r", 64, "-f", 1, "-s", 2],
                                  ["-l", 1000, "-r", 64, "-u", 1, "-f", 1, "-s", 2]],
                         ids=["small", "small_2", "small_3"])
def test_multicore_small_csv(tmp_path, opts):
    csv = tmp_path / "small.csv"
    csv.write_bytes(_fixture("small.csv"))
    prefix = str(tmp_path / "res")
    print(f"case small.csv {opts}", flush=True)
    p = subprocess.run([CONSOLE, "-i", str(csv), "-o", prefix, "-v", "true"] + [str(a) for a in opts],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, (p.stdout, p.stderr)
    A = _read_csv_bytes(_fixture("small.csv")).astype(np.float32)
    m, n = A.shape
    kw = dict(zip(opts[::2], opts[1::2]))
    total, batch = int(kw.get("-l", 64)), int(kw.get("-r", 64))
    s = min(int(kw["-s"]), n)
    otf = bool(int(kw.get("-u", 0)))
    total = max(total, batch)
    total = -(-total // batch) * batch
    X0 = _reference_starts(n, s, True, total, np.float32)  # multicore_console draws in float without -d
    o = O.solve(A.astype(np.float64), 1, s=s, tol=0.01, max_iter=20, total_sp=total, batch=batch, otf=otf,
                X0=X0.astype(np.float64), lambda_rule=O.LAMBDA_GPU)
    x = np.zeros(n)
    for line in open(prefix + "_x"):
        i, v = line.split(",")
        x[int(i)] = float(v)
    lines = dict(l.split(": ", 1) for l in open(prefix + "_optimizationStatistics").read().splitlines() if ": " in l)
    fval = float(lines["Objective value"])
    print(f"  x {x} oracle {o['x']}; fval {fval:.9g} oracle {o['fval']:.9g}; it {lines['Elapsed it (total)']} "
          f"oracle {o['it']}", flush=True)
    assert int(lines["Formulation"]) == 1 and int(lines["Starting points"]) == total
    assert abs(fval - o["fval"]) <= 1e-6 * abs(o["fval"])
    np.testing.assert_allclose(x, o["x"], atol=1e-6)
    assert np.count_nonzero(x) == np.count_nonzero(o["x"]) <= s
