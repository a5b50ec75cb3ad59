"""spca_console, the command-line drop-in for the reference's gpu_console / multicore_console
(c++/src/frontends/gpu_console.cu, c++/src/utils/option_console_parser.h,
c++/src/utils/file_reader.h).

CPU tests run the binary in its echo mode (-e: parse, read the CSV, print the settings, no
GPU call): the usage exit, the -f 1..8 numbering, the reference defaults, the batch rule, the
clamp and the CSV shape.  The GPU test runs a solve and checks the two output files against
the same solve through the C ABI.
"""
import os
import subprocess

import numpy as np
import pytest

import spca_mi355x as S

CONSOLE = os.path.join(os.path.dirname(S.LIB_PATH), "spca_console")

# reference -f numbering -> formulation enum (option_console_parser.h:103-137)
FLAG_TO_ENUM = {1: 1, 2: 0, 3: 3, 4: 2, 5: 5, 6: 4, 7: 7, 8: 6}


def _run(args, check=True):
    p = subprocess.run([CONSOLE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    if check:
        assert p.returncode == 0, (p.returncode, p.stdout, p.stderr)
    return p


def _echo(args):
    out = _run(list(args) + ["-e"]).stdout.split()
    return dict(zip(out[0::2], out[1::2]))


def _write_csv(path, A, trailing_newline=True):
    rows = [",".join(repr(float(v)) for v in row) for row in A.astype(np.float32)]
    with open(path, "w") as f:
        f.write("\n".join(rows) + ("\n" if trailing_newline else ""))


@pytest.fixture(scope="module")
def small_csv(tmp_path_factory):
    p = tmp_path_factory.mktemp("console") / "data.csv"
    A = np.arange(15, dtype=np.float32).reshape(3, 5) / 7
    _write_csv(p, A)
    return str(p)


def test_usage_when_required_options_missing(small_csv):
    p = _run(["-i", small_csv, "-o", "x"], check=False)  # no -f
    assert p.returncode == 1 and "Usage" in p.stdout
    p = _run(["-i", small_csv, "-f", "9", "-o", "x"], check=False)  # unknown formulation number
    assert p.returncode == 1


@pytest.mark.parametrize("flag", range(1, 9))
def test_formulation_numbering(small_csv, flag):
    e = _echo(["-i", small_csv, "-o", "x", "-f", flag])
    assert int(e["formulation"]) == FLAG_TO_ENUM[flag]


def test_reference_defaults_and_shape(small_csv):
    e = _echo(["-i", small_csv, "-o", "x", "-f", 2])
    assert (int(e["m"]), int(e["n"])) == (3, 5)
    # s = 10 clamped to n = 5 (chceckInputAndModifyIt); gamma 0, 20 iterations, tol 0.01,
    # 64 starting points in one batch of 64, no OTF, single precision
    assert int(e["s"]) == 5 and float(e["gamma"]) == 0.0 and int(e["max_iter"]) == 20
    assert float(e["tol"]) == 0.01 and int(e["total_sp"]) == 64 and int(e["batch"]) == 64
    assert int(e["otf"]) == 0 and int(e["double"]) == 0 and int(e["precision"]) == 0


def test_options_batch_rule_and_rounding(small_csv):
    e = _echo(["-i", small_csv, "-o", "x", "-f", 6, "-g", 0.25, "-m", 7, "-t", 1e-4, "-l", 10, "-r", 16, "-u", 1,
               "-p", "bf16x3", "-s", 3])
    # a batch larger than the total raises the total to the batch (option_console_parser.h:148-150)
    assert int(e["total_sp"]) == 16 and int(e["batch"]) == 16
    assert float(e["gamma"]) == 0.25 and int(e["max_iter"]) == 7 and float(e["tol"]) == 1e-4
    assert int(e["otf"]) == 1 and int(e["double"]) == 0 and int(e["precision"]) == 1 and int(e["s"]) == 3
    # the total is rounded up to whole batches
    e = _echo(["-i", small_csv, "-o", "x", "-f", 2, "-l", 70, "-r", 32])
    assert int(e["total_sp"]) == 96 and int(e["batch"]) == 32


@pytest.mark.parametrize("extra", [[], ["-e"]])
def test_double_precision_is_rejected(small_csv, tmp_path, extra):
    """-d asks the reference for a double solve (gpu_console.cu:97-101); this solver has no
    double products, so the console exits with status 2 and says so, before any GPU call and
    without writing result files."""
    prefix = str(tmp_path / "res")
    p = _run(["-i", small_csv, "-o", prefix, "-f", 2, "-d", "double"] + extra, check=False)
    assert p.returncode == 2 and "double precision" in p.stderr
    assert not os.path.exists(prefix + "_x") and not os.path.exists(prefix + "_optimizationStatistics")


def test_csv_without_trailing_newline(tmp_path):
    p = tmp_path / "d.csv"
    _write_csv(p, np.ones((4, 7), dtype=np.float32), trailing_newline=False)
    e = _echo(["-i", p, "-o", "x", "-f", 1])
    assert (int(e["m"]), int(e["n"])) == (4, 7)


@pytest.mark.gpu
@pytest.mark.parametrize("flag,prec", [(2, "fp32"), (3, "fp32"), (6, "bf16x3")])
def test_console_solve_matches_c_abi(ctx, tmp_path, flag, prec):
    """The console's <prefix>_x and <prefix>_optimizationStatistics carry the same solve as
    spca_solve with the same settings (reference starting points, FP32 data from the CSV)."""
    rng = np.random.default_rng(7)
    m, n = 96, 300
    A = rng.standard_normal((m, n)).astype(np.float32)
    A[:, :6] += 3.0 * rng.standard_normal((m, 1)).astype(np.float32)
    csv = tmp_path / "a.csv"
    _write_csv(csv, A)
    prefix = str(tmp_path / "res")
    gamma = 0.5 if flag == 6 else 0.0
    _run(["-i", csv, "-o", prefix, "-f", flag, "-s", 8, "-g", gamma, "-m", 40, "-t", 1e-6, "-l", 64, "-r", 32,
          "-p", prec])
    f = FLAG_TO_ENUM[flag]
    st = S.OptimizationSettings(formulation=f, constraintParameter=8, penaltyParameter=gamma, tolerance=1e-6,
                                maximumIterations=40, totalStartingPoints=64, batchSize=32,
                                precision={"fp32": S.PREC_FP32, "bf16x3": S.PREC_BF16X3}[prec],
                                start=S.START_REFERENCE)
    stats = S.OptimizationStatistics()
    M = ctx.upload(A, storage=S.STORE_FP32 | S.STORE_BF16X3)
    try:
        x = M.solve(st, stats=stats)
    finally:
        M.release()
    xs = np.zeros(n, dtype=np.float64)
    with open(prefix + "_x") as fx:
        for line in fx:
            i, v = line.split(",")
            xs[int(i)] = float(v)
    print(f"case flag={flag} prec={prec}: nnz {np.count_nonzero(xs)}, fval {stats.fval}", flush=True)
    np.testing.assert_array_equal(xs.astype(np.float32), x)
    lines = dict(l.split(": ", 1) for l in open(prefix + "_optimizationStatistics").read().splitlines() if ": " in l)
    assert int(lines["Formulation"]) == f
    assert float(lines["Objective value"]) == pytest.approx(stats.fval, rel=1e-12)
    assert int(lines["Elapsed it (total)"]) == stats.it
