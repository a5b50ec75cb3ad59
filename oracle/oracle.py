"""ctypes binding of the CPU oracle (oracle/liboracle_spca.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.  The oracle
restates the reference's CPU method; see spca_oracle.h for the file:line map.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_spca.so")

FORMULATIONS = [
    "L0_constrained_L2_PCA", "L0_constrained_L1_PCA", "L1_constrained_L2_PCA", "L1_constrained_L1_PCA",
    "L0_penalized_L2_PCA", "L0_penalized_L1_PCA", "L1_penalized_L2_PCA", "L1_penalized_L1_PCA",
]
HARD_SORT, HARD_SET = 0, 1
TERM_BATCH_MAX, TERM_PER_SP = 0, 1
LAMBDA_CPU, LAMBDA_GPU = 0, 1


class Settings(ctypes.Structure):
    _fields_ = [
        ("formulation", ctypes.c_int), ("constraint", ctypes.c_uint), ("penalty", ctypes.c_double),
        ("tolerance", ctypes.c_double), ("max_iter", ctypes.c_int), ("total_sp", ctypes.c_int),
        ("batch", ctypes.c_int), ("use_otf", ctypes.c_int), ("termination", ctypes.c_int),
        ("hard_rule", ctypes.c_int), ("threads", ctypes.c_int), ("gemm_backend", ctypes.c_int),
        ("lambda_rule", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("it", ctypes.c_uint), ("fval", ctypes.c_double), ("best_sp", ctypes.c_int),
        ("loop_s", ctypes.c_double), ("wall_s", ctypes.c_double), ("threads_used", ctypes.c_int),
        ("total_sp", ctypes.c_int), ("constraint", ctypes.c_uint),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        for sfx in ("f32", "f64"):
            real = ctypes.c_float if sfx == "f32" else ctypes.c_double
            getattr(L, f"oracle_solve_{sfx}").restype = real
            getattr(L, f"oracle_solve_{sfx}").argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                          ctypes.POINTER(Settings), P, P,
                                                          ctypes.POINTER(Stats), P, P, P, P, P]
            getattr(L, f"oracle_hard_threshold_{sfx}").restype = real
            getattr(L, f"oracle_hard_threshold_{sfx}").argtypes = [P, ctypes.c_uint, ctypes.c_uint,
                                                                   ctypes.c_int, P]
            for nm in ("soft_threshold", "soft_threshold_gpu"):
                getattr(L, f"oracle_{nm}_{sfx}").restype = real
                getattr(L, f"oracle_{nm}_{sfx}").argtypes = [P, ctypes.c_uint, ctypes.c_uint, P, P]
            for nm in ("l0", "l1"):
                fn = getattr(L, f"oracle_{nm}_pen_threshold_{sfx}")
                fn.restype = real
                fn.argtypes = [P, ctypes.c_uint, ctypes.c_double]
            getattr(L, f"oracle_generate_problem_{sfx}").argtypes = [ctypes.c_int, ctypes.c_int, P,
                                                                     ctypes.c_int, ctypes.c_int]
            getattr(L, f"oracle_starting_point_{sfx}").argtypes = [P, ctypes.c_uint, ctypes.c_uint,
                                                                   ctypes.c_int, ctypes.c_uint]
            getattr(L, f"oracle_explained_variance_{sfx}").restype = ctypes.c_double
            getattr(L, f"oracle_explained_variance_{sfx}").argtypes = [P, ctypes.c_int, ctypes.c_int,
                                                                       ctypes.c_int, P, ctypes.c_int]
        L.oracle_rand_r.restype = ctypes.c_uint
        L.oracle_rand_r.argtypes = [ctypes.POINTER(ctypes.c_uint)]
        L.oracle_cblas_name.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def rand_r_sequence(seed, count):
    s = ctypes.c_uint(seed)
    return np.array([lib().oracle_rand_r(ctypes.byref(s)) for _ in range(count)], dtype=np.uint32)


def generate_problem(m, n, uniform_scale=False, dtype=np.float64, ld=None):
    ld = m if ld is None else ld
    A = np.zeros((n, ld), dtype=dtype)  # column-major storage as rows of a C array
    getattr(lib(), f"oracle_generate_problem_{_sfx(dtype)}")(m, n, _ptr(A), ld, int(uniform_scale))
    return A.T  # (ld x n) view, Fortran order


def starting_point(n, s, constrained, k, dtype=np.float64):
    v = np.zeros(n, dtype=dtype)
    getattr(lib(), f"oracle_starting_point_{_sfx(dtype)}")(_ptr(v), n, s, int(constrained), k)
    return v


def hard_threshold(x, s, rule=HARD_SORT):
    x = np.array(x, copy=True)
    tau = np.zeros(1, dtype=x.dtype)
    nrm = getattr(lib(), f"oracle_hard_threshold_{_sfx(x.dtype)}")(_ptr(x), x.size, s, rule, _ptr(tau))
    return x, float(nrm), float(tau[0])


def soft_threshold(x, s, rule=LAMBDA_CPU):
    x = np.array(x, copy=True)
    lam = np.zeros(1, dtype=x.dtype)
    found = ctypes.c_int(0)
    nm = "soft_threshold_gpu" if rule == LAMBDA_GPU else "soft_threshold"
    nrm = getattr(lib(), f"oracle_{nm}_{_sfx(x.dtype)}")(_ptr(x), x.size, s, _ptr(lam), ctypes.byref(found))
    return x, float(nrm), float(lam[0]), bool(found.value)


def pen_threshold(x, gamma, l1):
    x = np.array(x, copy=True)
    f = getattr(lib(), f"oracle_{'l1' if l1 else 'l0'}_pen_threshold_{_sfx(x.dtype)}")(_ptr(x), x.size,
                                                                                       float(gamma))
    return x, float(f)


def solve(A, formulation, s=10, gamma=0.0, tol=0.01, max_iter=20, total_sp=64, batch=None, otf=False,
          termination=TERM_BATCH_MAX, hard_rule=HARD_SORT, X0=None, threads=8, gemm_backend=0,
          want_X=False, lambda_rule=LAMBDA_CPU):
    """Solve with the oracle.  A: (m x n) array (any order; copied to column-major).
    Returns dict(fval, x, it, best_sp, sp_val, sp_iters, sp_nnz, sp_err, X (optional), loop_s, wall_s)."""
    A = np.asfortranarray(A)
    dtype = A.dtype
    m, n = A.shape
    st = Settings(formulation, s, gamma, tol, max_iter, total_sp, total_sp if batch is None else batch,
                  int(otf), termination, hard_rule, threads, gemm_backend, lambda_rule)
    b = st.batch if st.batch else total_sp
    tb = -(-total_sp // b)
    tot = tb * b
    x = np.zeros(n, dtype=dtype)
    sp_val = np.zeros(tot, dtype=dtype)
    sp_iters = np.zeros(tot, dtype=np.int32)
    sp_nnz = np.zeros(tot, dtype=np.int32)
    sp_err = np.zeros(tot, dtype=dtype)
    X = np.zeros((tot, n), dtype=dtype) if want_X else None
    if X0 is not None:
        X0 = np.ascontiguousarray(np.asarray(X0, dtype=dtype).T)  # (B x n) rows = columns
        assert X0.shape == (tot, n)
    stats = Stats()
    fval = getattr(lib(), f"oracle_solve_{_sfx(dtype)}")(_ptr(A), A.strides[1] // A.itemsize, m, n,
                                                       ctypes.byref(st), _ptr(X0), _ptr(x),
                                                       ctypes.byref(stats), _ptr(sp_val), _ptr(sp_iters),
                                                       _ptr(sp_nnz), _ptr(sp_err), _ptr(X))
    out = dict(fval=float(fval), x=x, it=stats.it, best_sp=stats.best_sp, sp_val=sp_val, sp_iters=sp_iters,
               sp_nnz=sp_nnz, sp_err=sp_err, loop_s=stats.loop_s, wall_s=stats.wall_s, threads=stats.threads_used,
               total_sp=stats.total_sp, constraint=stats.constraint)
    if want_X:
        out["X"] = X.T
    return out


def explained_variance(A, x, l2=True):
    A = np.asfortranarray(A)
    x = np.ascontiguousarray(x, dtype=A.dtype)
    return getattr(lib(), f"oracle_explained_variance_{_sfx(A.dtype)}")(
        _ptr(A), A.strides[1] // A.itemsize, A.shape[0], A.shape[1], _ptr(x), int(l2))


def cblas_name():
    n = lib().oracle_cblas_name()
    return n.decode() if n else None
