"""Python mirror of the reference solver interface over the C ABI of libspca_mi355x.so.

The names follow the reference's C++ API so that code written against it reads the same:

* ``SPCA_Formulation``        c++/src/class/optimization_settings.h:27-37
* ``OptimizationSettings``    c++/src/class/optimization_settings.h:71-168
* ``OptimizationStatistics``  c++/src/class/optimization_statistics.h:27-46
* ``GPUSolver.denseDataSolver``  c++/src/gpugpower/gpu_sparse_PCA_solver.h:29-211

Every solve runs on the MI355X through HIP kernels; there is no CPU fallback.  Loading
raises if the library has not been built (``python 24am_amd/spca_mi355x/build.py``).
"""
import ctypes
import enum
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspca_mi355x.so")

SPCA_OK, SPCA_EINVAL, SPCA_ENOMEM, SPCA_EBUDGET, SPCA_EDEVICE, SPCA_ECANCELLED, SPCA_ENOTFOUND, SPCA_EAGAIN = range(8)
STATUS_NAMES = {0: "SPCA_OK", 1: "SPCA_EINVAL", 2: "SPCA_ENOMEM", 3: "SPCA_EBUDGET", 4: "SPCA_EDEVICE",
                5: "SPCA_ECANCELLED", 6: "SPCA_ENOTFOUND", 7: "SPCA_EAGAIN"}

STORE_FP32, STORE_BF16X3, STORE_FP16, STORE_FIX8 = 1, 2, 4, 8
PREC_FP32, PREC_BF16X3, PREC_FP16 = 0, 1, 2
TERM_BATCH_MAX, TERM_PER_SP = 0, 1
START_EXPLICIT, START_REFERENCE, START_REFERENCE_F64, START_GAUSSIAN = 0, 1, 2, 3

# every symbol include/spca.h declares
EXPORTS = [
    "spca_ctx_create", "spca_ctx_destroy", "spca_last_error", "spca_version", "spca_matrix_upload",
    "spca_matrix_from_device", "spca_matrix_add_storage", "spca_matrix_release", "spca_matrix_bytes", "spca_solve",
    "spca_dense_solver_device", "spca_dense_solver_host", "spca_settings_default", "spca_settings_check",
    "spca_is_constrained", "spca_is_l1_constrained", "spca_is_l1_penalized", "spca_is_l1_variance",
    "spca_resolve_support", "spca_server_start", "spca_server_stop", "spca_submit", "spca_poll", "spca_cancel",
    "spca_wait", "spca_solve_fused", "spca_set_tuning", "spca_get_plan", "spca_matrix_generate",
    "spca_matrix_download", "spca_ctx_memory", "spca_solve_components", "spca_matrix_upload_csc",
    "spca_product_forms", "spca_solve_noalloc", "spca_server_set_batching", "spca_server_get_info",
]
NOALLOC_SCRATCH_BYTES = 128 << 20  # the fixed device array of spca_solve_noalloc
FORM_FP32, FORM_BF16X3, FORM_FP16, FORM_FIX8 = 0, 1, 2, 4
FUSED_GRAPH = 1
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_CANCELLED = 0, 1, 2, 3


class SPCA_Formulation(enum.IntEnum):
    L0_constrained_L2_PCA = 0
    L0_constrained_L1_PCA = 1
    L1_constrained_L2_PCA = 2
    L1_constrained_L1_PCA = 3
    L0_penalized_L2_PCA = 4
    L0_penalized_L1_PCA = 5
    L1_penalized_L2_PCA = 6
    L1_penalized_L1_PCA = 7


class SpcaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class _Settings(ctypes.Structure):
    _fields_ = [
        ("formulation", ctypes.c_int), ("constraint", ctypes.c_uint), ("penalty", ctypes.c_double),
        ("tolerance", ctypes.c_double), ("max_iter", ctypes.c_int), ("total_sp", ctypes.c_int),
        ("batch", ctypes.c_int), ("use_otf", ctypes.c_int), ("termination", ctypes.c_int),
        ("precision", ctypes.c_int), ("start", ctypes.c_int), ("check_every", ctypes.c_int),
        ("resolve_iters", ctypes.c_int), ("resolve_tol", ctypes.c_double), ("seed", ctypes.c_uint64),
        ("sparse_products", ctypes.c_int),
        ("exact_sign_products", ctypes.c_int),
        ("exact_ties", ctypes.c_int),
    ]


class _Stats(ctypes.Structure):
    _fields_ = [
        ("it", ctypes.c_uint), ("fval", ctypes.c_double), ("error", ctypes.c_double),
        ("loop_ms", ctypes.c_double), ("wall_ms", ctypes.c_double), ("n", ctypes.c_uint),
        ("best_sp", ctypes.c_int), ("total_sp", ctypes.c_int), ("constraint", ctypes.c_uint),
        ("ev_fp64", ctypes.c_double), ("gemm_ax_ms", ctypes.c_double), ("gemm_atz_ms", ctypes.c_double),
        ("step_ms", ctypes.c_double), ("sp_iterations", ctypes.c_ulonglong),
        ("ev_resolve_fp64", ctypes.c_double), ("resolve_iterations", ctypes.c_int),
        ("sparse_iterations", ctypes.c_uint),
        ("union_iterations", ctypes.c_uint),
        ("list_tries", ctypes.c_uint), ("list_steps", ctypes.c_uint), ("step_kernel_ms", ctypes.c_double),
    ]


class _JobTimes(ctypes.Structure):
    _fields_ = [("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("latency_ms", ctypes.c_double)]


class _ServerInfo(ctypes.Structure):
    _fields_ = [("groups", ctypes.c_int), ("joined_running", ctypes.c_int), ("max_members", ctypes.c_int),
                ("solo_jobs", ctypes.c_int)]


class _SpResult(ctypes.Structure):
    _fields_ = [("value", ctypes.c_float), ("iters", ctypes.c_int), ("nnz", ctypes.c_int),
                ("converged", ctypes.c_int), ("ev_fp64", ctypes.c_double)]


class _FusedJob(ctypes.Structure):
    _fields_ = [("matrix", ctypes.c_uint64), ("settings", ctypes.POINTER(_Settings)), ("X0", ctypes.c_void_p),
                ("x_out", ctypes.c_void_p), ("stats", ctypes.POINTER(_Stats)), ("per_sp", ctypes.c_void_p),
                ("status", ctypes.c_int)]


_lib = None


def lib():
    """Load libspca_mi355x.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run python 24am_amd/spca_mi355x/build.py")
        L = ctypes.CDLL(LIB_PATH)
        P, I, U, U64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_uint64
        sig = {
            "spca_ctx_create": (I, [I, U64, ctypes.POINTER(P)]),
            "spca_ctx_destroy": (I, [P]),
            "spca_last_error": (ctypes.c_char_p, [P]),
            "spca_version": (ctypes.c_char_p, []),
            "spca_matrix_upload": (I, [P, P, I, I, I, U, U64, ctypes.POINTER(U64)]),
            "spca_matrix_from_device": (I, [P, P, I, I, I, U, U64, ctypes.POINTER(U64)]),
            "spca_matrix_add_storage": (I, [P, U64, U]),
            "spca_matrix_release": (I, [P, U64]),
            "spca_matrix_bytes": (I, [P, U64, ctypes.POINTER(U64)]),
            "spca_solve": (I, [P, U64, ctypes.POINTER(_Settings), P, P, ctypes.POINTER(_Stats), P, P]),
            "spca_solve_noalloc": (I, [P, U64, ctypes.POINTER(_Settings), P, P, ctypes.POINTER(_Stats), P, P]),
            "spca_solve_components": (I, [P, U64, ctypes.POINTER(_Settings), P, I, P, P]),
            "spca_matrix_upload_csc": (I, [P, P, P, P, I, I, P, P, U64, ctypes.POINTER(U64)]),
            "spca_dense_solver_device": (I, [P, U, U, P, P, ctypes.POINTER(_Settings), ctypes.POINTER(_Stats), U, U]),
            "spca_dense_solver_host": (I, [P, P, I, P, U, U, ctypes.POINTER(_Settings), ctypes.POINTER(_Stats),
                                           ctypes.POINTER(ctypes.c_double)]),
            "spca_settings_default": (None, [ctypes.POINTER(_Settings)]),
            "spca_settings_check": (I, [ctypes.POINTER(_Settings), U]),
            "spca_is_constrained": (I, [I]),
            "spca_is_l1_constrained": (I, [I]),
            "spca_is_l1_penalized": (I, [I]),
            "spca_is_l1_variance": (I, [I]),
            "spca_resolve_support": (I, [P, U64, I, P, I, ctypes.c_double, P, P, ctypes.POINTER(I)]),
            "spca_server_start": (I, [P, I]),
            "spca_server_stop": (I, [P]),
            "spca_server_set_batching": (I, [P, I]),
            "spca_server_get_info": (I, [P, ctypes.POINTER(_ServerInfo)]),
            "spca_submit": (I, [P, U64, ctypes.POINTER(_Settings), P, I, ctypes.POINTER(U64)]),
            "spca_poll": (I, [P, U64, ctypes.POINTER(I)]),
            "spca_cancel": (I, [P, U64]),
            "spca_wait": (I, [P, U64, P, ctypes.POINTER(_Stats), P, ctypes.POINTER(_JobTimes)]),
            "spca_solve_fused": (I, [P, ctypes.POINTER(_FusedJob), I, U]),
            "spca_set_tuning": (I, [P, I, ctypes.c_char_p]),
            "spca_matrix_generate": (I, [P, I, I, U64, I, P, I, U, U64, P, ctypes.POINTER(U64)]),
            "spca_matrix_download": (I, [P, U64, P, I]),
            "spca_ctx_memory": (I, [P, ctypes.POINTER(U64), ctypes.POINTER(U64), ctypes.POINTER(U64)]),
            "spca_get_plan": (I, [P, I, I, I, I, I, ctypes.POINTER(I), ctypes.POINTER(I),
                                  ctypes.POINTER(ctypes.c_double)]),
            "spca_product_forms": (I, [ctypes.POINTER(_Settings), I, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def version():
    return lib().spca_version().decode()


class OptimizationSettings:
    """Reference OptimizationSettings (defaults from optimization_settings.h:101-123) plus
    the MI355X knobs: precision, termination rule, starting-point source, check cadence."""

    def __init__(self, **kw):
        self.tolerance = 0.01
        self.constraintParameter = 10
        self.penaltyParameter = 0.0
        self.totalStartingPoints = 64
        self.batchSize = 64
        self.useOTF = False
        self.maximumIterations = 20
        self.formulation = SPCA_Formulation.L0_constrained_L2_PCA
        self.precision = PREC_FP32
        self.termination = TERM_BATCH_MAX
        self.start = START_EXPLICIT
        self.check_every = 1
        self.resolve_iters = 0  # > 0: FP64 re-solve on the support of the result
        self.resolve_tol = 1e-12
        self.seed = 0  # START_GAUSSIAN: key of the device-generated starting points
        # 1: gather A*X over the nonzeros of sparse iterates; 2: also the FP32 product over the union
        # of the other iterates' supports when it is small (batch mode); 0: always dense
        self.sparse_products = 1
        # L1 variance, FP32 and bf16x3: A^T sgn(z) from the exact bf16 split of A (0: the precision's product)
        self.exact_sign_products = 1
        self.exact_ties = 1
        # reference useDoublePrecision: this solver has no double products, so True is
        # rejected (SPCA_EINVAL) when the settings reach the library, not run in FP32
        self.useDoublePrecision = False
        for k, v in kw.items():
            if not hasattr(self, k):
                raise AttributeError(k)
            setattr(self, k, v)

    def isConstrainedProblem(self):
        return bool(lib().spca_is_constrained(int(self.formulation)))

    def isL1ConstrainedProblem(self):
        return bool(lib().spca_is_l1_constrained(int(self.formulation)))

    def isL1PenalizedProblem(self):
        return bool(lib().spca_is_l1_penalized(int(self.formulation)))

    def chceckInputAndModifyIt(self, n):
        """Reference clamp: s <= n, total SPs rounded up to a batch multiple; returns batches."""
        c = self._c()
        tb = lib().spca_settings_check(ctypes.byref(c), n)
        self.constraintParameter, self.batchSize, self.totalStartingPoints = c.constraint, c.batch, c.total_sp
        return tb

    def product_forms(self, m, n):
        """(A*X form, A^T*Z form) a dense m x n solve of these settings runs (FORM_*), decided on
        the host; raises SpcaError (SPCA_EINVAL) for settings a solve rejects (n >= 2^23)."""
        c = self._c()
        ax, atz = ctypes.c_int(), ctypes.c_int()
        rc = lib().spca_product_forms(ctypes.byref(c), int(m), int(n), ctypes.byref(ax), ctypes.byref(atz))
        if rc != SPCA_OK:
            raise SpcaError(rc, f"no solve of {m} x {n} with these settings")
        return ax.value, atz.value

    def _c(self):
        if self.useDoublePrecision:
            raise SpcaError(SPCA_EINVAL, "useDoublePrecision: the products run in FP32, bf16x3 or fp16 "
                                         "(precision); there is no double solve")
        return _Settings(int(self.formulation), int(self.constraintParameter), float(self.penaltyParameter),
                         float(self.tolerance), int(self.maximumIterations), int(self.totalStartingPoints),
                         int(self.batchSize), int(bool(self.useOTF)), int(self.termination), int(self.precision),
                         int(self.start), int(self.check_every), int(self.resolve_iters), float(self.resolve_tol),
                         int(self.seed), int(self.sparse_products), int(self.exact_sign_products),
                         int(self.exact_ties))


class OptimizationStatistics:
    """Reference OptimizationStatistics plus timing and FP64 explained variance."""

    def __init__(self):
        self.it = 0
        self.fval = 0.0
        self.error = 0.0
        self.totalTrueComputationTime = 0.0  # seconds, iteration loop (hipEvents)
        self.totalElapsedTime = 0.0          # seconds, whole call
        self.n = 0
        self.values = []
        self.iters = []
        self.cardinalities = []
        self.ev_fp64 = None
        self.sp_ev_fp64 = []
        self.converged = []
        self.best_sp = -1
        self.gemm_ax_ms = 0.0
        self.gemm_atz_ms = 0.0
        self.step_ms = 0.0
        self.step_kernel_ms = 0.0  # of step_ms, the formulation step kernel alone
        self.sp_iterations = 0
        self.ev_resolve_fp64 = None
        self.resolve_iterations = 0
        self.sparse_iterations = 0
        self.union_iterations = 0
        self.list_tries = 0  # constrained steps that started from the previous threshold's list
        self.list_steps = 0  # of those, the ones the list decided

    def _fill(self, s, per_sp=None):
        self.it, self.fval, self.error = s.it, s.fval, s.error
        self.totalTrueComputationTime = s.loop_ms / 1e3
        self.totalElapsedTime = s.wall_ms / 1e3
        self.n, self.best_sp, self.ev_fp64 = s.n, s.best_sp, s.ev_fp64
        self.gemm_ax_ms, self.gemm_atz_ms, self.step_ms = s.gemm_ax_ms, s.gemm_atz_ms, s.step_ms
        self.step_kernel_ms = s.step_kernel_ms
        self.sp_iterations = s.sp_iterations
        self.ev_resolve_fp64, self.resolve_iterations = s.ev_resolve_fp64, s.resolve_iterations
        self.sparse_iterations = s.sparse_iterations
        self.union_iterations = s.union_iterations
        self.list_tries, self.list_steps = s.list_tries, s.list_steps
        if per_sp is not None:
            self.values = np.array([r.value for r in per_sp], dtype=np.float32)
            self.iters = np.array([r.iters for r in per_sp], dtype=np.int32)
            self.cardinalities = np.array([r.nnz for r in per_sp], dtype=np.int32)
            self.converged = np.array([r.converged for r in per_sp], dtype=bool)
            self.sp_ev_fp64 = np.array([r.ev_fp64 for r in per_sp], dtype=np.float64)


def _check(ctx_ptr, rc):
    if rc != SPCA_OK:
        msg = lib().spca_last_error(ctx_ptr).decode() if ctx_ptr else ""
        raise SpcaError(rc, msg)


class Context:
    """Owns the device, the solver stream, resident matrices and the HBM budget."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self._p = ctypes.c_void_p()
        rc = lib().spca_ctx_create(device, hbm_budget_bytes, ctypes.byref(self._p))
        if rc != SPCA_OK:
            raise SpcaError(rc, "spca_ctx_create failed")

    @property
    def ptr(self):
        return self._p

    def close(self):
        if self._p:
            lib().spca_ctx_destroy(self._p)
            self._p = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def start_server(self, workers=4, batching=64):
        """Start the request server: `workers` worker threads, each on its own stream.  A worker
        that takes a dense one-batch job runs a fused group that queued jobs of its precision
        join at chunk boundaries (at most `batching` jobs at once; 0: every job alone)."""
        _check(self._p, lib().spca_server_start(self._p, workers))
        _check(self._p, lib().spca_server_set_batching(self._p, batching))

    def server_info(self):
        """Counters of the running server: groups, joined_running, max_members, solo_jobs."""
        i = _ServerInfo()
        _check(self._p, lib().spca_server_get_info(self._p, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in _ServerInfo._fields_}

    def stop_server(self):
        _check(self._p, lib().spca_server_stop(self._p))

    def memory(self):
        """(used, cached, budget) bytes of device memory held by the context."""
        u, c, b = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        _check(self._p, lib().spca_ctx_memory(self._p, ctypes.byref(u), ctypes.byref(c), ctypes.byref(b)))
        return u.value, c.value, b.value

    def set_tuning(self, mode, cache_path=None):
        """mode 1: time the product plans of each new solve shape and keep the fastest
        (cached in memory and, with cache_path, in a JSON file); mode 0: fixed heuristic."""
        _check(self._p, lib().spca_set_tuning(self._p, int(mode), cache_path.encode() if cache_path else None))

    def plan(self, precision, m, n, batch, fused=False):
        """(splits, big, tuned_ms) of the plan a solve of this shape uses now."""
        sp, bg, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(self._p, lib().spca_get_plan(self._p, int(precision), int(m), int(n), int(batch), int(fused),
                                            ctypes.byref(sp), ctypes.byref(bg), ctypes.byref(ms)))
        return sp.value, bg.value, ms.value

    def solve_fused(self, jobs, graph=True, per_sp=True):
        """Solve uneven jobs together, one fused launch per phase and iteration.
        jobs: list of (matrix, settings, X0 or None).  Returns a list of
        (x, OptimizationStatistics) in the order given."""
        keep, cj = [], (_FusedJob * len(jobs))()
        outs = []
        for k, (M, settings, X0) in enumerate(jobs):
            c = settings._c()
            total = _total_after_check(c, M.n)
            X0p = None
            if c.start == START_EXPLICIT:
                X0 = np.asfortranarray(X0, dtype=np.float32)
                if X0.shape != (M.n, total):
                    raise ValueError(f"job {k}: X0 must be {M.n} x {total}")
                X0p = X0.ctypes.data_as(ctypes.c_void_p)
            x = np.zeros(M.n, dtype=np.float32)
            st = _Stats()
            res = (_SpResult * total)() if per_sp else None
            keep += [c, X0, x, st, res]
            cj[k].matrix = M.handle
            cj[k].settings = ctypes.pointer(c)
            cj[k].X0 = X0p
            cj[k].x_out = x.ctypes.data_as(ctypes.c_void_p)
            cj[k].stats = ctypes.pointer(st)
            cj[k].per_sp = ctypes.cast(res, ctypes.c_void_p) if res is not None else None
            outs.append((x, st, res))
        _check(self._p, lib().spca_solve_fused(self._p, cj, len(jobs), FUSED_GRAPH if graph else 0))
        result = []
        for x, st, res in outs:
            S = OptimizationStatistics()
            S._fill(st, res)
            result.append((x, S))
        return result

    def generate(self, m, n, seed, theta=(), spike_nnz=0, storage=STORE_FP32, key=0, want_loadings=False):
        """Spiked m x n matrix generated on the device (Philox noise + planted loadings).
        Returns the Matrix, and with want_loadings also the n x len(theta) loadings."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        L = np.zeros((n, len(th)), dtype=np.float32, order="F") if want_loadings else None
        h = ctypes.c_uint64()
        _check(self._p, lib().spca_matrix_generate(self._p, m, n, int(seed), len(th),
                                                   th.ctypes.data_as(ctypes.c_void_p) if len(th) else None,
                                                   int(spike_nnz), storage, key,
                                                   L.ctypes.data_as(ctypes.c_void_p) if want_loadings else None,
                                                   ctypes.byref(h)))
        M = Matrix(self, h.value, m, n)
        return (M, L) if want_loadings else M

    def upload(self, A, storage=STORE_FP32, key=0):
        """A: (m x n) float32 array (any order); copied column major to the device."""
        A = np.asfortranarray(A, dtype=np.float32)
        m, n = A.shape
        h = ctypes.c_uint64()
        _check(self._p, lib().spca_matrix_upload(self._p, A.ctypes.data_as(ctypes.c_void_p), m, n, m, storage, key,
                                                 ctypes.byref(h)))
        return Matrix(self, h.value, m, n)

    def upload_csc(self, col_ptr, row_idx, vals, m, n, col_means=None, row_means=None, key=0):
        """Sparse A from CSC arrays (zero-based), centred implicitly by the optional column and
        row means: A_c = A - 1 col_means^T - row_means 1^T (the reference sparse solver's
        doMean / doRowMean).  Solves run in FP32 on the CSR/CSC copies."""
        cp = np.ascontiguousarray(col_ptr, dtype=np.int32)
        ri = np.ascontiguousarray(row_idx, dtype=np.int32)
        va = np.ascontiguousarray(vals, dtype=np.float32)
        if cp.size != n + 1 or ri.size != cp[-1] or va.size != cp[-1]:
            raise ValueError("col_ptr must have n+1 entries ending at len(row_idx) == len(vals)")
        cm = None if col_means is None else np.ascontiguousarray(col_means, dtype=np.float32)
        rm = None if row_means is None else np.ascontiguousarray(row_means, dtype=np.float32)
        ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
        h = ctypes.c_uint64()
        _check(self._p, lib().spca_matrix_upload_csc(self._p, ptr(cp), ptr(ri), ptr(va), int(m), int(n), ptr(cm),
                                                     ptr(rm), int(key), ctypes.byref(h)))
        return Matrix(self, h.value, m, n)


class Matrix:
    def __init__(self, ctx, handle, m, n):
        self.ctx, self.handle, self.m, self.n = ctx, handle, m, n

    def download(self):
        """The resident FP32 matrix as an (m x n) array."""
        A = np.zeros((self.m, self.n), dtype=np.float32, order="F")
        _check(self.ctx.ptr, lib().spca_matrix_download(self.ctx.ptr, self.handle, A.ctypes.data_as(ctypes.c_void_p),
                                                        self.m))
        return A

    def nbytes(self):
        b = ctypes.c_uint64()
        _check(self.ctx.ptr, lib().spca_matrix_bytes(self.ctx.ptr, self.handle, ctypes.byref(b)))
        return b.value

    def add_storage(self, storage):
        _check(self.ctx.ptr, lib().spca_matrix_add_storage(self.ctx.ptr, self.handle, storage))

    def resolve_support(self, x, formulation, max_iter=1000, tol=1e-12):
        """FP64 re-solve on the support of x; returns (y (n doubles), ev, iterations or -1 if converged)."""
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(self.n, dtype=np.float64)
        ev = ctypes.c_double()
        its = ctypes.c_int()
        _check(self.ctx.ptr, lib().spca_resolve_support(self.ctx.ptr, self.handle, int(formulation),
                                                        x.ctypes.data_as(ctypes.c_void_p), max_iter, tol,
                                                        y.ctypes.data_as(ctypes.c_void_p), ctypes.byref(ev),
                                                        ctypes.byref(its)))
        return y, ev.value, its.value

    def release(self):
        if self.handle:
            _check(self.ctx.ptr, lib().spca_matrix_release(self.ctx.ptr, self.handle))
            self.handle = 0

    def solve_components(self, settings, k, X0=None):
        """k sparse components by deflation (reference SparseDeflationCollection,
        c++/src/gpower/sparse_PCA_solver_for_CSC.h:56-106): component c is solved with every
        iterate projected off components 0 .. c-1.  Returns X (n x k, unit columns) and a list
        of k OptimizationStatistics."""
        c = settings._c()
        total = _total_after_check(c, self.n)
        X0p = None
        if c.start == START_EXPLICIT:
            if X0 is None:
                raise ValueError("X0 required for START_EXPLICIT")
            X0 = np.asfortranarray(X0, dtype=np.float32)
            if X0.shape != (self.n, total):
                raise ValueError(f"X0 must be {self.n} x {total}")
            X0p = X0.ctypes.data_as(ctypes.c_void_p)
        X = np.zeros((self.n, k), dtype=np.float32, order="F")
        st = (_Stats * k)()
        _check(self.ctx.ptr, lib().spca_solve_components(self.ctx.ptr, self.handle, ctypes.byref(c), X0p, int(k),
                                                         X.ctypes.data_as(ctypes.c_void_p),
                                                         ctypes.cast(st, ctypes.c_void_p)))
        out = []
        for i in range(k):
            S = OptimizationStatistics()
            S._fill(st[i])
            out.append(S)
        return X, out

    def solve_noalloc(self, settings, X0=None, stats=None, want_X=False, per_sp=True):
        """solve() through spca_solve_noalloc: the device workspace is the library's fixed array
        (no device allocation); raises SpcaError (SPCA_EINVAL) for shapes that do not fit it."""
        return self.solve(settings, X0=X0, stats=stats, want_X=want_X, per_sp=per_sp, _entry="spca_solve_noalloc")

    def solve(self, settings, X0=None, stats=None, want_X=False, per_sp=True, _entry="spca_solve"):
        """Returns x (unit norm, length n); fills `stats`; with want_X also returns X (n x total)."""
        c = settings._c()
        total = _total_after_check(c, self.n)
        x = np.zeros(self.n, dtype=np.float32)
        X0p = None
        if c.start == START_EXPLICIT:
            if X0 is None:
                raise ValueError("X0 required for START_EXPLICIT")
            X0 = np.asfortranarray(X0, dtype=np.float32)
            if X0.shape != (self.n, total):
                raise ValueError(f"X0 must be {self.n} x {total}")
            X0p = X0.ctypes.data_as(ctypes.c_void_p)
        res = (_SpResult * total)() if per_sp else None
        Xf = np.zeros((self.n, total), dtype=np.float32, order="F") if want_X else None
        s = _Stats()
        _check(self.ctx.ptr, getattr(lib(), _entry)(self.ctx.ptr, self.handle, ctypes.byref(c), X0p,
                                                    x.ctypes.data_as(ctypes.c_void_p), ctypes.byref(s),
                                                    ctypes.cast(res, ctypes.c_void_p) if res is not None else None,
                                                    Xf.ctypes.data_as(ctypes.c_void_p) if want_X else None))
        if stats is not None:
            stats._fill(s, res)
        return (x, Xf) if want_X else x


class Job:
    """A submitted solve (continuous batching through the context's request server)."""

    def __init__(self, matrix, handle, total):
        self.matrix, self.handle, self.total = matrix, handle, total

    def poll(self):
        st = ctypes.c_int()
        _check(self.matrix.ctx.ptr, lib().spca_poll(self.matrix.ctx.ptr, self.handle, ctypes.byref(st)))
        return st.value

    def cancel(self):
        _check(self.matrix.ctx.ptr, lib().spca_cancel(self.matrix.ctx.ptr, self.handle))

    def wait(self, stats=None, per_sp=True):
        """Blocks until done; returns (x, times dict).  Raises SpcaError on failure/cancel."""
        x = np.zeros(self.matrix.n, dtype=np.float32)
        s = _Stats()
        res = (_SpResult * self.total)() if per_sp else None
        t = _JobTimes()
        rc = lib().spca_wait(self.matrix.ctx.ptr, self.handle, x.ctypes.data_as(ctypes.c_void_p), ctypes.byref(s),
                             ctypes.cast(res, ctypes.c_void_p) if res is not None else None, ctypes.byref(t))
        _check(self.matrix.ctx.ptr, rc)
        if stats is not None:
            stats._fill(s, res)
        return x, {"queue_ms": t.queue_ms, "run_ms": t.run_ms, "latency_ms": t.latency_ms}


def _submit(matrix, settings, X0=None, per_sp=True):
    c = settings._c()
    total = _total_after_check(c, matrix.n)
    X0p = None
    if c.start == START_EXPLICIT:
        X0 = np.asfortranarray(X0, dtype=np.float32)
        if X0.shape != (matrix.n, total):
            raise ValueError(f"X0 must be {matrix.n} x {total}")
        X0p = X0.ctypes.data_as(ctypes.c_void_p)
    h = ctypes.c_uint64()
    _check(matrix.ctx.ptr, lib().spca_submit(matrix.ctx.ptr, matrix.handle, ctypes.byref(c), X0p, int(per_sp),
                                             ctypes.byref(h)))
    return Job(matrix, h.value, total)


Matrix.submit = _submit


def _total_after_check(c, n):
    c2 = _Settings.from_buffer_copy(c)
    lib().spca_settings_check(ctypes.byref(c2), n)
    return c2.total_sp


class GPUSolver:
    """Mirror of SPCASolver::GPUSolver (c++/src/gpugpower/gpu_sparse_PCA_solver.h)."""

    @staticmethod
    def denseDataSolver(ctx, A, settings, stats, X0=None):
        """Solve on host matrix A (m x n); returns the unit-norm x of length n.  With X0=None
        the reference starting-point generator is used (SP k seeded with k)."""
        if X0 is None and settings.start == START_EXPLICIT:
            settings = OptimizationSettings(**{k: v for k, v in vars(settings).items()})
            settings.start = START_REFERENCE
        need = {PREC_FP32: STORE_FP32, PREC_BF16X3: STORE_BF16X3, PREC_FP16: STORE_FP16}[settings.precision]
        M = ctx.upload(A, storage=need)
        try:
            return M.solve(settings, X0=X0, stats=stats)
        finally:
            M.release()
