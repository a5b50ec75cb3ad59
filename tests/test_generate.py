"""Device spiked-matrix generator (spca_matrix_generate) against a NumPy restatement of its
stated construction: Philox4x32-10 counters + Box-Muller.  The Philox restatement is pinned
by the published known-answer vectors (Salmon et al., SC 2011; the Random123 KAT set)."""
import numpy as np
import pytest

import spca_mi355x as S

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def philox(c, k):
    """Philox4x32-10 on NumPy uint64 arrays (one lane per element)."""
    c = [np.asarray(x, dtype=np.uint64) for x in c]
    k0, k1 = np.uint64(k[0]), np.uint64(k[1])
    msk = np.uint64(0xffffffff)
    for r in range(10):
        if r:
            k0, k1 = (k0 + np.uint64(W0)) & msk, (k1 + np.uint64(W1)) & msk
        p0, p1 = np.uint64(M0) * c[0], np.uint64(M1) * c[2]
        c = [(p1 >> np.uint64(32)) ^ c[1] ^ k0, p1 & msk, (p0 >> np.uint64(32)) ^ c[3] ^ k1, p0 & msk]
    return c


def normal(a, b, stream, seed):
    a = np.asarray(a, dtype=np.uint64)
    b = np.broadcast_to(np.asarray(b, dtype=np.uint64), a.shape)
    z = np.zeros(a.shape, dtype=np.uint64)
    c = philox([a, b, z + np.uint64(stream), z], (seed & 0xffffffff, seed >> 32))
    u0 = (c[0].astype(np.float64) + 0.5) * 2.0**-32
    u1 = (c[1].astype(np.float64) + 0.5) * 2.0**-32
    return np.sqrt(-2 * np.log(u0)) * np.cos(2 * np.pi * u1)


def reference_matrix(m, n, seed, theta, nnz):
    r, c = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
    A = normal(r.ravel(), c.ravel(), 0, seed).reshape(m, n).astype(np.float32)
    L = np.zeros((n, len(theta)))
    for k, th in enumerate(theta):
        idx, t = [], 0
        while len(idx) < min(nnz, n):
            col = int(philox([np.uint64(t), np.uint64(k), np.uint64(3), np.uint64(0)],
                             (seed & 0xffffffff, seed >> 32))[0]) % n
            if col not in idx:
                idx.append(col)
            t += 1
        v = normal(np.arange(len(idx)), k, 2, seed)
        v /= np.linalg.norm(v)
        g = normal(np.arange(m), k, 1, seed).astype(np.float32)
        coef = (np.sqrt(th) * v).astype(np.float32)
        for t, col in enumerate(idx):
            A[:, col] = A[:, col] + coef[t] * g
        L[idx, k] = v
    return A, L


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for c, k, want in kat:
        got = philox([np.uint64(x) for x in c], k)
        assert tuple(int(x) for x in got) == want


def test_reference_matrix_statistics():
    A, L = reference_matrix(64, 500, 7, [50.0], 20)
    noise = np.delete(A, np.nonzero(L[:, 0])[0], axis=1)
    assert abs(noise.mean()) < 0.02 and abs(noise.std() - 1) < 0.02
    np.testing.assert_allclose(np.linalg.norm(L, axis=0), 1.0)


@pytest.mark.gpu
def test_device_generator_matches_restatement(ctx):
    m, n, seed, theta, nnz = 96, 700, 0x1234_5678_9abc, [30.0, 12.0], 25
    want, Lw = reference_matrix(m, n, seed, theta, nnz)
    M, L = ctx.generate(m, n, seed, theta=theta, spike_nnz=nnz, want_loadings=True)
    try:
        A = M.download()
        np.testing.assert_array_equal(L != 0, Lw != 0)
        np.testing.assert_allclose(L, Lw, rtol=1e-6, atol=1e-7)
        # device logf/cosf against float64 NumPy: a few ulp
        np.testing.assert_allclose(A, want, rtol=1e-5, atol=1e-5)
    finally:
        M.release()


@pytest.mark.gpu
def test_device_gaussian_starts_match_restatement(ctx):
    """START_GAUSSIAN draws SP k from counters (i, k, 4); the same solve from the NumPy
    restatement's starting points (explicit X0) agrees to float rounding."""
    m, n, B, seed = 80, 500, 40, 99
    A, _ = reference_matrix(m, n, 5, [25.0], 20)
    M = ctx.upload(A)
    try:
        X0 = np.stack([normal(np.arange(n), k, 4, seed) for k in range(B)], axis=1)
        X0 = (X0 / np.linalg.norm(X0, axis=0)).astype(np.float32)
        for f in (0, 6):
            st = S.OptimizationSettings(formulation=f, constraintParameter=10, penaltyParameter=2.0, tolerance=0.0,
                                        maximumIterations=10, totalStartingPoints=B, batchSize=B, seed=seed,
                                        start=S.START_GAUSSIAN)
            sg = S.OptimizationStatistics()
            M.solve(st, stats=sg)
            st.start = S.START_EXPLICIT
            se = S.OptimizationStatistics()
            M.solve(st, X0=X0, stats=se)
            np.testing.assert_allclose(sg.values, se.values, rtol=1e-5)
            np.testing.assert_array_equal(sg.cardinalities, se.cardinalities)
    finally:
        M.release()
