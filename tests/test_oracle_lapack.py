"""Pins the oracle's linear-algebra restatement (CPU only).

- LAPACK factorisations and solves vs scipy.linalg.lapack (an independent
  implementation of the same netlib algorithms the reference calls through
  LAPACKE): dgbtrf/dgbtrs, dpbtrf/dpbtrs, dpttrf/dpttrs, dgetrf/dgetrs.
- The reference's own solver tests (tests/splines/splines_linear_problem.cpp:
  49-472): A X = I and A X^T = I to 1e-10 for every problem kind, on the same
  matrices, plus agreement with numpy.linalg.solve.
"""
import numpy as np
import pytest
import scipy.linalg.lapack as la

import oracle as O
from matrices import nonsym, offset_banded, pds, periodic_band


def load(prob, A):
    N = A.shape[0]
    for i in range(N):
        for j in range(N):
            if A[i, j] != 0.0:
                prob.set(i, j, A[i, j])
    return prob


def solve_and_validate(prob, A):
    """splines_linear_problem.cpp:49-142."""
    N = A.shape[0]
    prob.setup()
    rows = prob.rhs_rows
    inv = np.zeros((rows, N))
    inv[:N] = np.eye(N)
    inv = prob.solve(inv, transpose=False)[:N]
    inv_tr = np.zeros((rows, N))
    inv_tr[:N] = np.eye(N)
    inv_tr = prob.solve(inv_tr, transpose=True)[:N]
    assert np.abs(A @ inv - np.eye(N)).max() <= 1e-10
    assert np.abs(A @ inv_tr.T - np.eye(N)).max() <= 1e-10
    # and against an independent dense solve
    ref = np.linalg.solve(A, np.eye(N))
    assert np.abs(inv - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(inv_tr - np.linalg.solve(A.T, np.eye(N))).max() <= 1e-12 * max(1.0, np.abs(ref).max())


SIZES = [(10, k) for k in range(1, 7)] + [(20, k) for k in range(1, 7)]


def test_dense():
    A = nonsym(10, 10)
    solve_and_validate(load(O.Problem("dense", 10), A), A)


def test_band():
    A = nonsym(10, 3)
    p = O.Problem("band", 10, 3, 3, False)
    solve_and_validate(load(p, A), A)


def test_pds_band():
    A = pds(10, 3)
    p = O.Problem("band", 10, 3, 3, True)
    assert p.type == "pds_band"
    solve_and_validate(load(p, A), A)


def test_pds_tridiag():
    A = pds(10, 1)
    p = O.Problem("band", 10, 1, 1, True)
    assert p.type == "pds_tridiag"
    solve_and_validate(load(p, A), A)


def test_2x2_dense_main():
    A = nonsym(10, 10)
    p = O.Problem("dense2x2", 10, bottom=3)
    solve_and_validate(load(p, A), A)


def test_3x3_dense_center():
    A = nonsym(10, 10)
    p = O.Problem("dense3x3", 10, bottom=2, top=3)
    solve_and_validate(load(p, A), A)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_nonsymmetric(N, k):
    A = nonsym(N, k)
    solve_and_validate(load(O.Problem("band", N, k, k, False), A), A)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_pds(N, k):
    A = pds(N, k)
    solve_and_validate(load(O.Problem("band", N, k, k, True), A), A)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_offset_banded(N, k):
    A = offset_banded(N, k)
    solve_and_validate(load(O.Problem("band", N, 0, 2 * k, True), A), A)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_2x2(N, k):
    A = nonsym(N, k)
    solve_and_validate(load(O.Problem("block", N, k, k, False, bottom=3), A), A)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_3x3(N, k):
    A = nonsym(N, k)
    solve_and_validate(load(O.Problem("block", N, k, k, False, bottom=3, top=2), A), A)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_periodic_band(N, k):
    """splines_linear_problem.cpp:404-433: periodic band shifted by s."""
    for s in range(-k + k // 2 + 1, k - k // 2):
        A = periodic_band(N, k, s)
        solve_and_validate(load(O.Problem("periodic", N, k - s, k + s, False), A), A)


# ---------------------------------------------------------------------------
# Factor-level agreement with scipy's LAPACK
# ---------------------------------------------------------------------------


def _band_storage(A, kl, ku):
    """LAPACK band storage with kl extra rows for dgbtrf: AB[kl+ku+i-j, j] = A[i, j]."""
    N = A.shape[0]
    ab = np.zeros((2 * kl + ku + 1, N))
    for j in range(N):
        for i in range(max(0, j - ku), min(N, j + kl + 1)):
            ab[kl + ku + i - j, j] = A[i, j]
    return ab


@pytest.mark.parametrize("N,k", [(20, 2), (20, 4), (37, 3), (50, 1)])
def test_gbtrf_matches_scipy(N, k):
    rng = np.random.default_rng(N * 10 + k)
    A = nonsym(N, k) * (1 + 0.1 * rng.standard_normal((N, N)))
    A = np.where(np.abs(np.subtract.outer(np.arange(N), np.arange(N))) <= k, A, 0.0)
    p = load(O.Problem("band", N, k, k, False), A)
    assert p.type == "band"
    p.setup()
    lu, piv, info = la.dgbtrf(_band_storage(A, k, k), k, k)
    assert info == 0
    q = np.ctypeslib.as_array(O.lib().or_problem_q(p.h), shape=((2 * k + k + 1) * N,)).reshape(2 * k + k + 1, N)
    ipiv = np.ctypeslib.as_array(O.lib().or_problem_ipiv(p.h), shape=(N,))
    assert np.array_equal(ipiv, piv)
    assert np.abs(q - lu).max() <= 1e-15 * np.abs(lu).max()
    b = rng.standard_normal((N, 3))
    for trans in (0, 1):
        x_ref, info = la.dgbtrs(lu, k, k, b, piv, trans=trans)
        x = p.solve(b.copy(), transpose=bool(trans))
        assert np.abs(x - x_ref).max() <= 1e-14 * np.abs(x_ref).max()


@pytest.mark.parametrize("N,k", [(20, 2), (31, 3), (16, 1)])
def test_pbtrf_matches_scipy(N, k):
    A = pds(N, k)
    p = load(O.Problem("band", N, k, k, True), A)
    p.setup()
    ab = np.zeros((k + 1, N))
    for j in range(N):
        for i in range(j, min(N, j + k + 1)):
            ab[i - j, j] = A[i, j]
    if k == 1:
        d, e, info = la.dpttrf(np.diag(A).copy(), np.diag(A, -1).copy())
        q = np.ctypeslib.as_array(O.lib().or_problem_q(p.h), shape=(2 * N,))
        assert np.abs(q[:N] - d).max() <= 1e-15 * np.abs(d).max()
        assert np.abs(q[N:2 * N - 1] - e).max() <= 1e-15 * max(1.0, np.abs(e).max())
    else:
        c, info = la.dpbtrf(ab, lower=1)
        assert info == 0
        q = np.ctypeslib.as_array(O.lib().or_problem_q(p.h), shape=((k + 1) * N,)).reshape(k + 1, N)
        assert np.abs(q - c).max() <= 1e-15 * np.abs(c).max()
    b = np.random.default_rng(1).standard_normal((N, 2))
    x = p.solve(b.copy())
    assert np.abs(A @ x - b).max() <= 1e-13 * np.abs(b).max()
