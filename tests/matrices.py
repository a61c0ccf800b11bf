"""The matrices of the reference's solver tests
(tests/splines/splines_linear_problem.cpp:172-433), as dense numpy arrays."""
import numpy as np


def nonsym(N, k):
    """splines_linear_problem.cpp:172-181: non-symmetric full-rank band matrix."""
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = 3.0 / 4 * ((N + 1) * i + 1)
        for j in range(max(0, i - k), i):
            A[i, j] = -(1.0 / 4) / k * (N * i + j + 1)
        for j in range(i + 1, min(N, i + k + 1)):
            A[i, j] = -(1.0 / 4) / k * (N * i + j + 1)
    return A


def pds(N, k):
    """splines_linear_problem.cpp:193-202: SPD band matrix."""
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = 2.0 * k + 1
        for j in range(max(0, i - k), min(N, i + k + 1)):
            if j != i:
                A[i, j] = -1.0
    return A


def offset_banded(N, k):
    """splines_linear_problem.cpp:214-225: band shifted to the upper side."""
    A = np.zeros((N, N))
    for i in range(N):
        A[i, i] = -1.0
        if i + k < N:
            A[i, i + k] = 2.0 * k + 1
    return A


def periodic_band(N, k, s):
    """splines_linear_problem.cpp:404-433: periodic band shifted by s."""
    A = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            diag = (j - i) % N
            if diag == s or diag == N + s:
                A[i, j] = 2.0 * k + 1
            elif diag <= s + k or diag >= N + s - k:
                A[i, j] = -1.0
    return A


# (N, k) of INSTANTIATE_TEST_SUITE_P (splines_linear_problem.cpp:469-472)
SIZES = [(10, k) for k in range(1, 7)] + [(20, k) for k in range(1, 7)]
