"""GPU parity of SplinesLinearProblem::solve(b, transpose) through the C ABI
(ddcx_problem_*), on the reference's own solver tests
(tests/splines/splines_linear_problem.cpp:49-472): every matrix kind the
maker produces, both directions, with the reference's check A X = I (and
A X^T = I for the transposed solve) to 1e-10, plus agreement with the
oracle's restated solve on the same matrix to <= 1e-13 relative.  The
non-symmetric band matrices pivot, so these runs cover the row-interchange
sweeps of the build kernels (forward and transposed)."""
import numpy as np
import pytest

import oracle as O
from conftest import rel_err
from matrices import SIZES, nonsym, offset_banded, pds, periodic_band

pytestmark = pytest.mark.gpu


def _load(setter, A):
    """Sets the nonzero entries of A through setter(i, j, a_ij)."""
    N = A.shape[0]
    for i in range(N):
        for j in range(N):
            if A[i, j] != 0.0:
                setter(i, j, A[i, j])


def solve_and_validate(prob, orc, A, cuda, nrhs_extra=300):
    """splines_linear_problem.cpp:82-142 on the device."""
    import torch

    N = A.shape[0]
    _load(prob.set_element, A)
    for i in range(N):
        for j in range(N):
            assert prob.get_element(i, j) == A[i, j]
    prob.setup_solver()
    _load(orc.set, A)
    orc.setup()
    rows = prob.required_number_of_rhs_rows()
    assert rows == orc.rhs_rows and prob.size() == N
    rng = np.random.default_rng(N)
    extra = rng.standard_normal((rows, nrhs_extra))
    for transpose in (False, True):
        b = np.zeros((rows, N))
        b[:N] = np.eye(N)
        b = np.concatenate([b, extra], axis=1)
        tb = torch.from_numpy(b).to(cuda)
        prob.solve(tb, transpose=transpose)
        got = tb.cpu().numpy()[:N]
        inv = got[:, :N]
        if transpose:
            assert np.abs(A @ inv.T - np.eye(N)).max() <= 1e-10
        else:
            assert np.abs(A @ inv - np.eye(N)).max() <= 1e-10
        ref = orc.solve(b.copy(), transpose=transpose)[:N]
        assert rel_err(got, ref) <= 1e-13, (transpose, rel_err(got, ref))
    # elements read back after setup are the assembled ones
    for i in range(N):
        assert prob.get_element(i, (i + 1) % N) == A[i, (i + 1) % N]


def test_dense(cuda):
    import ddc_amd.linear_problem as LP

    A = nonsym(10, 10)
    solve_and_validate(LP.make_new_dense(10), O.Problem("dense", 10), A, cuda)


def test_band_class(cuda):
    """SplinesLinearProblemBand(10, 3, 3) built directly (the maker would pick dense storage)."""
    import ddc_amd.linear_problem as LP

    A = nonsym(10, 3)
    p = LP.SplinesLinearProblem(LP.BAND_STORAGE, 10, 3, 3)
    assert p.main_kind() == "band"
    solve_and_validate(p, O.Problem("band_storage", 10, 3, 3), A, cuda)
    assert p.pivoted()


def test_pds_band(cuda):
    import ddc_amd.linear_problem as LP

    p = LP.make_new_band(10, 3, 3, True)
    assert p.main_kind() == "pds_band"
    solve_and_validate(p, O.Problem("band", 10, 3, 3, True), pds(10, 3), cuda)


def test_pds_tridiag(cuda):
    import ddc_amd.linear_problem as LP

    p = LP.make_new_band(10, 1, 1, True)
    assert p.main_kind() == "pds_tridiag"
    solve_and_validate(p, O.Problem("band", 10, 1, 1, True), pds(10, 1), cuda)


def test_2x2_blocks_dense_main(cuda):
    """SplinesLinearProblem2x2Blocks(10, Dense(7)) (splines_linear_problem.cpp:246-268)."""
    import ddc_amd.linear_problem as LP

    p = LP.make_new_dense_blocks(10, 3)
    assert p.blocked() == 1
    solve_and_validate(p, O.Problem("dense2x2", 10, bottom=3), nonsym(10, 10), cuda)


def test_3x3_blocks_dense_center(cuda):
    """SplinesLinearProblem3x3Blocks(10, 2, Dense(5)) (splines_linear_problem.cpp:270-293)."""
    import ddc_amd.linear_problem as LP

    p = LP.make_new_dense_blocks(10, 3, 2)
    assert p.blocked() == 2
    solve_and_validate(p, O.Problem("dense3x3", 10, bottom=3, top=2), nonsym(10, 10), cuda)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_nonsymmetric(N, k, cuda):
    import ddc_amd.linear_problem as LP

    p = LP.make_new_band(N, k, k, False)
    solve_and_validate(p, O.Problem("band", N, k, k, False), nonsym(N, k), cuda)
    assert p.pivoted()


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_pds(N, k, cuda):
    import ddc_amd.linear_problem as LP

    solve_and_validate(LP.make_new_band(N, k, k, True), O.Problem("band", N, k, k, True), pds(N, k), cuda)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_offset_banded(N, k, cuda):
    import ddc_amd.linear_problem as LP

    solve_and_validate(LP.make_new_band(N, 0, 2 * k, True), O.Problem("band", N, 0, 2 * k, True),
                       offset_banded(N, k), cuda)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_2x2(N, k, cuda):
    import ddc_amd.linear_problem as LP

    p = LP.make_new_block_matrix_with_band_main_block(N, k, k, False, 3)
    solve_and_validate(p, O.Problem("block", N, k, k, False, bottom=3), nonsym(N, k), cuda)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_3x3(N, k, cuda):
    import ddc_amd.linear_problem as LP

    p = LP.make_new_block_matrix_with_band_main_block(N, k, k, False, 3, 2)
    solve_and_validate(p, O.Problem("block", N, k, k, False, bottom=3, top=2), nonsym(N, k), cuda)


@pytest.mark.parametrize("N,k", SIZES)
def test_sizes_periodic_band(N, k, cuda):
    import ddc_amd.linear_problem as LP

    for s in range(-k + k // 2 + 1, k - k // 2):
        p = LP.make_new_periodic_band_matrix(N, k - s, k + s, False)
        solve_and_validate(p, O.Problem("periodic", N, k - s, k + s, False), periodic_band(N, k, s), cuda)
