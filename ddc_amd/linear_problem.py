"""Host-side mirror of DDC's SplinesLinearProblem and SplinesLinearProblemMaker
(src/ddc/kernels/splines/splines_linear_problem.hpp:26-127,
splines_linear_problem_maker.hpp:25-170) over the libddcx C ABI.

Elements are set on the host; ``setup_solver()`` factors in FP64 on the host
(the LAPACK algorithms the reference calls) and uploads the factors;
``solve(b, transpose)`` then runs in place on the device, one right-hand side
per column of ``b`` (the reference's LayoutRight MultiRHS view, shape
(required_number_of_rhs_rows(), nrhs)), on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import call, ddcx_problem_desc, ddcx_problem_info, lib

DENSE, BAND, BLOCK_BAND, PERIODIC_BAND, BLOCK_DENSE, BAND_STORAGE = 0, 1, 2, 3, 4, 5
MAIN_KINDS = ("dense", "band", "pds_band", "pds_tridiag")


class SplinesLinearProblem:
    """One matrix of the kinds SplinesLinearProblemMaker makes (use the maker
    functions below rather than this constructor)."""

    def __init__(self, kind, n, kl=0, ku=0, pds=False, bottom_size=0, top_size=0):
        desc = ddcx_problem_desc(int(kind), int(n), int(kl), int(ku), int(bool(pds)), int(bottom_size),
                                 int(top_size))
        h = C.c_void_p()
        call("ddcx_problem_create", C.byref(desc), C.byref(h))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                lib().ddcx_problem_destroy(self._h)
            except Exception:
                pass
            self._h = None

    def _info(self):
        info = ddcx_problem_info()
        call("ddcx_problem_get_info", self._h, C.byref(info))
        return info

    def size(self):
        return self._info().size

    def required_number_of_rhs_rows(self):
        return self._info().rhs_rows

    def main_kind(self):
        """Storage of the main block (the whole matrix when not blocked)."""
        return MAIN_KINDS[self._info().main_kind]

    def blocked(self):
        """0 plain, 1 2x2 blocks, 2 3x3 blocks."""
        return self._info().blocked

    def pivoted(self):
        return bool(self._info().pivoted)

    def set_element(self, i, j, aij):
        call("ddcx_problem_set_element", self._h, int(i), int(j), float(aij))

    def get_element(self, i, j):
        v = C.c_double()
        call("ddcx_problem_get_element", self._h, int(i), int(j), C.byref(v))
        return v.value

    def setup_solver(self):
        call("ddcx_problem_setup_solver", self._h)

    def solve(self, b, transpose=False):
        """In place: b is a float64 device tensor of shape
        (required_number_of_rhs_rows(), nrhs) whose rows have unit-stride
        columns (b.stride(1) == 1)."""
        if b.dtype != torch.float64 or not b.is_cuda:
            raise TypeError("b must be a float64 device tensor")
        if b.dim() != 2 or b.shape[0] != self.required_number_of_rhs_rows() or (b.shape[1] > 1 and b.stride(1) != 1):
            raise ValueError("b must be (required_number_of_rhs_rows(), nrhs) with unit-stride columns")
        call("ddcx_problem_solve", self._h, C.c_void_p(b.data_ptr()), b.stride(0) if b.shape[0] > 1 else b.shape[1],
             b.shape[1], int(bool(transpose)), C.c_void_p(torch.cuda.current_stream().cuda_stream))
        return b


def make_new_dense(n):
    """SplinesLinearProblemDense(n)."""
    return SplinesLinearProblem(DENSE, n)


def make_new_band(n, kl, ku, pds):
    """SplinesLinearProblemMaker::make_new_band (maker.hpp:56-77)."""
    return SplinesLinearProblem(BAND, n, kl, ku, pds)


def make_new_block_matrix_with_band_main_block(n, kl, ku, pds, bottom_right_size, top_left_size=0):
    """SplinesLinearProblemMaker::make_new_block_matrix_with_band_main_block (maker.hpp:93-112)."""
    return SplinesLinearProblem(BLOCK_BAND, n, kl, ku, pds, bottom_right_size, top_left_size)


def make_new_periodic_band_matrix(n, kl, ku, pds):
    """SplinesLinearProblemMaker::make_new_periodic_band_matrix (maker.hpp:131-146)."""
    return SplinesLinearProblem(PERIODIC_BAND, n, kl, ku, pds)


def make_new_dense_blocks(n, bottom_size, top_size=0):
    """SplinesLinearProblem2x2Blocks(n, Dense(n - bottom_size)) for top_size 0,
    else SplinesLinearProblem3x3Blocks(n, top_size, Dense(n - top_size -
    bottom_size)) (splines_linear_problem_2x2_blocks.hpp, _3x3_blocks.hpp)."""
    return SplinesLinearProblem(BLOCK_DENSE, n, 0, 0, False, bottom_size, top_size)
