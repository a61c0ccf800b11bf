"""Shared spline configurations and synthetic data for the tests.

Every case is built twice: by the oracle (``oracle.Spline``, the CPU
restatement) and by the product (``ddc_amd``), from the same breakpoints.
"""
from __future__ import annotations

import numpy as np

import oracle as O

P, H, HH, G = O.PERIODIC, O.HERMITE, O.HOMOGENEOUS_HERMITE, O.GREVILLE
BC_NAMES = {P: "P", H: "H", HH: "HH", G: "G"}


def breaks_for(kind, ncells, seed=0, ratio=1.05):
    if kind == "uniform":
        return np.linspace(0.0, 1.0, ncells + 1)
    if kind == "random":
        rng = np.random.default_rng(seed)
        inner = np.sort(rng.uniform(0.0, 1.0, ncells - 1))
        return np.concatenate([[0.0], inner, [1.0]])
    if kind == "jitter":
        rng = np.random.default_rng(seed)
        x = np.linspace(0.0, 1.0, ncells + 1)
        x[1:-1] += rng.uniform(-0.3, 0.3, ncells - 1) / ncells
        return x
    if kind == "geometric":
        k = np.arange(ncells + 1, dtype=np.float64)
        return (ratio ** k - 1.0) / (ratio ** ncells - 1.0)
    raise ValueError(kind)


class Case:
    def __init__(self, degree, uniform, bcl, bcu=None, ncells=10, mesh="random", seed=0, ratio=1.05):
        self.degree, self.uniform, self.bcl = degree, uniform, bcl
        self.bcu = bcl if bcu is None else bcu
        self.ncells, self.mesh, self.seed, self.ratio = ncells, mesh, seed, ratio
        self.periodic = bcl == P
        self.breaks = None if uniform else breaks_for(mesh, ncells, seed, ratio)

    @property
    def id(self):
        kind = "U" if self.uniform else f"NU{self.mesh[0]}"
        return f"d{self.degree}-{kind}-{BC_NAMES[self.bcl]}{BC_NAMES[self.bcu]}-n{self.ncells}"

    def oracle(self):
        if self.uniform:
            return O.Spline(self.degree, True, self.bcl, self.bcu, ncells=self.ncells, rmin=0.0, rmax=1.0)
        return O.Spline(self.degree, False, self.bcl, self.bcu, breaks=self.breaks)

    def product(self, precision=0, refine_steps=0):
        import ddc_amd as D

        if self.uniform:
            bs = D.UniformBSplines(self.degree, 0.0, 1.0, self.ncells, periodic=self.periodic)
        else:
            bs = D.NonUniformBSplines(self.degree, self.breaks, periodic=self.periodic)
        builder = D.SplineBuilder(bs, D.SplineBuilderClosure(self.bcl), D.SplineBuilderClosure(self.bcu),
                                  precision=D.Precision(precision), refine_steps=refine_steps)
        return bs, builder


def values(npts, nbatch, seed, x=None):
    """f(x, y) = sin(2 pi x k_y) + 0.1 N(0, 1), k_y in {1..8} (one line per y)."""
    rng = np.random.default_rng(seed)
    if x is None:
        x = np.linspace(0, 1, npts, endpoint=False)
    k = rng.integers(1, 9, size=nbatch)
    f = np.sin(2 * np.pi * x[:, None] * k[None, :]) + 0.1 * rng.standard_normal((npts, nbatch))
    return f  # (npts, nbatch): batch-contiguous


def derivs(nbv, nbatch, seed):
    rng = np.random.default_rng(seed + 1000)
    return rng.standard_normal((max(nbv, 0), nbatch))


# Cases covering every matrix shape the maker produces for degrees 1..5 and
# every closure kind (the reference tests degrees 1..5 in CI).
ALL_CASES = []
for d in (1, 2, 3, 4, 5):
    ALL_CASES.append(Case(d, True, P, ncells=16))
    ALL_CASES.append(Case(d, False, P, ncells=16, mesh="random"))
    ALL_CASES.append(Case(d, True, G, ncells=16))
    ALL_CASES.append(Case(d, False, G, ncells=16, mesh="random"))
    if d >= 2:
        ALL_CASES.append(Case(d, True, H, ncells=16))
        ALL_CASES.append(Case(d, False, H, ncells=16, mesh="random"))
        ALL_CASES.append(Case(d, False, HH, ncells=16, mesh="random"))
        ALL_CASES.append(Case(d, False, H, G, ncells=16, mesh="random"))
        ALL_CASES.append(Case(d, True, G, H, ncells=16))

# Larger cases that take the register-resident tiled build kernel (main block
# of 256..2048 rows): every sweep shape (tridiagonal, PDS band, unpivoted band)
# and both segment lengths.
TILED_CASES = [
    Case(3, True, P, ncells=1024),
    Case(5, True, P, ncells=1024),
    Case(2, True, P, ncells=600),
    Case(4, True, G, ncells=700),
    Case(3, True, H, ncells=300),
    Case(3, False, G, ncells=1024, mesh="random"),
    Case(2, False, G, ncells=500, mesh="random"),
    Case(4, False, G, ncells=400, mesh="random"),
    Case(5, False, G, ncells=1024, mesh="random", seed=3),
    Case(3, False, P, ncells=800, mesh="random"),
    Case(3, False, H, ncells=2048, mesh="geometric", ratio=1.002),
    Case(5, False, H, ncells=2048, mesh="geometric", ratio=1.002),
    Case(3, False, H, ncells=1500, mesh="random"),
    Case(5, False, HH, ncells=1200, mesh="uniform"),
    Case(3, True, P, ncells=1800),
]

# Short lines: main blocks of 64 to 255 rows on the short tilings (16 x 8 and
# 16 x 9 up to 144 rows, 32 x 8 above; FP64 build only; the mixed-precision
# and fused kernels start at 256 rows).
SHORT_TILED_CASES = [
    Case(3, True, P, ncells=128),
    Case(4, False, G, ncells=140, mesh="random"),
    Case(3, False, G, ncells=128, mesh="random"),
    Case(3, True, G, ncells=128),
    Case(5, True, P, ncells=200),
    Case(2, False, G, ncells=100, mesh="random"),
    Case(4, True, G, ncells=90),
    Case(3, False, H, ncells=180, mesh="random"),
    Case(5, False, HH, ncells=250, mesh="uniform"),
]

# Degree-6 non-uniform bases on random meshes, whose band factorisation
# (dgbtrf) interchanges rows: the pivoted sweeps of the generic kernels.
PIVOTED_CASES = [
    Case(6, False, G, ncells=128, mesh="random", seed=3),
    Case(6, False, H, ncells=128, mesh="random", seed=3),
    Case(6, False, H, ncells=300, mesh="random", seed=0),
    Case(6, False, G, ncells=1024, mesh="random", seed=1),
]

# The precision-form workload: Hermite closures on a geometrically stretched
# mesh (ratio 1.05 over 2048 cells, smallest cell ~2e-45 of the domain).
STRETCHED_CASES = [
    Case(3, False, H, ncells=2048, mesh="geometric", ratio=1.05),
    Case(5, False, H, ncells=2048, mesh="geometric", ratio=1.05),
]


def box_filtered_normal(npts, nbatch, seed, width=9):
    """N(0, 1) values smoothed along the interpolation axis by a width-point
    box filter with clamped ends (one line per column)."""
    rng = np.random.default_rng(seed)
    raw = rng.standard_normal((npts, nbatch))
    h = width // 2
    idx = np.clip(np.arange(npts)[:, None] + np.arange(-h, h + 1)[None, :], 0, npts - 1)
    return raw[idx].mean(axis=1)
