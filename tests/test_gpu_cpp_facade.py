"""The C++ facade (include/ddcx.hpp) on device arrays: tests/cpp/facade_device,
a compiled C++ program, builds and evaluates through SplineBuilder /
SplineEvaluator (both layouts, periodic and constant rules with eval_pos)
and SplineBuilder2D / SplineEvaluator2D (values, deriv_dim_1, adjoint, and
the builder with Hermite closures and its adjoint) and SplineBuilder3D with
Hermite closures, and
writes its inputs and outputs; they are checked here against the oracle at
the parity bars (1e-13; 1e-12 for accumulated coefficient gradients)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "facade_device")


@pytest.fixture(scope="module")
def outputs(cuda, tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (make -C tests/cpp)")
    d = tmp_path_factory.mktemp("facade")
    r = subprocess.run([EXE, str(d)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    return lambda name: np.fromfile(os.path.join(d, name + ".bin"), dtype=np.float64)


@pytest.mark.parametrize("tag", ["1d_bc_", "1d_ic_"])
def test_facade_1d(outputs, tag):
    nc, B, P = 64, 33, 100
    orc = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=nc)
    m, s = orc.npts, orc.size
    bc = tag == "1d_bc_"
    lay = "batch_contiguous" if bc else "interp_contiguous"
    shp = (lambda r: (r, B)) if bc else (lambda r: (B, r))
    vals, coords = outputs(tag + "vals").reshape(shp(m)), outputs(tag + "coords").reshape(shp(P))
    coef = orc.build(vals, layout=lay)
    assert rel_err(outputs(tag + "coef").reshape(shp(s)), coef) <= 1e-13
    assert rel_err(outputs(tag + "eval").reshape(shp(P)), orc.eval(coords, coef, layout=lay)) <= 1e-13
    d1 = orc.eval(coords, coef, deriv=1, layout=lay)
    assert rel_err(outputs(tag + "deriv1").reshape(shp(P)), d1) <= 1e-13
    assert rel_err(outputs(tag + "integrals"), orc.integrate(coef, layout=lay)) <= 1e-13
    if bc:  # the device-callable single-point form in the program's own kernel: the batched values bit for bit
        assert np.array_equal(outputs(tag + "point_eval"), outputs(tag + "eval"))
        assert np.array_equal(outputs(tag + "point_deriv1"), outputs(tag + "deriv1"))


def test_facade_constant_rules(outputs):
    B, P = 5, 200
    orc = O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=outputs("rules_breaks"))
    vals = outputs("rules_vals").reshape(orc.npts, B)
    coords = outputs("rules_coords").reshape(P, B)
    coef = orc.build(vals)
    ref = orc.eval(coords, coef, extrap=("constant", "constant"), eval_pos=(0.3, 0.8))
    assert rel_err(outputs("rules_eval").reshape(P, B), ref) <= 1e-13
    assert np.array_equal(outputs("rules_point_eval"), outputs("rules_eval"))
    inside = (coords >= orc.rmin) & (coords <= orc.rmax)  # deriv() applies no rule: compare inside the domain
    d2 = orc.eval(coords, coef, deriv=2)
    assert rel_err(outputs("rules_point_deriv2").reshape(P, B)[inside], d2[inside]) <= 1e-12


def test_facade_2d(outputs):
    NB, P = 2, 500
    o1 = O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=outputs("2d_breaks"))
    o2 = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=16)
    s1, s2 = o1.size, o2.size
    vals = outputs("2d_vals").reshape(NB, o1.npts, o2.npts)
    coords = outputs("2d_coords").reshape(NB, P, 2)
    w = outputs("2d_w").reshape(NB, P)
    coef = O.spline2d_build(o1, o2, vals)
    assert rel_err(outputs("2d_coef").reshape(NB, s1, s2), coef) <= 1e-13
    assert rel_err(outputs("2d_eval").reshape(NB, P), O.spline2d_eval(o1, o2, coords, coef)) <= 1e-13
    assert rel_err(outputs("2d_deriv1").reshape(NB, P), O.spline2d_eval(o1, o2, coords, coef, 1, 0)) <= 1e-13
    assert rel_err(outputs("2d_eval_packed").reshape(NB, P), O.spline2d_eval(o1, o2, coords, coef)) <= 1e-13
    # adjoint: gcoef against the Jacobian of the oracle's evaluation, gx
    # against w times the partials at the wrapped coordinates
    gc = outputs("2d_gcoef").reshape(NB, s1, s2)
    units = np.eye(s1 * s2).reshape(s1 * s2, s1, s2)
    for l in range(NB):
        J = O.spline2d_eval(o1, o2, np.broadcast_to(coords[l], (s1 * s2, P, 2)).copy(), units)
        assert rel_err(gc[l], (J @ w[l]).reshape(s1, s2)) <= 1e-12
    wrapped = coords.copy()
    wrapped[..., 1] -= np.floor(wrapped[..., 1])
    gx = outputs("2d_gx").reshape(NB, P, 2)
    assert rel_err(gx[..., 0], w * O.spline2d_eval(o1, o2, wrapped, coef, 1, 0)) <= 1e-13
    assert rel_err(gx[..., 1], w * O.spline2d_eval(o1, o2, wrapped, coef, 0, 1)) <= 1e-13


def test_facade_2d_hermite(outputs):
    """SplineBuilder2D with Hermite closures in both dimensions through the
    facade: coefficients against the oracle (1e-13), and build_vjp's gradient
    of every input block against the transposed Jacobian of the oracle's
    build for that block (1e-13 for the values, 1e-12 for the derivative
    blocks)."""
    NB, nc2 = 2, 10
    H = O.HERMITE
    o1 = O.Spline(3, False, H, H, breaks=outputs("2dh_breaks"))
    o2 = O.Spline(5, True, H, H, ncells=nc2, rmin=0.0, rmax=1.0)
    n1, n2, s1, s2 = o1.npts, o2.npts, o1.size, o2.size
    e1, e2 = max(o1.nbv_l, o1.nbv_r), max(o2.nbv_l, o2.nbv_r)
    shapes = {"vals": (n1, n2), "dmin1": (e1, n2), "dmax1": (e1, n2), "dmin2": (n1, e2), "dmax2": (n1, e2),
              "mm": (e1, e2), "Mm": (e1, e2), "mM": (e1, e2), "MM": (e1, e2)}
    order = list(shapes)
    x = {k: outputs("2dh_" + k).reshape((NB,) + sh) for k, sh in shapes.items()}
    ref = O.spline2d_build(o1, o2, *[x[k] for k in order])
    assert rel_err(outputs("2dh_coef").reshape(NB, s1, s2), ref) <= 1e-13
    gs = outputs("2dh_gs").reshape(NB, s1 * s2)
    for k, sh in shapes.items():
        m = int(np.prod(sh))
        args = {j: np.zeros((m,) + s) for j, s in shapes.items()}
        args[k] = np.eye(m).reshape((m,) + sh)
        J = O.spline2d_build(o1, o2, *[args[j] for j in order]).reshape(m, -1)
        got = outputs("2dh_g" + k).reshape(NB, m)
        assert rel_err(got, gs @ J.T) <= (1e-13 if k == "vals" else 1e-12), k


def test_facade_3d_hermite(outputs, cuda):
    """SplineBuilder3D with Hermite closures in all three dimensions through
    the facade (every derivative, edge and corner input): the coefficients
    equal the Python builder's on the same inputs bit for bit (the same 1D
    builds on the same C ABI) and the oracle's 3D restatement to 1e-13."""
    import torch
    import ddc_amd as D

    NB = 2
    H = O.HERMITE
    ss = [O.Spline(3, False, H, H, breaks=outputs("3dh_breaks")),
          O.Spline(5, True, H, H, ncells=8, rmin=0.0, rmax=1.0),
          O.Spline(4, True, H, H, ncells=7, rmin=0.0, rmax=1.0)]
    n = [s.npts for s in ss]
    e = [max(s.nbv_l, s.nbv_r) for s in ss]
    shape = {"vals": n}
    for k in range(3):
        for sd in ("min", "max"):
            sh = list(n)
            sh[k] = e[k]
            shape[f"derivs_{sd}{k + 1}"] = sh
    for (a, b) in ((0, 1), (1, 2), (0, 2)):
        for sa in ("min", "max"):
            for sb in ("min", "max"):
                sh = list(n)
                sh[a], sh[b] = e[a], e[b]
                shape[f"mixed_derivs_{sa}{a + 1}_{sb}{b + 1}"] = sh
    for s1 in ("min", "max"):
        for s2 in ("min", "max"):
            for s3 in ("min", "max"):
                shape[f"mixed_derivs_{s1}1_{s2}2_{s3}3"] = e
    x = {k: outputs("3dh_" + k).reshape([NB] + list(sh)) for k, sh in shape.items()}
    got = outputs("3dh_coef").reshape([NB] + [s.size for s in ss])
    derivs = {k: v for k, v in x.items() if k != "vals"}
    ref = O.spline3d_build(*ss, x["vals"], derivs)
    assert rel_err(got, ref) <= 1e-13
    G = D.SplineBuilderClosure.HERMITE
    bl = [D.SplineBuilder(D.NonUniformBSplines(3, outputs("3dh_breaks")), G, G),
          D.SplineBuilder(D.UniformBSplines(5, 0.0, 1.0, 8), G, G),
          D.SplineBuilder(D.UniformBSplines(4, 0.0, 1.0, 7), G, G)]
    coef = torch.full(got.shape, np.nan, dtype=torch.float64, device=cuda)
    t = {k: torch.from_numpy(v).to(cuda) for k, v in derivs.items()}
    b3 = D.SplineBuilder3D(*bl)
    b3(coef, torch.from_numpy(x["vals"]).to(cuda), **t)
    assert np.array_equal(coef.cpu().numpy(), got)
    # build_vjp: every gradient equals the Python adjoint's bit for bit, and
    # satisfies <gs, build(input)> = <input, gradient> block by block with the
    # oracle's build
    gs = outputs("3dh_gs").reshape(got.shape)
    gv = torch.full([NB] + n, np.nan, dtype=torch.float64, device=cuda)
    gd = {"g" + k: torch.full(v.shape, np.nan, dtype=torch.float64, device=cuda) for k, v in derivs.items()}
    b3.build_vjp(torch.from_numpy(gs).to(cuda), gv, **gd)
    blocks = {"vals": gv.cpu().numpy()}
    blocks.update({k: gd["g" + k].cpu().numpy() for k in derivs})
    for k, py in blocks.items():
        cpp = outputs("3dh_g" + k).reshape(py.shape)
        assert np.array_equal(cpp, py), k
        v0 = x["vals"] if k == "vals" else np.zeros_like(x["vals"])
        dk = {j: (x[j] if j == k else np.zeros_like(u)) for j, u in derivs.items()}
        c = O.spline3d_build(*ss, v0, dk)
        lhs, rhs = float((gs * c).sum()), float((x[k] * cpp).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(np.abs(gs * c).sum(), 1e-300), k


def test_facade_3d(outputs, cuda):
    """SplineBuilder3D without derivative inputs through the facade (its two
    workspaces) and its adjoint: coefficients against the oracle (1e-13), and
    both coefficients and value gradients bit for bit equal to the Python
    builder's and its adjoint's on the same inputs."""
    import torch
    import ddc_amd as D

    NB = 2
    br = outputs("3d_breaks")
    ss = [O.Spline(3, False, O.GREVILLE, O.GREVILLE, breaks=br),
          O.Spline(5, True, O.PERIODIC, O.PERIODIC, ncells=8, rmin=0.0, rmax=1.0),
          O.Spline(4, True, O.GREVILLE, O.GREVILLE, ncells=7, rmin=0.0, rmax=1.0)]
    vals = outputs("3d_vals").reshape([NB] + [s.npts for s in ss])
    got = outputs("3d_coef").reshape([NB] + [s.size for s in ss])
    assert rel_err(got, O.splinend_build(ss, vals)) <= 1e-13
    C = D.SplineBuilderClosure
    b3 = D.SplineBuilder3D(D.SplineBuilder(D.NonUniformBSplines(3, br), C.GREVILLE, C.GREVILLE),
                           D.SplineBuilder(D.UniformBSplines(5, 0.0, 1.0, 8, periodic=True)),
                           D.SplineBuilder(D.UniformBSplines(4, 0.0, 1.0, 7), C.GREVILLE, C.GREVILLE))
    coef = torch.full(got.shape, np.nan, dtype=torch.float64, device=cuda)
    b3(coef, torch.from_numpy(vals).to(cuda))
    assert np.array_equal(coef.cpu().numpy(), got)
    gv = torch.full(vals.shape, np.nan, dtype=torch.float64, device=cuda)
    b3.build_vjp(torch.from_numpy(outputs("3d_gs").reshape(got.shape)).to(cuda), gv)
    assert np.array_equal(gv.cpu().numpy(), outputs("3d_gvals").reshape(vals.shape))
