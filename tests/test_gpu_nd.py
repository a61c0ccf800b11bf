"""GPU parity of the N-dimensional path against the oracle's restatement on
the same inputs:
- SplineBuilder3D (spline_builder_3d.hpp) and its adjoint;
- SplineEvaluatorND / SplineEvaluator3D values and derivatives
  (spline_evaluator_nd.hpp:684-826), including wrap and Null rules, for 1 to
  4 dimensions and strided coefficient views;
- integrate (spline_evaluator_nd.hpp:597-632).

Bar: <= 1e-13 relative (max norm); the adjoint identity <A u, v> = <u, A^T v>
to 1e-12.
"""
import numpy as np
import pytest

from cases import G, P, breaks_for
from conftest import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _bases(spec):
    """spec: list of (degree, kind, ncells) with kind in {"up", "ug", "nug"}:
    uniform periodic, uniform Greville, non-uniform Greville."""
    import ddc_amd as D
    import oracle as O

    out = []
    for i, (d, kind, n) in enumerate(spec):
        if kind == "up":
            o = O.Spline(d, True, P, P, ncells=n, rmin=0.0, rmax=1.0)
            bs = D.UniformBSplines(d, 0.0, 1.0, n, periodic=True)
            b = D.SplineBuilder(bs)
            ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
        elif kind == "ug":
            o = O.Spline(d, True, G, G, ncells=n, rmin=-1.0, rmax=2.0)
            bs = D.UniformBSplines(d, -1.0, 2.0, n)
            b = D.SplineBuilder(bs, D.SplineBuilderClosure.GREVILLE, D.SplineBuilderClosure.GREVILLE)
            ev = D.SplineEvaluator(bs, D.NullExtrapolationRule(), D.NullExtrapolationRule())
        else:
            br = breaks_for("random", n, seed=10 + i)
            o = O.Spline(d, False, G, G, breaks=br)
            bs = D.NonUniformBSplines(d, br)
            b = D.SplineBuilder(bs, D.SplineBuilderClosure.GREVILLE, D.SplineBuilderClosure.GREVILLE)
            ev = D.SplineEvaluator(bs, D.NullExtrapolationRule(), D.NullExtrapolationRule())
        out.append((o, bs, b, ev))
    return out


SPECS_3D = [[(3, "nug", 12), (3, "up", 16), (3, "ug", 10)],
            [(2, "up", 9), (5, "nug", 14), (4, "ug", 7)],
            [(3, "up", 40), (3, "up", 32), (3, "nug", 36)]]


def _coef(ss, nb, rng):
    c = rng.standard_normal((nb,) + tuple(o.size for o in ss))
    for k, o in enumerate(ss):  # periodic dimensions repeat their first `degree` coefficients
        if o.nbasis < o.size:
            src = np.take(c, np.arange(o.degree), axis=k + 1)
            idx = [slice(None)] * c.ndim
            idx[k + 1] = slice(o.nbasis, o.size)
            c[tuple(idx)] = src
    return c


@pytest.mark.parametrize("spec", SPECS_3D, ids=["d3mix", "d2d5d4", "d3big"])
def test_build3d_matches_oracle(spec, cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(1)
    nb = 3
    vals = rng.standard_normal((nb,) + tuple(o.npts for o in ss))
    ref = O.splinend_build(ss, vals)
    coef = torch.full(ref.shape, np.nan, dtype=torch.float64, device=cuda)
    D.SplineBuilder3D(*[x[2] for x in bases])(coef, torch.from_numpy(vals).to(cuda))
    got = coef.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= TOL


def test_build3d_vjp_adjoint_identity(cuda):
    import torch
    import ddc_amd as D

    bases = _bases(SPECS_3D[0])
    b3 = D.SplineBuilder3D(*[x[2] for x in bases])
    rng = np.random.default_rng(2)
    nb = 2
    u = torch.from_numpy(rng.standard_normal((nb,) + tuple(x[0].npts for x in bases))).to(cuda)
    v = torch.from_numpy(rng.standard_normal((nb,) + tuple(x[0].size for x in bases))).to(cuda)
    Au = torch.empty_like(v)
    b3(Au, u)
    Atv = torch.empty_like(u)
    b3.build_vjp(v, Atv)
    lhs = float((Au * v).sum())
    rhs = float((u * Atv).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)


@pytest.mark.parametrize("spec", SPECS_3D[:2], ids=["d3mix", "d2d5d4"])
@pytest.mark.parametrize("orders", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (2, 0, 1), (0, 2, 2)])
def test_eval3d_matches_oracle(spec, orders, cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(3)
    nb, npts = 2, 400
    coef = _coef(ss, nb, rng)
    ext = 0.2 if orders == (0, 0, 0) else 0.0  # wrap and Null rules on the value form
    coords = np.stack([rng.uniform(o.rmin - ext * (o.rmax - o.rmin), o.rmax + ext * (o.rmax - o.rmin), (nb, npts))
                       for o in ss], axis=-1)
    ref = O.splinend_eval(ss, coords, coef, orders=list(orders))
    ev = D.SplineEvaluator3D(*[x[3] for x in bases])
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev.deriv(orders, out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda))
    assert rel_err(out.cpu().numpy(), ref) <= TOL * 10 ** sum(orders)


@pytest.mark.parametrize("spec,orders", [
    ([(2, "up", 9), (5, "nug", 14), (4, "ug", 7)], (2, 5, 4)),
    ([(2, "up", 9), (5, "nug", 14), (4, "ug", 7)], (1, 4, 3)),
    ([(5, "up", 12), (5, "nug", 14), (5, "ug", 10)], (5, 4, 5)),
    ([(4, "ug", 8), (4, "nug", 9)], (4, 4)),
], ids=["mixed_top", "mixed_high", "d5_fixed", "d4_2d"])
def test_evalnd_orders_up_to_degree(spec, orders, cuda):
    """deriv() with every dimension's order up to its degree
    (spline_evaluator_nd.hpp:735 asserts order <= degree), on the mixed-degree
    kernel and the fixed-degree one, against the oracle."""
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(21)
    nb, npts = 2, 300
    coef = _coef(ss, nb, rng)
    coords = np.stack([rng.uniform(o.rmin, o.rmax, (nb, npts)) for o in ss], axis=-1)
    ref = O.splinend_eval(ss, coords, coef, orders=list(orders))
    ev = D.SplineEvaluatorND(*[x[3] for x in bases])
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev.deriv(orders, out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda))
    err = rel_err(out.cpu().numpy(), ref)
    print(f"orders {orders}: rel err {err:.2e}")
    assert err <= 1e-12


@pytest.mark.parametrize("spec", [[(3, "nug", 20)], [(2, "up", 9), (5, "nug", 14)],
                                  [(1, "ug", 5), (3, "up", 8), (2, "nug", 6), (4, "up", 7)],
                                  [(4, "nug", 11), (4, "up", 9)], [(5, "up", 7), (5, "nug", 12)],
                                  [(2, "ug", 5), (2, "up", 6), (2, "nug", 7), (2, "up", 4)],
                                  [(3, "up", 4), (3, "nug", 5), (3, "ug", 6), (3, "up", 5)]],
                         ids=["1d", "2d", "4d", "2d_same_d4", "2d_same_d5", "4d_same_d2", "4d_same_d3"])
def test_evalnd_other_dimensions(spec, cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(4)
    nb, npts = 3, 300
    coef = _coef(ss, nb, rng)
    coords = np.stack([rng.uniform(o.rmin - 0.1, o.rmax + 0.1, (nb, npts)) for o in ss], axis=-1)
    ev = D.SplineEvaluatorND(*[x[3] for x in bases])
    out = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev(out, torch.from_numpy(coords).to(cuda), torch.from_numpy(coef).to(cuda))
    assert rel_err(out.cpu().numpy(), O.splinend_eval(ss, coords, coef)) <= TOL
    orders = [1] * len(ss)
    cin = np.stack([np.clip(coords[..., k], o.rmin, o.rmax) for k, o in enumerate(ss)], axis=-1)
    ev.deriv(orders, out, torch.from_numpy(cin).to(cuda), torch.from_numpy(coef).to(cuda))
    assert rel_err(out.cpu().numpy(), O.splinend_eval(ss, cin, coef, orders=orders)) <= 1e-12


def test_evalnd_strided_coefficients(cuda):
    """Coefficients as a permuted (non-contiguous) view: the strides are passed as is."""
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(SPECS_3D[0])
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(5)
    nb, npts = 2, 200
    coef = _coef(ss, nb, rng)
    coords = np.stack([rng.uniform(o.rmin, o.rmax, (nb, npts)) for o in ss], axis=-1)
    ref = O.splinend_eval(ss, coords, coef)
    # store as (size3, nb, size1, size2) and view it back as (nb, size1, size2, size3)
    store = torch.from_numpy(np.ascontiguousarray(np.transpose(coef, (3, 0, 1, 2)))).to(cuda)
    view = store.permute(1, 2, 3, 0)
    assert not view.is_contiguous()
    out = torch.empty((nb, npts), dtype=torch.float64, device=cuda)
    D.SplineEvaluator3D(*[x[3] for x in bases])(out, torch.from_numpy(coords).to(cuda), view)
    assert rel_err(out.cpu().numpy(), ref) <= TOL


@pytest.mark.parametrize("spec", SPECS_3D + [[(3, "nug", 20)], [(2, "up", 9), (5, "nug", 14)]],
                         ids=["d3mix", "d2d5d4", "d3big", "1d", "2d"])
def test_integratend_matches_oracle(spec, cuda):
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(6)
    nb = 5
    coef = _coef(ss, nb, rng)
    ref = O.splinend_integrate(ss, coef)
    out = torch.full((nb,), np.nan, dtype=torch.float64, device=cuda)
    D.SplineEvaluatorND(*[x[3] for x in bases]).integrate(out, torch.from_numpy(coef).to(cuda))
    # sums of random-sign terms: compare against the sum of magnitudes
    scale = O.splinend_integrate(ss, np.abs(coef))
    assert np.all(np.abs(out.cpu().numpy() - ref) <= TOL * scale)


@pytest.mark.parametrize("spec,nb", [([(3, "nug", 64), (3, "up", 64), (3, "ug", 64)], 1),
                                     ([(5, "up", 40), (2, "ug", 30)], 3), (SPECS_3D[0], 700)],
                         ids=["64cubed_nb1", "2d_nb3", "d3mix_nb700"])
def test_integratend_split_matches_single_pass(spec, nb, cuda):
    """The split reduction (workspace entry point, several workgroups per
    batch entry) against the one-workgroup-per-entry entry point and the
    oracle, on shapes where the split is 1, a few, and many workgroups."""
    import ctypes as C
    import torch
    import ddc_amd as D
    import oracle as O
    from ddc_amd._lib import call

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(16)
    coef = _coef(ss, nb, rng)
    ev = D.SplineEvaluatorND(*[x[3] for x in bases])
    tc = torch.from_numpy(coef).to(cuda)
    nd = len(spec)
    st = (C.c_int64 * nd)(*[tc.stride(k + 1) for k in range(nd)])
    need = C.c_int64()
    call("ddcx_evaluatornd_integrate_workspace", ev._hs, nd, nb, C.byref(need))
    rows = int(np.prod([o.size for o in ss[:-1]]))
    if nb < 8 and rows >= 128:
        assert need.value > 0  # the split is in use
    one = torch.full((nb,), np.nan, dtype=torch.float64, device=cuda)
    call("ddcx_evaluatornd_integrate", ev._hs, nd, C.c_void_p(one.data_ptr()), 1, C.c_void_p(tc.data_ptr()),
         tc.stride(0), st, nb, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    split = torch.full((nb,), np.nan, dtype=torch.float64, device=cuda)
    ev.integrate(split, tc)
    ref = O.splinend_integrate(ss, coef)
    scale = O.splinend_integrate(ss, np.abs(coef))
    assert np.all(np.abs(split.cpu().numpy() - ref) <= TOL * scale)
    assert np.all(np.abs(split.cpu().numpy() - one.cpu().numpy()) <= 1e-14 * scale)
    # a workspace one byte short is refused
    if need.value > 0:
        ws = torch.empty(need.value, dtype=torch.uint8, device=cuda)
        with pytest.raises(D.DDCXError, match="workspace"):
            call("ddcx_evaluatornd_integrate_ws", ev._hs, nd, C.c_void_p(split.data_ptr()), 1,
                 C.c_void_p(tc.data_ptr()), tc.stride(0), st, nb, C.c_void_p(ws.data_ptr()), need.value - 1,
                 C.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_build_eval3d_reproduces_interpolation_points(cuda):
    """Evaluating the 3D spline at its interpolation grid returns the values."""
    import torch
    import ddc_amd as D

    bases = _bases(SPECS_3D[0])
    g = [x[0].points() for x in bases]
    rng = np.random.default_rng(7)
    nb = 2
    vals = rng.standard_normal((nb,) + tuple(len(x) for x in g))
    coef = torch.empty((nb,) + tuple(x[1].size() for x in bases), dtype=torch.float64, device=cuda)
    D.SplineBuilder3D(*[x[2] for x in bases])(coef, torch.from_numpy(vals).to(cuda))
    mesh = np.stack(np.meshgrid(*g, indexing="ij"), axis=-1).reshape(-1, 3)
    coords = np.broadcast_to(mesh, (nb,) + mesh.shape).copy()
    out = torch.empty((nb, mesh.shape[0]), dtype=torch.float64, device=cuda)
    ev = D.SplineEvaluator3D(*[x[3] for x in bases])
    ev(out, torch.from_numpy(coords).to(cuda), coef)
    assert np.abs(out.cpu().numpy() - vals.reshape(nb, -1)).max() <= 1e-12 * np.abs(vals).max()
    # the named partial-derivative helpers are deriv() with unit orders
    d1 = torch.empty_like(out)
    ev.deriv_dim_1(d1, torch.from_numpy(coords).to(cuda), coef)
    d2 = torch.empty_like(out)
    ev.deriv((1, 0, 0), d2, torch.from_numpy(coords).to(cuda), coef)
    assert torch.equal(d1, d2)


@pytest.mark.parametrize("closures", [("H", "H", "H"), ("H", "G", "P"), ("G", "H", "H"), ("P", "H", "G")])
def test_build3d_hermite_matches_oracle(closures, cuda):
    """SplineBuilder3D with Hermite closures (spline_builder_3d.hpp:889-1090):
    the derivative, edge and corner cross-derivative inputs as the reference
    takes them, random data, against the oracle's restatement (pinned by
    tensor-polynomial reproduction in test_oracle_nd.py)."""
    import torch
    import ddc_amd as D
    import oracle as O
    from test_oracle_nd import hermite_3d_data

    CL = {"H": (O.HERMITE, D.SplineBuilderClosure.HERMITE), "G": (O.GREVILLE, D.SplineBuilderClosure.GREVILLE),
          "P": (O.PERIODIC, D.SplineBuilderClosure.PERIODIC)}
    specs = [(3, False, 9), (5, True, 8), (4, True, 7)]
    ss, bl = [], []
    for k, ((deg, uni, n), c) in enumerate(zip(specs, closures)):
        oc, dc = CL[c]
        per = c == "P"
        if uni:
            ss.append(O.Spline(deg, True, oc, oc, ncells=n, rmin=0.0, rmax=1.0))
            bs = D.UniformBSplines(deg, 0.0, 1.0, n, periodic=per)
        else:
            br = breaks_for("random", n, seed=30 + k)
            ss.append(O.Spline(deg, False, oc, oc, breaks=br))
            bs = D.NonUniformBSplines(deg, br, periodic=per)
        bl.append(D.SplineBuilder(bs, dc, dc))
    polys = [np.polynomial.Polynomial(np.random.default_rng(k).standard_normal(s.degree + 1))
             for k, s in enumerate(ss)]
    nb = 2
    vals, d = hermite_3d_data(ss, polys, nb)
    rng = np.random.default_rng(7)
    vals = vals + 0.1 * rng.standard_normal(vals.shape)  # not a polynomial: every input matters
    d = {k: v + 0.1 * rng.standard_normal(v.shape) for k, v in d.items()}
    # inputs the closures do not take are not passed (the reference passes nullopt)
    used = {k: v for k, v in d.items() if _takes(k, closures)}
    ref = O.spline3d_build(*ss, vals, used)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(cuda) for k, v in used.items()}
    coef = torch.full((nb,) + tuple(s.size for s in ss), np.nan, dtype=torch.float64, device=cuda)
    D.SplineBuilder3D(*bl)(coef, torch.from_numpy(vals).to(cuda), **t)
    got = coef.cpu().numpy()
    assert np.all(np.isfinite(got))
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("closures", [("H", "H", "H"), ("H", "G", "P"), ("G", "H", "H"), ("P", "H", "G")])
def test_build3d_hermite_vjp_adjoint_identity(closures, cuda):
    """SplineBuilder3D.build_vjp with Hermite closures, block by block: for
    each input the closures take (values, derivatives, edge and corner cross
    derivatives), a random input in that block alone and a random gradient
    g of the coefficients satisfy <g, build(input)> = <input, gradient of
    that block>, with build the oracle's 3D restatement (1e-12 relative)."""
    import torch
    import ddc_amd as D
    import oracle as O
    from test_oracle_nd import hermite_3d_data

    CL = {"H": (O.HERMITE, D.SplineBuilderClosure.HERMITE), "G": (O.GREVILLE, D.SplineBuilderClosure.GREVILLE),
          "P": (O.PERIODIC, D.SplineBuilderClosure.PERIODIC)}
    specs = [(3, False, 9), (5, True, 8), (4, True, 7)]
    ss, bl = [], []
    for k, ((deg, uni, n), c) in enumerate(zip(specs, closures)):
        oc, dc = CL[c]
        per = c == "P"
        if uni:
            ss.append(O.Spline(deg, True, oc, oc, ncells=n, rmin=0.0, rmax=1.0))
            bs = D.UniformBSplines(deg, 0.0, 1.0, n, periodic=per)
        else:
            br = breaks_for("random", n, seed=30 + k)
            ss.append(O.Spline(deg, False, oc, oc, breaks=br))
            bs = D.NonUniformBSplines(deg, br, periodic=per)
        bl.append(D.SplineBuilder(bs, dc, dc))
    polys = [np.polynomial.Polynomial([1.0] * (s.degree + 1)) for s in ss]
    nb = 2
    vals, d = hermite_3d_data(ss, polys, nb)  # for the block shapes only
    used = {k: v for k, v in d.items() if _takes(k, closures)}
    rng = np.random.default_rng(17)
    gs = rng.standard_normal((nb,) + tuple(s.size for s in ss))
    gv = torch.full(vals.shape, np.nan, dtype=torch.float64, device=cuda)
    gd = {"g" + k: torch.full(v.shape, np.nan, dtype=torch.float64, device=cuda) for k, v in used.items()}
    D.SplineBuilder3D(*bl).build_vjp(torch.from_numpy(gs).to(cuda), gv, **gd)
    blocks = {"vals": (vals.shape, gv)}
    blocks.update({k: (v.shape, gd["g" + k]) for k, v in used.items()})
    for k, (shape, grad) in blocks.items():
        x = rng.standard_normal(shape)
        v0 = x if k == "vals" else np.zeros(vals.shape)
        dk = {j: (x if j == k else np.zeros(u.shape)) for j, u in used.items()}
        coef = O.spline3d_build(*ss, v0, dk)
        lhs = float((gs * coef).sum())
        g = grad.cpu().numpy()
        assert np.all(np.isfinite(g)), k
        rhs = float((x * g).sum())
        assert abs(lhs - rhs) <= 1e-12 * max(np.abs(gs * coef).sum(), 1e-300), k


def _takes(name, closures):
    """Whether SplineBuilder3D reads argument `name` for these closures:
    every dimension the argument carries derivatives of must be Hermite."""
    dims = [int(part[-1]) for part in name.split("_")[1:] if part[:3] in ("min", "max")]
    return all(closures[k - 1] == "H" for k in dims)


@pytest.mark.parametrize("spec,orders", [
    ([(3, "nug", 40), (3, "up", 36), (3, "ug", 34)], (0, 0, 0)),
    ([(3, "nug", 40), (3, "up", 36), (3, "ug", 34)], (1, 0, 2)),
    ([(5, "up", 33), (5, "nug", 17), (5, "ug", 20)], (0, 0, 0)),
    ([(1, "ug", 50), (1, "up", 3), (1, "nug", 70)], (0, 1, 0)),
    ([(2, "nug", 18), (2, "nug", 21), (2, "up", 16)], (0, 0, 0)),
], ids=["d3", "d3_deriv", "d5", "d1_one_tile_dim", "d2_two_nonuniform"])
def test_evalnd_binned_matches_direct(spec, orders, cuda):
    """The binned 3D evaluation (ddcx_evaluatornd_eval_binned: points grouped
    by tile of cells, patch in LDS) against the one-thread-per-point kernel
    bit for bit, and against the oracle; points outside every side of the
    domain on the value form (periodic wrap, Null rule); several tiles per
    dimension; 3 batch entries of different point densities."""
    import torch
    import ddc_amd as D
    import oracle as O

    bases = _bases(spec)
    ss = [x[0] for x in bases]
    rng = np.random.default_rng(17)
    nb, npts = 3, 30000
    coef = _coef(ss, nb, rng)
    ext = 0.2 if orders == (0, 0, 0) else 0.0
    coords = np.stack([rng.uniform(o.rmin - ext * (o.rmax - o.rmin), o.rmax + ext * (o.rmax - o.rmin), (nb, npts))
                       for o in ss], axis=-1)
    coords[1, : npts // 2] = coords[1, :1]  # one batch entry with half its points in one cell
    ev = D.SplineEvaluator3D(*[x[3] for x in bases])
    tc, tx = torch.from_numpy(coef).to(cuda), torch.from_numpy(coords).to(cuda)
    direct = torch.full((nb, npts), np.nan, dtype=torch.float64, device=cuda)
    ev.deriv(orders, direct, tx, tc, binned=False)
    binned = torch.full_like(direct, np.nan)
    ev.deriv(orders, binned, tx, tc, binned=True)
    torch.cuda.synchronize()
    assert np.array_equal(binned.cpu().numpy(), direct.cpu().numpy())
    ref = O.splinend_eval(ss, coords, coef, orders=list(orders))
    assert rel_err(binned.cpu().numpy(), ref) <= TOL * 10 ** sum(orders)
    # a second, smaller call reuses the workspace
    small = torch.full((nb, 100), np.nan, dtype=torch.float64, device=cuda)
    ev.deriv(orders, small, tx[:, :100].contiguous(), tc, binned=True)
    assert np.array_equal(small.cpu().numpy(), direct[:, :100].cpu().numpy())


def test_evalnd_binned_refuses_mixed_degrees(cuda):
    """Mixed degrees (or other than three dimensions) have no binned form:
    the workspace query gives 0 and binned=True raises; the default takes the
    direct kernel."""
    import ctypes as C
    import torch
    import ddc_amd as D
    from ddc_amd._lib import call

    bases = _bases(SPECS_3D[1])
    ev = D.SplineEvaluator3D(*[x[3] for x in bases])
    need = C.c_int64(-1)
    call("ddcx_evaluatornd_eval_workspace", ev._hs, 3, 1000, 2, C.byref(need))
    assert need.value == 0
    ss = [x[0] for x in bases]
    coef = torch.from_numpy(_coef(ss, 2, np.random.default_rng(1))).to(cuda)
    x = torch.rand((2, 1000, 3), dtype=torch.float64, device=cuda)
    out = torch.empty((2, 1000), dtype=torch.float64, device=cuda)
    with pytest.raises(NotImplementedError):
        ev(out, x, coef, binned=True)
    ev(out, x, coef)
