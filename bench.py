"""Benchmark of the batched spline build + evaluate hot path on one MI355X.

Workload (one step): SplineBuilder on a uniform periodic cubic basis with 1024
cells, batched over 65536 lines (values (1024, 65536) FP64, batch-contiguous,
512 MiB), then SplineEvaluator at 2^26 uniform random points (1024 per line).
Inputs are synthetic and resident in HBM before timing starts.

Prints ONE JSON line (see the keys below).  `roofline` is for the dominant
kernel, timed with HIP events on the stream it runs on; `cpu_baseline` is the
CPU oracle timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(m, B, nc, P, nbe=0):
    build = 8 * B * (m + nbe) + 8 * B * nc
    ev = 8 * P + 8 * P + 8 * B * nc
    return build, ev


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` on the headline workload, measured with
    rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel
    calibration) and recorded in profiles/<kernel>__uniform_periodic_d3_1024x65536.md
    as a line 'traffic_bytes_per_launch: <int>'."""
    path = os.path.join(ROOT, "profiles", kernel + "__uniform_periodic_d3_1024x65536.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return int(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def physical_cores():
    """Physical cores among the CPUs this process may run on (one per
    (package, core) pair of the allowed logical CPUs)."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    seen = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            key = (open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip())
        except OSError:
            key = ("?", str(c))
        seen.add(key)
    return max(1, len(seen))


def _median_runs(fn, min_runs=5, min_seconds=4.0):
    """1 warm-up call, then at least min_runs timed calls (more until
    min_seconds have passed); returns (median seconds, runs)."""
    fn()
    times = []
    t_start = time.perf_counter()
    while len(times) < min_runs or time.perf_counter() - t_start < min_seconds:
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2], len(times)


def cpu_baseline_run(seed):
    """The oracle (CPU restatement, OpenMP over lines; its threads pinned by
    the environment this process was started with) on
      - config 1 at its own shape: uniform periodic cubic, 1000 cells x 1000
        lines, interpolation-contiguous values [y][x], 1000 uniform points per
        line, f = sin(2 pi x k_y) + 0.1 N(0, 1);
      - a 4096-line sample of the headline workload (1024 cells,
        batch-contiguous, 1024 points per line).
    Build and eval are timed separately: 1 warm-up, median of >= 5 runs."""
    import numpy as np

    import oracle as O

    res = {}
    for name, ncells, B, layout in (("config1", 1000, 1000, "interp_contiguous"),
                                    ("headline_sample", 1024, 4096, "batch_contiguous")):
        orc = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=ncells)
        rng = np.random.default_rng(seed)
        x = orc.points()
        k = rng.integers(1, 9, B)
        f = np.sin(2 * np.pi * x[:, None] * k[None, :]) + 0.1 * rng.standard_normal((x.size, B))
        xe = rng.uniform(0, 1, (ncells, B))
        if layout == "interp_contiguous":
            f, xe = np.ascontiguousarray(f.T), np.ascontiguousarray(xe.T)
        coef = orc.build(f, layout=layout)
        t_b, n_b = _median_runs(lambda: orc.build(f, layout=layout))
        t_e, n_e = _median_runs(lambda: orc.eval(xe, coef, layout=layout))
        m, P = x.size, xe.size
        res[name] = {"ncells": ncells, "lines": B, "layout": layout, "eval_points": P,
                     "build_ms": round(t_b * 1e3, 3), "eval_ms": round(t_e * 1e3, 3), "runs": [n_b, n_e],
                     "build_eval_Mpts_s": round((m * B + P) / (t_b + t_e) / 1e6, 2)}
    return res


def cpu_baseline(seed):
    """Times cpu_baseline_run in a child process with OMP_NUM_THREADS = the
    physical cores (or the box's OMP_NUM_THREADS share, if smaller),
    OMP_PROC_BIND=close and OMP_PLACES=cores, so that the OpenMP runtime is
    started with those settings (this process has already loaded torch's)."""
    import subprocess

    phys = physical_cores()
    share = int(os.environ.get("OMP_NUM_THREADS", phys) or phys)
    cores = max(1, min(phys, share))
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-baseline-only", "--seed", str(seed)],
                       env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed: {r.stderr[-2000:]}")
    detail = json.loads(r.stdout.strip().splitlines()[-1])
    c1 = detail["config1"]
    return {"value": c1["build_eval_Mpts_s"], "unit": "Mpts/s", "cores": cores, "physical_cores": phys,
            "kind": "port",
            "sample": "oracle (OpenMP, OMP_PROC_BIND=close, OMP_PLACES=cores), config 1 at its shape: uniform "
                      "periodic cubic, 1000 cells x 1000 lines interp-contiguous, 1e6 eval points; build and eval "
                      "medians of >= 5 runs after 1 warm-up; value = (m*B + P) / (t_build + t_eval)",
            "detail": detail}


def _timer(stream):
    import torch

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / n * 1e-3

    return timed


def pattern_ceiling(src, dst, lb, timed, n):
    """GB/s of a plain copy of src -> dst ((rows, lines) batch-contiguous
    FP64, same shape) in tiles of lb lines by all rows: the access pattern of
    the tiled kernels, measured by the probe library built beside libddcx
    (None when it is not built)."""
    import ctypes

    import torch

    path = os.path.join(ROOT, "ddc_amd", "libddcx_probe.so")
    if not os.path.exists(path):
        return None
    fn = ctypes.CDLL(path).ddcx_probe_tile_copy
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                   ctypes.c_int, ctypes.c_void_p]
    rows, lines = src.shape
    assert dst.shape == src.shape and src.is_contiguous() and dst.is_contiguous()
    grid = min((lines + lb - 1) // lb, 2 * torch.cuda.get_device_properties(src.device).multi_processor_count)
    stream = torch.cuda.current_stream().cuda_stream

    def run():
        if fn(src.data_ptr(), dst.data_ptr(), rows, lines, lines, lb, grid, stream) != 0:
            raise RuntimeError("ddcx_probe_tile_copy failed")

    run()
    return 2 * src.numel() * 8 / timed(run, n) / 1e9


def _ceiling(gbps, lb, achieved):
    if gbps is None:
        return None
    return {"pattern": f"copy in batch-contiguous tiles of {lb} lines x all rows (row stride = the batch), "
                       "the dominant kernel's access pattern", "GBps": round(gbps, 1),
            "frac": round(achieved / gbps, 4)}


def _line_values(xg, B, gen, axis_interp_first=True):
    """f(x, y) = sin(2 pi x k_y) + 0.1 N(0, 1), k_y in {1..8}."""
    import numpy as np
    import torch

    dev = xg.device
    k = torch.randint(1, 9, (B,), device=dev, generator=gen).to(torch.float64)
    v = torch.sin(2 * np.pi * xg[:, None] * k[None, :]) + 0.1 * torch.randn((xg.numel(), B), device=dev,
                                                                           dtype=torch.float64, generator=gen)
    return v if axis_interp_first else v.T.contiguous()


def measure_config(D, bs, builder, B, layout, timed, n, gen, derivs=False):
    """Build, eval and deriv(1) times for one 1D configuration (1024 uniform
    random points per line, the interpolation extent)."""
    import torch

    dev = torch.device("cuda:0")
    m, nc = builder.npoints(), bs.size()
    xg = torch.from_numpy(builder.interpolation_points()).to(dev)
    bc = layout == "batch_contiguous"
    ax = 0 if bc else 1
    vals = _line_values(xg, B, gen, bc)
    coef = torch.empty((nc, B) if bc else (B, nc), device=dev, dtype=torch.float64)
    P_line = 1024
    lo, hi = bs.rmin(), bs.rmax()
    coords = lo + (hi - lo) * torch.rand((P_line, B) if bc else (B, P_line), device=dev, dtype=torch.float64,
                                         generator=gen)
    out = torch.empty_like(coords)
    ev = D.SplineEvaluator(bs, D.ConstantExtrapolationRule(), D.ConstantExtrapolationRule())
    for _ in range(2):
        builder(coef, vals, axis=ax)
        ev(out, coords, coef, axis=ax)
    t_b = timed(lambda: builder(coef, vals, axis=ax), n)
    t_e = timed(lambda: ev(out, coords, coef, axis=ax), n)
    t_d = timed(lambda: ev.deriv(1, out, coords, coef, axis=ax), n)
    # adjoints: evaluator (w = out as weights) then builder
    gx, gc, gv = torch.empty_like(coords), torch.empty_like(coef), torch.empty_like(vals)
    ev.eval_vjp(coords, out, coef, gx, gc, axis=ax)
    builder.build_vjp(gc, gv, axis=ax)
    t_ev = timed(lambda: ev.eval_vjp(coords, out, coef, gx, gc, axis=ax), n)
    t_bv = timed(lambda: builder.build_vjp(gc, gv, axis=ax), n)
    P = P_line * B
    bb, be = algorithmic_bytes(m, B, nc, P)
    bev = 24 * P + 16 * B * nc  # coords, w, gx; coef read, gcoef written
    return {"layout": layout, "build_us": round(t_b * 1e6, 1), "eval_us": round(t_e * 1e6, 1),
            "deriv_us": round(t_d * 1e6, 1), "build_GBps": round(bb / t_b / 1e9, 1),
            "eval_GBps": round(be / t_e / 1e9, 1), "deriv_GBps": round(be / t_d / 1e9, 1),
            "eval_vjp_us": round(t_ev * 1e6, 1), "eval_vjp_GBps": round(bev / t_ev / 1e9, 1),
            "build_vjp_us": round(t_bv * 1e6, 1), "build_vjp_GBps": round(bb / t_bv / 1e9, 1),
            "build_eval_Mpts_s": round((m * B + P) / (t_b + t_e) / 1e6, 1)}


def measure_fused(D, bs, builder, B, timed, n, gen):
    """The headline workload through ddcx_build_eval: build + eval in one pass,
    with the coefficients stored and kept on chip."""
    import torch

    dev = torch.device("cuda:0")
    m, nc = builder.npoints(), bs.size()
    xg = torch.from_numpy(builder.interpolation_points()).to(dev)
    vals = _line_values(xg, B, gen)
    coords = torch.rand((m, B), device=dev, dtype=torch.float64, generator=gen)
    out = torch.empty_like(coords)
    coef = torch.empty((nc, B), device=dev, dtype=torch.float64)
    ev = D.SplineEvaluator(bs)
    D.build_eval(builder, ev, out, coords, vals, coef)
    D.build_eval(builder, ev, out, coords, vals, None)
    t_s = timed(lambda: D.build_eval(builder, ev, out, coords, vals, coef), n)
    t_c = timed(lambda: D.build_eval(builder, ev, out, coords, vals, None), n)
    P = m * B
    by_s = 8 * B * m + 8 * B * nc + 16 * P
    by_c = 8 * B * m + 16 * P
    return {"with_coef_us": round(t_s * 1e6, 1), "with_coef_GBps": round(by_s / t_s / 1e9, 1),
            "with_coef_Mpts_s": round((m * B + P) / t_s / 1e6, 1),
            "on_chip_us": round(t_c * 1e6, 1), "on_chip_GBps": round(by_c / t_c / 1e9, 1),
            "on_chip_Mpts_s": round((m * B + P) / t_c / 1e6, 1)}


def measure_precision_form(D, degree, B, timed, n, gen):
    """Hermite builder on the geometric mesh x_k = (1.05^k - 1)/(1.05^2048 - 1):
    FP64 and FP32-sweep + FP64-refinement builds on the same inputs.  The
    error is measured against the FP64 GPU build, line by line."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    nc = 2048
    br = (1.05 ** np.arange(nc + 1, dtype=np.float64) - 1.0) / (1.05 ** nc - 1.0)
    bs = D.NonUniformBSplines(degree, br)
    H = D.SplineBuilderClosure.HERMITE
    b64 = D.SplineBuilder(bs, H, H)
    b32 = D.SplineBuilder(bs, H, H, precision=D.Precision.FP32_REFINE)
    m, nbv, size = b64.npoints(), b64.s_nbv_xmin, bs.size()
    raw = torch.randn((m + 8, B), device=dev, dtype=torch.float64, generator=gen)
    vals = torch.nn.functional.avg_pool1d(raw.T.unsqueeze(1), 9, stride=1).squeeze(1).T.contiguous()
    dl = torch.randn((nbv, B), device=dev, dtype=torch.float64, generator=gen)
    dr = torch.randn((nbv, B), device=dev, dtype=torch.float64, generator=gen)
    c64 = torch.empty((size, B), device=dev, dtype=torch.float64)
    c32 = torch.empty_like(c64)
    for _ in range(2):
        b64(c64, vals, dl, dr)
        b32(c32, vals, dl, dr)
    t64 = timed(lambda: b64(c64, vals, dl, dr), n)
    t32 = timed(lambda: b32(c32, vals, dl, dr), n)
    err = ((c32 - c64).abs().amax(0) / c64.abs().amax(0)).max().item()
    by = 8 * B * (m + 2 * nbv) + 8 * B * size
    return {"degree": degree, "fp64_build_us": round(t64 * 1e6, 1), "fp32_refine_build_us": round(t32 * 1e6, 1),
            "fp64_GBps": round(by / t64 / 1e9, 1), "fp32_refine_GBps": round(by / t32 / 1e9, 1),
            "speedup_fp32_refine_vs_fp64": round(t64 / t32, 3), "refine_steps": b32.info.refine_steps,
            "max_line_rel_err_vs_fp64": float(f"{err:.3e}"), "fp32_refine_kernel": b32.info.kernel}


def measure_advection(D, timed, n):
    """One periodic advection step on 512 x 512 (uniform periodic cubic along
    x, batched over y): build, eval at the feet x - sin(2 pi y) dt, and the
    adjoint (eval_vjp then build_vjp); for FP64 also the step as one fused
    kernel with the feet computed in it (ddcx_advect).  The data (~6 MB)
    stays in cache, so the figures are microseconds, not a fraction of HBM
    bandwidth."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    N = 512
    bs = D.UniformBSplines(3, 0.0, 1.0, N, periodic=True)
    ev = D.SplineEvaluator(bs, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule())
    x = torch.from_numpy(D.SplineBuilder(bs).interpolation_points()).to(dev)
    y = torch.arange(N, device=dev, dtype=torch.float64) / N
    vals = torch.sin(2 * np.pi * x)[:, None] * torch.cos(2 * np.pi * y)[None, :]
    shift = (torch.sin(2 * np.pi * y) * 0.01).contiguous()
    feet = x[:, None] - shift[None, :]
    w = torch.ones_like(feet)
    coef = torch.empty((bs.size(), N), dtype=torch.float64, device=dev)
    out, gx, gc, gv = torch.empty_like(feet), torch.empty_like(feet), torch.empty_like(coef), torch.empty_like(vals)
    res = {}
    for name, prec in (("fp64", D.Precision.FP64), ("fp32_refine", D.Precision.FP32_REFINE)):
        b = D.SplineBuilder(bs, precision=prec)
        b(coef, vals)
        ev(out, feet, coef)
        ev.eval_vjp(feet, w, coef, gx, gc)
        b.build_vjp(gc, gv)
        t_b = timed(lambda: b(coef, vals), n)
        t_e = timed(lambda: ev(out, feet, coef), n)
        t_a = timed(lambda: (ev.eval_vjp(feet, w, coef, gx, gc), b.build_vjp(gc, gv)), n)
        res[name] = {"build_us": round(t_b * 1e6, 1), "eval_us": round(t_e * 1e6, 1),
                     "adjoint_us": round(t_a * 1e6, 1), "step_us": round((t_b + t_e) * 1e6, 1)}
        if prec == D.Precision.FP64:
            D.advect(b, ev, out, vals, shift)
            res[name]["fused_step_us"] = round(timed(lambda: D.advect(b, ev, out, vals, shift), n) * 1e6, 1)
    return res


def measure_2d(D, timed, gen):
    """Tensor-product 2D: x non-uniform cubic (1024 cells, Greville closures,
    1027 points), y uniform periodic cubic (1024 cells), batch 16; build, and
    eval plus its adjoint at 2^24 scattered points per batch."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    nb, P = 16, 1 << 24
    rng = np.random.default_rng(0xDDC0005)
    br = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, 1023)), [1.0]])
    bs1 = D.NonUniformBSplines(3, br)
    bs2 = D.UniformBSplines(3, 0.0, 1.0, 1024, periodic=True)
    G = D.SplineBuilderClosure.GREVILLE
    b2d = D.SplineBuilder2D(D.SplineBuilder(bs1, G, G), D.SplineBuilder(bs2))
    ev = D.SplineEvaluator2D(D.SplineEvaluator(bs1, D.NullExtrapolationRule(), D.NullExtrapolationRule()),
                             D.SplineEvaluator(bs2, D.PeriodicExtrapolationRule(), D.PeriodicExtrapolationRule()))
    n1, n2 = b2d.builder1.npoints(), b2d.builder2.npoints()
    vals = torch.randn((nb, n1, n2), dtype=torch.float64, device=dev, generator=gen)
    coef = torch.empty((nb, bs1.size(), bs2.size()), dtype=torch.float64, device=dev)
    b2d(coef, vals)
    t_b = timed(lambda: b2d(coef, vals), 5)
    # the same build with FP32 sweeps + FP64 refinement in both passes
    F = D.Precision.FP32_REFINE
    b2f = D.SplineBuilder2D(D.SplineBuilder(bs1, G, G, precision=F), D.SplineBuilder(bs2, precision=F))
    coef32 = torch.empty_like(coef)
    b2f(coef32, vals)
    t_bf = timed(lambda: b2f(coef32, vals), 5)
    err32 = ((coef32 - coef).abs().amax(2) / coef.abs().amax(2)).max().item()
    del coef32
    coords = torch.rand((nb, P, 2), dtype=torch.float64, device=dev, generator=gen)
    out = torch.empty((nb, P), dtype=torch.float64, device=dev)
    ev(out, coords, coef)
    t_e = timed(lambda: ev(out, coords, coef), 3)
    w = torch.randn((nb, P), dtype=torch.float64, device=dev, generator=gen)
    gx, gc = torch.empty_like(coords), torch.empty_like(coef)
    ev.eval_vjp(coords, w, coef, gx, gc)
    t_a = timed(lambda: ev.eval_vjp(coords, w, coef, gx, gc), 2)
    by_b = 8 * nb * (n1 * n2 + bs1.size() * bs2.size())
    # the two 1D passes also write and read the dimension-1 spline once each
    by_b2 = by_b + 2 * 8 * nb * bs1.size() * n2
    by_e = 24 * nb * P + 8 * coef.numel()
    return {"build_us": round(t_b * 1e6, 1), "build_GBps": round(by_b / t_b / 1e9, 1),
            "build_GBps_with_intermediate": round(by_b2 / t_b / 1e9, 1),
            "fp32_refine_build_us": round(t_bf * 1e6, 1), "fp32_refine_build_GBps": round(by_b / t_bf / 1e9, 1),
            "fp32_refine_max_line_rel_err_vs_fp64": float(f"{err32:.3e}"),
            "eval_us": round(t_e * 1e6, 1), "eval_GBps": round(by_e / t_e / 1e9, 1),
            "eval_Mpts_s": round(nb * P / t_e / 1e6, 1), "eval_vjp_us": round(t_a * 1e6, 1),
            "points": nb * P}


def measure_3d(D, timed, gen):
    """Tensor-product 3D (SplineBuilder3D / SplineEvaluator3D): x non-uniform
    cubic (128 cells, Greville closures), y uniform periodic cubic (128
    cells), z uniform cubic (128 cells, Greville closures), batch 4; build,
    eval and deriv (1, 1, 1) at 2^22 scattered points per batch, integrate."""
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    nb, P, n = 4, 1 << 22, 128
    rng = np.random.default_rng(0xDDC0006)
    br = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, n - 1)), [1.0]])
    G = D.SplineBuilderClosure.GREVILLE
    bs = [D.NonUniformBSplines(3, br), D.UniformBSplines(3, 0.0, 1.0, n, periodic=True),
          D.UniformBSplines(3, 0.0, 1.0, n)]
    b3 = D.SplineBuilder3D(D.SplineBuilder(bs[0], G, G), D.SplineBuilder(bs[1]), D.SplineBuilder(bs[2], G, G))
    null, per = D.NullExtrapolationRule(), D.PeriodicExtrapolationRule()
    ev = D.SplineEvaluator3D(D.SplineEvaluator(bs[0], null, null), D.SplineEvaluator(bs[1], per, per),
                             D.SplineEvaluator(bs[2], null, null))
    npts = [b.npoints() for b in b3.builders]
    vals = torch.randn([nb] + npts, dtype=torch.float64, device=dev, generator=gen)
    coef = torch.empty([nb] + [b.size() for b in bs], dtype=torch.float64, device=dev)
    b3(coef, vals)
    t_b = timed(lambda: b3(coef, vals), 5)
    coords = torch.rand((nb, P, 3), dtype=torch.float64, device=dev, generator=gen)
    out = torch.empty((nb, P), dtype=torch.float64, device=dev)
    ev(out, coords, coef)
    t_e = timed(lambda: ev(out, coords, coef), 3)
    t_d = timed(lambda: ev.deriv((1, 1, 1), out, coords, coef), 3)
    integ = torch.empty(nb, dtype=torch.float64, device=dev)
    t_i = timed(lambda: ev.integrate(integ, coef), 5)
    by_b = 8 * (vals.numel() + coef.numel())
    by_e = 32 * nb * P + 8 * coef.numel()
    return {"build_us": round(t_b * 1e6, 1), "build_GBps": round(by_b / t_b / 1e9, 1),
            "eval_us": round(t_e * 1e6, 1), "eval_GBps": round(by_e / t_e / 1e9, 1),
            "eval_Mpts_s": round(nb * P / t_e / 1e6, 1), "deriv111_us": round(t_d * 1e6, 1),
            "integrate_us": round(t_i * 1e6, 1), "integrate_GBps": round(8 * coef.numel() / t_i / 1e9, 1),
            "points": nb * P}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncells", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="skip the other workloads")
    ap.add_argument("--cpu-baseline-only", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--seed", type=int, default=0xDDC0001, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_only:  # child process of cpu_baseline(): no GPU use
        print(json.dumps(cpu_baseline_run(args.seed)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the spline path shards by batch line with no communication; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import ddc_amd as D

    dev = torch.device("cuda:0")
    nc, B = args.ncells, args.batch
    bs = D.UniformBSplines(3, 0.0, 1.0, nc, periodic=True)
    builder = D.SplineBuilder(bs, D.SplineBuilderClosure.PERIODIC, D.SplineBuilderClosure.PERIODIC)
    ev = D.SplineEvaluator(bs)
    m = builder.npoints()
    ncoef = bs.size()
    g = torch.Generator(device=dev)
    g.manual_seed(0xDDC0002)
    xg = torch.from_numpy(builder.interpolation_points()).to(dev)
    vals = _line_values(xg, B, g)
    coef = torch.empty((ncoef, B), device=dev, dtype=torch.float64)
    P_line = m
    coords = torch.rand((P_line, B), device=dev, dtype=torch.float64, generator=g)
    out = torch.empty_like(coords)
    P = P_line * B
    torch.cuda.synchronize()

    def step():
        builder(coef, vals)
        ev(out, coords, coef)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # per-kernel durations with HIP events on the launch stream
    timed = _timer(torch.cuda.current_stream())
    n_k = max(5, args.steps)
    t_build = timed(lambda: builder(coef, vals), n_k)
    t_eval = timed(lambda: ev(out, coords, coef), n_k)
    # calibration on this box: a plain device copy of the same size as vals
    t_copy = timed(lambda: out.copy_(vals), n_k)
    copy_gbps = 2 * vals.numel() * 8 / t_copy / 1e9
    # the tiled kernels' access pattern (16-line tiles for the evaluator,
    # 32-line tiles for the build) in a plain copy of the same array
    tile_gbps = {lb: pattern_ceiling(vals, out, lb, timed, n_k) for lb in (16, 32)}
    b_build, b_eval = algorithmic_bytes(m, B, ncoef, P)
    dominant = "build" if t_build >= t_eval else "eval"
    t_dom = max(t_build, t_eval)
    ach = (b_build if dominant == "build" else b_eval) / t_dom / 1e9
    traffic = read_traffic("build_tiled_kernel" if dominant == "build" else "eval_stream_kernel")
    value = (m * B + P) / (ms * 1e-3) / 1e6
    res = {
        "metric": "spline build+eval Mpts/s, %HBM peak; FP32+refine vs FP64 at <=1e-12 rel err",
        "value": round(value, 2),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"uniform periodic cubic spline build + eval, {nc} cells x {B} lines "
                               f"(batch-contiguous), {P} eval points", "ncells": nc, "batch": B,
                   "eval_points": P, "precision": "fp64"},
        "kernels": {"build_us": round(t_build * 1e6, 2), "eval_us": round(t_eval * 1e6, 2),
                    "build_GBps": round(b_build / t_build / 1e9, 1), "eval_GBps": round(b_eval / t_eval / 1e9, 1),
                    "device_copy_GBps": round(copy_gbps, 1),
                    "tile_copy_16_lines_GBps": None if tile_gbps[16] is None else round(tile_gbps[16], 1),
                    "tile_copy_32_lines_GBps": None if tile_gbps[32] is None else round(tile_gbps[32], 1)},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "pattern_ceiling": _ceiling(tile_gbps[16 if dominant == "eval" else 32],
                                                 16 if dominant == "eval" else 32, ach)},
        "cpu_baseline": None,
    }
    if not args.headline_only:
        del vals, coef, coords, out
        torch.cuda.empty_cache()
        gen = torch.Generator(device=dev)
        gen.manual_seed(0xDDC0003)
        other = {}
        other["2a_fused_build_eval"] = measure_fused(D, bs, builder, B, timed, n_k, gen)
        torch.cuda.empty_cache()
        other["2a_batch_contiguous"] = measure_config(D, bs, builder, B, "batch_contiguous", timed, n_k, gen)
        other["2a_interp_contiguous"] = measure_config(D, bs, builder, B, "interp_contiguous", timed, n_k, gen)
        bs5 = D.UniformBSplines(5, 0.0, 1.0, nc, periodic=True)
        other["2b_uniform_periodic_d5"] = measure_config(D, bs5, D.SplineBuilder(bs5), B, "batch_contiguous", timed,
                                                         n_k, gen)
        rng = np.random.default_rng(0xDDC0002)
        br = np.concatenate([[0.0], np.sort(rng.uniform(0.0, 1.0, nc - 1)), [1.0]])
        bsn = D.NonUniformBSplines(3, br)
        G = D.SplineBuilderClosure.GREVILLE
        other["2c_nonuniform_greville_d3"] = measure_config(D, bsn, D.SplineBuilder(bsn, G, G), B,
                                                            "batch_contiguous", timed, n_k, gen)
        torch.cuda.empty_cache()
        other["3_precision_form_d3"] = measure_precision_form(D, 3, 32768, timed, n_k, gen)
        other["3_precision_form_d5"] = measure_precision_form(D, 5, 32768, timed, n_k, gen)
        torch.cuda.empty_cache()
        other["4_advection_512x512"] = measure_advection(D, timed, n_k)
        other["5_tensor_2d_1024x1024x16"] = measure_2d(D, timed, gen)
        torch.cuda.empty_cache()
        other["6_tensor_3d_128x128x128x4"] = measure_3d(D, timed, gen)
        res["workloads"] = other
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(0xDDC0001)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
