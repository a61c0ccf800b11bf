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


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy-kernel calibration), recorded in profiles/<name>.md as
    a line 'traffic_bytes_per_launch: <int>'."""
    path = os.path.join(ROOT, "profiles", name + ".md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return int(line.split(":", 1)[1].strip())
            except ValueError:
                return None
    return None


def cpu_baseline(ncells, B_sample, P_line, seed):
    """Oracle (CPU restatement, OpenMP over lines) on a bounded sample."""
    import numpy as np

    import oracle as O

    orc = O.Spline(3, True, O.PERIODIC, O.PERIODIC, ncells=ncells)
    rng = np.random.default_rng(seed)
    x = orc.points()
    f = np.sin(2 * np.pi * x[:, None] * rng.integers(1, 9, B_sample)[None, :])
    xe = rng.uniform(0, 1, (P_line, B_sample))
    orc.build(f)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while True:
        coef = orc.build(f)
        orc.eval(xe, coef)
        reps += 1
        if time.perf_counter() - t0 > 5.0 and reps >= 2:
            break
    dt = (time.perf_counter() - t0) / reps
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    mpts = (x.size * B_sample + P_line * B_sample) / dt / 1e6
    return {"value": round(mpts, 2), "unit": "Mpts/s", "cores": cores, "kind": "port",
            "sample": f"oracle build+eval, uniform periodic cubic, {ncells} cells x {B_sample} lines, "
                      f"{P_line} points/line, {reps} reps over {dt * reps:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ncells", type=int, default=1024)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
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
    k = torch.randint(1, 9, (B,), device=dev, generator=g).to(torch.float64)
    vals = torch.sin(2 * np.pi * xg[:, None] * k[None, :]) + 0.1 * torch.randn((m, B), device=dev, dtype=torch.float64,
                                                                                 generator=g)
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
    stream = torch.cuda.current_stream()

    def timed(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / n * 1e-3

    n_k = max(5, args.steps)
    t_build = timed(lambda: builder(coef, vals), n_k)
    t_eval = timed(lambda: ev(out, coords, coef), n_k)
    # calibration on this box: a plain device copy of the same size as vals
    t_copy = timed(lambda: out.copy_(vals), n_k)
    copy_gbps = 2 * vals.numel() * 8 / t_copy / 1e9
    b_build, b_eval = algorithmic_bytes(m, B, ncoef, P)
    dominant = "build" if t_build >= t_eval else "eval"
    t_dom = max(t_build, t_eval)
    ach = (b_build if dominant == "build" else b_eval) / t_dom / 1e9
    traffic = read_traffic(f"{dominant}_uniform_periodic_d3_1024x65536")
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
                    "device_copy_GBps": round(copy_gbps, 1)},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(nc, 4096, P_line, 0xDDC0002)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
