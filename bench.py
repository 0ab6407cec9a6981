#!/usr/bin/env python3
"""Benchmark: gridded-field ensemble Boris on one MI355X.

Workload (default, the gridded headline): 256^3 periodic turbulent box of side
1e7 m (B = B0 z + k^-5/3 random-phase dB, dB_rms = 0.5 B0, B0 = 5 nT),
E = -u x B with u = (4e5, 0, 0) m/s derived in-kernel, 4,194,304 protons
(1-100 keV, isotropic), MultistepBoris4(n=2), dt = T_g(B0)/40, 5,000 steps,
start and end records saved.

One "step" of this script = one full solve (all particles, all 5,000 Boris
steps) with initial states and records resident in HBM.  `value` is the
FP32-compensated pusher's particle-steps/s; the FP64 pusher is timed the same
way beside it, and the FP32C result's error against FP64 is reported.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "testparticle.jl_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
FP64_PEAK = 78.6e12        # FLOP/s, vector FP64
FP32_PEAK = 157.3e12       # FLOP/s, vector FP32 (packed)
# algorithmic bytes per particle-step of the gather: one vector field,
# 8 nodes x 3 components (E = -u x B derived in-kernel)
GATHER_BYTES = {"f32": 8 * 3 * 4, "f64": 8 * 3 * 8}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def timed_runs(ens, torch, warmup, steps):
    for _ in range(warmup):
        ens.run()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = ens.run()
        kms.append(st.kernel_ms)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(kms)), st


def cpu_baseline(w, A64, ax, nthreads, target_s=12.0):
    """The oracle (CPU restatement of the reference's EnsembleThreads solve:
    FP64, reference-form MultistepBoris4 with atan/sincos) on all host cores,
    over a bounded sample of the same particles."""
    import oracle as O

    B = O.grid(np.ascontiguousarray(A64.transpose(3, 2, 1, 0)), [a[0] for a in ax],
               [a[1] - a[0] for a in ax], O.EXTRAP_PERIODIC)
    E = O.uxb(w.extra["u"], B)
    kw = dict(q2m=w.q2m, tspan=w.tspan, dt=w.dt, alg=O.ALG_MULTISTEP, order_N=4, nsub=2,
              form=O.FORM_REFERENCE, save_everystep=False, nthreads=nthreads)
    probe = max(nthreads * 4, 64)
    r = O.solve(w.xv0[:probe], E, B, **kw)
    rate = probe * w.nsteps / max(r.elapsed, 1e-6)
    n = int(min(len(w.xv0), max(probe, rate * target_s / w.nsteps)))
    r = O.solve(w.xv0[:n], E, B, **kw)
    return {"value": n * w.nsteps / r.elapsed, "unit": "particle-steps/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{n} of {len(w.xv0)} particles x {w.nsteps} steps, FP64 "
                      f"MultistepBoris4(n=2) reference form, {r.elapsed:.1f} s wall"}


def bench_turbulent(args):
    import torch

    import testparticle_amd as TP
    from testparticle_amd import workloads as W
    from testparticle_amd.device import DeviceEnsemble

    N, n, nsteps = args.particles, args.grid, args.nsteps
    log(f"[bench] generating {n}^3 turbulent field and {N} protons")
    w, A64 = W.turbulent_box(N=N, n=n, nsteps=nsteps, dtype=np.float64)
    ax = W.grid_axes(n, w.extra["L"])
    alg = TP.MultistepBoris(4, n=2)
    res = {}
    for prec, storage in (("fp32c", "f32"), ("fp64", "f64")):
        A = A64 if storage == "f64" else A64.astype(np.float32)
        B = TP.GridField(A, *ax, bc=TP.PeriodicBC(), layout=args.layout)
        E = TP.UxBField(w.extra["u"], B)
        prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, E, B, TP.ZeroField()))
        log(f"[bench] {prec}: uploading and warming up")
        ens = DeviceEnsemble(prob, alg, w.xv0, dt=w.dt, save_everystep=False, precision=prec)
        wall, kms, st = timed_runs(ens, torch, args.warmup, args.steps)
        assert st.particle_steps == N * nsteps, st.particle_steps
        res[prec] = dict(wall=wall, kernel_ms=kms, final=ens.final_states(), storage=storage)
        log(f"[bench] {prec}: {wall * 1e3:.1f} ms/solve, kernel {kms:.1f} ms, "
            f"{N * nsteps / wall:.3e} particle-steps/s")
        del ens, B, E, prob
        torch.cuda.empty_cache()
    x0 = w.xv0[:, :3]
    x64, x32 = res["fp64"]["final"][:, :3], res["fp32c"]["final"][:, :3]
    d = np.linalg.norm(x32 - x64, axis=1)
    err_disp = d / np.maximum(np.linalg.norm(x64 - x0, axis=1), 1.0)
    err_pos = d / np.linalg.norm(x64, axis=1)
    ps = N * nsteps
    r32 = res["fp32c"]
    value = ps / r32["wall"]
    bytes_launch = ps * GATHER_BYTES["f32"] + N * (48 + 2 * 56)
    achieved = bytes_launch / (r32["kernel_ms"] * 1e-3)
    out = {
        "metric": "particle-steps/s (gridded Boris) FP32-comp vs FP64 at <=1e-5 err; adjoint cost",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["wall"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded k^-5/3 turbulent field and particle states)",
        "config": {"workload": f"turbulent_box_{n}^3_multistep4_n2", "particles": N,
                   "grid": n, "steps_per_solve": nsteps, "alg": "MultistepBoris4(n=2)",
                   "E": "-u x B in-kernel", "saves": "start+end", "field_layout": args.layout,
                   "parallelism": "single GPU, one thread per particle"},
        "fp64_value": ps / res["fp64"]["wall"],
        "fp32c_over_fp64": res["fp64"]["wall"] / r32["wall"],
        "fp32c_err_vs_fp64": {"metric": "|x32 - x64| / |x64 - x0| at the final step",
                              "median": float(np.median(err_disp)),
                              "p95": float(np.percentile(err_disp, 95)),
                              "max": float(np.max(err_disp)),
                              "median_rel_position": float(np.median(err_pos))},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_fixed<S32G<UxB, n=2, periodic, cells>>", "kernel_ms": r32["kernel_ms"],
                     "algorithmic_bytes_per_particle_step": GATHER_BYTES["f32"]},
    }
    # HBM traffic per launch: FETCH_SIZE + WRITE_SIZE per particle-step from the
    # rocprofv3 --pmc passes recorded under profiles/ (scaled by the copy-probe
    # calibration), times this launch's particle-steps
    prof = os.path.join(ROOT, "profiles", f"traffic_turbulent_box_{n}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            t = json.load(f)
        if t.get("layout") == args.layout:
            out["roofline"]["traffic"] = ps * (t["fetch_bytes_per_particle_step"] +
                                               t["write_bytes_per_particle_step"])
            out["roofline"]["traffic_source"] = os.path.relpath(prof, ROOT)
    if args.cpu_baseline:
        log("[bench] CPU baseline (oracle, all host cores)")
        out["cpu_baseline"] = cpu_baseline(w, A64, ax, cpu_threads())
    else:
        out["cpu_baseline"] = None
    return out


def copy_probe(nbytes=1 << 30, reps=3):
    """A plain device copy of known size: under rocprofv3 --pmc it gives the
    counter reading (FETCH_SIZE / WRITE_SIZE) per byte actually moved, which
    scales the pusher's counters to HBM bytes."""
    import torch

    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=1 << 22)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--nsteps", type=int, default=5000)
    ap.add_argument("--layout", choices=("cells", "nodes"), default="cells",
                    help="grid storage on the device (cells: one cache line per gather)")
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the ensemble pusher runs on a single GPU "
              "(--gpus 1); multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    out = bench_turbulent(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
