#!/usr/bin/env python3
"""Benchmark: ensemble Boris pushers on one MI355X.

Headline (`value`): gridded turbulent box, 256^3 periodic grid of side 1e7 m
(B = B0 z + k^-5/3 random-phase dB, dB_rms = 0.5 B0, B0 = 5 nT), E = -u x B
with u = (4e5, 0, 0) m/s derived in-kernel, 4,194,304 protons (1-100 keV,
isotropic), MultistepBoris4(n=2), dt = T_g(B0)/40, 5,000 steps, start and end
records saved.  One "step" of this script = one full solve (all particles,
all 5,000 Boris steps) with initial states and records resident in HBM.
`value` is the FP64 pusher's particle-steps/s in the reference's own
multistep arithmetic (atan, sincos), bit-identical to the oracle; the
FP32-compensated pusher and the rational-coefficient form are timed the same
way beside it ("variants"), and the FP32C final positions are compared with
FP64 at several horizons next to the FP64 chaos floor (FP64 against FP64
started one ulp away).

Secondary sections (same JSON line, key "secondary"):
  gyration 1,024 protons in a uniform B, standard Boris, 6,400 steps, a record
           every 10 (the reference's CPU-runnable workload), FP64 and FP32C,
           with the oracle over the whole ensemble beside them
  dipole   analytic Earth dipole, 1,048,576 protons, standard Boris, per-particle
           dt = T_g/32, 10,000 steps, removal inside 1 R_E (FP64 and FP32C)
  batch    64 uneven turbulent-box instances (32^3..128^3, ~2e7 protons),
           AdaptiveBoris(0.1), exits at the box faces, <= 2,000 steps, one
           persistent graph-captured launch after autotuning (FP32C and FP64)
  adjoint  64^3 grid, dB = 0.3 B0, 262,144 protons at 10 keV, standard Boris,
           2,048 steps, checkpoints every 64: gradient cost against the plain
           forward solve and dL/dB against central differences of the FP64 solve

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
# per particle of a solve: its initial state (6 doubles) read, its length and
# return code (2 x int32) written; per record: x, v (6 doubles) and t
STATE_BYTES, LEN_BYTES, RECORD_BYTES = 6 * 8, 2 * 4, 7 * 8
# adjoint reverse sweep per particle-step: 2 gathers + segment state written
# and read + dL/dB record (key + Bbar + weights) written, read, written, read
ADJ_REVERSE_BYTES = {"fp32c": 2 * 96 + 2 * 36 + 4 * 32, "fp64": 2 * 192 + 2 * 48 + 4 * 52}
# algorithmic FLOPs per particle-step (FMA = 2; div, sqrt = 1): standard Boris
# with E = 0 (37) + position update (6) + dipole evaluation (27)
DIPOLE_FLOP = 37 + 6 + 27
QI, MI = 1.602176634e-19, 1.67262192369e-27   # proton charge and mass (constants.jl:4-20)
SECTIONS = ("headline", "gyration", "dipole", "batch", "adjoint")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def timed_runs(run, torch, warmup, steps):
    """warmup + `steps` timed calls of run() bracketed by device syncs;
    returns (wall seconds per call, mean kernel ms, last stats)."""
    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    st = None
    for _ in range(steps):
        st = run()
        kms.append(st.kernel_ms)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    return wall, float(np.mean(kms)), st


def rel_pos_err(xa, xb):
    d = np.linalg.norm(xa - xb, axis=1) / np.linalg.norm(xb, axis=1)
    d = d[np.isfinite(d)]
    return {"median": float(np.median(d)), "p95": float(np.percentile(d, 95)),
            "max": float(np.max(d))}


def host_info():
    """The host cores the CPU baselines run on: the threads used (the
    oracle's OpenMP team is sized to min(16, affinity): 16 host cores is the
    share of one GPU on the machines this runs on), the affinity count, the
    CPU model and the SMT state."""
    info = {"threads": cpu_threads()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/smt/active") as f:
            info["smt_active"] = f.read().strip() == "1"
    except OSError:
        info["smt_active"] = None
    return info


def oracle_calibration():
    """One thread on the reference's serial benchmark problem
    (benchmark/bench_serial.jl:10-45: proton, uniform B = (0, 0, 1e-8) T,
    E = 0, v0 = (1e5, 0, 0), dt = 1 ns, 1e8 steps, a record every 1e5
    steps), published at 1.08e8 particle-steps/s per core
    (docs/src/index.md:48-54)."""
    import oracle as O

    cal = O.solve(np.array([[0.0, 0.0, 0.0, 1.0e5, 0.0, 0.0]]), O.zero(), O.uniform([0.0, 0.0, 1.0e-8]),
                  q2m=QI / MI, tspan=(0.0, 0.1), dt=1.0e-9, savestepinterval=100000,
                  maxiters=200_000_000, nthreads=1)
    return {"problem": "bench_serial.jl: 1 proton, uniform B, 1e8 Boris steps", "threads": 1,
            "value": 1.0e8 / cal.elapsed, "published_per_core": 1.08e8}


def oracle_rate(run, n_total, nthreads, target_s, what):
    """Times run(n) (the oracle over a sample of n particles; returns its
    elapsed seconds and the Boris steps it advanced) on a probe, then on a
    sample sized to ~target_s of CPU work, and returns the cpu_baseline
    record."""
    probe = min(max(nthreads * 4, 64), n_total)
    el, steps = run(probe)
    per_particle_s = max(el, 1e-6) / probe
    n = int(min(n_total, max(probe, target_s / per_particle_s)))
    el, steps = run(n)
    return {"value": steps / el, "unit": "particle-steps/s", "cores": nthreads,
            "kind": "port", "host": host_info(),
            "sample": f"{n} of {n_total} particles, {steps} particle-steps, {what}, {el:.1f} s wall"}


def cpu_baseline(w, A64, ax, nthreads, target_s=12.0):
    """The oracle (CPU restatement of the reference's EnsembleThreads solve:
    FP64, reference-form MultistepBoris4 with the FDLIBM atan and sin/cos) on
    the host cores, over a bounded sample of the same particles: the same
    arithmetic as the headline value."""
    import oracle as O

    B = O.grid(np.ascontiguousarray(A64.transpose(3, 2, 1, 0)), [a[0] for a in ax],
               [a[1] - a[0] for a in ax], O.EXTRAP_PERIODIC)
    E = O.uxb(w.extra["u"], B)
    kw = dict(q2m=w.q2m, tspan=w.tspan, dt=w.dt, alg=O.ALG_MULTISTEP, order_N=4, nsub=2,
              form=O.FORM_REFERENCE, save_everystep=False, nthreads=nthreads)
    def run(n):
        r = O.solve(w.xv0[:n], E, B, **kw)
        return r.elapsed, int(r.steps.sum())

    out = oracle_rate(run, len(w.xv0), nthreads, target_s,
                      f"FP64 MultistepBoris4(n=2) reference form, {w.nsteps} steps each")
    out["calibration"] = oracle_calibration()
    return out


# ---------------------------------------------------------------------------
# headline: gridded turbulent box
# ---------------------------------------------------------------------------
# (precision, multistep form) variants; the first is `value`: FP64 in the
# reference's own arithmetic (atan, sincos), bit-identical to the oracle
HEADLINE_VARIANTS = (("fp64", "reference"), ("fp32c", "reference"), ("fp64", "rational"),
                     ("fp32c", "rational"))


def _kernel_name(prec, form, layout):
    ns = "n=2" if form == "rational" else "reference form n=2"
    lay = "cells" if layout == "cells" else "nodes"
    if prec == "fp64":
        return f"k_fixed<S64G<UxB, {ns}, periodic, double, {lay}>>"
    return f"k_fixed<S32G<UxB, {ns}, periodic, {lay}>>"


def bench_turbulent(args, torch, TP, W, DeviceEnsemble):
    N, n, nsteps = args.particles, args.grid, args.nsteps
    log(f"[bench] generating {n}^3 turbulent field and {N} protons")
    w, A64 = W.turbulent_box(N=N, n=n, nsteps=nsteps, dtype=np.float64)
    ax = W.grid_axes(n, w.extra["L"])
    alg = TP.MultistepBoris(4, n=2)
    horizons = [h for h in (10, 100, 1000) if h < nsteps] + [nsteps]
    x0p = w.xv0.copy()
    x0p[:, :3] = np.nextafter(x0p[:, :3], np.inf)        # one ulp away, for the chaos floor
    res = {}
    for prec, storage in (("fp64", "f64"), ("fp32c", "f32")):
        A = A64 if storage == "f64" else A64.astype(np.float32)
        B = TP.GridField(A, *ax, bc=TP.PeriodicBC(), layout=args.layout)
        E = TP.UxBField(w.extra["u"], B)
        param = (w.q2m, TP.mi, E, B, TP.ZeroField())
        prob = TP.TraceProblem(w.xv0, w.tspan, param)
        for form in ("reference", "rational"):
            key = (prec, form)
            if args.variants and f"{prec}_{form}" not in args.variants and key != HEADLINE_VARIANTS[0]:
                continue
            first = key == HEADLINE_VARIANTS[0]
            log(f"[bench] {prec} {form}: uploading and warming up")
            ens = DeviceEnsemble(prob, alg, w.xv0, dt=w.dt, save_everystep=False, precision=prec, form=form)
            # the value's variant takes the full count; the others at most 5 timed solves
            steps = args.steps if first else min(args.steps, 5)
            wall, kms, st = timed_runs(ens.run, torch, args.warmup if first else 1, steps)
            assert st.particle_steps == N * nsteps, st.particle_steps
            r = dict(wall=wall, kernel_ms=kms, storage=storage, steps=steps,
                     final={nsteps: ens.final_states()[:, :3]})
            log(f"[bench] {prec} {form}: {wall * 1e3:.1f} ms/solve, kernel {kms:.1f} ms, "
                f"{N * nsteps / wall:.3e} particle-steps/s")
            del ens
            if first and args.accuracy:
                # the host-buffer entry point: PCIe copies in and out included
                t0 = time.perf_counter()
                TP.solve(prob, alg, dt=w.dt, save_everystep=False, precision=prec, form=form)
                r["host_wall"] = time.perf_counter() - t0
            if args.accuracy and form == "reference":
                for h in horizons:
                    starts = [("x0", w.xv0)] + ([("x0p", x0p)] if prec == "fp64" else [])
                    for tag, xv in starts:
                        if h == nsteps and tag == "x0":
                            continue
                        pr = TP.TraceProblem(xv, (0.0, h * w.dt), param)
                        e = DeviceEnsemble(pr, alg, xv, dt=w.dt, save_everystep=False, precision=prec,
                                           form=form)
                        e.run()
                        r.setdefault("final" if tag == "x0" else "final_p", {})[h] = e.final_states()[:, :3]
                        del e
            res[key] = r
        if prec == "fp64" and args.layout == "cells" and (not args.variants or
                                                          "fp64_reference_nodes" in args.variants):
            # the plain FP64 pusher: FP64 on the node layout the caller's array
            # is in (4-8 lines per gather), reference form
            Bn = TP.GridField(A, *ax, bc=TP.PeriodicBC(), layout="nodes")
            pn = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, TP.UxBField(w.extra["u"], Bn), Bn,
                                                  TP.ZeroField()))
            log("[bench] fp64 reference, node layout: uploading and warming up")
            ens = DeviceEnsemble(pn, alg, w.xv0, dt=w.dt, save_everystep=False, precision=prec,
                                 form="reference")
            wall, kms, st = timed_runs(ens.run, torch, 1, 1)
            assert st.particle_steps == N * nsteps, st.particle_steps
            log(f"[bench] fp64 reference, node layout: {wall * 1e3:.1f} ms/solve")
            res[("fp64", "reference_nodes")] = dict(wall=wall, kernel_ms=kms, storage=storage, steps=1,
                                                    final={nsteps: ens.final_states()[:, :3]})
            del ens, Bn, pn
        del B, E, prob
        torch.cuda.empty_cache()
    ps = N * nsteps
    vkey = HEADLINE_VARIANTS[0]
    rv = res[vkey]
    value = ps / rv["wall"]
    # algorithmic bytes of one launch: the gather per particle-step, the state
    # read and the two records written per particle
    bytes_launch = ps * GATHER_BYTES["f64"] + N * (STATE_BYTES + LEN_BYTES + 2 * RECORD_BYTES)
    achieved = bytes_launch / (rv["kernel_ms"] * 1e-3)
    variants = {}
    for key, r in res.items():
        variants[f"{key[0]}_{key[1]}"] = {
            "value": ps / r["wall"], "ms_per_solve": r["wall"] * 1e3, "kernel_ms": r["kernel_ms"],
            "timed_solves": r["steps"], "storage": r["storage"],
            "kernel": (_kernel_name("fp64", "reference", "nodes") if key[1] == "reference_nodes"
                       else _kernel_name(*key, args.layout)),
            "algorithmic_bytes_per_particle_step": GATHER_BYTES[r["storage"]],
            "hbm_frac": ps * GATHER_BYTES[r["storage"]] / (r["kernel_ms"] * 1e-3) / HBM_PEAK}
    out = {
        "metric": "particle-steps/s (gridded Boris) FP32-comp vs FP64 at <=1e-5 err; adjoint cost",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rv["wall"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^-5/3 turbulent field and particle states)",
        "config": {"workload": f"turbulent_box_{n}^3_multistep4_n2", "particles": N,
                   "grid": n, "steps_per_solve": nsteps, "alg": "MultistepBoris4(n=2)",
                   "multistep_form": "reference (atan, sincos, Taylor branch: multistep_boris.jl:94-120)",
                   "E": "-u x B in-kernel", "saves": "start+end", "field_layout": args.layout,
                   "field_storage": "f64",
                   "parallelism": "single GPU, one thread per particle"},
        "value_note": "FP64 in the reference's arithmetic, bit-identical to the oracle (the <=1e-5 "
                      "qualifier holds for FP64 only: the flow is chaotic, see fp32c_error)",
        "variants": variants,
    }
    if ("fp32c", "reference") in res:
        out["fp32c_over_fp64"] = res[("fp64", "reference")]["wall"] / res[("fp32c", "reference")]["wall"]
        if ("fp64", "reference_nodes") in res:
            out["fp32c_over_plain_fp64"] = {
                "value": res[("fp64", "reference_nodes")]["wall"] / res[("fp32c", "reference")]["wall"],
                "note": "FP32C (float cells) against the plain FP64 pusher: FP64 storage and arithmetic "
                        "on the node layout the caller's array is in, reference form; fp32c_over_fp64 "
                        "compares both on the cell layout"}
    if "host_wall" in rv:
        out["host_buffers"] = {"value": ps / rv["host_wall"], "ms_per_solve": rv["host_wall"] * 1e3,
                               "note": "tpamd_solve from host arrays (FP64, reference form): PCIe "
                                       "copies and host output allocation included"}
    fin = {k: r["final"][nsteps] for k, r in res.items()}
    if ("fp64", "rational") in fin:
        out["rational_vs_reference_fp64"] = rel_pos_err(fin[("fp64", "rational")], fin[("fp64", "reference")])
    if args.accuracy and ("fp32c", "reference") in res:
        acc = {}
        x64 = res[("fp64", "reference")]
        x32 = res[("fp32c", "reference")]
        for h in horizons:
            acc[str(h)] = {"fp32c_vs_fp64": rel_pos_err(x32["final"][h], x64["final"][h]),
                           "fp64_one_ulp_floor": rel_pos_err(x64["final_p"][h], x64["final"][h])}
        le = [h for h in horizons if acc[str(h)]["fp32c_vs_fp64"]["median"] <= 1e-5]
        out["fp32c_error"] = {
            "metric": "|x_fp32c - x_fp64| / |x_fp64| of the final positions, per horizon (steps), "
                      "reference form",
            "chaos_floor": "fp64_one_ulp_floor = FP64 started one ulp away in position vs FP64",
            "longest_horizon_with_median_le_1e-5": max(le) if le else None,
            "horizons": acc}
    out["roofline"] = {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                       "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None,
                       "kernel": _kernel_name(*vkey, args.layout),
                       "kernel_ms": rv["kernel_ms"],
                       "algorithmic_bytes_per_particle_step": GATHER_BYTES["f64"],
                       "algorithmic_bytes_per_launch": bytes_launch}
    # HBM traffic per launch from the rocprofv3 --pmc passes recorded under
    # profiles/ (scaled by the copy-probe calibration): FETCH_SIZE per
    # particle-step times this launch's particle-steps, plus WRITE_SIZE per
    # particle (the two records) times its particles
    prof = os.path.join(ROOT, "profiles", f"traffic_turbulent_box_{n}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            t = json.load(f)
        v = t.get("variants", {}).get(f"{vkey[0]}_{vkey[1]}")
        if t.get("layout") == args.layout and v:
            out["roofline"]["traffic"] = (ps * v["fetch_bytes_per_particle_step"] +
                                          N * v["write_bytes_per_particle"])
            out["roofline"]["traffic_source"] = os.path.relpath(prof, ROOT)
    if args.cpu_baseline:
        log("[bench] CPU baseline (oracle, host cores)")
        out["cpu_baseline"] = cpu_baseline(w, A64, ax, cpu_threads())
    else:
        out["cpu_baseline"] = None
    return out


# ---------------------------------------------------------------------------
# secondary: plain Boris in a uniform field (the CPU-runnable workload)
# ---------------------------------------------------------------------------
def bench_gyration(args, torch, TP, W, DeviceEnsemble):
    """1,024 protons, uniform B = (0, 0, 1e-8) T, E = 0, positions in
    [-1, 1]^3 m, Maxwellian velocities (thermal speed 1e5 m/s), dt =
    gyroperiod/64, 6,400 steps, a record every 10 steps: FP64 and FP32C on the
    device, the oracle over the whole ensemble on the host cores, FP64
    bit-identity and |v|^2 drift."""
    import oracle as O

    w = W.uniform_gyration(N=1024)
    N = len(w.xv0)
    prob = TP.TraceProblem(w.xv0, w.tspan, (w.q2m, TP.mi, TP.ZeroField(), TP.UniformField(w.extra["B"]),
                                            TP.ZeroField()))
    out = {"workload": "uniform_gyration_1024", "particles": N, "steps": w.nsteps,
           "savestepinterval": w.savestepinterval, "alg": "Boris"}
    res = {}
    for prec in ("fp64", "fp32c"):
        ens = DeviceEnsemble(prob, TP.Boris(), w.xv0, dt=w.dt, savestepinterval=w.savestepinterval,
                             precision=prec)
        wall, kms, st = timed_runs(ens.run, torch, 2, max(args.aux_steps, 5))
        u, t, ln, _r = ens.host()
        res[prec] = (u, t, ln)
        out[prec] = {"value": st.particle_steps / wall, "unit": "particle-steps/s", "ms_per_solve": wall * 1e3,
                     "kernel_ms": kms, "records_per_particle": int(ln[0]),
                     "note": "6.6e6 particle-steps: launch and record copy dominate at this size"}
        del ens
    o = O.solve(w.xv0, O.zero(), O.uniform(w.extra["B"]), q2m=w.q2m, tspan=w.tspan, dt=w.dt,
                savestepinterval=w.savestepinterval, nthreads=cpu_threads())
    u64, t64, ln64 = res["fp64"]
    out["fp64_bit_identical_to_oracle"] = bool(np.array_equal(ln64, o.len) and
                                               np.array_equal(u64[:, :o.u.shape[1]], o.u) and
                                               np.array_equal(t64[:, :o.t.shape[1]], o.t))
    v0 = np.sum(w.xv0[:, 3:] ** 2, axis=1)
    for prec, (u, _t, ln) in res.items():
        vend = u[np.arange(N), ln - 1, 3:]
        d = np.abs(np.sum(vend ** 2, axis=1) - v0) / v0
        out[prec]["energy_drift"] = {"median": float(np.median(d)), "max": float(np.max(d))}
    out["fp32c_vs_fp64"] = rel_pos_err(res["fp32c"][0][np.arange(N), res["fp32c"][2] - 1, :3],
                                       u64[np.arange(N), ln64 - 1, :3])
    steps = int(o.steps.sum())
    out["cpu_baseline"] = {"value": steps / o.elapsed, "unit": "particle-steps/s", "cores": cpu_threads(),
                           "kind": "port", "host": host_info(),
                           "sample": f"the whole ensemble: {N} particles x {w.nsteps} steps, FP64 Boris, "
                                     f"{o.elapsed:.3f} s wall"}
    log(f"[bench] gyration: fp64 {out['fp64']['value']:.3e}, fp32c {out['fp32c']['value']:.3e}, "
        f"cpu {out['cpu_baseline']['value']:.3e} particle-steps/s")
    return out


# ---------------------------------------------------------------------------
# secondary: analytic dipole (FLOP-bound)
# ---------------------------------------------------------------------------
def bench_dipole(args, torch, TP, W, DeviceEnsemble):
    w = W.dipole_belt(N=args.dipole_particles, nsteps=10000)
    log(f"[bench] dipole: {len(w.xv0)} protons x {w.nsteps} steps")
    prob = TP.TraceProblem(w.xv0, (0.0, 0.0), (w.q2m, TP.mi, TP.ZeroField(),
                                               TP.DipoleField(w.extra["b0"], w.extra["R"]),
                                               TP.ZeroField()))
    out = {"workload": "dipole_belt_L4-7_10keV-10MeV", "particles": len(w.xv0),
           "steps": w.nsteps, "alg": "Boris", "dt": "T_g(x0)/32 per particle",
           "boundary": "removed when |x| < 1 R_E", "flop_per_particle_step": DIPOLE_FLOP}
    fin = {}
    for prec, peak in (("fp64", FP64_PEAK), ("fp32c", FP32_PEAK)):
        ens = DeviceEnsemble(prob, TP.Boris(), w.xv0, dts=w.dts, nsteps=w.nsteps,
                             save_everystep=False, isoutside=TP.InsideSphere(TP.Re),
                             precision=prec)
        wall, kms, st = timed_runs(ens.run, torch, 1, args.aux_steps)
        u, tt, ln, ret = ens.host()
        last = np.maximum(ln - 1, 0)
        ar = np.arange(len(ln))
        # With save_everystep = false a removed particle still gets its closing
        # record and returns Success (boris.jl:523-551): it is removed when that
        # record's time is before its t1 (a survivor's closing record is at t1)
        steps_done = np.rint(tt[ar, last] / w.dts).astype(np.int64)
        removed = steps_done < w.nsteps
        fin[prec] = (u[ar, last], removed)
        rate = st.particle_steps / wall
        flops = st.particle_steps * DIPOLE_FLOP / (kms * 1e-3)
        life = steps_done[removed]
        out[prec] = {"value": rate, "unit": "particle-steps/s", "ms_per_solve": wall * 1e3,
                     "kernel_ms": kms, "particle_steps": int(st.particle_steps),
                     "particle_steps_from_closing_records": int(np.sum(steps_done)),
                     "removed": {"count": int(removed.sum()), "fraction": float(removed.mean()),
                                 "lifetime_steps": ({"min": int(life.min()), "median": float(np.median(life)),
                                                     "max": int(life.max())} if life.size else None),
                                 "rule": "closing record before t1 (retcode stays Success with "
                                         "save_everystep = false, boris.jl:523-551)"},
                     "roofline": {"bound": "valu", "achieved": flops / 1e12, "peak": peak / 1e12,
                                  "unit": "TFLOP/s", "frac": flops / peak,
                                  "kernel": "k_fixed<DipoleOnly<S64A>, Boris>" if prec == "fp64"
                                  else "k_fixed<DipoleOnly<S32CA>, Boris>"}}
        if args.valu:
            # against the FMA rate measured on this GPU (FP32 without packed math)
            meas = args.valu["fp64_fma_flops_per_s" if prec == "fp64" else "fp32_fma_flops_per_s"]
            out[prec]["roofline"]["measured_fma_peak"] = meas / 1e12
            out[prec]["roofline"]["frac_of_measured_peak"] = flops / meas
        log(f"[bench] dipole {prec}: {rate:.3e} particle-steps/s ({flops / 1e12:.1f} TFLOP/s)")
        del ens
    u64, rm64 = fin["fp64"]
    u32, rm32 = fin["fp32c"]
    both = ~rm64 & ~rm32                      # compared only where both precisions survive
    out["removed_in_one_precision_only"] = int(np.sum(rm64 != rm32))
    v0 = np.sum(w.xv0[:, 3:] ** 2, axis=1)
    for prec, (u, _r) in fin.items():
        drift = np.abs(np.sum(u[:, 3:] ** 2, axis=1) - v0) / v0
        out[prec]["energy_drift"] = {"median": float(np.median(drift[both])),
                                     "max": float(np.max(drift[both]))}
    out["fp32c_vs_fp64"] = rel_pos_err(u32[both, :3], u64[both, :3])
    out["fp32c_vs_fp64"]["particles"] = int(both.sum())
    out["fp32c_over_fp64"] = out["fp32c"]["value"] / out["fp64"]["value"]
    if args.cpu_baseline:
        import oracle as O

        Bo = O.dipole(w.extra["b0"], w.extra["R"])

        def run(n):
            r = O.solve(w.xv0[:n], O.zero(), Bo, q2m=w.q2m, tspan=(0.0, 0.0), dt_pp=w.dts[:n],
                        nsteps_pp=w.nsteps, save_everystep=False, bnd_flags=2, r_in=TP.Re,
                        nthreads=cpu_threads())
            return r.elapsed, int(r.steps.sum())

        log("[bench] dipole CPU baseline (oracle)")
        out["cpu_baseline"] = oracle_rate(run, len(w.xv0), cpu_threads(), args.cpu_target_s,
                                          f"FP64 Boris, per-particle dt, {w.nsteps} steps, removal "
                                          f"inside 1 R_E")
    return out


# ---------------------------------------------------------------------------
# secondary: batched uneven ensemble in one persistent graph launch
# ---------------------------------------------------------------------------
def bench_batch(args, torch, TP, W):
    from testparticle_amd.device import DeviceBatch

    log(f"[bench] batch: generating 64 instances, ~{args.batch_total:.2e} protons")
    inst, tspan = W.batched_turbulence(n_inst=64, total=args.batch_total)
    probs = [TP.TraceProblem(xv, tspan, (TP.qi / TP.mi, TP.mi, TP.ZeroField(),
                                         TP.GridField(A, *axes, layout=args.layout), TP.ZeroField()))
             for A, axes, xv in inst]
    N = sum(len(xv) for _, _, xv in inst)
    out = {"workload": "batched_turbulence_64", "instances": 64, "particles": N,
           "grid_edges": sorted({int(A.shape[1]) for A, _, _ in inst}),
           "grid_bytes_f32": int(sum(A.size * 4 for A, _, _ in inst)),
           "alg": "AdaptiveBoris(safety=0.1)", "max_steps": 2000, "boundary": "exit at box faces",
           "field_layout": args.layout}
    fin = {}
    for prec in ("fp32c", "fp64"):
        b = DeviceBatch(probs, TP.AdaptiveBoris(safety=0.1), precision=prec,
                        isoutside=TP.OutsideGrid(), max_steps=2000, save_everystep=False)
        b.run()
        bs, ppt = b.autotune()
        wall, kms, st = timed_runs(b.run, torch, 1, args.aux_steps)
        fin[prec] = b.final_states()[:, :3]
        rate = st.particle_steps / wall
        # particles live ~1.4 steps, so the per-particle state and record
        # bytes weigh as much as the per-step gathers (float grids: 96 B).
        # AdaptiveBoris gathers once at the start (dt from |B(x0)|) and once
        # after every completed step: particle_steps + particles gathers.
        nrec = int(b.len.sum().item())
        algo = ((st.particle_steps + N) * GATHER_BYTES["f32"] + N * (STATE_BYTES + LEN_BYTES)
                + nrec * RECORD_BYTES)
        achieved = algo / (kms * 1e-3)
        out[prec] = {"value": rate, "unit": "particle-steps/s", "ms_per_run": wall * 1e3,
                     "kernel_ms": kms, "particle_steps": int(st.particle_steps),
                     "particles_per_s": N / wall, "records": nrec,
                     "autotuned": {"block": bs, "particles_per_lane": ppt},
                     "roofline": {"bound": "hbm", "achieved": achieved / 1e9,
                                  "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                                  "frac": achieved / HBM_PEAK, "kernel": "k_batch (adaptive)",
                                  "algorithmic_bytes_per_launch": int(algo),
                                  "bytes": "96 B per gather (float cell; particle_steps + particles "
                                           "gathers) + 56 B per particle (state in, length and return "
                                           "code) + 56 B per record"}}
        log(f"[bench] batch {prec}: {rate:.3e} particle-steps/s, block {bs}, ppt {ppt}")
        b.close()
        del b
        torch.cuda.empty_cache()
    out["fp32c_vs_fp64"] = rel_pos_err(fin["fp32c"], fin["fp64"])
    out["fp32c_over_fp64"] = out["fp32c"]["value"] / out["fp64"]["value"]
    if args.cpu_baseline:
        import oracle as O

        def run(frac):
            # the first ceil(frac N_k) particles of every instance, each
            # instance solved by the oracle on its own grid (fill-NaN
            # interpolant, exits at the grid's node box)
            el, steps, n = 0.0, 0, 0
            for A, axes, xv in inst:
                k = max(1, int(math.ceil(frac * len(xv))))
                data = np.ascontiguousarray(A.transpose(3, 2, 1, 0))
                org = [a[0] for a in axes]
                sp = [(a[-1] - a[0]) / (len(a) - 1) for a in axes]
                g = O.grid(data, org, sp, O.EXTRAP_FILL_NAN)
                hi = [org[d] + float(len(axes[d]) - 1) / (1.0 / sp[d]) for d in range(3)]
                r = O.solve(xv[:k], O.zero(), g, q2m=TP.qi / TP.mi, tspan=tspan, adaptive=True,
                            safety=0.1, max_steps=2000, nout_cap=4, save_everystep=False,
                            bnd_flags=8, box_lo=org, box_hi=hi, nthreads=cpu_threads())
                el += r.elapsed
                steps += int(r.steps.sum())
                n += k
            return el, steps, n

        log("[bench] batch CPU baseline (oracle)")
        el, steps, n = run(1e-4)
        frac = min(1.0, 1e-4 * args.cpu_target_s / max(el, 1e-6))
        el, steps, n = run(frac)
        if frac < 1.0 and el < 0.25 * args.cpu_target_s:
            # the tiny calibration run is dominated by per-instance set-up
            frac = min(1.0, frac * args.cpu_target_s / max(el, 1e-6))
            el, steps, n = run(frac)
        out["cpu_baseline"] = {"value": steps / el, "unit": "particle-steps/s", "cores": cpu_threads(),
                               "kind": "port", "host": host_info(), "particles_per_s": n / el,
                               "sample": f"the first {frac:.2e} of each instance's particles: {n} of "
                                         f"{N}, {steps} particle-steps, FP64 AdaptiveBoris(0.1), "
                                         f"{el:.1f} s wall"}
    return out


# ---------------------------------------------------------------------------
# secondary: adjoint cost and gradient check
# ---------------------------------------------------------------------------
def bench_adjoint(args, torch, TP, W, DeviceEnsemble):
    import oracle as O

    n, L, B0, N, nsteps, K = 64, 1.0e7, 5.0e-9, args.adjoint_particles, 2048, 64
    A = W.turbulent_field(n, L, B0, 0.3, 55, dtype=np.float64)
    A32 = A.astype(np.float32).astype(np.float64)       # the values float storage holds
    ax = W.grid_axes(n, L)
    rng = np.random.default_rng(56)
    x = rng.uniform(0, L, size=(N, 3))
    v = W.isotropic_directions(rng, N) * math.sqrt(2 * 1.0e4 * TP.qi / TP.mi)
    xv0 = np.ascontiguousarray(np.hstack([x, v]))
    tg = rng.uniform(0, L, size=(N, 3))
    dt = 2 * math.pi * TP.mi / (TP.qi * B0) / 40
    q2m = TP.qi / TP.mi
    log(f"[bench] adjoint: {n}^3 grid, {N} protons x {nsteps} steps, checkpoints every {K}")

    def prob_of(Av, storage, h=nsteps, layout="cells"):
        B = TP.GridField(Av, *ax, bc=TP.PeriodicBC(), storage=storage, layout=layout)
        return TP.TraceProblem(xv0, (0.0, h * dt), (q2m, TP.mi, TP.ZeroField(), B, TP.ZeroField()))

    def grad(prec, h):
        storage = "f32" if prec == "fp32c" else "f64"
        return TP.loss_gradient(prob_of(A, storage, h), dt=dt, targets=tg, precision=prec, ckpt_interval=K)

    out = {"workload": "adjoint_turbulent_64^3", "particles": N, "steps": nsteps,
           "checkpoint_interval": K, "dt": "T_g(B0)/40", "box": "1e7 m periodic",
           "loss": "mean |x(T) - x*|^2",
           "precisions": {"fp32c": "float storage, forward / recompute / reverse steps in float, FP64 "
                                   "checkpoints, float adjoint state with a per-particle power-of-two "
                                   "exponent, dL/dB summed in double",
                          "fp64": "double storage and arithmetic"}}
    for prec, storage in (("fp32c", "f32"), ("fp64", "f64")):
        # plain forward solve of the same problem, best of both layouts
        fwd = {}
        for layout in ("nodes", "cells"):
            ens = DeviceEnsemble(prob_of(A, storage, layout=layout), TP.Boris(), xv0, dt=dt,
                                 save_everystep=False, save_start=False, precision=prec)
            _w, kms, _st = timed_runs(ens.run, torch, 1, args.aux_steps)
            fwd[layout] = kms
            del ens
        rs = []
        err = None
        for _ in range(1 + args.aux_steps):
            try:
                rs.append(grad(prec, nsteps))
            except FloatingPointError as e:          # reported, never returned as a result
                err = str(e)
                break
        if err:
            out[prec] = {"error": err}
            continue
        rs = rs[1:]
        f_ms = float(np.mean([r.forward_ms for r in rs]))
        r_ms = float(np.mean([r.reverse_ms for r in rs]))
        best = min(fwd.values())
        g = rs[-1]
        # the reverse sweep's roofline: per particle-step two gathers
        # (recompute and reverse), the segment state written and read back,
        # and the dL/dB record written, sorted (read + write) and summed (read)
        abytes = ADJ_REVERSE_BYTES[prec]
        psteps = N * nsteps
        roof = {"bound": "hbm", "kernel": "reverse sweep: k_adj_reverse + k_adj_localsort + k_adj_bins_seg",
                "achieved": psteps * abytes / (r_ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": psteps * abytes / (r_ms * 1e-3) / HBM_PEAK,
                "algorithmic_bytes_per_particle_step": abytes, "traffic": None,
                "forward_frac": psteps * GATHER_BYTES["f32" if prec == "fp32c" else "f64"]
                                / (best * 1e-3) / HBM_PEAK}
        prof = os.path.join(ROOT, "profiles", "traffic_adjoint_turbulent_64.json")
        if os.path.exists(prof):
            with open(prof) as f:
                t = json.load(f)
            roof["traffic"] = psteps * t["kernels"][prec]["total_bytes_per_particle_step"]
            roof["traffic_source"] = os.path.relpath(prof, ROOT)
        out[prec] = {"forward_solve_ms": fwd, "adjoint_forward_ms": f_ms, "adjoint_reverse_ms": r_ms,
                     "cost_over_forward": (f_ms + r_ms) / best, "roofline": roof,
                     "cost_floor_by_bytes": (GATHER_BYTES["f32" if prec == "fp32c" else "f64"] + abytes)
                                            / GATHER_BYTES["f32" if prec == "fp32c" else "f64"],
                     "loss": g.loss,
                     "max_abs_dL_dB": float(np.max(np.abs(g.dL_dB))),
                     "max_abs_dL_dxv0": float(np.max(np.abs(g.dL_dxv0))),
                     "nonfinite_gradients": int(np.sum(~np.isfinite(g.dL_dB)) + np.sum(~np.isfinite(g.dL_dxv0)))}
        log(f"[bench] adjoint {prec}: forward solve {best:.2f} ms, adjoint {f_ms:.2f} + {r_ms:.2f} ms "
            f"= {(f_ms + r_ms) / best:.2f}x")

    # Gradient accuracy by horizon (steps), for both precisions:
    #   dxv0_vs_duals: dL/d(x0, v0) of 64 sampled particles against forward-mode
    #     duals through the oracle (max over particles of |g - g*| / |g*|);
    #   dB_dir_vs_duals: <dL/dB, d> along two dense random directions against
    #     the oracle's duals over all particles (max relative error);
    #   dB_vs_fd: 32 touched grid values against central differences of the
    #     FP64 solve, h = 1e-6 B0, formed per particle (median / max relative).
    #     The FP64 adjoint equals the duals (above), so its column is the error
    #     of the differences themselves: past ~200 steps the chaotic flow is
    #     too nonlinear for any difference step, and only the duals pin the
    #     gradient.
    # FP32C is compared with the duals of the float-stored field values.
    grids = {"fp64": O.grid(np.ascontiguousarray(A.transpose(3, 2, 1, 0)), [a[0] for a in ax],
                            [a[1] - a[0] for a in ax], O.EXTRAP_PERIODIC),
             "fp32c": O.grid(np.ascontiguousarray(A32.transpose(3, 2, 1, 0)), [a[0] for a in ax],
                             [a[1] - a[0] for a in ax], O.EXTRAP_PERIODIC)}
    idx = np.linspace(0, N - 1, 64).astype(int)
    nvals = 3 * n ** 3
    dirs = np.random.default_rng(60).normal(size=(2, nvals))
    acc = {"fp32c": {}, "fp64": {}}
    nth = cpu_threads()
    for H in args.adjoint_horizons:
        tspan = (0.0, H * dt)
        gs = {}
        for prec in ("fp32c", "fp64"):
            try:
                gs[prec] = grad(prec, H)
            except FloatingPointError as e:
                acc[prec][str(H)] = {"error": str(e)}
        picks = None
        fd = None
        if H <= args.adjoint_fd_max and "fp64" in gs:
            gabs = np.abs(gs["fp64"].dL_dB)
            cand = np.argwhere(gabs > 1e-3 * gabs.max())
            picks = cand[np.random.default_rng(57).choice(len(cand), 32, replace=False)]
            hB = 1e-6 * B0
            fd = []
            for c, i, j, k in picks:
                xs = []
                for sgn in (1.0, -1.0):
                    Ap = A.copy()
                    Ap[c, i, j, k] += sgn * hB
                    xs.append(TP.solve(prob_of(Ap, "f64", H), TP.Boris(), dt=dt, save_everystep=False,
                                       save_start=False).final_states()[:, :3])
                fd.append(float(np.mean(np.sum((xs[0] - xs[1]) * (xs[0] + xs[1] - 2 * tg), axis=1)))
                          / (2 * hB))
            fd = np.array(fd)
        for prec, r in gs.items():
            row = {}
            Bo = grids[prec]
            ref = np.zeros((len(idx), 6))
            for q, i in enumerate(idx):
                xT, J = O.boris_tangent(xv0[i], Bo, q2m=q2m, tspan=tspan, dt=dt)
                ref[q] = 2.0 / N * (xT - tg[i]) @ J
            e = np.linalg.norm(r.dL_dxv0[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
            row["dxv0_vs_duals"] = {"median": float(np.median(e)), "max": float(np.max(e))}
            if prec == "fp64" or H <= 256:
                gnode = r.dL_dB.transpose(3, 2, 1, 0).reshape(-1)
                _l, dd = O.loss_tangent(xv0, tg, Bo, q2m=q2m, tspan=tspan, dt=dt, dense=dirs, nthreads=nth)
                row["dB_dir_vs_duals_max"] = float(np.max(np.abs(dirs @ gnode - dd) / np.abs(dd)))
            if fd is not None:
                ga = np.array([r.dL_dB[c, i, j, k] for c, i, j, k in picks])
                rel = np.abs(ga - fd) / np.abs(fd)
                row["dB_vs_fd"] = {"median": float(np.median(rel)), "max": float(np.max(rel)), "nodes": 32,
                                   "h": 1e-6 * B0}
            row["max_abs_dL_dB"] = float(np.max(np.abs(r.dL_dB)))
            acc[prec][str(H)] = row
            log(f"[bench] adjoint {prec} {H} steps: {json.dumps(row)}")
    for prec in acc:
        if prec in out and "error" not in out[prec]:
            out[prec]["accuracy_by_horizon"] = acc[prec]
    return out


def _ceiling_program(name):
    """Runs one of the measured-ceiling programs built beside the library
    (gather_ceiling, valu_ceiling) as a child process and returns its JSON
    line.  None if it is not built or fails."""
    import subprocess

    exe = os.path.join(ROOT, "testparticle.jl_amd", "_build", name)
    if not os.path.exists(exe):
        return None
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    if r.returncode != 0:
        log(f"[bench] {name} failed: {r.stderr.strip()}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def gather_ceiling():
    """Measured ceiling of the random cell-record gather."""
    return _ceiling_program("gather_ceiling")


def valu_ceiling():
    """Measured FMA rate of the vector ALU, FP64 and FP32 (no packed math)."""
    return _ceiling_program("valu_ceiling")


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
    ap.add_argument("--cpu-target-s", type=float, default=8.0,
                    help="seconds of oracle work per secondary section's CPU baseline")
    ap.add_argument("--no-ceiling", dest="ceiling", action="store_false",
                    help="skip the measured gather ceiling")
    ap.add_argument("--no-accuracy", dest="accuracy", action="store_false",
                    help="skip the FP32C-vs-FP64 horizon runs")
    ap.add_argument("--variants", type=lambda t: [v for v in t.split(",") if v], default=None,
                    help="headline variants besides the value's (fp32c_reference, fp64_rational, "
                         "fp32c_rational, fp64_reference_nodes; default all; 'none' for none)")
    ap.add_argument("--sections", default=",".join(SECTIONS),
                    help="comma-separated subset of " + ",".join(SECTIONS))
    ap.add_argument("--aux-steps", type=int, default=2, help="timed runs per secondary section")
    ap.add_argument("--dipole-particles", type=int, default=1 << 20)
    ap.add_argument("--batch-total", type=int, default=20_000_000)
    ap.add_argument("--adjoint-particles", type=int, default=1 << 18)
    ap.add_argument("--adjoint-horizons", type=lambda t: [int(x) for x in t.split(",") if x],
                    default=[16, 64, 256, 2048], help="steps at which gradient accuracy is reported")
    ap.add_argument("--adjoint-fd-max", type=int, default=256,
                    help="longest horizon with a central-difference check")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: the ensemble pusher runs on a single GPU "
              "(--gpus 1); runs over several GPUs are not supported", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    sections = [s for s in args.sections.split(",") if s]
    bad = [s for s in sections if s not in SECTIONS]
    if bad:
        ap.error(f"unknown sections {bad}")

    ceiling = gather_ceiling() if args.ceiling and "headline" in sections else None
    if ceiling:
        log(f"[bench] measured gather ceiling: {ceiling['float_cell_gathers_per_s']:.3e} float-cell, "
            f"{ceiling['double_cell_gathers_per_s']:.3e} double-cell gathers/s")
    args.valu = valu_ceiling() if args.ceiling and "dipole" in sections else None
    if args.valu:
        log(f"[bench] measured FMA rate: {args.valu['fp64_fma_flops_per_s'] / 1e12:.1f} TFLOP/s FP64, "
            f"{args.valu['fp32_fma_flops_per_s'] / 1e12:.1f} TFLOP/s FP32")

    import torch

    import testparticle_amd as TP
    from testparticle_amd import workloads as W
    from testparticle_amd.device import DeviceEnsemble

    out = bench_turbulent(args, torch, TP, W, DeviceEnsemble) if "headline" in sections else {}
    if out and ceiling and args.layout == "cells":
        # one cell-record gather per particle-step
        out["roofline"]["measured_gather_ceiling"] = ceiling
        out["roofline"]["frac_of_measured_ceiling"] = {
            f"{k}": v["value"] / ceiling["float_cell_gathers_per_s" if v["storage"] == "f32"
                                         else "double_cell_gathers_per_s"]
            for k, v in out["variants"].items() if not k.endswith("_nodes")}   # cell gathers only
    if out:
        log("[bench] headline " + json.dumps(out))
    sec = {}
    runners = {"gyration": lambda: bench_gyration(args, torch, TP, W, DeviceEnsemble),
               "dipole": lambda: bench_dipole(args, torch, TP, W, DeviceEnsemble),
               "batch": lambda: bench_batch(args, torch, TP, W),
               "adjoint": lambda: bench_adjoint(args, torch, TP, W, DeviceEnsemble)}
    for name in SECTIONS[1:]:
        if name in sections:
            sec[name] = runners[name]()
            log(f"[bench] {name} " + json.dumps(sec[name]))
    if sec:
        out["secondary"] = sec
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
