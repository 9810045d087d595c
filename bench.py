#!/usr/bin/env python3
"""Benchmark of the force-evaluation + velocity-Verlet hot path on one MI355X.

Headline workload (one JSON line on stdout): all-pairs gravity, N = 65,536 equal-mass
bodies drawn from a seeded Plummer sphere (G = M = a = 1), velocity Verlet dt = 1e-3,
FP32 pair forces with Kahan accumulation, FP64 positions / velocities / integrator.
A "step" is one velocity-Verlet step (one force evaluation over all N(N-1) ordered
pairs plus the integrator update).  value = pair interactions/s = N(N-1) * steps / s,
with the state resident in HBM before the timed region starts.

The line also carries:
  roofline     -- the dominant kernel (k_grav_tiled): 20 flop per interaction divided by
                  its average duration, timed with HIP events on the library's stream
  cpu_baseline -- the FP64 oracle (a restatement of the reference's per-atom loop) on
                  one host core, timed on a bounded sample of targets
  secondary    -- Lennard-Jones N = 32,000 (FCC, rho* = 0.8442, T* = 0.9, r_c = 2.5) and
                  N = 864 (T* = 0.72) steps/s in the mixed and FP64 paths, with their
                  CPU-oracle baselines, force errors and energy drift

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-secondary]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nbodysimulator.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

PEAK_FP32 = 157.3     # TFLOP/s, MI355X published FP32 vector/matrix peak
PEAK_HBM = 8000.0     # GB/s
FLOP_GRAV = 20        # per ordered interaction (GPU Gems 3 ch. 31 convention)
FLOP_LJ = 29          # per ordered in-cutoff pair


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def cpu_gravity_sample(u, m, targets):
    """FP64 oracle accelerations of `targets` against all sources, and its wall time."""
    import oracle_lib as O
    lib = O.load()
    n = len(m)
    uu = np.ascontiguousarray(u)
    mm = np.ascontiguousarray(m)
    out = np.zeros((len(targets), 3))
    t0 = time.perf_counter()
    for k, i in enumerate(targets):
        lib.orc_grav_accel(O.dp(out[k]), O.dp(uu), int(i), n, O.dp(mm), 1.0)
    return out, time.perf_counter() - t0


def cpu_lj_sample(u, L, targets):
    import oracle_lib as O
    lib = O.load()
    n = len(u)
    uu = np.ascontiguousarray(u)
    ms = np.ones(n)
    idx = np.arange(n, dtype=np.int64)
    bc = O.make_bc(O.BC_CUBIC, L)
    out = np.zeros((len(targets), 3))
    t0 = time.perf_counter()
    for k, i in enumerate(targets):
        lib.orc_lj_accel(O.dp(out[k]), O.dp(uu), int(i), idx.ctypes.data_as(C.POINTER(C.c_int64)),
                         n, O.dp(ms), 1.0, 1.0, 2.5, C.byref(bc))
    return out, time.perf_counter() - t0


def mixed_accel(ctx, u):
    """a(u) of the mixed path, read back from one VV step of dt = 2^-60 from rest."""
    ctx.set_state(u, np.zeros_like(u))
    dt = 2.0 ** -60
    _, vf = ctx.vv_run(dt, 1, save_every=1)
    return (vf[1] - vf[0]) / dt


def bench_gravity(args, lib, gen):
    n = 65536
    u, v, m = gen.plummer(n, seed=2024)
    ctx = lib.Context(n, lib.SYS_POTENTIAL, m)
    ctx.add_gravitational(1.0)
    ctx.set_precision(lib.PREC_MIXED)
    dt = 1e-3
    # force error on 4096 sampled targets against the FP64 oracle (also the CPU sample)
    targets = np.random.default_rng(7).choice(n, 4096, replace=False)
    a_gpu = mixed_accel(ctx, u)[targets]
    a_ref, cpu_s = cpu_gravity_sample(u, m, targets)
    force_err = float(np.abs(a_gpu - a_ref).max() / np.abs(a_ref).max())
    # timed run
    ctx.set_state(u, v)
    ke0, _, pg0 = ctx.energies()
    ctx.vv_run(dt, args.warmup, seed=1)
    t0 = time.perf_counter()
    ctx.vv_run(dt, args.steps, seed=1)          # synchronous: returns after the stream drains
    elapsed = time.perf_counter() - t0
    ke1, _, pg1 = ctx.energies()
    e0, e1 = ke0 + pg0, ke1 + pg1
    pairs = n * (n - 1)
    value = pairs * args.steps / elapsed
    kms, kname = ctx.time_force_kernel(20)
    achieved = FLOP_GRAV * pairs / (kms * 1e-3) / 1e12
    sample_pairs = len(targets) * (n - 1)
    cpu_rate = sample_pairs / cpu_s
    return {
        "value": value,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "config": {"workload": "gravity, Plummer sphere N=65536 (G=M=a=1, equal masses), "
                               "velocity Verlet dt=1e-3, FP32 pair forces with Kahan sums, "
                               "FP64 state",
                   "n": n, "dt": dt, "steps_in_config": 100,
                   "force_error_vs_fp64_oracle": force_err,
                   "energy_drift_rel": abs(e1 - e0) / abs(e0),
                   "energy_drift_steps": args.warmup + args.steps},
        "roofline": {"bound": "valu", "kernel": kname, "achieved": achieved,
                     "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": traffic_from_profiles(kname),
                     "kernel_avg_ms": kms,
                     "flop_per_launch": FLOP_GRAV * pairs},
        "cpu_baseline": {"value": cpu_rate, "unit": "pair interactions/s", "cores": 1,
                         "kind": "port",
                         "sample": f"FP64 oracle (reference per-atom loop restated in C, "
                                   f"-O2 -ffp-contract=off) on {len(targets)} targets x {n} "
                                   f"sources, {cpu_s:.2f} s",
                         "steps_per_s": cpu_rate / pairs},
    }


def bench_lj(args, lib, gen, cells, T, steps, cpu_targets, fp64=False):
    s = gen.lj_liquid(cells, rho=0.8442, T=T, seed=5)
    n = s.u.shape[0]
    out = {}
    for mode in (("mixed", lib.PREC_MIXED),) + ((("fp64", lib.PREC_FP64_REF),) if fp64 else ()):
        ctx = lib.Context(n, lib.SYS_POTENTIAL, s.masses)
        ctx.set_boundary(lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
        ctx.add_lennard_jones(1.0, 1.0, 2.5)
        ctx.set_kb(1.0)
        ctx.set_precision(mode[1])
        ctx.set_state(s.u, s.v)
        ke0, pe0, _ = ctx.energies()
        nst = steps if mode[0] == "mixed" else max(steps // 10, 10)
        ctx.vv_run(0.005, 20)
        t0 = time.perf_counter()
        ctx.vv_run(0.005, nst)
        el = time.perf_counter() - t0
        ke1, pe1, _ = ctx.energies()
        pairs = ctx.count_pairs() if mode[0] == "mixed" else None
        rec = {"steps": nst, "steps_per_s": nst / el, "ms_per_step": 1e3 * el / nst,
               "energy_drift_rel": abs((ke1 + pe1) - (ke0 + pe0)) / abs(ke0 + pe0)}
        if pairs:
            rec["pairs_per_step"] = pairs
            rec["pair_interactions_per_s"] = pairs * nst / el
            kms, kname = ctx.time_force_kernel(50)
            rec["force_kernel"] = {"name": kname, "avg_ms": kms,
                                   "achieved_tflops": FLOP_LJ * pairs / (kms * 1e-3) / 1e12}
        out[mode[0]] = rec
    # force error on jittered positions (a perfect lattice has zero net force, so the
    # relative error there would measure rounding noise only)
    uj = s.u + 0.02 * np.random.default_rng(17).standard_normal(s.u.shape)
    targets = np.random.default_rng(3).choice(n, min(cpu_targets, n), replace=False)
    ref, cpu_s = cpu_lj_sample(uj, s.L, targets)
    ctx = lib.Context(n, lib.SYS_POTENTIAL, s.masses)
    ctx.set_boundary(lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_precision(lib.PREC_MIXED)
    a = mixed_accel(ctx, uj)[targets]
    cpu_step_s = cpu_s * n / len(targets)
    out["force_error_vs_fp64_oracle"] = float(np.abs(a - ref).max() / np.abs(ref).max())
    out["cpu_baseline"] = {"steps_per_s": 1.0 / cpu_step_s, "cores": 1, "kind": "port",
                           "sample": f"FP64 oracle LJ accelerations of {len(targets)} of {n} "
                                     f"atoms (all-pairs, as the reference), {cpu_s:.2f} s, "
                                     f"extrapolated to one step"}
    out["gpu_vs_cpu_steps"] = out["mixed"]["steps_per_s"] * cpu_step_s
    out["n"] = n
    return out


def profile_kernels(lib, gen):
    """Short launch sequence for hardware-counter runs: the gravity force kernel at the
    headline size and a plain copy kernel of known size (the byte-counter calibration)."""
    n = 65536
    u, v, m = gen.plummer(n, seed=2024)
    ctx = lib.Context(n, lib.SYS_POTENTIAL, m)
    ctx.add_gravitational(1.0)
    ctx.set_precision(lib.PREC_MIXED)
    ctx.set_state(u, v)
    kms, kname = ctx.time_force_kernel(10)
    nbytes = 1 << 30
    cms = ctx.copy_probe(nbytes, 5)
    print(json.dumps({"kernel": kname, "kernel_avg_ms": kms, "copy_bytes": nbytes,
                      "copy_avg_ms": cms, "copy_GBps": 2 * nbytes / (cms * 1e-3) / 1e9}),
          flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the dominant kernels (gravity N=65536 force, 10 launches; "
                         "a 1 GiB copy-kernel calibration, 5 launches) for rocprofv3 counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; the multi-GPU sharding is a design only "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    import nbodysimulator as nb
    from nbodysimulator import _lib as lib
    from nbodysimulator import generators as gen
    if nb.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        return 1
    if args.profile_kernels:
        return profile_kernels(lib, gen)
    log("[bench] gravity N=65536")
    g = bench_gravity(args, lib, gen)
    res = {
        "metric": "pair interactions/s; LJ steps/s learned-stride vs plain VV at equal energy drift",
        "value": g["value"],
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair forces, f64 state",
        "data": "synthetic (seeded Plummer sphere)",
        "config": g["config"],
        "roofline": g["roofline"],
        "cpu_baseline": g["cpu_baseline"],
    }
    if not args.no_secondary:
        log("[bench] LJ N=32000")
        res["secondary"] = {"lj_n32000_T0.9": bench_lj(args, lib, gen, 20, 0.9, 2000, 600)}
        log("[bench] LJ N=864")
        res["secondary"]["lj_n864_T0.72"] = bench_lj(args, lib, gen, 6, 0.72, 1000, 864,
                                                     fp64=True)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
