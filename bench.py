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
                  its average duration over the timed steps (a HIP event pair around each
                  of its launches, on the library's stream)
  cpu_baseline -- the FP64 oracle (a restatement of the reference's per-atom loop) on
                  one host core, timed on a bounded sample of targets
  secondary    -- Lennard-Jones N = 32,000 (FCC, rho* = 0.8442, T* = 0.9, r_c = 2.5) and
                  N = 864 (T* = 0.72) steps/s in the mixed and FP64 paths; SPC/Fw water
                  (4,096 molecules, reaction field 0.9 nm, Berendsen 300 K); 512 batched
                  LJ replicas; the learned 8-dt stride on LJ N = 32,000 (training, held-out
                  error, simulated time per second against plain velocity Verlet at equal
                  drift, the stride sweep 4 and 2, and each stride without the network);
                  each with its CPU-oracle baseline, force error and 10k-step energy drift

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-secondary] [--learn-traj T]
    python bench.py --profile-kernels     (short launch sequence for rocprofv3 counters)
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
PEAK_BF16_MFMA = 2500.0  # TFLOP/s, dense BF16 MFMA
FLOP_GRAV = 20        # per ordered interaction (GPU Gems 3 ch. 31 convention)
FLOP_LJ = 29          # per ordered in-cutoff pair


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_from_profiles(kernel: str, shape: str):
    """HBM bytes per launch of `kernel` at tile shape `shape` ("<targets per thread>x
    <source chunks>") from the committed rocprofv3 counter summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("shapes", {}).get(shape, {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def host_cores():
    """Cores of the host (os.cpu_count) and those this process may run on."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"host_cores_total": os.cpu_count(), "host_cores_available": avail}


def workload_traffic(key: str, field: str = "hbm_bytes_per_launch"):
    """A field of workload `key` (lj32000, water, batch, mlp_infer, lr_chain) in the
    committed rocprofv3 counter summary of the benchmark workloads
    (profiles/workload_traffic.json; None if absent)."""
    path = os.path.join(ROOT, "profiles", "workload_traffic.json")
    if not key or not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get("kernels", {}).get(key, {}).get(field)
    except (OSError, ValueError):
        return None


def roofline(bound, kernel, work, kernel_ms, unit="TFLOP/s", peak=PEAK_FP32, traffic_key=None,
             **extra):
    """Roofline object: `work` algorithmic flop (or bytes, unit GB/s) per launch divided by
    the kernel's average launch duration, against the MI355X peak of that roof; traffic =
    HBM bytes per launch from the committed counter summary of workload `traffic_key`."""
    scale = 1e12 if unit == "TFLOP/s" else 1e9
    achieved = work / (kernel_ms * 1e-3) / scale
    out = {"bound": bound, "kernel": kernel, "achieved": achieved, "peak": peak, "unit": unit,
           "frac": achieved / peak, "traffic": workload_traffic(traffic_key),
           "kernel_avg_ms": kernel_ms, "work_per_launch": work}
    out.update(extra)
    return out


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


def gravity_issue_model(ctx, pairs, kernel_ms, equal_masses):
    """Issue-limited time of k_grav_tiled from the measured issue rates of its
    instruction kinds (nbx_valu_probe, wave64 instructions per SIMD per ns): per pair
    3 v_sub + 3 v_mul (2 when every body has the same mass: the kernel then leaves the
    mass out of the pair term) at the v_mul rate, 6 FMA at the v_fmac rate, one
    v_rsq_f32."""
    r_fma, r_mul, r_rsq = (ctx.valu_probe(k, 5) for k in (0, 1, 2))
    simds = 4 * 256
    muls = 2 if equal_masses else 3
    ns_per_wave_pair = (3.0 + muls) / r_mul + 6.0 / r_fma + 1.0 / r_rsq  # per 64 pairs, per SIMD
    bound_ms = pairs / 64.0 * ns_per_wave_pair / simds * 1e-6
    return {"fmac_per_simd_ns": r_fma, "mul_per_simd_ns": r_mul, "rsq_per_simd_ns": r_rsq,
            "instructions_per_pair": f"3 sub, {muls} mul, 6 fma, 1 v_rsq_f32",
            "issue_bound_ms": bound_ms, "frac_of_issue_bound": bound_ms / kernel_ms}


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
    # timed run, with the tile shape (targets per thread, source chunks) from the autotuner
    ctx.set_state(u, v)
    tuned = ctx.autotune(dt, 20)
    # the configured 100 steps from the initial state: energy drift over exactly them
    ke_a, _, pg_a = ctx.energies()
    ctx.vv_run(dt, 100, seed=1)
    ke_b, _, pg_b = ctx.energies()
    drift100 = abs((ke_b + pg_b) - (ke_a + pg_a)) / abs(ke_a + pg_a)
    ctx.set_state(u, v)
    ke0, _, pg0 = ctx.energies()
    ctx.vv_run(dt, args.warmup, seed=1)
    # HIP event pairs around every k_grav_tiled launch of the timed steps (on the
    # library's stream; the step graphs are off while they record: at 1 ms per kernel the
    # host stays ahead of the GPU either way)
    ctx.set_kernel_timing(True)
    t0 = time.perf_counter()
    ctx.vv_run(dt, args.steps, seed=1)          # synchronous: returns after the stream drains
    elapsed = time.perf_counter() - t0
    kms, klaunches, kname = ctx.kernel_timing()
    ctx.set_kernel_timing(False)
    ke1, _, pg1 = ctx.energies()
    e0, e1 = ke0 + pg0, ke1 + pg1
    pairs = n * (n - 1)
    value = pairs * args.steps / elapsed
    kms_b2b, _ = ctx.time_force_kernel(20)
    # the same run continued to 10k steps for the long-run energy drift
    drift10k = None
    if not args.no_secondary and not args.no_long_drift:
        ctx.vv_run(dt, 10000 - args.steps - args.warmup)
        ke2, _, pg2 = ctx.energies()
        drift10k = abs((ke2 + pg2) - e0) / abs(e0)
    achieved = FLOP_GRAV * pairs / (kms * 1e-3) / 1e12
    issue = gravity_issue_model(ctx, pairs, kms, bool(np.all(m == m[0])))
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
                   "energy_drift_rel_100_steps": drift100,
                   "energy_drift_rel_timed_run": abs(e1 - e0) / abs(e0),
                   "energy_drift_timed_run_steps": args.warmup + args.steps,
                   "energy_drift_rel_10k": drift10k,
                   "autotuned": {"targets_per_thread": int(tuned["shape0"]),
                                 "source_chunks": int(tuned["shape1"]),
                                 "ms_per_step_default": tuned["entry_ms_per_step"]}},
        "roofline": {"bound": "valu", "kernel": kname, "achieved": achieved,
                     "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32,
                     "traffic": traffic_from_profiles(
                         kname, f"{int(tuned['shape0'])}x{int(tuned['shape1'])}"),
                     "kernel_avg_ms": kms,
                     "kernel_launches_timed": klaunches,
                     "kernel_avg_ms_back_to_back": kms_b2b,
                     "flop_per_launch": FLOP_GRAV * pairs,
                     "issue_model": issue},
        "cpu_baseline": {"value": cpu_rate, "unit": "pair interactions/s", "cores": 1,
                         "kind": "port", **host_cores(),
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
        tuned = None
        if mode[0] == "mixed":
            # launch shape (lanes per atom, list skin) picked by the autotuner; 1,000
            # trial steps per candidate so that small skins pay their rebuilds in the trial
            tuned = ctx.autotune(0.005, 1000)
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
        if tuned:
            rec["autotuned"] = {"lanes": int(tuned["shape0"]), "skin": tuned["shape1"],
                                "build_every": tuned["build_every"],
                                "ms_per_step_default": tuned["entry_ms_per_step"]}
        if pairs:
            rec["pairs_per_step"] = pairs
            rec["pair_interactions_per_s"] = pairs * nst / el
            kms, kname = ctx.time_force_kernel(50)
            rec["roofline"] = roofline("valu", kname, FLOP_LJ * pairs, kms,
                                       traffic_key="lj32000" if n == 32000 else None,
                                       flop_per_pair=FLOP_LJ)
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
                           **host_cores(),
                           "sample": f"FP64 oracle LJ accelerations of {len(targets)} of {n} "
                                     f"atoms (all-pairs, as the reference), {cpu_s:.2f} s, "
                                     f"extrapolated to one step"}
    out["gpu_vs_cpu_steps"] = out["mixed"]["steps_per_s"] * cpu_step_s
    out["n"] = n
    return out


def oracle_system_water(w, Rc):
    import oracle_lib as O
    from nbodysimulator import generators as gen
    P = gen.SPCFW
    orc = O.System(w.masses, charges=w.charges, water=True, bc=O.make_bc(O.BC_CUBIC, w.L),
                   kb=P["kb"])
    orc.water_params((P["eps"], P["sigma"], Rc), O.TERM_RF, (P["k_el"], Rc, math.inf),
                     P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    return orc


def water_context(lib, gen, w, Rc, thermostat):
    P = gen.SPCFW
    n = len(w.masses)
    ctx = lib.Context(n, lib.SYS_WATER, w.masses, w.charges)
    ctx.set_boundary(lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
    ctx.set_kb(P["kb"])
    ctx.add_lennard_jones(P["eps"], P["sigma"], Rc)
    ctx.add_reaction_field(P["k_el"], Rc, math.inf)
    ctx.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    ctx.set_precision(lib.PREC_MIXED)
    if thermostat:
        ctx.set_thermostat(lib.THERM_BERENDSEN, 300.0, 0.1)
    ctx.set_state(w.u, w.v)
    return ctx


def water_pair_counts(u, L, rc):
    """Ordered pairs within rc that the non-bonded kernel evaluates (every atom pair minus
    the 6 ordered intramolecular pairs of each molecule) and ordered O-O pairs within rc
    (the Lennard-Jones ones), from a periodic k-d tree on the host."""
    from scipy.spatial import cKDTree
    w = np.mod(np.asarray(u), L)
    w[w >= L] -= L
    n = len(w)
    all_pairs = 2 * len(cKDTree(w, boxsize=L).query_pairs(rc))
    oo = 2 * len(cKDTree(w[0::3], boxsize=L).query_pairs(rc))
    return all_pairs - 6 * (n // 3), oo


def bench_water(args, lib, gen):
    """SPC/Fw water, 4096 molecules (16^3 lattice, random orientations, 0.997 g/cm^3),
    reaction field and LJ cutoff 0.9 nm, Berendsen 300 K (tau 0.1 ps), dt 0.5 fs."""
    w = gen.water_box(per_side=16, seed=7)
    n = len(w.masses)
    Rc = 0.9
    dt, steps = 0.0005, 2000
    ctx = water_context(lib, gen, w, Rc, True)
    tuned = ctx.autotune(dt, 400)
    ctx.vv_run(dt, 20)
    t0 = time.perf_counter()
    ctx.vv_run(dt, steps)
    el = time.perf_counter() - t0
    T1 = ctx.temperature()
    pairs = ctx.count_pairs()
    kms, kname = ctx.time_force_kernel(50)
    uw, _ = ctx.get_state()
    p_el, p_oo = water_pair_counts(uw, w.L, Rc)
    # NVE drift of the same system over 10k steps (no thermostat)
    nve = water_context(lib, gen, w, Rc, False)
    k0, p0, _ = nve.energies()
    nve.vv_run(dt, 10000)
    k1, p1, _ = nve.energies()
    # force error on the initial state and the CPU baseline (one full RHS of the oracle)
    orc = oracle_system_water(w, Rc)
    t0 = time.perf_counter()
    ref = orc.rhs(w.u)
    cpu_s = time.perf_counter() - t0
    fe = water_context(lib, gen, w, Rc, False)
    a = mixed_accel(fe, w.u)
    return {"n_atoms": n, "molecules": n // 3, "steps": steps, "dt_ps": dt,
            "steps_per_s": steps / el, "ms_per_step": 1e3 * el / steps,
            "pairs_per_step": pairs, "pair_interactions_per_s": pairs * steps / el,
            "temperature_after_K": T1,
            "autotuned": {"lanes": int(tuned["shape0"]), "skin_nm": tuned["shape1"],
                          "build_every": tuned["build_every"],
                          "ms_per_step_default": tuned["entry_ms_per_step"]},
            "roofline": roofline("valu", kname, 27 * p_el + 10 * p_oo, kms, traffic_key="water",
                                 flop_per_pair="27 per non-excluded pair within r_c "
                                               "(reaction field) + 10 per O-O pair (LJ)",
                                 pairs_electrostatic=p_el, pairs_oo=p_oo),
            "nve_energy_drift_rel_10k_steps": abs((k1 + p1) - (k0 + p0)) / abs(k0 + p0),
            "force_error_vs_fp64_oracle": float(np.abs(a - ref).max() / np.abs(ref).max()),
            "cpu_baseline": {"steps_per_s": 1.0 / cpu_s, "cores": 1, "kind": "port",
                             **host_cores(),
                             "sample": f"one FP64 oracle RHS of all {n} atoms (all pairs with "
                                       f"exclusions, as the reference), {cpu_s:.2f} s"},
            "gpu_vs_cpu_steps": (steps / el) * cpu_s}


def bench_batch(args, lib, gen):
    """512 LJ replicas of 108..4000 atoms (N = 4 n^3, n ~ U{3..10}), T* ~ U[0.7, 1.5],
    Langevin (gamma 1) or Andersen (nu 0.01), dt 0.005, 5000 steps in persistent launches."""
    import oracle_lib as O
    reps = gen.lj_batch(512, seed=0)
    systems = [gen.lj_liquid(r.cells, rho=0.8442, T=r.T, seed=k) for k, r in enumerate(reps)]
    natoms = [s.u.shape[0] for s in systems]
    b = lib.Batch(natoms, [s.L for s in systems], np.concatenate([s.masses for s in systems]))
    for k, r in enumerate(reps):
        if r.thermostat == "langevin":
            b.set_thermostat(k, lib.THERM_LANGEVIN, r.T, 1.0)
        else:
            b.set_thermostat(k, lib.THERM_ANDERSEN, r.T, 0.01)
    u0 = np.concatenate([s.u for s in systems])
    v0 = np.concatenate([s.v for s in systems])
    b.set_state(u0, v0)
    tuned = b.autotune(0.005, 200)
    steps = 5000
    b.run(0.005, 10)
    t0 = time.perf_counter()
    b.run(0.005, steps)
    el = time.perf_counter() - t0
    ke, pe, pairs = b.energies()
    T = 2.0 * ke / (3.0 * np.asarray(natoms))
    Tt = np.array([r.T for r in reps])
    # CPU baseline: the oracle's all-pairs RHS of 8 replicas, scaled by sum N^2
    pick = np.random.default_rng(1).choice(512, 8, replace=False)
    t0 = time.perf_counter()
    for k in pick:
        s = systems[k]
        O.System(s.masses, bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5).rhs(s.u)
    cpu_s = time.perf_counter() - t0
    n2 = np.asarray(natoms, dtype=np.float64) ** 2
    cpu_step_s = cpu_s * n2.sum() / n2[pick].sum()
    # force error: the same replicas without thermostats, one step of dt = 2^-60 from rest
    # gives a = (v1 - v0) / dt; checked against the oracle on 8 replicas (jittered lattice)
    offs = np.concatenate([[0], np.cumsum(natoms)])
    uj = u0 + 0.02 * np.random.default_rng(17).standard_normal(u0.shape)
    fb = lib.Batch(natoms, [s.L for s in systems], np.concatenate([s.masses for s in systems]))
    fb.set_state(uj, np.zeros_like(uj))
    dt_tiny = 2.0 ** -60
    fb.run(dt_tiny, 1)
    _, v1 = fb.get_state()
    a_gpu = v1 / dt_tiny
    errs = []
    for k in pick:
        s = systems[k]
        ref = O.System(s.masses, bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5) \
            .rhs(uj[offs[k]:offs[k + 1]])
        errs.append(float(np.abs(a_gpu[offs[k]:offs[k + 1]] - ref).max() / np.abs(ref).max()))
    # NVE energy drift: the same replicas without thermostats, 10k steps
    nve = lib.Batch(natoms, [s.L for s in systems], np.concatenate([s.masses for s in systems]))
    nve.set_skin(tuned["skin"])
    nve.set_state(u0, v0)
    k0, p0, _ = nve.energies()
    nve.run(0.005, 10000)
    k1, p1, _ = nve.energies()
    nat = np.asarray(natoms, dtype=np.float64)
    tail = 0.5 * nat * (nat - 1) * 4.0 * (2.5 ** -12 - 2.5 ** -6)   # n(n-1)/2 U(R)
    e0, e1 = k0 + p0, k1 + p1
    drift_ref = np.abs(e1 - e0) / np.abs(e0)
    drift_sh = np.abs(e1 - e0) / np.abs(e0 - tail)
    # critical path by replica size: 256 copies of one replica (one per CU, Langevin
    # T* = 1.0, the tuned skin); a batch's step takes as long as its slowest CU, which
    # holds one of its largest replicas
    per_size = {}
    for cells in sorted({r.cells for r in reps}):
        s1 = gen.lj_liquid(cells, rho=0.8442, T=1.0, seed=cells)
        n1 = s1.u.shape[0]
        ub = lib.Batch([n1] * 256, [s1.L] * 256, np.tile(s1.masses, 256))
        ub.set_skin(tuned["skin"])
        for k in range(256):
            ub.set_thermostat(k, lib.THERM_LANGEVIN, 1.0, 1.0)
        ub.set_state(np.tile(s1.u, (256, 1)), np.tile(s1.v, (256, 1)))
        ub.run(0.005, 100)
        t0 = time.perf_counter()
        ub.run(0.005, 500)
        per_size[str(n1)] = 1e3 * (time.perf_counter() - t0) / 500
        ub.close()
    ms = 1e3 * el / steps
    atoms = int(sum(natoms))
    return {"replicas": 512, "atoms": atoms, "steps": steps,
            "roofline": roofline("valu", "k_batch_steps", FLOP_LJ * float(pairs.sum()), ms,
                                 flop_per_pair=FLOP_LJ,
                                 traffic_per_step=workload_traffic("batch", "hbm_bytes_per_step"),
                                 note="one persistent launch covers the steps: kernel time "
                                      "per step = wall time per step",
                                 hbm_view={"algorithmic_bytes_per_step": 120 * atoms,
                                           "achieved_GBps": 120 * atoms / (ms * 1e-3) / 1e9,
                                           "peak_GBps": PEAK_HBM,
                                           "frac": 120 * atoms / (ms * 1e-3) / 1e9 / PEAK_HBM}),
            "force_error_vs_fp64_oracle": max(errs),
            "force_error_replicas": [int(natoms[k]) for k in pick],
            "nve_energy_drift_rel_10k": {"median_ref_pe": float(np.median(drift_ref)),
                                         "max_ref_pe": float(drift_ref.max()),
                                         "median_shifted_pe": float(np.median(drift_sh)),
                                         "max_shifted_pe": float(drift_sh.max())},
            "autotuned": {"skin": tuned["skin"], "ms_per_step_default": tuned["entry_ms_per_step"]},
            "ms_per_step_by_replica_size": per_size,
            "ms_per_step": 1e3 * el / steps, "atom_steps_per_s": sum(natoms) * steps / el,
            "pairs_per_step": int(pairs.sum()),
            "pair_interactions_per_s": float(pairs.sum()) * steps / el,
            "temperature_rel_error_median": {
                th: float(np.median(np.abs(T / Tt - 1.0)[[r.thermostat == th for r in reps]]))
                for th in ("langevin", "andersen")},
            "cpu_baseline": {"steps_per_s": 1.0 / cpu_step_s, "cores": 1, "kind": "port",
                             **host_cores(),
                             "sample": f"FP64 oracle all-pairs RHS of 8 replicas ({cpu_s:.2f} s), "
                                       f"scaled by sum N^2 to all 512"},
            "gpu_vs_cpu_steps": (steps / el) * cpu_step_s}


def bench_learned(args, lib, gen, ntraj=200, traj_steps=2000, iters=4000,
                  sweep=(4, 2), sweep_traj=20):
    """Learned stride on LJ N=32,000 (FCC 20^3, rho* 0.8442, T* 0.9, NVE), dt = 0.005.

    For the configured stride 8: a training set from `ntraj` plain trajectories of
    `traj_steps` steps (256 atoms per strided frame) plus held-out trajectories, Adam
    training, then 10k dt-steps of simulated time with the learned stride (energy monitor
    every 25 strided steps, fallback to plain velocity Verlet at a relative deviation of
    1e-4), timed against plain velocity Verlet at dt from the same state.  The same is done
    for the strides in `sweep` with `sweep_traj` trajectories each, and every stride is
    also run with a zero network (the strided step without the learned correction)."""
    dt = 0.005
    nsteps_dt = 10000
    sim_time = nsteps_dt * dt

    def ctx_for(seed):
        s = gen.lj_liquid(20, rho=0.8442, T=0.9, seed=seed)
        ctx = lib.Context(s.u.shape[0], lib.SYS_POTENTIAL, s.masses)
        ctx.set_boundary(lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
        ctx.add_lennard_jones(1.0, 1.0, 2.5)
        ctx.set_kb(1.0)
        ctx.set_precision(lib.PREC_MIXED)
        ctx.set_state(s.u, s.v)
        return ctx, s

    def train(stride, ntr):
        model = lib.Model(1)
        heldm = lib.Model(1)
        nheld = max(2, ntr // 20)
        t0 = time.perf_counter()
        gctx, _ = ctx_for(0)
        for tr in range(ntr + nheld):
            s = gen.lj_liquid(20, rho=0.8442, T=0.9, seed=1000 + tr)
            gctx.set_state(s.u, s.v)
            gctx.learn_collect(model if tr < ntr else heldm, dt, stride, traj_steps // stride,
                               256, seed=tr)
        t_collect = time.perf_counter() - t0
        model.fit_normalization()
        t0 = time.perf_counter()
        losses = model.train(iters, batch=8192, lr=1e-3, seed=7)
        t_train = time.perf_counter() - t0
        # the plain integrator scores the strided step on trajectories never trained on
        heldm.set(*model.get())
        mse, base = heldm.evaluate(0, heldm.num_samples)
        info = {"trajectories": ntr, "heldout_trajectories": nheld, "steps_each": traj_steps,
                "samples": model.num_samples, "collect_s": t_collect, "iters": iters,
                "batch": 8192, "train_s": t_train,
                "final_loss": float(np.mean(losses[-50:])),
                "heldout_mse_over_uncorrected": [float(x) for x in mse / base]}
        return model, info

    def run(model, stride):
        ctx, _ = ctx_for(4242)
        ctx.autotune(dt, 1000)    # the launch shape of the plain reference run below
        t0 = time.perf_counter()
        rep = ctx.learned_run(model, dt, stride, nsteps_dt // stride, drift_tol=1e-4,
                              check_every=25)
        el = time.perf_counter() - t0
        u, _ = ctx.get_state()
        return {"sim_time_per_s": sim_time / el, "wall_s": el, "report": rep,
                "energy_drift_rel_10k": abs(rep["E"] - rep["E0"]) / abs(rep["E0"]),
                "fraction_of_time_strided": rep["strided_steps"] * stride / nsteps_dt}, u

    def plain_stride_sweep(strides=(1, 2, 3, 4, 5, 6, 8), check_dt=500):
        """Plain velocity Verlet at stride * dt (no network) through the production path,
        from an equilibrated state (4,000 steps at dt from the FCC start), over the same
        10k dt of simulated time: the largest relative deviation of the shifted total
        energy (list energy every ~check_dt dt-steps) and simulated time per second."""
        eq, _ = ctx_for(4242)
        eq.set_skin(0.5)
        eq.vv_run(dt, 4000)
        ue, ve = eq.get_state()
        rows = {}
        for st in strides:
            c, _ = ctx_for(4242)
            c.set_skin(0.5)
            c.set_state(ue, ve)
            k0, p0 = c.learn_energy()
            e0 = k0 + p0
            every = max(1, check_dt // st)
            total = nsteps_dt // st
            done, maxdev, wall, err = 0, 0.0, 0.0, None
            while done < total:
                m = min(every, total - done)
                t0 = time.perf_counter()
                try:
                    c.vv_run(dt * st, m)
                except lib.NBXError as e:   # an exploding stride overflows its lists
                    err = str(e)
                    break
                wall += time.perf_counter() - t0
                done += m
                k, p = c.learn_energy()
                dev = abs(k + p - e0) / abs(e0)
                maxdev = max(maxdev, dev if math.isfinite(dev) else math.inf)
                if not maxdev < 0.5:
                    break
            rows[str(st)] = {"steps": done, "of": total, "max_rel_dev": maxdev, "error": err,
                             "sim_time_per_s": done * st * dt / wall if wall > 0 else None}
        # unit masses, kb = 1: T = sum v^2 / 3N
        return {"equilibrated_T": float((ve ** 2).sum()) / (3.0 * ue.shape[0]), "strides": rows}

    # plain velocity Verlet at dt: the reference point and the yardstick
    plain, s = ctx_for(4242)
    plain.autotune(dt, 1000)
    ke0, pe0 = plain.learn_energy()
    plain.vv_run(dt, 20)
    plain.set_state(s.u, s.v)
    t0 = time.perf_counter()
    plain.vv_run(dt, nsteps_dt)
    t_plain = time.perf_counter() - t0
    ke1, pe1 = plain.learn_energy()
    up, _ = plain.get_state()
    e0 = ke0 + pe0
    out = {"n": 32000, "dt": dt, "simulated_time": sim_time,
           "plain": {"sim_time_per_s": sim_time / t_plain, "steps_per_s": nsteps_dt / t_plain,
                     "energy_drift_rel_10k": abs((ke1 + pe1) - e0) / abs(e0)},
           "strides": {}}
    out["plain_stride_sweep"] = plain_stride_sweep()
    # MFMA rooflines of the learned kernels, from the committed kernel trace of the workload
    # run (their launches are not timed inside this line)
    rl = {}
    for key, kname, work, what in (
            ("mlp_infer", "k_lr_mlp_infer", 76032 * 32000,
             "76,032 flop per atom (35-128-128-128-6 forward), 32,000 atoms"),
            ("lr_chain", "k_lr_chain", 2 * 76032 * 8192,
             "2 x 76,032 flop per sample (forward and backward chain), 8,192 samples")):
        us = workload_traffic(key, "avg_us_traced")
        if us:
            rl[kname] = roofline("mfma", kname, work, us * 1e-3, peak=PEAK_BF16_MFMA,
                                 traffic_key=key, work_counted=what,
                                 timing="kernel-trace average, profiles/workload_kernel_counters.md")
    out["rooflines"] = rl
    zero = lib.Model(0)
    zero.set(np.zeros(lib.MLP_PARAMS), np.concatenate([np.zeros(35), np.ones(35), np.ones(6)]))
    rs, g_plain = plain.rdf(up[None], 200)
    for stride in (8,) + tuple(sweep):
        model, info = train(stride, ntraj if stride == 8 else sweep_traj)
        learned, ul = run(model, stride)
        _, g_l = plain.rdf(ul[None], 200)
        learned["rdf_max_abs_diff_vs_plain"] = float(np.abs(g_l - g_plain).max())
        without, _ = run(zero, stride)
        out["strides"][str(stride)] = {
            "training": info, "learned": learned, "without_network": without,
            "speedup_sim_time_vs_plain": learned["sim_time_per_s"] / out["plain"]["sim_time_per_s"]}
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
    tuned = ctx.autotune(1e-3, 20)          # the same tile shape as the headline run
    kms, kname = ctx.time_force_kernel(10)
    nbytes = 1 << 30
    cms = ctx.copy_probe(nbytes, 5)
    print(json.dumps({"kernel": kname, "kernel_avg_ms": kms, "copy_bytes": nbytes,
                      "targets_per_thread": int(tuned["shape0"]), "source_chunks": int(tuned["shape1"]),
                      "copy_avg_ms": cms, "copy_GBps": 2 * nbytes / (cms * 1e-3) / 1e9}),
          flush=True)
    return 0


def profile_learned(lib, gen, stride=2, nsteps=200):
    """Strided steps of LJ N=32,000 with a zero network (every kernel of the strided step
    runs; the corrections are zero), for a kernel trace."""
    s = gen.lj_liquid(20, rho=0.8442, T=0.9, seed=4242)
    ctx = lib.Context(s.u.shape[0], lib.SYS_POTENTIAL, s.masses)
    ctx.set_boundary(lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(lib.PREC_MIXED)
    ctx.set_state(s.u, s.v)
    zero = lib.Model(0)
    zero.set(np.zeros(lib.MLP_PARAMS), np.concatenate([np.zeros(35), np.ones(35), np.ones(6)]))
    t0 = time.perf_counter()
    rep = ctx.learned_run(zero, 0.005, stride, nsteps, drift_tol=1.0, check_every=25)
    el = time.perf_counter() - t0
    print(json.dumps({"stride": stride, "strided_steps": nsteps,
                      "us_per_strided_step": 1e6 * el / nsteps, "report": rep}), flush=True)
    return 0


def profile_workloads(lib, gen):
    """Every benchmark workload briefly, without step graphs (each kernel dispatch is a
    separate counter sample), for rocprofv3 counter and trace runs: gravity N=65,536
    (autotuned tile shape, 5 steps), LJ N=32,000 (8 lanes, skin 0.6, 300 steps), water
    4,096 molecules (autotuned, 100 steps), 512 batched LJ replicas (one 100-step launch),
    the learned stride (200 collected frames, 50 training iterations, 50 strided steps
    with a zero network), and a 1 GiB copy kernel (the byte-counter calibration)."""
    os.environ["NBX_GRAPH_STEPS"] = "0"
    out = {}
    n = 65536
    u, v, m = gen.plummer(n, seed=2024)
    g = lib.Context(n, lib.SYS_POTENTIAL, m)
    g.add_gravitational(1.0)
    g.set_precision(lib.PREC_MIXED)
    g.set_state(u, v)
    tuned = g.autotune(1e-3, 20)
    g.vv_run(1e-3, 5)
    out["gravity"] = {"targets_per_thread": int(tuned["shape0"]),
                      "source_chunks": int(tuned["shape1"])}
    s = gen.lj_liquid(20, rho=0.8442, T=0.9, seed=5)
    lj = lib.Context(s.u.shape[0], lib.SYS_POTENTIAL, s.masses)
    lj.set_boundary(lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    lj.add_lennard_jones(1.0, 1.0, 2.5)
    lj.set_kb(1.0)
    lj.set_precision(lib.PREC_MIXED)
    lj.set_skin(0.6)
    lj.set_state(s.u, s.v)
    lj.vv_run(0.005, 300)
    out["lj32000"] = lj.stats()
    w = gen.water_box(per_side=16, seed=7)
    wc = water_context(lib, gen, w, 0.9, True)
    tuned = wc.autotune(0.0005, 50)
    wc.vv_run(0.0005, 100)
    out["water"] = {"lanes": int(tuned["shape0"]), "skin_nm": tuned["shape1"]}
    reps = gen.lj_batch(512, seed=0)
    systems = [gen.lj_liquid(r.cells, rho=0.8442, T=r.T, seed=k) for k, r in enumerate(reps)]
    natoms = [x.u.shape[0] for x in systems]
    b = lib.Batch(natoms, [x.L for x in systems], np.concatenate([x.masses for x in systems]))
    for k, r in enumerate(reps):
        b.set_thermostat(k, lib.THERM_LANGEVIN if r.thermostat == "langevin" else lib.THERM_ANDERSEN,
                         r.T, 1.0 if r.thermostat == "langevin" else 0.01)
    b.set_state(np.concatenate([x.u for x in systems]), np.concatenate([x.v for x in systems]))
    b.run(0.005, 100, steps_per_launch=1000)   # one direct 100-step launch
    out["batch"] = {"replicas": 512, "atoms": int(sum(natoms))}
    model = lib.Model(0)
    lj.learn_collect(model, 0.005, 2, 200, 256, seed=1)
    model.fit_normalization()
    model.train(50, batch=8192)
    zero = lib.Model(0)
    zero.set(np.zeros(lib.MLP_PARAMS), np.concatenate([np.zeros(35), np.ones(35), np.ones(6)]))
    lj.learned_run(zero, 0.005, 2, 50, drift_tol=1.0, check_every=25)
    out["learned"] = {"collected_frames": 200, "train_iters": 50, "strided_steps": 50}
    out["copy_probe_ms"] = g.copy_probe(1 << 30, 3)   # byte-counter calibration, 1 GiB
    print(json.dumps(out), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--profile-learned", action="store_true",
                    help="only run 200 strided steps (stride 2, zero network) of LJ N=32000 "
                         "for a kernel trace of the strided step")
    ap.add_argument("--no-long-drift", action="store_true",
                    help="skip the 10k-step gravity drift continuation (kernel-trace runs)")
    ap.add_argument("--learn-traj", type=int, default=200,
                    help="training trajectories of the learned-stride line")
    ap.add_argument("--profile-kernels", action="store_true",
                    help="only launch the dominant kernels (gravity N=65536 force, 10 launches; "
                         "a 1 GiB copy-kernel calibration, 5 launches) for rocprofv3 counters")
    ap.add_argument("--profile-workloads", action="store_true",
                    help="only run every workload briefly without step graphs, for rocprofv3 "
                         "counter and trace runs")
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
    if args.profile_workloads:
        return profile_workloads(lib, gen)
    if args.profile_learned:
        return profile_learned(lib, gen)
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
        res["secondary"] = {"lj_n32000_T0.9": bench_lj(args, lib, gen, 20, 0.9, 10000, 600)}
        log("[bench] LJ N=864")
        res["secondary"]["lj_n864_T0.72"] = bench_lj(args, lib, gen, 6, 0.72, 10000, 864,
                                                     fp64=True)
        log("[bench] water 4096 molecules")
        res["secondary"]["water_spcfw_4096"] = bench_water(args, lib, gen)
        log("[bench] batch 512 replicas")
        res["secondary"]["lj_batch_512"] = bench_batch(args, lib, gen)
        log("[bench] learned stride N=32000")
        res["secondary"]["lj_n32000_learned_stride"] = bench_learned(args, lib, gen,
                                                                     ntraj=args.learn_traj)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
