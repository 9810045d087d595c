"""Batched replicas (one persistent launch, one workgroup per replica) against the FP64
oracle and the single-system path.

Forces: norm-wise error <= 2e-5 per replica (FP32 pair forces from LDS positions, FP32
sums of 8 pairs, Kahan sums of those).  Trajectories: replica r with thermostat seed
seed + r * 0x9E3779B97F4A7C15 draws the Philox words of a single-system run with that seed
(its normals come from FP32 hardware transcendentals instead of FP64 library calls), so
short runs agree with the single-system mixed path to 1e-6."""
import numpy as np
import pytest

import oracle_lib as O
from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu

GOLD = 0x9E3779B97F4A7C15


def make_batch(cells_list, T=0.9, jitter=0.02, seed=1):
    systems = [gen.lj_liquid(c, rho=0.8442, T=T, seed=seed + k, jitter=jitter)
               for k, c in enumerate(cells_list)]
    natoms = [s.u.shape[0] for s in systems]
    b = _lib.Batch(natoms, [s.L for s in systems], np.concatenate([s.masses for s in systems]),
                   1.0, 1.0, 2.5, 1.0)
    return b, systems


def test_batch_forces_match_oracle(nbx):
    print("[case] batch forces 3..7 cells", flush=True)
    b, systems = make_batch([3, 4, 5, 6, 7])
    U = np.concatenate([s.u for s in systems])
    b.set_state(U, np.zeros_like(U))
    dt = 2.0 ** -60
    b.run(dt, 1)
    _, V = b.get_state()
    A = V / dt
    o = 0
    for s in systems:
        n = s.u.shape[0]
        ref = O.System(s.masses, bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1, 1, 2.5).rhs(s.u)
        err = np.abs(A[o:o + n] - ref).max() / np.abs(ref).max()
        assert err <= 2e-5, (n, err)
        o += n


def test_batch_nve_energy_and_determinism(nbx):
    print("[case] batch NVE 2000 steps", flush=True)
    b, systems = make_batch([3, 5, 8], jitter=0.0)
    U = np.concatenate([s.u for s in systems])
    V = np.concatenate([s.v for s in systems])
    b.set_state(U, V)
    ke0, pe0, pairs0 = b.energies()
    b.run(0.005, 700, steps_per_launch=250)
    b.run(0.005, 1300, steps_per_launch=250)
    ke1, pe1, _ = b.energies()
    e0, e1 = ke0 + pe0, ke1 + pe1
    assert np.all(np.abs(e1 - e0) / np.abs(e0) < 1e-3), (e0, e1)
    assert np.all(pairs0 > 0)
    U1, V1 = b.get_state()
    b.set_state(U, V)
    b.run(0.005, 2000, steps_per_launch=1000)
    U2, V2 = b.get_state()
    assert np.array_equal(U1, U2) and np.array_equal(V1, V2)
    assert b.stats()["steps_taken"] == 2000


def test_batch_thermostats_reach_targets(nbx):
    print("[case] batch thermostats", flush=True)
    b, systems = make_batch([3, 4, 5, 4], T=0.8)
    targets = [(_lib.THERM_LANGEVIN, 1.3, 1.0), (_lib.THERM_ANDERSEN, 1.1, 5.0),
               (_lib.THERM_LANGEVIN, 0.9, 2.0), (_lib.THERM_ANDERSEN, 1.5, 10.0)]
    for r, t in enumerate(targets):
        b.set_thermostat(r, *t)
    U = np.concatenate([s.u for s in systems])
    V = np.concatenate([s.v for s in systems])
    b.set_state(U, V)
    b.run(0.005, 2000, seed=11)
    # kinetic temperature averaged over 10 frames 100 steps apart (one frame of a
    # 108..500-atom replica fluctuates by ~5 %)
    ke = np.zeros(len(systems))
    for _ in range(10):
        b.run(0.005, 100, seed=11)
        ke += b.energies()[0] / 10
    for r, (s, t) in enumerate(zip(systems, targets)):
        n = s.u.shape[0]
        T = 2 * ke[r] / (3 * n)
        assert T == pytest.approx(t[1], rel=0.12), (r, T, t)


def test_batch_replica_matches_single_system_path(nbx):
    print("[case] batch vs single-system Langevin", flush=True)
    b, systems = make_batch([4, 5])
    for r in range(2):
        b.set_thermostat(r, _lib.THERM_LANGEVIN, 1.2, 1.0)
    U = np.concatenate([s.u for s in systems])
    V = np.concatenate([s.v for s in systems])
    b.set_state(U, V)
    seed = 77
    b.run(0.005, 20, seed=seed)
    Ub, Vb = b.get_state()
    o = 0
    for r, s in enumerate(systems):
        n = s.u.shape[0]
        ctx = _lib.Context(n, _lib.SYS_POTENTIAL, s.masses)
        ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
        ctx.add_lennard_jones(1, 1, 2.5)
        ctx.set_kb(1.0)
        ctx.set_precision(_lib.PREC_MIXED)
        ctx.set_thermostat(_lib.THERM_LANGEVIN, 1.2, 1.0)
        ctx.set_state(s.u, s.v)
        ctx.vv_run(0.005, 20, seed=(seed + r * GOLD) % (1 << 64))
        u, v = ctx.get_state()
        assert np.abs(Ub[o:o + n] - u).max() < 1e-6
        assert np.abs(Vb[o:o + n] - v).max() < 1e-5
        o += n


def test_batch_argument_errors(nbx):
    with pytest.raises(ValueError):
        _lib.Batch([5000], [20.0], np.ones(5000))
    b, _ = make_batch([3])
    with pytest.raises(IndexError):
        b.set_thermostat(1, _lib.THERM_LANGEVIN, 1.0, 1.0)
    with pytest.raises(NotImplementedError):
        b.set_thermostat(0, _lib.THERM_BERENDSEN, 1.0, 1.0)
    with pytest.raises(ValueError):
        b.run(0.005, 1)   # no state


def test_batch_autotune_keeps_state_and_is_reproducible(nbx):
    """The skin autotuner restores the state; a run after it equals a run of a fresh batch
    given the chosen skin, bit for bit."""
    print("[case] batch autotune", flush=True)
    b, systems = make_batch([3, 5, 6])
    for r in range(3):
        b.set_thermostat(r, _lib.THERM_LANGEVIN, 1.0, 1.0)
    U = np.concatenate([s.u for s in systems])
    V = np.concatenate([s.v for s in systems])
    b.set_state(U, V)
    tuned = b.autotune(0.005, 50)
    u0, v0 = b.get_state()
    assert np.array_equal(u0, U) and np.array_equal(v0, V)
    assert b.stats()["steps_taken"] == 0
    assert tuned["skin"] in (0.2, 0.3, 0.4, 0.5, 0.6, 0.8) and tuned["ms_per_step"] > 0
    b.run(0.005, 300, seed=4)
    c, _ = make_batch([3, 5, 6])
    for r in range(3):
        c.set_thermostat(r, _lib.THERM_LANGEVIN, 1.0, 1.0)
    c.set_skin(tuned["skin"])
    c.set_state(U, V)
    c.run(0.005, 300, seed=4)
    ub, vb = b.get_state()
    uc, vc = c.get_state()
    assert np.array_equal(ub, uc) and np.array_equal(vb, vc)


def test_batch_list_overflow_replays_the_call(nbx):
    # imploding dilute replicas: the per-atom list capacity sized for the initial density
    # is exceeded inside the launch; the call is restored from its device backup and
    # replayed with larger lists.  The replay is deterministic (two batches agree bit for
    # bit) and every replica follows the single-system path.  Bound of the latter: the
    # truncated LJ force jumps by |F(R)| = 0.039 at the cutoff, and the two paths round r^2
    # differently (box-unit FP32 coordinates against FP32 coordinates in sigma), so a pair
    # that sits at R within rounding may count in one path one step earlier: a velocity
    # difference of up to |F(R)| dt / m = 7.8e-5, i.e. up to 7.8e-5 in position per such
    # event over the remaining t <= 1.  A replay that lost or repeated steps, or kept
    # truncated lists, is off by 1e-3 or more.
    print("[case] batch implosion", flush=True)
    systems = [gen.lj_liquid(c, rho=0.25, T=0.1, seed=5 + c, jitter=0.05) for c in (5, 6)]
    vs = []
    for s in systems:
        centre = 0.5 * s.L
        vs.append(s.v - 2.0 * (s.u - centre) / centre)
    natoms = [s.u.shape[0] for s in systems]
    runs = []
    for _ in range(2):
        b = _lib.Batch(natoms, [s.L for s in systems],
                       np.concatenate([s.masses for s in systems]), 1.0, 1.0, 2.5, 1.0)
        b.set_state(np.concatenate([s.u for s in systems]), np.concatenate(vs))
        slots0 = b.stats()["list_slots"]
        b.run(0.002, 500)
        slots1 = b.stats()["list_slots"]
        runs.append(b.get_state())
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    Ub = runs[0][0]
    o, devs = 0, []
    for s, v in zip(systems, vs):
        n = s.u.shape[0]
        ctx = _lib.Context(n, _lib.SYS_POTENTIAL, s.masses)
        ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
        ctx.add_lennard_jones(1, 1, 2.5)
        ctx.set_kb(1.0)
        ctx.set_precision(_lib.PREC_MIXED)
        ctx.set_state(s.u, v)
        ctx.vv_run(0.002, 500)
        u, _ = ctx.get_state()
        devs.append(float(np.abs(Ub[o:o + n] - u).max()))
        o += n
    print(f"[case] batch implosion: list slots {slots0} -> {slots1}, max |du| {max(devs):.2e}",
          flush=True)
    assert slots1 > slots0
    assert max(devs) < 4e-4
