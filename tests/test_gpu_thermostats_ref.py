"""Thermostats on the GPU: the Nose-Hoover law (src/thermostats.jl:116-122) through the
FP64 path bit for bit against the oracle (LJ) and at 1e-12 (water, acos), the mixed path
against the FP64 path, and the reference's own thermostat functional tests
(test/thermostat_test.jl: liquid argon, N = 125, 200 steps) through the Python mirror of
the reference API, for all four thermostats the GPU path provides."""
import math

import numpy as np
import pytest

import oracle_lib as O
import nbodysimulator as nbs
from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def lj_nh(n_cells, prec, T0=0.5, tau=0.1, seed=11):
    s = gen.lj_liquid(n_cells, rho=0.8442, T=0.72, seed=seed, jitter=0.02)
    n = s.u.shape[0]
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, np.ones(n))
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(prec)
    ctx.set_thermostat(_lib.THERM_NOSEHOOVER, T0, tau)
    orc = O.System(np.ones(n), bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5)
    orc.thermostat(O.THERM_NOSEHOOVER, T0, tau)
    return ctx, orc, s


def test_fp64_nosehoover_trajectory_bitwise(nbx):
    print("[case] fp64 Nose-Hoover LJ N=864, 100 steps", flush=True)
    ctx, orc, s = lj_nh(6, _lib.PREC_FP64_REF)
    ctx.set_state(s.u, s.v)
    uf, vf = ctx.vv_run(0.005, 100, save_every=50)
    _, _, (uo, vo) = orc.vv_run(s.u, s.v, 0.005, 100, save_every=50)
    assert np.array_equal(uf, uo) and np.array_equal(vf, vo)
    # a split run continues the thermostat state (zeta and its rate stay on the device)
    ctx.set_state(s.u, s.v)
    ctx.vv_run(0.005, 40)
    ctx.vv_run(0.005, 60)
    u2, v2 = ctx.get_state()
    assert np.array_equal(u2, uo[-1]) and np.array_equal(v2, vo[-1])


def test_fp64_nosehoover_water(nbx):
    print("[case] fp64 Nose-Hoover water 27 molecules", flush=True)
    w = gen.water_box(per_side=3, seed=5)
    P = gen.SPCFW
    n = len(w.masses)
    Rc = min(0.9, 0.45 * w.L)
    ctx = _lib.Context(n, _lib.SYS_WATER, w.masses, w.charges)
    ctx.set_boundary(_lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
    ctx.set_kb(P["kb"])
    ctx.add_lennard_jones(P["eps"], P["sigma"], Rc)
    ctx.add_electrostatic(P["k_el"], Rc)
    ctx.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    ctx.set_thermostat(_lib.THERM_NOSEHOOVER, 300.0, 0.05)
    orc = O.System(w.masses, charges=w.charges, water=True, bc=O.make_bc(O.BC_CUBIC, w.L),
                   kb=P["kb"])
    orc.water_params((P["eps"], P["sigma"], Rc), O.TERM_ELEC, (P["k_el"], Rc), P["rOH"],
                     P["aHOH"], P["kbond"], P["kangle"])
    orc.thermostat(O.THERM_NOSEHOOVER, 300.0, 0.05)
    ctx.set_state(w.u, w.v)
    ctx.vv_run(0.0005, 50)
    u, v = ctx.get_state()
    uo, vo, _ = orc.vv_run(w.u, w.v, 0.0005, 50)
    assert np.abs(u - uo).max() <= 1e-12 * np.abs(uo).max()
    assert np.abs(v - vo).max() <= 1e-10 * np.abs(vo).max()


def test_mixed_nosehoover_tracks_fp64(nbx):
    print("[case] mixed vs fp64 Nose-Hoover N=864", flush=True)
    c64, _, s = lj_nh(6, _lib.PREC_FP64_REF)
    cmx, _, _ = lj_nh(6, _lib.PREC_MIXED)
    temps = {}
    for name, c in (("fp64", c64), ("mixed", cmx)):
        c.set_state(s.u, s.v)
        T = [c.temperature()]
        for _ in range(4):
            c.vv_run(0.005, 100)
            T.append(c.temperature())
        temps[name] = np.array(T)
    # the same thermostat dynamics (the FP32 pair forces only perturb the chaotic
    # trajectory): temperature curves within 2 %, and both leave T0 (ndf+1)/ndf closer
    # than they started
    assert np.abs(temps["mixed"] / temps["fp64"] - 1).max() <= 0.02, temps
    target = 0.5 * (3 * 864 + 1) / (3 * 864)
    for T in temps.values():
        assert abs(T[-1] - target) < abs(T[0] - target)


def test_mixed_nosehoover_split_runs_bitwise(nbx):
    """The mixed path redoes the last drift of a run (or of a chunk) at the start of the
    next one; the thermostat variable zeta must advance once per step all the same: runs of
    40 + 60 (and 700 steps, two internal chunks) equal unsplit ones bit for bit, LJ and
    SPC/Fw water."""
    print("[case] mixed Nose-Hoover split runs N=864", flush=True)
    a, _, s = lj_nh(6, _lib.PREC_MIXED)
    b, _, _ = lj_nh(6, _lib.PREC_MIXED)
    a.set_state(s.u, s.v)
    b.set_state(s.u, s.v)
    a.vv_run(0.005, 40)
    a.vv_run(0.005, 60)
    b.vv_run(0.005, 100)
    ua, va = a.get_state()
    ub, vb = b.get_state()
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)
    print("[case] mixed Nose-Hoover 700 steps against 300 + 400, water", flush=True)
    w = gen.water_box(per_side=8, seed=3)
    P = gen.SPCFW
    ctxs = []
    for _ in range(2):
        c = _lib.Context(len(w.masses), _lib.SYS_WATER, w.masses, w.charges)
        c.set_boundary(_lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
        c.set_kb(P["kb"])
        c.add_lennard_jones(P["eps"], P["sigma"], 0.9)
        c.add_reaction_field(P["k_el"], 0.9, math.inf)
        c.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
        c.set_precision(_lib.PREC_MIXED)
        c.set_thermostat(_lib.THERM_NOSEHOOVER, 300.0, 0.05)
        c.set_state(w.u, w.v)
        ctxs.append(c)
    ctxs[0].vv_run(0.0005, 700)
    ctxs[1].vv_run(0.0005, 300)
    ctxs[1].vv_run(0.0005, 400)
    u0, v0 = ctxs[0].get_state()
    u1, v1 = ctxs[1].get_state()
    assert np.array_equal(u0, u1) and np.array_equal(v0, v1)


def test_mixed_nosehoover_reports_the_state_pair(nbx):
    """The mixed path's last drift of a run belongs to the next step and advances zeta
    already; nbx_get_thermostat_state reports the pair of the state all the same.  One step
    from a fresh state: zeta_1 = dt xi_0, xi_0 = (1/tau)^2 (T(v_0)/T0 - (ndf+1)/ndf) (the
    first right-hand side sees v_0; src/thermostats.jl:116-122).  A following FP64 run (a
    new solve: its first right-hand side sees v) advances the reported zeta once per step:
    zeta' = zeta + dt xi(T(v)) after one step."""
    print("[case] mixed Nose-Hoover reported zeta, LJ N=864", flush=True)
    T0, tau, dt, n = 0.5, 0.1, 0.005, 864
    ctx, _, s = lj_nh(6, _lib.PREC_MIXED, T0=T0, tau=tau)
    ctx.set_state(s.u, s.v)

    def rate(T):
        return (1.0 / tau) ** 2 * (T / T0 - (3 * n + 1) / (3 * n))

    xi0 = rate(ctx.temperature())
    ctx.vv_run(dt, 1)
    z1, _ = ctx.thermostat_state()
    assert abs(z1 - dt * xi0) <= 1e-12 * abs(dt * xi0), (z1, dt * xi0)
    ctx.vv_run(dt, 39)
    z40, _ = ctx.thermostat_state()
    T40 = ctx.temperature()
    ctx.set_precision(_lib.PREC_FP64_REF)
    ctx.vv_run(dt, 1)
    z41, _ = ctx.thermostat_state()
    step = dt * rate(T40)
    assert abs(z41 - (z40 + step)) <= 1e-9 * abs(step), (z41, z40, step)


def water_berendsen(w, prec, T0, tau):
    P = gen.SPCFW
    c = _lib.Context(len(w.masses), _lib.SYS_WATER, w.masses, w.charges)
    c.set_boundary(_lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
    c.set_kb(P["kb"])
    c.add_lennard_jones(P["eps"], P["sigma"], 0.9)
    c.add_reaction_field(P["k_el"], 0.9, math.inf)
    c.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    c.set_precision(prec)
    c.set_thermostat(_lib.THERM_BERENDSEN, T0, tau)
    c.set_state(w.u, w.v)
    return c


def test_mixed_water_berendsen_follows_fp64_and_splits_bitwise(nbx):
    """The mixed water step takes the Berendsen temperature of the drift before it from
    that drift's block partials (no kernel runs between the steps).  Strong coupling
    (tau 0.01 ps, T0 150 K from 300 K: the coefficient gamma (T0/T - 1) starts at -25/ps and
    T drops by ~7 K per step at first) makes the trajectory depend on that temperature
    step by step; the mixed trajectory stays within 1e-5 nm of the FP64 path's (the bar of the full-size water test), its
    temperature within 1e-5, and runs of 300 + 400 steps (chunks of at most 500, graphs
    of 50) equal a 700-step run bit for bit."""
    print("[case] mixed water Berendsen vs FP64, 512 molecules", flush=True)
    w = gen.water_box(per_side=8, seed=3)
    mx = water_berendsen(w, _lib.PREC_MIXED, 150.0, 0.01)
    rf = water_berendsen(w, _lib.PREC_FP64_REF, 150.0, 0.01)
    mx.vv_run(0.0005, 200)
    rf.vv_run(0.0005, 200)
    um, _ = mx.get_state()
    ur, _ = rf.get_state()
    dev = float(np.abs(um - ur).max())
    Tm, Tr = mx.temperature(), rf.temperature()
    print(f"[case] max |du| {dev:.2e} nm, T {Tm:.4f} / {Tr:.4f} K", flush=True)
    assert dev < 1e-5
    assert abs(Tm - Tr) <= 1e-5 * Tr
    assert Tr < 280.0   # the coupling is strong enough to matter
    print("[case] mixed water Berendsen 700 steps against 300 + 400", flush=True)
    a = water_berendsen(w, _lib.PREC_MIXED, 150.0, 0.01)
    b = water_berendsen(w, _lib.PREC_MIXED, 150.0, 0.01)
    a.vv_run(0.0005, 700)
    b.vv_run(0.0005, 300)
    b.vv_run(0.0005, 400)
    ua, va = a.get_state()
    ub, vb = b.get_state()
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)


def argon(thermostat_factory, precision, alg_kind="vv"):
    T = 120.0
    T0 = 90
    kb = 8.3144598e-3
    eps = T * kb
    sigma = 0.34
    rho = 1374 / 1.6747
    m = 39.95
    N = 125
    L = (m * N / rho) ** (1 / 3)
    R = 0.5 * L
    v_dev = math.sqrt(kb * T / m)
    bodies = nbs.generate_bodies_in_cell_nodes(N, m, v_dev, L, seed=7)
    tau = 0.5e-3
    params = nbs.LennardJonesParameters(eps, sigma, R)
    system = nbs.PotentialNBodySystem(bodies, {"lennard_jones": params})
    sim = nbs.NBodySimulation(system, (0.0, 200 * tau), nbs.CubicPeriodicBoundaryConditions(L),
                              thermostat_factory(T0, tau), kb)
    res = nbs.run_simulation(sim, nbs.VelocityVerlet(precision=precision, seed=3), dt=tau)
    return nbs.temperature(res, 200 * tau), T0, T


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("kind", ["andersen", "berendsen", "nosehoover", "langevin"])
def test_reference_argon_thermostats(nbx, kind, precision):
    """test/thermostat_test.jl:15-100 with the GPU integrator (the Langevin case runs the
    GPU's Ornstein-Uhlenbeck splitting instead of the reference's EM solver)."""
    print(f"[case] reference argon thermostat {kind} {precision}", flush=True)
    factories = {
        "andersen": lambda T0, tau: nbs.AndersenThermostat(T0, 0.1 / tau),
        "berendsen": lambda T0, tau: nbs.BerendsenThermostat(T0, 10 * tau),
        "nosehoover": lambda T0, tau: nbs.NoseHooverThermostat(T0, 20 * tau),
        "langevin": lambda T0, tau: nbs.LangevinThermostat(T0, 10),
    }
    T2, T0, T = argon(factories[kind], precision)
    if kind == "berendsen":
        assert abs(T2 - T0) / T <= 0.1
    else:
        assert abs(T2 - T0) / T0 <= 0.5
