"""Euler-Maruyama steps of the Langevin SDE (nbx_em_run; SciMLBase.SDEProblem for a
potential system, src/nbody_to_ode.jl:567-598) against the oracle's orc_em_run, which
draws the same Philox normals.

Bars: the FP64 path within 1e-12 (positions) / 1e-11 (velocities) relative -- the
device's log/sin/cos behind the Box-Muller normals may differ from the host libm in the
last ulp, as for the other Langevin parity tests; the mixed path (FP32 pair forces,
5e-7 relative) within 1e-6 / 1e-4 after 20 steps; split runs bit for bit; the noise
variance of free particles with sigma from the first body's mass (5 standard errors)."""
import math

import numpy as np
import pytest

import oracle_lib as O
from nbodysimulator import _lib
from nbodysimulator import generators as gen
import nbodysimulator as nbs

pytestmark = pytest.mark.gpu

LANG = (_lib.THERM_LANGEVIN, 1.3, 1.0)


def relerr(a, ref):
    return float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))


def lj_ctx(s, prec, therm=LANG):
    n = s.u.shape[0]
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, np.ones(n))
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(prec)
    if therm:
        ctx.set_thermostat(*therm)
    orc = O.System(np.ones(n), bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5)
    if therm:
        orc.thermostat(*therm)
    return ctx, orc


@pytest.fixture(scope="module")
def c1(nbx):
    return gen.lj_liquid(6, rho=0.8442, T=0.72, seed=11, jitter=0.02)


def test_fp64_em_matches_oracle(c1):
    print("[case] fp64 EM N=864 40 steps", flush=True)
    ctx, orc = lj_ctx(c1, _lib.PREC_FP64_REF)
    ctx.set_state(c1.u, c1.v)
    uf, vf = ctx.em_run(0.005, 40, seed=1234, save_every=10)
    u, v = ctx.get_state()
    uo, vo, (ufo, vfo) = orc.em_run(c1.u, c1.v, 0.005, 40, seed=1234, save_every=10)
    assert relerr(u, uo) <= 1e-12 and relerr(v, vo) <= 1e-11
    assert uf.shape == (5, 864, 3)
    assert relerr(uf, ufo) <= 1e-12 and relerr(vf, vfo) <= 1e-11
    assert ctx.steps_taken() == 40


@pytest.mark.parametrize("prec", [_lib.PREC_FP64_REF, _lib.PREC_MIXED])
def test_em_split_runs_are_bitwise(c1, prec):
    print(f"[case] EM split runs prec={prec}", flush=True)
    a, _ = lj_ctx(c1, prec)
    a.set_state(c1.u, c1.v)
    a.em_run(0.005, 20, seed=7)
    a.em_run(0.005, 80, seed=7)
    b, _ = lj_ctx(c1, prec)
    b.set_state(c1.u, c1.v)
    b.em_run(0.005, 100, seed=7)
    ua, va = a.get_state()
    ub, vb = b.get_state()
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)


def test_mixed_em_tracks_fp64(c1):
    print("[case] mixed EM vs fp64 EM 20 steps", flush=True)
    m, _ = lj_ctx(c1, _lib.PREC_MIXED)
    f, _ = lj_ctx(c1, _lib.PREC_FP64_REF)
    for ctx in (m, f):
        ctx.set_state(c1.u, c1.v)
        ctx.em_run(0.005, 20, seed=3)
    um, vm = m.get_state()
    uf, vf = f.get_state()
    assert relerr(um, uf) <= 1e-6 and relerr(vm, vf) <= 1e-4


def test_em_then_vv_continues_from_the_em_state(c1):
    """Velocity Verlet (NVE) after an Euler-Maruyama run in the same mixed context
    continues from the EM state: it matches a fresh context given that state (to the
    FP32 summation-order level over 10 steps), so no stale half-step velocity or
    acceleration survives the switch of integrator."""
    print("[case] mixed EM then VV", flush=True)
    m, _ = lj_ctx(c1, _lib.PREC_MIXED)
    m.set_state(c1.u, c1.v)
    m.em_run(0.005, 30, seed=2)
    u1, v1 = m.get_state()
    m.set_thermostat(_lib.THERM_NONE, 0.0, 1.0)
    m.vv_run(0.005, 10)
    u2, v2 = m.get_state()
    g, _ = lj_ctx(c1, _lib.PREC_MIXED, therm=None)
    g.set_state(u1, v1)
    g.vv_run(0.005, 10)
    u3, v3 = g.get_state()
    assert relerr(u2, u3) <= 1e-9 and relerr(v2, v3) <= 1e-6


@pytest.mark.parametrize("prec", [_lib.PREC_FP64_REF, _lib.PREC_MIXED])
def test_em_free_particle_noise_uses_first_body_mass(prec):
    """No forces (G = 0), v(0) = 0: after N steps each velocity component has variance
    sigma^2 dt (1 - r^(2N)) / (1 - r^2), r = 1 - gamma dt, sigma^2 = 2 gamma kb T / m_1."""
    print(f"[case] EM free-particle variance prec={prec}", flush=True)
    n, gamma, kT, dt, N = 4001, 1.0, 1.5, 0.01, 400
    m = np.ones(n)
    m[0] = 2.0
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, m)
    ctx.add_gravitational(0.0)
    ctx.set_kb(1.0)
    ctx.set_precision(prec)
    ctx.set_thermostat(_lib.THERM_LANGEVIN, kT, gamma)
    rng = np.random.default_rng(1)
    ctx.set_state(rng.uniform(-50, 50, (n, 3)), np.zeros((n, 3)))
    ctx.em_run(dt, N, seed=21)
    _, v = ctx.get_state()
    r = 1.0 - gamma * dt
    want = 2 * gamma * kT / m[0] * dt * (1 - r ** (2 * N)) / (1 - r * r)
    x = v[1:].ravel()
    assert abs(x.var() - want) < 5 * want * math.sqrt(2.0 / x.size), (x.var(), want)


def test_em_argument_errors(c1):
    ctx, _ = lj_ctx(c1, _lib.PREC_FP64_REF, therm=None)
    ctx.set_state(c1.u, c1.v)
    with pytest.raises(ValueError):
        ctx.em_run(0.005, 1)                     # no Langevin thermostat
    ctx.set_thermostat(_lib.THERM_BERENDSEN, 1.0, 0.1)
    with pytest.raises(ValueError):
        ctx.em_run(0.005, 1)
    w = gen.water_box(per_side=2, seed=1)
    cw = _lib.Context(len(w.masses), _lib.SYS_WATER, w.masses, w.charges)
    cw.set_boundary(_lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
    P = gen.SPCFW
    cw.add_lennard_jones(P["eps"], P["sigma"], 0.45 * w.L)
    cw.add_electrostatic(P["k_el"], 0.45 * w.L)
    cw.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    cw.set_thermostat(_lib.THERM_LANGEVIN, 300.0, 1.0)
    cw.set_state(w.u, w.v)
    with pytest.raises(NotImplementedError):
        cw.em_run(0.0005, 1)


def test_run_simulation_with_em(c1):
    """run_simulation(sim, EM(); dt) -- the reference's SDE path through the Python
    mirror: frames every step, stats.naccept, and the same final state as em_run."""
    bodies = [nbs.MassBody(c1.u[i], c1.v[i], 1.0) for i in range(864)]
    system = nbs.PotentialNBodySystem(bodies, {"lennard_jones": nbs.LennardJonesParameters(1.0, 1.0, 2.5)})
    sim = nbs.NBodySimulation(system, (0.0, 0.05), nbs.CubicPeriodicBoundaryConditions(c1.L),
                              nbs.LangevinThermostat(1.3, 1.0), 1.0)
    res = nbs.run_simulation(sim, nbs.EM(seed=5), dt=0.005)
    assert res.solution.stats.naccept == 10 and res.solution.u.shape == (11, 864, 3)
    ctx, _ = lj_ctx(c1, _lib.PREC_FP64_REF)
    ctx.set_state(c1.u, c1.v)
    ctx.em_run(0.005, 10, seed=5)
    u, v = ctx.get_state()
    assert np.array_equal(res.solution.u[-1], u) and np.array_equal(res.solution.v[-1], v)
