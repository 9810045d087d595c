"""Magnetic dipoles (src/basic_potentials.jl:333-365) on the GPU: the FP64 reference-order
kernel against the oracle bit for bit, the mixed FP32 tiled kernel at 2e-5, the closure
protocol per atom range, and the reference's own functional test (repelling dipoles,
test/magnetostaic_test.jl:5-32) through the Python mirror of the reference API."""
import math

import numpy as np
import pytest

import oracle_lib as O
import nbodysimulator as nbs
from nbodysimulator import _lib

pytestmark = pytest.mark.gpu


def dipoles(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-3.0, 3.0, size=(n, 3))
    mm = rng.normal(size=(n, 3))
    ms = rng.uniform(0.5, 2.0, n)
    return u, mm, ms


def mag_ctx(u, mm, ms, mu, prec):
    n = len(ms)
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, ms)
    ctx.set_moments(mm)
    ctx.add_magnetostatic(mu)
    ctx.set_precision(prec)
    orc = O.System(ms).moments(mm).add(O.TERM_MAG, mu)
    return ctx, orc


def test_fp64_magnetic_forces_bitwise(nbx):
    print("[case] fp64 magnetic dipoles N=700", flush=True)
    u, mm, ms = dipoles(700, 1)
    ctx, orc = mag_ctx(u, mm, ms, 0.7, _lib.PREC_FP64_REF)
    a = ctx.rhs(u)
    assert np.array_equal(a, orc.rhs(u))
    # the closure protocol on a sub-range adds exactly those rows
    dv = np.zeros_like(u)
    ctx.accelerations(0, u, 100, 250, dv)
    ref = orc.rhs(u)
    assert np.array_equal(dv[100:250], ref[100:250])
    assert not dv[:100].any() and not dv[250:].any()


def test_mixed_magnetic_forces(nbx):
    print("[case] mixed magnetic dipoles N=2048", flush=True)
    u, mm, ms = dipoles(2048, 2)
    ctx, orc = mag_ctx(u, mm, ms, 0.7, _lib.PREC_MIXED)
    ctx.set_state(u, np.zeros_like(u))
    dt = 2.0 ** -60
    _, vf = ctx.vv_run(dt, 1, save_every=1)
    a = (vf[1] - vf[0]) / dt
    ref = orc.rhs(u)
    err = np.abs(a - ref).max() / np.abs(ref).max()
    assert err <= 2e-5, err
    assert ctx.count_pairs() == 2048 * 2047


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_reference_repelling_dipoles(nbx, precision):
    print(f"[case] reference repelling dipoles {precision}", flush=True)
    d1 = 0.01
    m1 = m2 = 5.0e-6
    rho = 7800
    M = 1.2e6
    mm1 = (0.0, 0.0, M * m1 / rho)
    mm2 = (0.0, 0.0, M * m2 / rho)
    mu_4pi = 1.0e-7
    t1, t2 = 0.0, 1.0
    tau = (t2 - t1) / 100
    p1 = nbs.MagneticParticle((-d1 / 2, 0.0, 0.0), (0.0, 0.0, 0.0), m1, mm1)
    p2 = nbs.MagneticParticle((d1 / 2, 0.0, 0.0), (0.0, 0.0, 0.0), m2, mm2)
    system = nbs.PotentialNBodySystem([p1, p2], {"magnetic": nbs.MagnetostaticParameters(mu_4pi)})
    sim = nbs.NBodySimulation(system, (t1, t2))
    res = nbs.run_simulation(sim, nbs.VelocityVerlet(precision=precision), dt=tau)
    d2 = np.linalg.norm(nbs.get_position(res, t2, 2) - nbs.get_position(res, t2, 1))
    v_expected = math.sqrt(mu_4pi * (np.dot(mm1, mm2) * (1 / d1 ** 3 - 1 / d2 ** 3)) / m1)
    v_actual = np.linalg.norm(nbs.get_velocity(res, t2, 2))
    assert abs(v_expected - v_actual) <= 0.001
    if precision == "fp64":
        # the FP64 path is the oracle's velocity Verlet bit for bit
        orc = O.System([m1, m2]).moments(np.array([mm1, mm2])).add(O.TERM_MAG, mu_4pi)
        u0 = np.array([[-d1 / 2, 0.0, 0.0], [d1 / 2, 0.0, 0.0]])
        uo, vo, _ = orc.vv_run(u0, np.zeros_like(u0), tau, 100)
        assert np.array_equal(nbs.get_position(res, t2).T, uo)
