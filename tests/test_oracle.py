"""Pins the FP64 CPU oracle (oracle/nbx_oracle.c) without running Julia.

Each check is chosen so that a plausible restatement mistake (dropped term, wrong sign,
index or operand order) fails it: textbook closed forms, F = -grad U against the
oracle's own energy functions (a separate code path), Newton's third law, extended-
precision brute force, the reference's own closed-form tests (test/*.jl) and
published known-answer vectors for Philox4x32-10.
"""
import math

import numpy as np
import pytest

import oracle_lib as O


def lj_system(n, bc=None, eps=1.0, sigma=1.0, R=2.5, m=1.0):
    s = O.System(np.full(n, m), bc=bc)
    return s.add(O.TERM_LJ, eps, sigma, R)


# ---------------------------------------------------------------------------
# Boundary conditions (src/boundary_conditions.jl:111-172)
# ---------------------------------------------------------------------------

def test_periodic_bc_wraps_into_lo_hi_not_minimum_image():
    bc = O.make_bc(O.BC_PERIODIC, [0, 1, 0, 1, 0, 1])
    rij, r, r2 = O.interparticle_distance([0.1, 0.5, 0.95], [0.4, 0.3, 0.05], bc)
    np.testing.assert_allclose(rij, [0.7, 0.2, 0.9], rtol=0, atol=1e-15)
    assert r2 == pytest.approx(0.49 + 0.04 + 0.81, rel=1e-15)


def test_cubic_bc_is_half_open_minimum_image():
    bc = O.make_bc(O.BC_CUBIC, 1.0)
    rij, _, _ = O.interparticle_distance([0.6, -0.5, 0.5], [0, 0, 0], bc)
    # x >= L/2 wraps down, x = -L/2 stays, x = +L/2 wraps to -L/2
    np.testing.assert_array_equal(rij, [0.6 - 1.0, -0.5, -0.5])
    rij, _, _ = O.interparticle_distance([100.25, -3.25, 0.0], [0, 0, 0], bc)
    np.testing.assert_allclose(rij, [0.25, -0.25, 0.0], atol=1e-12)


def test_infinite_box_plain_difference():
    rij, r, r2 = O.interparticle_distance([3.0, 4.0, 0.0], [0, 0, 0], O.make_bc())
    assert r == 5.0 and r2 == 25.0
    np.testing.assert_array_equal(rij, [3.0, 4.0, 0.0])


# ---------------------------------------------------------------------------
# Force laws against closed forms
# ---------------------------------------------------------------------------

def test_lj_two_body_matches_minus_dU_dr():
    r = 1.3
    s = lj_system(2, R=math.inf)
    u = np.array([[-r / 2, 0, 0], [r / 2, 0, 0]])
    a = s.rhs(u)
    # U = 4 eps ((s/r)^12 - (s/r)^6);  F_0 = dU/dr * x_hat (pulls atom 0 towards +x if attractive)
    dUdr = 4.0 * (-12.0 / r ** 13 + 6.0 / r ** 7)
    np.testing.assert_allclose(a[0], [dUdr, 0, 0], rtol=1e-14)
    np.testing.assert_allclose(a[1], [-dUdr, 0, 0], rtol=1e-14)


def test_lj_cutoff_and_mass_scaling():
    s = O.System(np.array([2.0, 1.0])).add(O.TERM_LJ, 1.5, 1.1, 1.2)
    a = s.rhs(np.array([[0, 0, 0], [1.25, 0, 0]]))
    np.testing.assert_array_equal(a, 0.0)  # beyond R
    a = s.rhs(np.array([[0, 0, 0], [1.15, 0, 0]]))
    r = 1.15
    F = 4 * 1.5 * (12 * 1.1 ** 12 / r ** 13 - 6 * 1.1 ** 6 / r ** 7)
    np.testing.assert_allclose(a[:, 0], [-F / 2.0, F / 1.0], rtol=1e-13)


def test_coulomb_two_body():
    k, q1, q2, m1, m2 = 9.0e9, 1e-3, -2e-3, 1.0, 3.0
    s = O.System(np.array([m1, m2]), charges=np.array([q1, q2])).add(O.TERM_ELEC, k, math.inf)
    a = s.rhs(np.array([[0, 0, 0], [0, 2.0, 0]]))
    F = k * q1 * q2 / 4.0  # along -y for atom 0 when attractive
    np.testing.assert_allclose(a[0], [0, -F / m1, 0], rtol=1e-14)
    np.testing.assert_allclose(a[1], [0, F / m2, 0], rtol=1e-14)


def test_gravity_two_body():
    s = O.System(np.array([2.0, 5.0])).add(O.TERM_GRAV, 0.7)
    a = s.rhs(np.array([[0, 0, 0], [0, 0, 3.0]]))
    np.testing.assert_allclose(a[0], [0, 0, 0.7 * 5.0 / 9.0], rtol=1e-14)
    np.testing.assert_allclose(a[1], [0, 0, -0.7 * 2.0 / 9.0], rtol=1e-14)


def _fd_forces(energy, u, h=1e-6):
    g = np.zeros_like(u)
    for i in range(u.shape[0]):
        for k in range(3):
            up = u.copy(); up[i, k] += h
            dn = u.copy(); dn[i, k] -= h
            g[i, k] = -(energy(up) - energy(dn)) / (2 * h)
    return g


def _rand_config(n, L, seed, dmin=0.9):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, L, 3)
        ok = True
        for q in pts:
            d = p - q
            d -= L * np.round(d / L)
            if np.linalg.norm(d) < dmin:
                ok = False
        if ok:
            pts.append(p)
    return np.array(pts)


def _away_from_cutoff(u, L, R, margin=1e-3):
    n = len(u)
    for i in range(n):
        for j in range(i + 1, n):
            d = u[i] - u[j]
            d -= L * np.round(d / L)
            if abs(np.linalg.norm(d) - R) < margin:
                return False
    return True


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lj_force_is_minus_gradient_of_energy(seed):
    L, R = 5.0, 2.2
    u = _rand_config(12, L, seed)
    assert _away_from_cutoff(u, L, R)
    s = lj_system(12, bc=O.make_bc(O.BC_CUBIC, L), R=R, m=1.7)
    a = s.rhs(u)
    F = _fd_forces(s.potential_energy, u)
    np.testing.assert_allclose(a * 1.7, F, rtol=1e-5, atol=1e-6 * np.abs(F).max())


def test_coulomb_and_rf_forces_are_minus_gradient():
    L = 6.0
    u = _rand_config(10, L, 5)
    q = np.array([1, -1] * 5, dtype=float)
    m = np.linspace(1, 2, 10)
    bc = O.make_bc(O.BC_CUBIC, L)
    for kind, params in ((O.TERM_ELEC, (2.0, 2.5)), (O.TERM_RF, (2.0, 2.5, 78.5)),
                         (O.TERM_RF, (2.0, 2.5, math.inf))):
        assert _away_from_cutoff(u, L, 2.5)
        s = O.System(m, charges=q, bc=bc).add(kind, *params)
        a = s.rhs(u)
        F = _fd_forces(s.potential_energy, u)
        np.testing.assert_allclose(a * m[:, None], F, rtol=1e-5, atol=1e-7 * np.abs(F).max())


def test_gravity_force_is_minus_gradient():
    rng = np.random.default_rng(3)
    u = rng.normal(size=(9, 3))
    m = rng.uniform(0.5, 2, 9)
    s = O.System(m).add(O.TERM_GRAV, 1.3)
    a = s.rhs(u)
    F = _fd_forces(lambda x: O.grav_potential(x, m, 1.3), u)
    np.testing.assert_allclose(a * m[:, None], F, rtol=1e-6, atol=1e-8)


def water_system(nmol, L, seed, elec_kind=O.TERM_ELEC, jitter=0.02):
    from nbodysimulator.generators import SPCFW as P
    rng = np.random.default_rng(seed)
    side = int(round(nmol ** (1 / 3)))
    a = L / side
    g = (np.arange(side) + 0.5) * a
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    Opos = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1)[:nmol]
    from nbodysimulator.generators import random_rotations
    R = random_rotations(nmol, rng)
    h1 = np.array([P["rOH"] * 1.03, 0, 0])
    th = P["aHOH"] * 0.95
    h2 = np.array([P["rOH"] * 0.97 * math.cos(th), P["rOH"] * 0.97 * math.sin(th), 0])
    u = np.empty((3 * nmol, 3))
    u[0::3] = Opos
    u[1::3] = Opos + R @ h1
    u[2::3] = Opos + R @ h2
    u += jitter * 0.1 * rng.standard_normal(u.shape)
    ms = np.tile([P["mO"], P["mH"], P["mH"]], nmol)
    qs = np.tile([P["qO"], P["qH"], P["qH"]], nmol)
    elec = (P["k_el"], 0.45 * L) if elec_kind == O.TERM_ELEC else (P["k_el"], 0.45 * L, math.inf)
    s = O.System(ms, charges=qs, water=True, bc=O.make_bc(O.BC_CUBIC, L), kb=P["kb"])
    s.water_params((P["eps"], P["sigma"], 0.45 * L), elec_kind, elec, P["rOH"], P["aHOH"],
                   P["kbond"], P["kangle"])
    return s, u, ms


@pytest.mark.parametrize("elec_kind", [O.TERM_ELEC, O.TERM_RF])
def test_water_rhs_is_minus_gradient_of_water_energy(elec_kind):
    s, u, ms = water_system(8, 1.0, 11, elec_kind)
    a = s.rhs(u)
    F = _fd_forces(s.potential_energy, u, h=1e-7)
    np.testing.assert_allclose(a * ms[:, None], F, rtol=2e-5, atol=2e-5 * np.abs(F).max())


def test_newton_third_law_cubic_and_infinite():
    for bc in (O.make_bc(O.BC_CUBIC, 7.0), O.make_bc()):
        u = _rand_config(40, 7.0, 9, dmin=0.95)
        m = np.random.default_rng(1).uniform(1, 3, 40)
        s = O.System(m, bc=bc).add(O.TERM_LJ, 1.0, 1.0, 2.5)
        tot = (s.rhs(u) * m[:, None]).sum(axis=0)
        scale = np.abs(s.rhs(u) * m[:, None]).sum()
        assert np.abs(tot).max() <= 1e-13 * scale


def test_lj_matches_extended_precision_brute_force():
    L, R = 6.0, 2.5
    u = _rand_config(24, L, 21, dmin=0.85)
    s = lj_system(24, bc=O.make_bc(O.BC_CUBIC, L), R=R)
    a = s.rhs(u)
    ul = u.astype(np.longdouble)
    ref = np.zeros_like(ul)
    for i in range(24):
        d = ul[i] - ul
        d -= np.longdouble(L) * np.round(d / np.longdouble(L))
        r2 = (d * d).sum(axis=1)
        r2[i] = np.inf
        inr = r2 < R * R
        s6 = (1 / r2[inr]) ** 3
        ref[i] = (24 * ((2 * s6 * s6 - s6) / r2[inr])[:, None] * d[inr]).sum(axis=0)
    err = np.abs(a - ref.astype(np.float64)).max() / np.abs(ref).max()
    assert err < 1e-13


def test_angle_force_alone_is_minus_gradient():
    # a single molecule: the only forces are bond + angle; remove the bond term by
    # putting both O-H lengths at rOH, so only the angle acts
    from nbodysimulator.generators import SPCFW as P
    ms = np.array([P["mO"], P["mH"], P["mH"]])
    qs = np.zeros(3)
    s = O.System(ms, charges=qs, water=True, kb=P["kb"])
    s.water_params((P["eps"], P["sigma"], 1.0), O.TERM_ELEC, (0.0, 1.0), P["rOH"], P["aHOH"],
                   P["kbond"], P["kangle"])
    th = P["aHOH"] + 0.2
    r0 = P["rOH"]
    u = np.array([[0, 0, 0], [r0, 0, 0], [r0 * math.cos(th), r0 * math.sin(th), 0.0]])
    a = s.rhs(u)
    # analytic: U = ka/2 (th - th0)^2; torque-free in-plane forces of magnitude dU/dth / r0
    dU = P["kangle"] * 0.2
    f1 = a[1] * ms[1]
    assert np.linalg.norm(f1) == pytest.approx(dU / r0, rel=1e-12)
    assert f1[1] > 0  # H1 on +x is pushed towards H2 (angle shrinks)
    np.testing.assert_allclose((a * ms[:, None]).sum(axis=0), 0, atol=1e-9 * dU)
    F = _fd_forces(s.potential_energy, u, h=1e-8)
    np.testing.assert_allclose(a * ms[:, None], F, rtol=1e-5, atol=1e-6 * np.abs(F).max())


# ---------------------------------------------------------------------------
# Velocity Verlet: the reference's closed-form tests
# ---------------------------------------------------------------------------

def test_lj_two_particles_speed_from_energy():  # test/lennard_jones_test.jl:4-34
    r1 = 1.3
    s = lj_system(2, R=math.inf)
    u0 = np.array([[-r1 / 2, 0, 0], [r1 / 2, 0, 0]])
    u, v, _ = s.vv_run(u0, np.zeros((2, 3)), 1e-3, 1000)
    r2 = np.linalg.norm(u[1] - u[0])
    v_exp = math.sqrt(4 * ((1 / r1) ** 12 - (1 / r2) ** 12 - ((1 / r1) ** 6 - (1 / r2) ** 6)))
    assert np.linalg.norm(v[1]) == pytest.approx(v_exp, rel=1e-3)


FIG8_U = np.array([[-0.995492, 0, 0], [0.995492, 0, 0], [0, 0, 0.0]])
FIG8_V = np.array([[-0.347902, -0.53393, 0], [-0.347902, -0.53393, 0], [0.695804, 1.06786, 0.0]])


def test_figure_eight_returns_after_period():  # test/gravitational_test.jl:233-240
    s = O.System(np.ones(3)).add(O.TERM_GRAV, 1.0)
    u, v, _ = s.vv_run(FIG8_U, FIG8_V, math.pi / 130, 260)
    np.testing.assert_allclose(u, FIG8_U, atol=1e-3)
    assert s.kinetic_energy(FIG8_V) == pytest.approx(1.218, abs=1e-3)


def test_two_charges_repel_speed_from_energy():  # test/electrostatics_test.jl:42-68
    k, q = 9.0e9, 1.0e-3
    s = O.System(np.ones(2), charges=np.array([q, q])).add(O.TERM_ELEC, k, math.inf)
    u0 = np.array([[-0.5, 0, 0], [0.5, 0, 0]])
    u, v, _ = s.vv_run(u0, np.zeros((2, 3)), 1e-3, 1000)
    r2 = np.linalg.norm(u[1] - u[0])
    v_exp = math.sqrt(k * q * q * (1 / 1.0 - 1 / r2))
    assert np.linalg.norm(v[1]) == pytest.approx(v_exp, rel=1e-3)


def test_eight_charges_energy_conservation():  # test/electrostatics_test.jl:84-119
    n, L, m, q, k = 8, 1.0, 1.0, 1.0, 9.0e9
    dL = L / (math.ceil(n ** (1 / 3)) + 1)
    pts = []
    xs = np.arange(dL / 2, L + 1e-12, dL)
    for x in xs:
        for y in xs:
            for z in xs:
                if len(pts) < n:
                    pts.append([x, y, z])
    u0 = np.array(pts)
    tau = 0.01 * dL / math.sqrt(2 * k * q * q / (dL * m))
    s = O.System(np.full(n, m), charges=np.full(n, q), bc=O.make_bc(O.BC_CUBIC, L))
    s.add(O.TERM_ELEC, k, 0.45 * L)
    u, v, _ = s.vv_run(u0, np.zeros((n, 3)), tau, 1000)
    e1 = s.potential_energy(u0)
    e2 = s.potential_energy(u) + s.kinetic_energy(v)
    assert abs((e2 - e1) / e1) < 1e-3


def _kepler_error(dt):
    s = O.System(np.array([1.0, 1e-9])).add(O.TERM_GRAV, 1.0)
    u0 = np.array([[0, 0, 0], [1.0, 0, 0]])
    v0 = np.array([[0, 0, 0], [0, 1.0, 0]])
    T = 2 * math.pi
    n = int(round(T / dt))
    u, _, _ = s.vv_run(u0, v0, T / n, n)
    return np.linalg.norm(u[1] - u0[1])


def test_vv_is_second_order():
    e1, e2 = _kepler_error(2 * math.pi / 200), _kepler_error(2 * math.pi / 400)
    assert 3.6 < e1 / e2 < 4.4


def test_berendsen_relaxation_of_free_particles():
    n, T0, Ti, tau = 200, 2.0, 1.0, 0.05
    rng = np.random.default_rng(0)
    v0 = rng.standard_normal((n, 3))
    m = np.ones(n)
    v0 *= math.sqrt(Ti / ((v0 * v0).sum() / (3 * n)))
    s = O.System(m).thermostat(O.THERM_BERENDSEN, T0, tau)
    dt, steps = 1e-4, 1000
    u, v, _ = s.vv_run(np.zeros((n, 3)), v0, dt, steps)
    T = (v * v).sum() / (3 * n)
    expect = T0 + (Ti - T0) * math.exp(-steps * dt / tau)
    assert T == pytest.approx(expect, rel=2e-3)


@pytest.mark.parametrize("kind,param", [(O.THERM_ANDERSEN, 50.0), (O.THERM_LANGEVIN, 5.0)])
def test_stochastic_thermostats_reach_target_temperature(kind, param):
    n, T0 = 400, 1.7
    m = np.full(n, 2.0)
    s = O.System(m).thermostat(kind, T0, param)
    u, v, _ = s.vv_run(np.zeros((n, 3)), np.zeros((n, 3)), 0.01, 400, seed=5)
    T = (m[:, None] * v * v).sum() / (3 * n)
    assert T == pytest.approx(T0, rel=0.08)
    # Maxwellian: <v^4> / <v^2>^2 = 3 per component
    k4 = (v ** 4).mean() / (v ** 2).mean() ** 2
    assert k4 == pytest.approx(3.0, rel=0.15)


def test_kinetic_energy_and_temperature_of_argon_triplet():  # test/lennard_jones_test.jl:47-82
    T, kb, m = 120.0, 8.3144598e-3, 39.95
    vd = math.sqrt(3 * kb * T / m)
    v = np.array([[0, 0, -vd], [0, -vd, 0], [-vd, 0, 0]])
    ms = np.full(3, m)
    lib = O.load()
    assert lib.orc_md_temperature(O.dp(v), O.dp(ms), kb, 3, 0) == pytest.approx(120.0, abs=1e-6)
    ke = O.System(ms).kinetic_energy(v)
    assert ke == pytest.approx(m * 3 * vd * vd / 2, rel=1e-15)


# ---------------------------------------------------------------------------
# RDF (src/nbody_simulation_result.jl:664-709)
# ---------------------------------------------------------------------------

def test_rdf_bin_convention():
    L = 1.0
    f = np.array([[[0.1, 0.1, 0.1], [0.35, 0.1, 0.1]]])   # r = 0.25 -> bin ceil(2.5) = 3
    rs, gr = O.rdf(f, L, maxbin=10)
    nideal = 4 / 3 * math.pi * 2 / L ** 3 * (0.3 ** 3 - 0.2 ** 3)
    assert gr[2] == pytest.approx((2 / (1 * 2)) / nideal, rel=1e-12)
    assert np.count_nonzero(gr) == 1
    np.testing.assert_allclose(rs, (np.arange(10) + 0.5) * 0.1, rtol=1e-12)


def test_rdf_of_ideal_gas_is_one():
    rng = np.random.default_rng(0)
    frames = rng.uniform(0, 4.0, (20, 300, 3))
    rs, gr = O.rdf(frames, 4.0, maxbin=20)
    sel = (rs > 0.8) & (rs < 1.9)
    assert np.abs(gr[sel] - 1).max() < 0.08


# ---------------------------------------------------------------------------
# Philox4x32-10 known-answer vectors (Random123 kat_vectors)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(ctr, key, expect):
    assert tuple(O.philox(ctr, key)) == expect


# ---------------------------------------------------------------------------------
# Magnetic dipoles (src/basic_potentials.jl:333-365)
# ---------------------------------------------------------------------------------

def test_mag_two_dipoles_closed_form(oracle):
    # parallel moments along z, separation along x: F_1 = -3 mu m^2 / d^4 x_hat
    # (U = mu m^2 / d^3 falls with d, so the dipoles repel)
    d, m, mu, mass = 0.7, 1.3, 0.9, 2.0
    u = np.array([[-d / 2, 0.0, 0.0], [d / 2, 0.0, 0.0]])
    mm = np.array([[0.0, 0.0, m], [0.0, 0.0, m]])
    s = oracle.System([mass, mass]).moments(mm).add(oracle.TERM_MAG, mu)
    a = s.rhs(u)
    F = 3 * mu * m * m / d ** 4
    assert np.allclose(a[0], [-F / mass, 0, 0], rtol=1e-14)
    assert np.allclose(a[1], [F / mass, 0, 0], rtol=1e-14)
    # head-to-tail along the separation: attraction of magnitude 6 mu m^2 / d^4
    mm2 = np.array([[m, 0.0, 0.0], [m, 0.0, 0.0]])
    a2 = oracle.System([mass, mass]).moments(mm2).add(oracle.TERM_MAG, mu).rhs(u)
    assert np.allclose(a2[0], [6 * mu * m * m / d ** 4 / mass, 0, 0], rtol=1e-14)


def test_mag_force_is_minus_energy_gradient(oracle):
    rng = np.random.default_rng(31)
    n = 6
    u = rng.uniform(-1.5, 1.5, size=(n, 3))
    mm = rng.normal(size=(n, 3))
    ms = rng.uniform(0.5, 2.0, n)
    mu = 0.37
    a = oracle.System(ms).moments(mm).add(oracle.TERM_MAG, mu).rhs(u)
    h = 1e-6
    for i in range(n):
        for k in range(3):
            up, um = u.copy(), u.copy()
            up[i, k] += h
            um[i, k] -= h
            g = (oracle.mag_potential(up, mm, mu) - oracle.mag_potential(um, mm, mu)) / (2 * h)
            assert abs(-g / ms[i] - a[i, k]) <= 1e-6 * max(1.0, abs(a[i, k])), (i, k)
    # Newton's third law: total force vanishes
    assert np.abs((a * ms[:, None]).sum(0)).max() <= 1e-10 * np.abs(a * ms[:, None]).max()


# ---------------------------------------------------------------------------------
# Nose-Hoover (src/thermostats.jl:116-122)
# ---------------------------------------------------------------------------------

def test_nosehoover_free_particles_follow_the_thermostat_ode(oracle):
    """Free particles under the reference's Nose-Hoover law: dT/dt = -2 zeta T and
    dzeta/dt = (T/T0 - (ndf+1)/ndf) / tau^2 (zeta(0) = 0).  The oracle's velocity
    Verlet at small dt follows a high-accuracy solution of that ODE; a wrong sign, a
    missing 1/tau^2 or the plain T/T0 - 1 would not."""
    from scipy.integrate import solve_ivp
    n, m, kb, T0, tau = 50, 2.0, 1.0, 1.0, 0.5
    rng = np.random.default_rng(41)
    v = rng.normal(size=(n, 3))
    u = rng.uniform(-5, 5, size=(n, 3))
    ndf = 3 * n
    Ti = m * (v * v).sum() / (kb * ndf)
    s = oracle.System(np.full(n, m), kb=kb).thermostat(oracle.THERM_NOSEHOOVER, T0, tau)
    r = (ndf + 1) / ndf
    tend = 4.0
    sol = solve_ivp(lambda t, y: [-2 * y[1], (np.exp(y[0]) / T0 - r) / tau ** 2],
                    (0, tend), [np.log(Ti), 0.0], t_eval=np.linspace(0, tend, 11),
                    rtol=1e-11, atol=1e-12)
    Te = np.exp(sol.y[0])
    errs = []
    for dt in (2e-3, 1e-3, 5e-4):
        steps = int(round(tend / dt))
        _, _, (uf, vf) = s.vv_run(u, v, dt, steps, save_every=steps // 10)
        Tg = np.array([m * (w * w).sum() / (kb * ndf) for w in vf])
        errs.append(np.abs(Tg / Te - 1).max())
    # the reference's coupling (zeta advanced with the previous rate) is first order in dt:
    # the error halves with dt and goes to zero, i.e. the scheme converges to this ODE
    assert errs[2] < 6e-3, errs
    assert 1.6 < errs[0] / errs[1] < 2.5 and 1.6 < errs[1] / errs[2] < 2.5, errs
    dt = 1e-3
    # equilibrium point: T = T0 (ndf+1)/ndf keeps zeta at zero and the velocities fixed
    v_eq = v * np.sqrt(T0 * r / Ti)
    _, v1, _ = s.vv_run(u, v_eq, dt, 50)
    assert np.allclose(v1, v_eq, rtol=1e-12, atol=1e-14)


# ---------------------------------------------------------------------------
# Euler-Maruyama for the Langevin SDE (src/nbody_to_ode.jl:567-598)
# ---------------------------------------------------------------------------

def test_em_free_particles_without_noise_closed_form():
    """T = 0 (sigma = 0), no forces: v_n = r^n v_0 and x_n = x_0 + v_0 (1 - r^n) / gamma
    with r = 1 - gamma dt, the exact Euler-Maruyama solution of dv = -gamma v dt."""
    rng = np.random.default_rng(3)
    n, gamma, dt, steps = 6, 0.7, 0.01, 300
    u0, v0 = rng.uniform(-3, 3, (n, 3)), rng.normal(size=(n, 3))
    s = O.System(np.linspace(1.0, 2.0, n)).thermostat(O.THERM_LANGEVIN, 0.0, gamma)
    u, v, _ = s.em_run(u0, v0, dt, steps, seed=5)
    r = 1.0 - gamma * dt
    assert np.allclose(v, r ** steps * v0, rtol=1e-12, atol=0)
    assert np.allclose(u, u0 + v0 * (1 - r ** steps) / gamma, rtol=1e-11, atol=1e-13)


def test_em_one_step_two_body_gravity():
    """One step from (x, v) with sigma = 0: x' = x + dt v and v' = v + dt (a(x) - gamma v),
    with a(x) the two-body Newton acceleration written out here; a dropped force, a
    force taken at x' or a sign error would not match."""
    m = np.array([2.0, 5.0])
    G, gamma, dt = 0.7, 0.3, 0.05
    u0 = np.array([[0.0, 0.0, 0.0], [1.5, 0.5, -0.2]])
    v0 = np.array([[0.1, -0.2, 0.3], [-0.4, 0.2, 0.1]])
    s = O.System(m).add(O.TERM_GRAV, G).thermostat(O.THERM_LANGEVIN, 0.0, gamma)
    u, v, _ = s.em_run(u0, v0, dt, 1)
    d = u0[1] - u0[0]
    r3 = np.linalg.norm(d) ** 3
    a = np.array([G * m[1] * d / r3, -G * m[0] * d / r3])
    assert np.allclose(u, u0 + dt * v0, rtol=1e-14, atol=1e-15)
    assert np.allclose(v, v0 + dt * (a - gamma * v0), rtol=1e-13, atol=1e-15)


def test_em_noise_variance_uses_first_body_mass():
    """Free particles from v = 0: v_{k+1} = r v_k + sigma sqrt(dt) xi, so after N steps
    every velocity component has variance sigma^2 dt (1 - r^(2N)) / (1 - r^2) with
    sigma^2 = 2 gamma kb T / m_1 -- the first body's mass for every atom, as in the
    reference (the light atoms below then sit at about half the target temperature).
    Tolerance: 5 standard errors of a variance estimated from ~6,000 normal samples."""
    n, gamma, kT, dt, N = 2001, 1.0, 1.5, 0.01, 400
    m = np.ones(n)
    m[0] = 2.0
    s = O.System(m, kb=1.0).thermostat(O.THERM_LANGEVIN, kT, gamma)
    u0 = np.zeros((n, 3))
    _, v, _ = s.em_run(u0, np.zeros((n, 3)), dt, N, seed=11)
    r = 1.0 - gamma * dt
    sig2 = 2 * gamma * kT / m[0]
    want = sig2 * dt * (1 - r ** (2 * N)) / (1 - r * r)
    x = v[1:].ravel()
    se = want * math.sqrt(2.0 / x.size)
    assert abs(x.var() - want) < 5 * se, (x.var(), want)
    assert abs(x.mean()) < 5 * math.sqrt(want / x.size)


# ---------------------------------------------------------------------------
# Energy conventions (src/nbody_simulation_result.jl:293-351): closed forms that fix the
# constants, so that dropping the Lennard-Jones far-pair tail or the reaction-field
# shift fails a test (finite-difference checks cannot see a constant)
# ---------------------------------------------------------------------------

def _U_lj(r2, eps=1.0, s2=1.0):
    q = s2 / r2
    return 4.0 * eps * (q ** 6 - q ** 3)


def test_lj_energy_two_atoms_inside_and_beyond_cutoff():
    R = 2.5
    s = lj_system(2, R=R, eps=1.3, sigma=1.1)
    for r in (1.05, 2.0, 2.4999):
        u = np.array([[0.0, 0, 0], [r, 0, 0]])
        assert s.potential_energy(u) == pytest.approx(_U_lj(r * r, 1.3, 1.21), rel=1e-14)
    # at and beyond R every pair adds U(R), the reference's constant tail (:308-310)
    for r in (R, 2.6, 7.0, 1e3):
        u = np.array([[0.0, 0, 0], [0, 0, r]])
        assert s.potential_energy(u) == pytest.approx(_U_lj(R * R, 1.3, 1.21), rel=1e-14)


def test_lj_energy_three_atoms_near_and_far_pairs():
    R = 2.5
    s = lj_system(3, R=R)
    u = np.array([[0.0, 0, 0], [1.2, 0, 0], [0, 0, 9.0]])   # one pair inside, two far
    expect = _U_lj(1.44) + 2.0 * _U_lj(R * R)
    assert s.potential_energy(u) == pytest.approx(expect, rel=1e-14)


def test_lj_energy_is_continuous_across_the_cutoff():
    R = 2.5
    s = lj_system(2, R=R)
    e_in = s.potential_energy(np.array([[0.0, 0, 0], [R * (1 - 1e-12), 0, 0]]))
    e_out = s.potential_energy(np.array([[0.0, 0, 0], [R * (1 + 1e-12), 0, 0]]))
    assert e_in == pytest.approx(e_out, rel=1e-9)


def test_lj_energy_of_a_liquid_is_shifted_sum_plus_tail():
    """E_ref = sum over pairs within R of (U(r) - U(R)) + n(n-1)/2 U(R), with the shifted
    sum taken independently here (numpy, minimum image, all pairs at once)."""
    from nbodysimulator import generators as gen
    st = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=4, jitter=0.02)   # N = 864
    n, L, R = st.u.shape[0], st.L, 2.5
    s = lj_system(n, bc=O.make_bc(O.BC_CUBIC, L), R=R)
    d = st.u[:, None, :] - st.u[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.einsum("ijk,ijk->ij", d, d)
    iu = np.triu_indices(n, 1)
    r2 = r2[iu]
    inside = r2 < R * R
    e_shift = np.sum(_U_lj(r2[inside]) - _U_lj(R * R))
    tail = n * (n - 1) / 2 * _U_lj(R * R)
    # both sums cancel to ~1e-12 relative rounding; the tail is ~half the total here, so
    # a dropped or doubled tail misses by ~50 %
    assert s.potential_energy(st.u) == pytest.approx(e_shift + tail, rel=1e-10)
    assert abs(tail) > 0.3 * abs(e_shift + tail)


def _krf(R, eps_rf):
    return 1.0 / (2 * R ** 3) if math.isinf(eps_rf) else (eps_rf - 1) / ((2 * eps_rf + 1) * R ** 3)


@pytest.mark.parametrize("eps_rf", [78.5, math.inf])
def test_reaction_field_pair_energy_closed_form(eps_rf):
    k, R, q1, q2 = 138.935458, 0.9, -0.82, 0.41
    s = O.System(np.array([15.999, 1.008]), charges=np.array([q1, q2])).add(O.TERM_RF, k, R, eps_rf)
    krf = _krf(R, eps_rf)
    crf = 1.0 / R + krf * R * R
    for r in (0.25, 0.5, 0.89):
        u = np.array([[0.0, 0, 0], [0, r, 0]])
        expect = k * q1 * q2 * (1.0 / r + krf * r * r - crf)
        assert s.potential_energy(u) == pytest.approx(expect, rel=1e-13)
    # shifted to zero at the cutoff: continuous there, zero beyond
    e_in = s.potential_energy(np.array([[0.0, 0, 0], [0, R * (1 - 1e-10), 0]]))
    assert abs(e_in) < 1e-8 * k * abs(q1 * q2) / R
    assert s.potential_energy(np.array([[0.0, 0, 0], [0, 1.5 * R, 0]])) == 0.0


def test_truncated_coulomb_energy_uses_q_over_R_beyond_cutoff():
    k, R, q1, q2 = 2.0, 1.5, 0.7, -0.3
    s = O.System(np.array([1.0, 1.0]), charges=np.array([q1, q2])).add(O.TERM_ELEC, k, R)
    for r, expect in ((1.0, k * q1 * q2 / 1.0), (3.0, k * q1 * q2 / R)):
        u = np.array([[0.0, 0, 0], [r, 0, 0]])
        assert s.potential_energy(u) == pytest.approx(expect, rel=1e-14)
