"""Host-side mirror of the reference API (no GPU): types, defaults, show strings,
step planning and the seeded generators."""
import math

import numpy as np
import pytest

import nbodysimulator as nb
from nbodysimulator import generators as gen
from nbodysimulator.simulation import _step_plan


def test_lennard_jones_defaults_and_show():  # test/lennard_jones_test.jl:152-166
    p = nb.LennardJonesParameters()
    assert (p.ϵ, p.σ, p.R, p.σ2, p.R2) == (1.0, 1.0, 2.5, 1.0, 6.25)
    assert p.show() == "Lennard-Jones:\n\tϵ:1.0\n\tσ:1.0\n\tR:2.5\n"
    assert nb.LennardJonesParameters(2, 5, 10).show() == "Lennard-Jones:\n\tϵ:2\n\tσ:5\n\tR:10\n"


def test_other_parameter_defaults_and_show():
    assert nb.GravitationalParameters().G == 6.67408e-11
    assert nb.GravitationalParameters().show() == "Gravitational:\n\tG:6.67408e-11\n"
    assert nb.GravitationalParameters(35.67).show() == "Gravitational:\n\tG:35.67\n"
    assert nb.ElectrostaticParameters().show() == "Electrostatic:\n\tk:9.0e9\n"
    assert nb.ElectrostaticParameters(1.0).show() == "Electrostatic:\n\tk:1.0\n"
    assert nb.ElectrostaticParameters(9.0e9, 8.0).R2 == 64.0
    assert nb.BerendsenThermostat(300.0, 0.1).γ == 0.5 / 0.1


def test_potential_system_from_symbols_show():  # test/shared/custom_potential_body.jl:42-52
    bodies = [nb.MassBody([1, 0, 0], [0, 0, 0], 1.0), nb.MassBody([-1, 0, 0], [0, 0, 0], 1.0)]
    s = nb.PotentialNBodySystem(bodies, [":lennard_jones", ":electrostatic", ":gravitational",
                                         ":magnetostatic"])
    assert s.show() == ("Potentials: \nLennard-Jones:\n\tϵ:1.0\n\tσ:1.0\n\tR:2.5\n"
                        "Electrostatic:\n\tk:9.0e9\nMagnetostatic:\n\tμ/4π:1.0e-7\n"
                        "Gravitational:\n\tG:6.67408e-11\n")


def test_simulation_show():  # test/lennard_jones_test.jl:37-41
    p1 = nb.MassBody([-0.65, 0, 0], [0, 0, 0], 1.0)
    p2 = nb.MassBody([0.65, 0, 0], [0, 0, 0], 1.0)
    system = nb.PotentialNBodySystem([p1, p2], {"lennard_jones": nb.LennardJonesParameters(1.0, 1.0, math.inf)})
    assert system.show() == "Potentials: \nLennard-Jones:\n\tϵ:1.0\n\tσ:1.0\n\tR:Inf\n"
    sim = nb.NBodySimulation(system, (0.0, 1.0))
    assert sim.show() == ("Timespan: (0.0, 1.0)\nBoundary conditions: InfiniteBox{Float64}("
                          "[-Inf, Inf, -Inf, Inf, -Inf, Inf])\nPotentials: \nLennard-Jones:\n"
                          "\tϵ:1.0\n\tσ:1.0\n\tR:Inf\n")


def test_periodic_boundary_iteration_and_indexing():  # test/lennard_jones_test.jl:168-179
    pbc = nb.PeriodicBoundaryConditions(15.0e-6)
    assert list(pbc) == [0, 15.0e-6, 0, 15.0e-6, 0, 15.0e-6]
    boundary = [0.0, 15.0e-6, 0.0, 15.0e-6, 0.0, 15.0e-6]
    for i in range(1, 7):
        assert pbc[i] == boundary[i - 1]
    with pytest.raises(IndexError):
        pbc[7]


def test_masses_and_degrees_of_freedom():
    mols = [nb.MassBody([0, 0, 0], [0, 0, 0], 18.0)] * 2
    w = nb.WaterSPCFw(mols, 1.0, 16.0, 0.41, -0.82, nb.LennardJonesParameters(),
                      nb.ElectrostaticParameters(), nb.SPCFwParameters(0.1012, 1.9764, 1.0, 1.0))
    np.testing.assert_array_equal(nb.get_masses(w), [16, 1, 1, 16, 1, 1])
    from nbodysimulator.types import get_degrees_of_freedom
    assert get_degrees_of_freedom(w) == (6, 4, 14)   # test/water_test.jl:333-336


def test_water_massbody_expansion():  # src/nbody_to_ode.jl:37-55
    p = nb.SPCFwParameters(0.1012, 1.9764, 1.0, 1.0)
    w = nb.WaterSPCFw([nb.MassBody([1, 2, 3], [0.1, 0, 0], 18.0)], 1.0, 16.0, 0.41, -0.82,
                      nb.LennardJonesParameters(), nb.ElectrostaticParameters(), p)
    sim = nb.NBodySimulation(w, (0.0, 1.0))
    from nbodysimulator.types import gather_bodies_initial_coordinates
    u, v, n = gather_bodies_initial_coordinates(sim)
    np.testing.assert_allclose(u[1] - u[0], [0.1012, 0, 0])
    np.testing.assert_allclose(u[2] - u[0], [math.cos(1.9764) * 0.1012, 0, math.sin(1.9764) * 0.1012])
    np.testing.assert_array_equal(v, [[0.1, 0, 0]] * 3)


def test_step_plan_matches_fixed_dt_semantics():
    assert _step_plan(0.0, 1.0, 1e-3) == (1000, 1e-3, None)       # "Time steps: 1000"
    n, dt, last = _step_plan(0.0, 1.0, 0.3)
    assert n == 3 and last == pytest.approx(0.1)


def test_fcc_generator_counts_and_density():
    s = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=1)
    assert s.u.shape == (864, 3)
    assert s.L == pytest.approx((864 / 0.8442) ** (1 / 3))
    assert s.L == pytest.approx(10.077577, rel=1e-6)
    d = s.u[:, None, :] - s.u[None, :, :]
    d -= s.L * np.round(d / s.L)
    r = np.sqrt((d * d).sum(-1)) + np.eye(864) * 10
    assert r.min() == pytest.approx(s.L / 6 / math.sqrt(2), rel=1e-12)   # FCC nearest neighbour
    np.testing.assert_allclose(s.v.sum(0), 0, atol=1e-12)
    assert (s.v * s.v).sum() / (3 * 864) == pytest.approx(0.72, rel=1e-12)


def test_plummer_virial_equilibrium():
    u, v, m = gen.plummer(8192, seed=3)
    ke = 0.5 * (m[:, None] * v * v).sum()
    assert ke == pytest.approx(3 * math.pi / 64, rel=0.05)
    # W from a sub-sample of pairs (exact W is O(N^2))
    idx = np.random.default_rng(0).choice(8192, 1500, replace=False)
    d = u[idx, None, :] - u[None, idx, :]
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    W = -0.5 * (1 / 8192) ** 2 * (1 / r).sum() * (8192 / 1500) ** 2
    assert W == pytest.approx(-3 * math.pi / 32, rel=0.08)
    assert 2 * ke / -W == pytest.approx(1.0, rel=0.1)


def test_water_box_geometry():
    w = gen.water_box(per_side=4, seed=2)
    P = gen.SPCFW
    O, H1, H2 = w.u[0::3], w.u[1::3], w.u[2::3]
    np.testing.assert_allclose(np.linalg.norm(H1 - O, axis=1), P["rOH"], rtol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(H2 - O, axis=1), P["rOH"], rtol=1e-12)
    c = ((H1 - O) * (H2 - O)).sum(1) / P["rOH"] ** 2
    np.testing.assert_allclose(np.arccos(c), P["aHOH"], rtol=1e-10)
    full = gen.water_box(per_side=16, seed=2)
    assert full.L == pytest.approx(4.97182, rel=1e-4)   # 4096 molecules at 0.997 g/cm^3


def test_batch_table():
    reps = gen.lj_batch(512, seed=0)
    Ns = np.array([r.N for r in reps])
    assert set(Ns) <= {4 * n ** 3 for n in range(3, 11)}
    assert 108 <= Ns.min() and Ns.max() <= 4000
    assert {r.thermostat for r in reps} == {"langevin", "andersen"}
    assert all(0.7 <= r.T <= 1.5 for r in reps)
