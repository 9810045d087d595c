"""test/water_test.jl:155-262 through the GPU integrator: two SPC/Fw molecules run for
2 tau with velocity Verlet, written as PDB (one REMARK 250 per saved time, three HETATM
records per molecule and time), and the reference's inline PDB read back and run."""
import io
import math
import os

import numpy as np
import pytest

import nbodysimulator as nbs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "water_3mol_2models.pdb")

# test/water_test.jl:5-31
kb = 8.3144598e-3
T = 298.16
mO, mH = 15.999, 1.00794
mH2O = mO + 2 * mH
rho = 997 / 1.6747
L = (mH2O * 216 / rho) ** (1 / 3)
v_dev = math.sqrt(kb * T / mH2O)
tau = 0.5e-3


def water_system(bodies):
    return nbs.WaterSPCFw(bodies, mH, mO, 0.41, -0.84,
                          nbs.LennardJonesParameters(0.1554253 * 4.184, 0.3165492, 0.9),
                          nbs.ElectrostaticParameters(138.935458, 0.49 * L),
                          nbs.SPCFwParameters(0.1012, 113.24 * math.pi / 180,
                                              1059.162 * 4.184 * 1.0e2, 75.9 * 4.184))


def test_pdb_file_compilation(nbx):
    print("[case] water PDB compilation", flush=True)
    p1 = nbs.MassBody([L / 3, L / 3, 2 * L / 3], [0, 0, -v_dev], mH2O)
    p2 = nbs.MassBody([L / 3, 2 * L / 3, L / 3], [0, -v_dev, 0], mH2O)
    sim = nbs.NBodySimulation(water_system([p1, p2]), (0.0, 2 * tau),
                              nbs.CubicPeriodicBoundaryConditions(L), kb)
    result = nbs.run_simulation(sim, nbs.VelocityVerlet(), dt=tau)
    buf = io.StringIO()
    nbs.write_pdb_data(buf, result)
    lines = buf.getvalue().split("\n")
    steps = sum(ln[:10] == "REMARK 250" for ln in lines)
    hetatm = sum(ln[:6] == "HETATM" for ln in lines)
    assert steps == len(result.solution.t) == 3
    assert hetatm / steps / 3 == 2


def test_load_water_molecules_from_pdb_and_run(nbx):
    print("[case] water from PDB fixture", flush=True)
    with open(GOLDEN) as f:
        bodies = nbs.extract_from_pdb(f)
    sim = nbs.NBodySimulation(water_system(bodies), (0.0, 2 * tau),
                              nbs.CubicPeriodicBoundaryConditions(L), kb)
    result = nbs.run_simulation(sim, nbs.VelocityVerlet(), dt=tau)
    assert len(result.solution.t) == 3
    assert np.all(np.isfinite(result.solution.u))
    eps = 0.00005
    assert abs(bodies[0].O.r[0] - 1.8642) < eps and abs(bodies[1].H1.r[2] - 0.3107) < eps
