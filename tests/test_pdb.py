"""PDB trajectories (src/nbody_simulation_result.jl:786-1001) on the host: the reader
against the reference's own inline PDB (test/water_test.jl:196-262, kept as
tests/golden/water_3mol_2models.pdb), the writer's record layout and counts
(test/water_test.jl:155-193), and a write -> read round trip."""
import io
import os

import numpy as np
import pytest

import nbodysimulator as nbs
from nbodysimulator.pdb import _julia_float
from nbodysimulator.simulation import SimulationResult, Solution

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "water_3mol_2models.pdb")


def water_result(u_frames, ts, L):
    n = u_frames.shape[1] // 3
    bodies = [nbs.MassBody(np.zeros(3), np.zeros(3), 18.0) for _ in range(n)]
    water = nbs.WaterSPCFw(bodies, 1.00794, 15.999, 0.41, -0.82,
                           nbs.LennardJonesParameters(0.65, 0.316, 0.9),
                           nbs.ElectrostaticParameters(138.935458, 0.9),
                           nbs.SPCFwParameters(0.1012, 1.976, 1059.162, 75.9))
    sim = nbs.NBodySimulation(water, (ts[0], ts[-1]), nbs.CubicPeriodicBoundaryConditions(L),
                              nbs.NullThermostat(), 0.00831)
    sol = Solution(np.asarray(ts), u_frames, np.zeros_like(u_frames))
    return SimulationResult(sol, sim)


def test_reader_on_the_reference_fixture():
    with open(GOLDEN) as f:
        wms = nbs.extract_from_pdb(f)
    eps = 0.00005
    assert len(wms) == 3
    assert np.allclose(wms[0].O.r, [1.8642, 1.8642, 0.0], atol=eps)
    assert np.allclose(wms[1].H1.r, [0.0957, 0.0, 0.3107], atol=eps)
    assert np.allclose(wms[2].H2.r, [1.8882, 1.7715, 0.6214], atol=eps)
    assert np.allclose(wms[2].H2.v, 0.0) and wms[0].O.m == 15.999 and wms[1].H2.m == 1.00794
    assert len(nbs.load_water_molecules_from_pdb(GOLDEN)) == 3


def test_reader_needs_two_models():
    with open(GOLDEN) as f:
        first = "".join(f.readlines()[:12])
    with pytest.raises(ValueError):
        nbs.extract_from_pdb(io.StringIO(first))


def test_writer_layout_and_counts():
    """Two molecules, three frames: one REMARK 250 per frame and 3 HETATM records per
    molecule and frame (test/water_test.jl:176-192); the H atoms of a molecule that
    straddles the box face stay next to its wrapped O; fixed-column records."""
    L = 1.5
    u0 = np.array([[1.45, 0.2, 0.3], [1.55, 0.2, 0.3], [1.42, 0.29, 0.3],   # O near x = L
                   [0.5, 0.5, 0.5], [0.6, 0.5, 0.5], [0.48, 0.59, 0.5]])
    frames = np.stack([u0, u0 + 0.01, u0 + 0.1])
    sr = water_result(frames, [0.0, 1e-5, 0.0005], L)
    buf = io.StringIO()
    nbs.write_pdb_data(buf, sr)
    lines = buf.getvalue().split("\n")
    assert sum(ln.startswith("REMARK 250") for ln in lines) == 3
    assert sum(ln.startswith("HETATM") for ln in lines) == 3 * 2 * 3
    assert lines[0] == "MODEL     1"
    assert lines[1] == "REMARK 250 time=0.0 picoseconds"
    assert "REMARK 250 time=0.0005 picoseconds" in lines
    assert "REMARK 250 time=1.0e-5 picoseconds" in lines
    # frame 1, O of molecule 1 at 14.5 A and H1 at 15.5 A -> wrapped next to O: 15.5
    # "HETATM", lpad(1, 5), "  ", rpad("O", 4), "HOH", lpad(1, 6), "    ", 3 x %8.3f,
    # lpad("1.00", 6), lpad("0.00", 6), lpad("", 10), lpad("O", 2)
    assert lines[2] == ("HETATM" + "    1" + "  " + "O   " + "HOH" + "     1" + "    " +
                        "  14.500" + "   2.000" + "   3.000" + "  1.00" + "  0.00" + " " * 10 +
                        " O")
    assert lines[3][30:38] == "  15.500" and lines[3][13:17] == "H1  "
    assert lines[5][:26] == "HETATM    4  O   HOH     2"
    # frame 3: O of molecule 1 at 15.5 A wraps to 0.5 A, its H1 (16.5 A) follows to 1.5 A
    m3 = lines.index("MODEL     3")
    assert lines[m3 + 2][30:38] == "   0.500" and lines[m3 + 3][30:38] == "   1.500"
    # 6 + 5 + 2 + 4 + 3 + 6 + 4 + 3 x 8 + 6 + 6 + 10 + 2 columns
    assert all(len(ln) == 78 for ln in lines if ln.startswith("HETATM"))


def test_write_then_read_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    L = 2.0
    o = rng.uniform(0, L, (4, 3))
    u = np.concatenate([np.stack([p, p + [0.1, 0, 0], p + [0, 0.1, 0]]) for p in o])
    sr = water_result(np.stack([u, u + 0.001]), [0.0, 0.0005], L)
    path = tmp_path / "w.pdb"
    nbs.save_to_pdb(sr, path)
    wms = nbs.load_water_molecules_from_pdb(path)
    got = np.array([[w.O.r, w.H1.r, w.H2.r] for w in wms]).reshape(-1, 3)
    want = u.copy()
    for i in range(4):
        shift = want[3 * i] - (want[3 * i] - L * np.floor(want[3 * i] / L))
        want[3 * i:3 * i + 3] -= shift
    assert np.abs(got - want).max() <= 0.5e-4 + 1e-12


def test_writer_argon_records():
    L = 3.0
    u = np.array([[[0.1, 0.2, 3.1], [1.0, 2.0, 2.9]]])
    bodies = [nbs.MassBody(x, np.zeros(3), 1.0) for x in u[0]]
    system = nbs.PotentialNBodySystem(bodies, {"lennard_jones": nbs.LennardJonesParameters()})
    sim = nbs.NBodySimulation(system, (0.0, 1.0), nbs.CubicPeriodicBoundaryConditions(L))
    sr = SimulationResult(Solution(np.array([0.0]), u, np.zeros_like(u)), sim)
    buf = io.StringIO()
    nbs.write_pdb_data(buf, sr)
    lines = buf.getvalue().split("\n")
    assert lines[1] == "REMARK 250 time=0.0 steps"
    assert lines[2] == ("HETATM" + "    1" + "  " + "Ar  " + "Ar" + "     1" + "    " +
                        "   1.000" + "   2.000" + "   1.000" + "  1.00" + "  0.00" + " " * 10 +
                        "Ar")


def test_julia_float_strings():
    assert [_julia_float(x) for x in (0.0, 0.0005, 1e-5, 2.5, 1e-10, 123.0)] == \
        ["0.0", "0.0005", "1.0e-5", "2.5", "1.0e-10", "123.0"]
