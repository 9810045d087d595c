"""Protein Data Bank trajectories (src/nbody_simulation_result.jl:786-1001).

The writer formats the saved frames of a SimulationResult (positions x 10, i.e. nm -> A,
wrapped into the box); the reader builds SPC/Fw water molecules from the first model of
a trajectory written that way.  Host-side text I/O over frames the GPU run already
copied out at the save cadence."""
from __future__ import annotations

from typing import IO, List

import numpy as np

from .simulation import SimulationResult, get_position
from .types import MassBody, WaterMolecule, WaterSPCFw


def _julia_float(x: float) -> str:
    """Julia's `string(x::Float64)`: shortest round-trip digits, always with a decimal
    point, scientific as `1.0e-5` (Python: `1e-05`)."""
    s = repr(float(x))
    if s in ("inf", "-inf", "nan"):
        return {"inf": "Inf", "-inf": "-Inf", "nan": "NaN"}[s]
    if "e" in s:
        mant, ex = s.split("e")
        if "." not in mant:
            mant += ".0"
        return f"{mant}e{int(ex)}"
    return s


def _hetatm(index: int, name: str, resname: str, resid: int, xyz, element: str) -> str:
    # "HETATM", lpad(i, 5), "  ", rpad(name, 4), resname, lpad(resid, 6), "    ",
    # three %8.3f coordinates, lpad("1.00", 6), lpad("0.00", 6), lpad("", 10),
    # lpad(element, 2)  (src/nbody_simulation_result.jl:814-837, :853-858)
    return (f"HETATM{index:>5}  {name:<4}{resname}{resid:>6}    "
            f"{xyz[0]:8.3f}{xyz[1]:8.3f}{xyz[2]:8.3f}{'1.00':>6}{'0.00':>6}{'':>10}{element:>2}")


def _wrap(cc: np.ndarray, L: float) -> np.ndarray:
    return cc - L * np.floor(cc / L)


def write_pdb_data(f: IO[str], sr: SimulationResult) -> None:
    """One MODEL per saved time: water as O/H1/H2 of residue HOH (each H placed next to
    its wrapped O, :791-840), any other system as Ar atoms (:843-862)."""
    sim = sr.simulation
    L = 10 * sim.boundary_conditions.L
    n = len(sim.system.bodies)
    water = isinstance(sim.system, WaterSPCFw)
    for count, t in enumerate(sr.solution.t, start=1):
        cc0 = 10 * np.asarray(get_position(sr, t), dtype=np.float64)   # 3 x N
        cc = _wrap(cc0, L)
        f.write(f"{'MODEL':<10}{count}\n")
        f.write(f"REMARK 250 time={_julia_float(t)} {'picoseconds' if water else 'steps'}\n")
        if water:
            for i in range(n):
                o, h1, h2 = 3 * i, 3 * i + 1, 3 * i + 2
                cc[:, h1] = cc[:, o] + cc0[:, h1] - cc0[:, o]
                cc[:, h2] = cc[:, o] + cc0[:, h2] - cc0[:, o]
            for i in range(n):
                o, h1, h2 = 3 * i, 3 * i + 1, 3 * i + 2
                f.write(_hetatm(o + 1, "O", "HOH", i + 1, cc[:, o], "O") + "\n")
                f.write(_hetatm(h1 + 1, "H1", "HOH", i + 1, cc[:, h1], "H") + "\n")
                f.write(_hetatm(h2 + 1, "H2", "HOH", i + 1, cc[:, h2], "H") + "\n")
        else:
            for i in range(n):
                f.write(_hetatm(i + 1, "Ar", "Ar", i + 1, cc[:, i], "Ar") + "\n")
        f.write("ENDMDL\n")


def save_to_pdb(sr: SimulationResult, path) -> None:
    """Write a simulation result to a PDB file (save_to_pdb, :864-881)."""
    with open(path, "w") as f:
        write_pdb_data(f, sr)


def _three(f: IO[str], first: str, den: float):
    rows = [first, f.readline(), f.readline()]
    return [np.array([float(p) for p in r.split()[5:8]]) / den for r in rows]


def extract_from_pdb(f: IO[str]) -> List[WaterMolecule]:
    """Water molecules from the first MODEL of a trajectory (:929-1001): coordinates / 10
    (A -> nm), masses 15.999 and 1.00794, zero velocities.  The second MODEL must be
    present (the reference reads it for a time step, then discards the velocities)."""
    den = 10
    cc: list = []

    def next_model():
        for ln in f:
            if ln[:5] == "MODEL":
                return f.readline()
        raise ValueError("PDB data holds fewer than two MODEL records")

    def read_model(collect):
        for ln in f:
            if ln[:6] == "ENDMDL":
                return
            if len(ln) > 13 and ln[13] == "O":
                collect(_three(f, ln, den))

    float(next_model().split()[2][5:])
    read_model(cc.extend)
    float(next_model().split()[2][5:])
    wms: List[WaterMolecule] = []
    mc = 0

    def second(rows):
        nonlocal mc
        o = MassBody(cc[3 * mc], np.zeros(3), 15.999)
        h1 = MassBody(cc[3 * mc + 1], np.zeros(3), 1.00794)
        h2 = MassBody(cc[3 * mc + 2], np.zeros(3), 1.00794)
        wms.append(WaterMolecule(o, h1, h2))
        mc += 1

    read_model(second)
    return wms


def load_water_molecules_from_pdb(path) -> List[WaterMolecule]:
    """SPC/Fw water molecules from a PDB trajectory written by save_to_pdb (:901-927)."""
    with open(path) as f:
        return extract_from_pdb(f)


__all__ = ["write_pdb_data", "save_to_pdb", "extract_from_pdb", "load_water_molecules_from_pdb"]
