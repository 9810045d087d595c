"""Data model of NBodySimulator.jl, kept unchanged (names, fields, defaults, show).

Reference: src/bodies.jl, src/boundary_conditions.jl:1-109, src/thermostats.jl,
src/basic_potentials.jl:1-238, src/nbody_system.jl, src/nbody_simulation.jl.
"""
from __future__ import annotations

import decimal
import math
from dataclasses import dataclass, field
from typing import Any, Dict, List, Sequence

import numpy as np

kb_SI = 1.38e-23  # src/nbody_simulation.jl:7


def julia_repr(x) -> str:
    """Text of Julia's `show(io, x)` for Int and Float64 values: shortest round-trip
    digits, positional for 1e-4 <= |x| < 1e6 (decimal exponent -4..5), otherwise
    `d.ddde<exp>` (e.g. 9.0e9, 6.67408e-11)."""
    if isinstance(x, (bool, np.bool_)):
        return "true" if x else "false"
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    sign, digits, exp = decimal.Decimal(repr(x)).as_tuple()
    ds = "".join(map(str, digits)).rstrip("0") or "0"
    e10 = len(digits) + exp - 1
    if -5 < e10 < 6:
        r = repr(abs(x))
        if "e" in r:  # Python uses e-notation below 1e-4 only; not reached here
            r = format(abs(x), "f")
        body = r
    else:
        body = ds[0] + "." + (ds[1:] or "0") + "e" + str(e10)
    return ("-" if sign else "") + body


_show_float = julia_repr


# ---------------------------------------------------------------------------
# Bodies: src/bodies.jl:10-113
# ---------------------------------------------------------------------------

def _vec3(a) -> np.ndarray:
    v = np.asarray(a, dtype=np.float64).reshape(3)
    return v


@dataclass
class MassBody:
    r: Any
    v: Any
    m: float

    def __post_init__(self):
        self.r = _vec3(self.r)
        self.v = _vec3(self.v)


@dataclass
class ChargedParticle:
    r: Any
    v: Any
    m: float
    q: float

    def __post_init__(self):
        self.r = _vec3(self.r)
        self.v = _vec3(self.v)


@dataclass
class MagneticParticle:
    r: Any
    v: Any
    m: float
    mm: Any

    def __post_init__(self):
        self.r = _vec3(self.r)
        self.v = _vec3(self.v)
        self.mm = _vec3(self.mm)


@dataclass
class WaterMolecule:
    O: MassBody
    H1: MassBody
    H2: MassBody


# ---------------------------------------------------------------------------
# Boundary conditions: src/boundary_conditions.jl:1-109
# ---------------------------------------------------------------------------

class BoundaryConditions:
    pass


class PeriodicBoundaryConditions(BoundaryConditions):
    """Wraps each displacement component into [lo, hi) (not minimum image)."""

    def __init__(self, boundary):
        if np.ndim(boundary) == 0:
            L = boundary
            boundary = (0, L, 0, L, 0, L)
        if len(boundary) != 6:
            raise ValueError("PeriodicBoundaryConditions needs 6 bounds")
        self.boundary = tuple(boundary)

    def __iter__(self):
        return iter(self.boundary)

    def __getitem__(self, i: int):
        # Julia 1-based indexing, BoundsError outside 1..6
        if not 1 <= i <= len(self.boundary):
            raise IndexError(f"BoundsError: attempt to access PeriodicBoundaryConditions at {i}")
        return self.boundary[i - 1]


class InfiniteBox(BoundaryConditions):
    def __init__(self, boundary=None):
        self.boundary = tuple(boundary) if boundary is not None else (
            -math.inf, math.inf, -math.inf, math.inf, -math.inf, math.inf)


class CubicPeriodicBoundaryConditions(BoundaryConditions):
    def __init__(self, L):
        self.L = L


# ---------------------------------------------------------------------------
# Thermostats: src/thermostats.jl
# ---------------------------------------------------------------------------

class Thermostat:
    pass


class NullThermostat(Thermostat):
    pass


@dataclass
class AndersenThermostat(Thermostat):
    T: float
    ν: float


@dataclass
class BerendsenThermostat(Thermostat):
    T: float
    τ: float
    γ: float = None

    def __post_init__(self):
        if self.γ is None:
            self.γ = 0.5 / self.τ      # src/thermostats.jl:72-74


@dataclass
class NoseHooverThermostat(Thermostat):
    T: float
    τ: float


@dataclass
class LangevinThermostat(Thermostat):
    T: float
    γ: float


# ---------------------------------------------------------------------------
# Potential parameters: src/basic_potentials.jl:33-238
# ---------------------------------------------------------------------------

class PotentialParameters:
    pass


class LennardJonesParameters(PotentialParameters):
    def __init__(self, ϵ=1.0, σ=1.0, R=2.5, σ2=None, R2=None):
        self.ϵ, self.σ, self.R = ϵ, σ, R
        self.σ2 = σ * σ if σ2 is None else σ2
        self.R2 = R * R if R2 is None else R2

    def show(self) -> str:
        return (f"Lennard-Jones:\n\tϵ:{_show_float(self.ϵ)}\n\tσ:{_show_float(self.σ)}\n"
                f"\tR:{_show_float(self.R)}\n")


class GravitationalParameters(PotentialParameters):
    def __init__(self, G=6.67408e-11):
        self.G = G

    def show(self) -> str:
        return f"Gravitational:\n\tG:{_show_float(self.G)}\n"


class ElectrostaticParameters(PotentialParameters):
    def __init__(self, k=9.0e9, R=math.inf, R2=None):
        self.k, self.R = k, R
        self.R2 = R * R if R2 is None else R2

    def show(self) -> str:
        return f"Electrostatic:\n\tk:{_show_float(self.k)}\n"


class ReactionFieldParameters(PotentialParameters):
    """Reaction-field electrostatics (Tironi et al. 1995); no counterpart in the
    reference.  eps_rf = dielectric constant beyond the cutoff R (inf = conducting)."""

    def __init__(self, k, R, eps_rf=math.inf):
        self.k, self.R, self.eps_rf = k, R, eps_rf
        self.R2 = R * R

    def show(self) -> str:
        return (f"Reaction field:\n\tk:{_show_float(self.k)}\n\tR:{_show_float(self.R)}\n"
                f"\tϵ_rf:{_show_float(self.eps_rf)}\n")


class MagnetostaticParameters(PotentialParameters):
    def __init__(self, μ_4π=1.0e-7):
        self.μ_4π = μ_4π

    def show(self) -> str:
        return f"Magnetostatic:\n\tμ/4π:{_show_float(self.μ_4π)}\n"


class SPCFwParameters(PotentialParameters):
    def __init__(self, rOH, aHOH, kb, ka):
        self.rOH, self.aHOH, self.kb, self.ka = rOH, aHOH, kb, ka


# ---------------------------------------------------------------------------
# Systems: src/nbody_system.jl
# ---------------------------------------------------------------------------

class NBodySystem:
    bodies: list


class PotentialNBodySystem(NBodySystem):
    def __init__(self, bodies, potentials=None):
        self.bodies = list(bodies)
        if potentials is None:
            potentials = []
        if isinstance(potentials, dict):
            self.potentials: Dict[str, PotentialParameters] = dict(potentials)
        else:
            # PotentialNBodySystem(bodies; potentials = [:lennard_jones, ...]) (:102-122)
            p: Dict[str, PotentialParameters] = {}
            names = [str(s).lstrip(":") for s in potentials]
            if "lennard_jones" in names:
                p["lennard_jones"] = LennardJonesParameters()
            if "electrostatic" in names:
                p["electrostatic"] = ElectrostaticParameters()
            if "gravitational" in names:
                p["gravitational"] = GravitationalParameters()
            if "magnetostatic" in names:
                p["magnetostatic"] = MagnetostaticParameters()
            self.potentials = p

    def show(self) -> str:
        out = "Potentials: \n"
        for key in ("lennard_jones", "electrostatic", "magnetostatic", "gravitational"):
            if key in self.potentials:
                out += self.potentials[key].show()
        return out


class ChargedParticles(NBodySystem):
    def __init__(self, bodies, k):
        self.bodies = list(bodies)
        self.k = k


class GravitationalSystem(NBodySystem):
    def __init__(self, bodies, G):
        self.bodies = list(bodies)
        self.G = G


class WaterSPCFw(NBodySystem):
    def __init__(self, bodies, mH, mO, qH, qO, lj_parameters, e_parameters, scpfw_parameters):
        self.bodies = list(bodies)
        self.mH, self.mO, self.qH, self.qO = mH, mO, qH, qO
        self.lj_parameters = lj_parameters
        self.e_parameters = e_parameters
        self.scpfw_parameters = scpfw_parameters


def as_potential_system(system: NBodySystem) -> NBodySystem:
    """PotentialNBodySystem(::ChargedParticles / ::GravitationalSystem) (:136-150)."""
    if isinstance(system, ChargedParticles):
        return PotentialNBodySystem(system.bodies,
                                    {"electrostatic": ElectrostaticParameters(system.k)})
    if isinstance(system, GravitationalSystem):
        return PotentialNBodySystem(system.bodies,
                                    {"gravitational": GravitationalParameters(system.G)})
    return system


class NBodySimulation:
    """NBodySimulation(system, tspan, boundary_conditions=InfiniteBox(),
    thermostat=NullThermostat(), kb=kb_SI) (src/nbody_simulation.jl:233-316)."""

    def __init__(self, system, tspan, boundary_conditions=None, thermostat=None, kb=None):
        if isinstance(boundary_conditions, Thermostat) and thermostat is None:
            raise TypeError("boundary conditions come before the thermostat")
        if isinstance(thermostat, (int, float)) and kb is None:
            # NBodySimulation(system, tspan, bc, kb)
            kb, thermostat = thermostat, None
        self.system = as_potential_system(system)
        self.tspan = (tspan[0], tspan[1])
        self.boundary_conditions = boundary_conditions if boundary_conditions is not None \
            else InfiniteBox()
        self.thermostat = thermostat if thermostat is not None else NullThermostat()
        self.kb = kb_SI if kb is None else kb

    def show(self) -> str:
        bc = self.boundary_conditions
        if isinstance(bc, InfiniteBox):
            b = "InfiniteBox{Float64}([" + ", ".join(_show_float(float(x)) for x in bc.boundary) + "])"
        elif isinstance(bc, CubicPeriodicBoundaryConditions):
            b = f"CubicPeriodicBoundaryConditions{{Float64}}({_show_float(bc.L)})"
        else:
            b = ("PeriodicBoundaryConditions{Float64}([" +
                 ", ".join(_show_float(float(x)) for x in bc.boundary) + "])")
        t0, t1 = self.tspan
        sys = self.system.show() if hasattr(self.system, "show") else ""
        return f"Timespan: ({_show_float(t0)}, {_show_float(t1)})\nBoundary conditions: {b}\n{sys}"


def get_masses(system) -> np.ndarray:
    """src/nbody_simulation_result.jl:121-139"""
    if isinstance(system, WaterSPCFw):
        n = len(system.bodies)
        ms = np.empty(3 * n)
        ms[0::3] = system.mO
        ms[1::3] = system.mH
        ms[2::3] = system.mH
        return ms
    return np.array([b.m for b in system.bodies], dtype=np.float64)


def get_coordinate_vector_count(system) -> int:
    return 3 * len(system.bodies) if isinstance(system, WaterSPCFw) else len(system.bodies)


def get_degrees_of_freedom(system):
    """(n, nc, ndf): src/nbody_simulation_result.jl:152-164"""
    if isinstance(system, WaterSPCFw):
        n = 3 * len(system.bodies)
        nc = 2 * len(system.bodies)
    else:
        n = len(system.bodies)
        nc = 0
    return n, nc, 3 * n - nc


def gather_bodies_initial_coordinates(simulation: NBodySimulation):
    """(u0, v0, n) with u0, v0 as (n_coords, 3) arrays (Julia 3 x n column-major memory).
    src/nbody_to_ode.jl:1-75"""
    system = simulation.system
    if isinstance(system, WaterSPCFw):
        mols = system.bodies
        n = len(mols)
        u0 = np.zeros((3 * n, 3))
        v0 = np.zeros((3 * n, 3))
        p = system.scpfw_parameters
        for i, m in enumerate(mols):
            if isinstance(m, WaterMolecule):
                u0[3 * i] = m.O.r; u0[3 * i + 1] = m.H1.r; u0[3 * i + 2] = m.H2.r
                v0[3 * i] = m.O.v; v0[3 * i + 1] = m.H1.v; v0[3 * i + 2] = m.H2.v
            else:
                u0[3 * i] = m.r
                u0[3 * i + 1] = m.r + np.array([p.rOH, 0.0, 0.0])
                u0[3 * i + 2] = m.r + np.array([math.cos(p.aHOH) * p.rOH, 0.0,
                                                math.sin(p.aHOH) * p.rOH])
                v0[3 * i] = m.v; v0[3 * i + 1] = m.v; v0[3 * i + 2] = m.v
        return u0, v0, n
    bodies = system.bodies
    n = len(bodies)
    u0 = np.array([b.r for b in bodies], dtype=np.float64).reshape(n, 3)
    v0 = np.array([b.v for b in bodies], dtype=np.float64).reshape(n, 3)
    return u0, v0, n
