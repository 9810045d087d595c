"""run_simulation, SimulationResult and the analysis functions of NBodySimulator.jl,
computed by libnbx on the GPU.

Reference: src/nbody_simulation_result.jl (driver :468-540, accessors :1-188,
energies :209-446, rdf :664-709) and the closure protocol of src/nbody_to_ode.jl:110-261.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .types import (AndersenThermostat, BerendsenThermostat, ChargedParticle,
                    CubicPeriodicBoundaryConditions, ElectrostaticParameters,
                    GravitationalParameters, InfiniteBox, LangevinThermostat,
                    LennardJonesParameters, MagneticParticle, MagnetostaticParameters, NBodySimulation,
                    NoseHooverThermostat, NullThermostat, PeriodicBoundaryConditions,
                    PotentialNBodySystem, ReactionFieldParameters, SPCFwParameters,
                    WaterSPCFw, gather_bodies_initial_coordinates, get_coordinate_vector_count,
                    get_degrees_of_freedom, get_masses, julia_repr)


@dataclass
class VelocityVerlet:
    """The velocity-Verlet algorithm (OrdinaryDiffEqSymplecticRK.VelocityVerlet),
    executed on the GPU.  precision: "fp64" (reference arithmetic, bit for bit) or
    "mixed" (FP32 pair forces with Kahan sums, FP64 state, cell/Verlet lists)."""
    precision: str = "fp64"
    seed: int = 0
    skin: Optional[float] = None   # Verlet-list skin (mixed path); None: 0.12 x cutoff


MI355XVerlet = VelocityVerlet


@dataclass
class EM:
    """Euler-Maruyama (StochasticDiffEq.EM) for the Langevin SDE of a potential system
    with a LangevinThermostat (SciMLBase.SDEProblem, src/nbody_to_ode.jl:567-598),
    executed on the GPU: x' = x + dt v, v' = v + dt (a - gamma v) + sigma sqrt(dt) xi with
    sigma = sqrt(2 gamma kb T / m_1).  precision and skin as for VelocityVerlet."""
    precision: str = "fp64"
    seed: int = 0
    skin: Optional[float] = None


EulerMaruyama = EM


@dataclass
class _Stats:
    naccept: int = 0


@dataclass
class Solution:
    """Saved trajectory: t (nframes,), u and v (nframes, n, 3), i.e. Julia 3 x n frames."""
    t: np.ndarray
    u: np.ndarray
    v: np.ndarray
    stats: _Stats = field(default_factory=_Stats)

    def __call__(self, time: float):
        """State (u, v) at `time`: the saved frame, or linear interpolation between the
        neighbouring saved frames (the reference interpolates with its ODE solution)."""
        t = self.t
        k = int(np.searchsorted(t, time))
        tol = 1e-12 * max(1.0, abs(time))
        if k < len(t) and abs(t[k] - time) <= tol:
            return self.u[k], self.v[k]
        if k > 0 and abs(t[k - 1] - time) <= tol:
            return self.u[k - 1], self.v[k - 1]
        if k == 0 or k >= len(t):
            raise ValueError(f"time {time} outside the saved interval [{t[0]}, {t[-1]}]")
        w = (time - t[k - 1]) / (t[k] - t[k - 1])
        return ((1 - w) * self.u[k - 1] + w * self.u[k], (1 - w) * self.v[k - 1] + w * self.v[k])


class SimulationResult:
    def __init__(self, solution: Solution, simulation: NBodySimulation, context=None):
        self.solution = solution
        self.simulation = simulation
        self._ctx = context

    def __call__(self, time):
        return self.solution(time)

    def __iter__(self):
        for t in self.solution.t:
            yield (self, t)

    def show(self) -> str:
        t = self.solution.t
        return (f"N: {len(self.simulation.system.bodies)}\n{self.simulation.show()}"
                f"Time steps: {self.solution.stats.naccept}\n"
                f"t: {julia_repr(float(np.min(t)))}, {julia_repr(float(np.max(t)))}\n")


# ---------------------------------------------------------------------------
# Context construction: the host-side setup of src/nbody_to_ode.jl:263-433
# ---------------------------------------------------------------------------

def _set_bc(ctx: _lib.Context, bc):
    if isinstance(bc, CubicPeriodicBoundaryConditions):
        ctx.set_boundary(_lib.BC_CUBIC, [bc.L, 0, 0, 0, 0, 0])
    elif isinstance(bc, PeriodicBoundaryConditions):
        ctx.set_boundary(_lib.BC_PERIODIC, list(bc.boundary))
    elif isinstance(bc, InfiniteBox) or bc is None:
        ctx.set_boundary(_lib.BC_INFINITE, None)
    else:
        raise NotImplementedError(f"boundary condition {type(bc).__name__}")


def _add_potential(ctx: _lib.Context, p):
    if isinstance(p, LennardJonesParameters):
        ctx.add_lennard_jones(p.ϵ, p.σ, p.R)
    elif isinstance(p, ElectrostaticParameters):
        ctx.add_electrostatic(p.k, p.R)
    elif isinstance(p, ReactionFieldParameters):
        ctx.add_reaction_field(p.k, p.R, p.eps_rf)
    elif isinstance(p, GravitationalParameters):
        ctx.add_gravitational(p.G)
    elif isinstance(p, MagnetostaticParameters):
        ctx.add_magnetostatic(p.μ_4π)
    else:
        raise NotImplementedError(f"no GPU kernel for {type(p).__name__}; "
                                  "use a custom closure (get_accelerating_function)")


def _set_thermostat(ctx: _lib.Context, th):
    if th is None or isinstance(th, NullThermostat):
        ctx.set_thermostat(_lib.THERM_NONE, 0.0, 1.0)
    elif isinstance(th, AndersenThermostat):
        ctx.set_thermostat(_lib.THERM_ANDERSEN, th.T, th.ν)
    elif isinstance(th, BerendsenThermostat):
        ctx.set_thermostat(_lib.THERM_BERENDSEN, th.T, th.τ)
    elif isinstance(th, LangevinThermostat):
        ctx.set_thermostat(_lib.THERM_LANGEVIN, th.T, th.γ)
    elif isinstance(th, NoseHooverThermostat):
        ctx.set_thermostat(_lib.THERM_NOSEHOOVER, th.T, th.τ)
    else:
        raise NotImplementedError(f"thermostat {type(th).__name__}")


def _charges(system) -> Optional[np.ndarray]:
    if isinstance(system, WaterSPCFw):
        q = np.empty(3 * len(system.bodies))
        q[0::3] = system.qO
        q[1::3] = system.qH
        q[2::3] = system.qH
        return q
    if system.bodies and isinstance(system.bodies[0], ChargedParticle):
        return np.array([b.q for b in system.bodies], dtype=np.float64)
    return None


def build_context(simulation: NBodySimulation, precision: str = "fp64",
                  only=None) -> _lib.Context:
    """A libnbx context for `simulation`.  `only` restricts a potential system to one
    parameter object (used by the per-potential closure)."""
    system = simulation.system
    masses = get_masses(system)
    q = _charges(system)
    if isinstance(system, WaterSPCFw):
        ctx = _lib.Context(len(masses), _lib.SYS_WATER, masses, q)
        _set_bc(ctx, simulation.boundary_conditions)
        ctx.add_lennard_jones(system.lj_parameters.ϵ, system.lj_parameters.σ,
                              system.lj_parameters.R)
        _add_potential(ctx, system.e_parameters)
        p = system.scpfw_parameters
        ctx.set_spcfw(p.rOH, p.aHOH, p.kb, p.ka)
    else:
        ctx = _lib.Context(len(masses), _lib.SYS_POTENTIAL, masses, q)
        _set_bc(ctx, simulation.boundary_conditions)
        if system.bodies and all(isinstance(b, MagneticParticle) for b in system.bodies):
            ctx.set_moments(np.array([b.mm for b in system.bodies], dtype=np.float64))
        params = [only] if only is not None else list(system.potentials.values())
        for p in params:
            _add_potential(ctx, p)
    ctx.set_kb(simulation.kb)
    ctx.set_precision(_lib.PREC_MIXED if precision == "mixed" else _lib.PREC_FP64_REF)
    if only is None:
        _set_thermostat(ctx, simulation.thermostat)
    return ctx


# ---------------------------------------------------------------------------
# Driver: src/nbody_simulation_result.jl:468-540
# ---------------------------------------------------------------------------

def _step_plan(t1: float, t2: float, dt: float):
    span = t2 - t1
    n = int(round(span / dt))
    if n >= 1 and abs(n * dt - span) <= 1e-9 * abs(span):
        return n, dt, None
    n = int(math.floor(span / dt))
    last = span - n * dt
    return n, dt, last


def run_simulation(s: NBodySimulation, alg=None, *, dt: float = None,
                   save_everystep: Optional[bool] = None, seed: int = None,
                   precision: str = None, context=None) -> SimulationResult:
    """Integrate `s` with velocity Verlet on the GPU and return a SimulationResult.

    As in the reference, every step is saved unless save_everystep=False (then the
    first and last states only).  Thermostats: Andersen and Berendsen as in the
    reference; Langevin as Ornstein-Uhlenbeck half steps around each VV step.  With
    alg = EM() the Langevin SDE is integrated with Euler-Maruyama, as the reference's
    SDEProblem path does."""
    if alg is None:
        alg = VelocityVerlet()
    if not isinstance(alg, (VelocityVerlet, EM)):
        raise NotImplementedError("the GPU path integrates with VelocityVerlet / MI355XVerlet "
                                  "or EM (the Langevin SDE)")
    if dt is None:
        raise ValueError(f"{type(alg).__name__} needs a fixed dt")
    prec = precision or alg.precision
    rng_seed = alg.seed if seed is None else seed
    ctx = context or build_context(s, prec)
    if alg.skin is not None:
        ctx.set_skin(alg.skin)
    u0, v0, _ = gather_bodies_initial_coordinates(s)
    ctx.set_state(u0, v0)
    t1, t2 = float(s.tspan[0]), float(s.tspan[1])
    nsteps, dt, last = _step_plan(t1, t2, float(dt))
    step = ctx.em_run if isinstance(alg, EM) else ctx.vv_run
    save = True if save_everystep is None else bool(save_everystep)
    ts = [t1]
    us = [u0.copy()]
    vs = [v0.copy()]
    if save:
        if nsteps:
            uf, vf = step(dt, nsteps, rng_seed, save_every=1)
            us.extend(uf[1:]); vs.extend(vf[1:])
            ts.extend(t1 + dt * np.arange(1, nsteps + 1))
        if last:
            uf, vf = step(last, 1, rng_seed, save_every=1)
            us.append(uf[1]); vs.append(vf[1]); ts.append(t2)
    else:
        if nsteps:
            step(dt, nsteps, rng_seed)
        if last:
            step(last, 1, rng_seed)
        u, v = ctx.get_state()
        us.append(u); vs.append(v); ts.append(t2)
    steps = nsteps + (1 if last else 0)
    sol = Solution(np.asarray(ts), np.asarray(us), np.asarray(vs), _Stats(steps))
    return SimulationResult(sol, s, ctx)


def run_simulations(sims, alg=None, *, dt: float, seed: int = None,
                    steps_per_launch: int = 0):
    """Integrate many independent Lennard-Jones simulations together in one persistent
    GPU launch (one workgroup per replica).  Each simulation must hold a
    PotentialNBodySystem with the same LennardJonesParameters, a
    CubicPeriodicBoundaryConditions box, the same kb and tspan length, and a Null,
    Andersen or Langevin thermostat.  Returns one SimulationResult per simulation with
    the initial and final states (the reference runs each with run_simulation)."""
    if alg is None:
        alg = VelocityVerlet(precision="mixed")
    sims = list(sims)
    if not sims:
        return []
    first = sims[0]
    lj = None
    natoms, boxes, masses, us, vs = [], [], [], [], []
    for s in sims:
        sys_ = s.system
        if not isinstance(sys_, PotentialNBodySystem) or set(sys_.potentials) != {"lennard_jones"}:
            raise NotImplementedError("batched runs take Lennard-Jones potential systems")
        p = sys_.potentials["lennard_jones"]
        if lj is None:
            lj = p
        elif (p.ϵ, p.σ, p.R) != (lj.ϵ, lj.σ, lj.R):
            raise ValueError("every replica must use the same LennardJonesParameters")
        if not isinstance(s.boundary_conditions, CubicPeriodicBoundaryConditions):
            raise NotImplementedError("batched runs need CubicPeriodicBoundaryConditions")
        if s.kb != first.kb or (s.tspan[1] - s.tspan[0]) != (first.tspan[1] - first.tspan[0]):
            raise ValueError("replicas must share kb and the length of tspan")
        u0, v0, n = gather_bodies_initial_coordinates(s)
        natoms.append(n)
        boxes.append(float(s.boundary_conditions.L))
        masses.append(get_masses(sys_))
        us.append(u0)
        vs.append(v0)
    b = _lib.Batch(natoms, boxes, np.concatenate(masses), lj.ϵ, lj.σ, lj.R, first.kb)
    for r, s in enumerate(sims):
        th = s.thermostat
        if isinstance(th, AndersenThermostat):
            b.set_thermostat(r, _lib.THERM_ANDERSEN, th.T, th.ν)
        elif isinstance(th, LangevinThermostat):
            b.set_thermostat(r, _lib.THERM_LANGEVIN, th.T, th.γ)
        elif not (th is None or isinstance(th, NullThermostat)):
            raise NotImplementedError(f"batched runs take Andersen or Langevin, not {type(th).__name__}")
    if alg.skin is not None:
        b.set_skin(alg.skin)
    U0, V0 = np.concatenate(us), np.concatenate(vs)
    b.set_state(U0, V0)
    t1, t2 = float(first.tspan[0]), float(first.tspan[1])
    nsteps, dt, last = _step_plan(t1, t2, float(dt))
    if last:
        raise ValueError("batched runs need tspan to be a whole number of steps")
    b.run(dt, nsteps, alg.seed if seed is None else seed, steps_per_launch)
    U, V = b.get_state()
    out, o = [], 0
    for s, n in zip(sims, natoms):
        sol = Solution(np.array([t1, t2]), np.stack([U0[o:o + n], U[o:o + n]]),
                       np.stack([V0[o:o + n], V[o:o + n]]), _Stats(nsteps))
        out.append(SimulationResult(sol, s))
        o += n
    return out


# ---------------------------------------------------------------------------
# Accessors and analysis: src/nbody_simulation_result.jl:49-446
# ---------------------------------------------------------------------------

def get_velocity(sr: SimulationResult, time: float, i: int = 0):
    """3 x n velocities at `time`, or particle i's (1-based) when i > 0."""
    _, v = sr(time)
    n = get_coordinate_vector_count(sr.simulation.system)
    v = np.asarray(v)[:n]
    return v.T if i <= 0 else v[i - 1]


def get_position(sr: SimulationResult, time: float, i: int = 0):
    u, _ = sr(time)
    n = get_coordinate_vector_count(sr.simulation.system)
    u = np.asarray(u)[:n]
    return u.T if i <= 0 else u[i - 1]


def _state_ctx(sr: SimulationResult, time: float) -> _lib.Context:
    ctx = sr._ctx or build_context(sr.simulation)
    sr._ctx = ctx
    u, v = sr(time)
    ctx.set_state(u, v)
    return ctx


def temperature(result: SimulationResult, time: float) -> float:
    return _state_ctx(result, time).temperature()


def kinetic_energy(a, b=None):
    """kinetic_energy(result, time) or kinetic_energy(velocities (3 x n), masses)."""
    if isinstance(a, SimulationResult):
        return _state_ctx(a, b).energies()[0]
    v = np.asarray(a, dtype=np.float64)
    v = v.T if v.shape[0] == 3 and v.ndim == 2 and v.shape[1] != 3 else v
    ms = np.asarray(b, dtype=np.float64)
    # sum(dot(vec(sum(v .^ 2, dims = 1)), masses / 2)) on the host, sequential order
    s = (v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]) + v[:, 2] * v[:, 2]
    ke = 0.0
    for si, mi in zip(s, ms / 2):
        ke += si * mi
    return ke


def potential_energy(a, b):
    """potential_energy(result, time) or potential_energy(coordinates, simulation)."""
    if isinstance(a, SimulationResult):
        return _state_ctx(a, b).energies()[1]
    sim: NBodySimulation = b
    ctx = build_context(sim)
    u = np.asarray(a, dtype=np.float64)
    ctx.set_state(u.T if u.shape[0] == 3 and u.ndim == 2 and u.shape[1] != 3 else u)
    return ctx.energies()[1]


def gravitational_potential_energy(sr: SimulationResult, time: float) -> float:
    """-G sum_{i<j} m_i m_j / r_ij (the reference's potential_energy omits gravity)."""
    return _state_ctx(sr, time).energies()[2]


def total_energy(sr: SimulationResult, time: float) -> float:
    ke, pe, _ = _state_ctx(sr, time).energies()
    return ke + pe


def initial_energy(simulation: NBodySimulation) -> float:
    u0, v0, _ = gather_bodies_initial_coordinates(simulation)
    ctx = build_context(simulation)
    ctx.set_state(u0, v0)
    ke, pe, _ = ctx.energies()
    return pe + ke


def rdf(sr: SimulationResult, maxbin: int = 1000):
    """(rs, gr) over all saved frames, cubic box (src/nbody_simulation_result.jl:664-709)."""
    ctx = sr._ctx or build_context(sr.simulation)
    sr._ctx = ctx
    return ctx.rdf(sr.solution.u, maxbin)


# ---------------------------------------------------------------------------
# Closure protocol: get_accelerating_function (src/nbody_to_ode.jl:110-261)
# ---------------------------------------------------------------------------

class _Closure:
    """acceleration!(dv, u, v, t, i): adds potential p's acceleration of atom i
    (1-based) into the 3-vector dv.  All atoms are evaluated on the GPU in one call per
    distinct u and cached, so calling it for i = 1..n costs one kernel pass."""

    def __init__(self, ctx: _lib.Context, term: int, n: int):
        self.ctx, self.term, self.n = ctx, term, n
        self._u = None
        self._a = None

    def all_atoms(self, u) -> np.ndarray:
        u = _lib.as_c3n(u, self.n)
        if self._u is None or not np.array_equal(self._u, u):
            a = np.zeros((self.n, 3))
            self.ctx.accelerations(self.term, u, 0, self.n, a)
            self._u, self._a = u.copy(), a
        return self._a

    def __call__(self, dv, u, v, t, i: int):
        if not 1 <= i <= self.n:
            raise IndexError(f"BoundsError: atom {i} outside 1..{self.n}")
        dv += self.all_atoms(u)[i - 1]
        return dv


_CUSTOM = {}


def register_accelerating_function(param_type, factory):
    """Extension point: get_accelerating_function(p::param_type, sim) = factory(p, sim)."""
    _CUSTOM[param_type] = factory


def get_accelerating_function(parameters, simulation: NBodySimulation):
    for t, f in _CUSTOM.items():
        if isinstance(parameters, t):
            return f(parameters, simulation)
    system = simulation.system
    n = get_coordinate_vector_count(system)
    if isinstance(system, WaterSPCFw):
        ctx = build_context(simulation)
        if parameters is system.e_parameters:
            return _Closure(ctx, 0, n)
        if parameters is system.lj_parameters:
            return _Closure(ctx, 1, n)
        if isinstance(parameters, SPCFwParameters):
            return _Closure(ctx, 2, n)
        raise NotImplementedError(type(parameters).__name__)
    ctx = build_context(simulation, only=parameters)
    return _Closure(ctx, 0, n)


def get_group_accelerating_function(parameters: SPCFwParameters, simulation: NBodySimulation):
    """Angle closure (dv, u, v, t, i_mol) writing the three columns of molecule i_mol
    (1-based) into the whole (n, 3) dv (src/nbody_to_ode.jl:244-261)."""
    n = get_coordinate_vector_count(simulation.system)
    ctx = build_context(simulation)

    def acceleration(dv, u, v, t, i):
        o = 3 * (i - 1)
        ctx.accelerations(3, u, o, o + 1, dv)
        return dv
    return acceleration
