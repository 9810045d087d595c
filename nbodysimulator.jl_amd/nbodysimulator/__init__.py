"""NBodySimulator.jl's API over the gfx950 engine libnbx (include/nbx.h).

Same names, fields and argument meaning as the Julia package (module export list,
src/NBodySimulator.jl:26-36); the force evaluation, the velocity-Verlet step, the
thermostats (and the Euler-Maruyama step of the Langevin SDE), energies, temperature
and rdf run as HIP kernels on the GPU.
"""
from ._lib import LIB_PATH, Batch, Context, Model, NBXError, device_count, load
from .simulation import (EM, EulerMaruyama, MI355XVerlet, SimulationResult, Solution,
                         VelocityVerlet,
                         build_context, get_accelerating_function,
                         get_group_accelerating_function, get_position, get_velocity,
                         gravitational_potential_energy, initial_energy, kinetic_energy,
                         potential_energy, rdf, register_accelerating_function, run_simulation,
                         run_simulations,
                         temperature, total_energy)
from .types import (AndersenThermostat, BerendsenThermostat, ChargedParticle, ChargedParticles,
                    CubicPeriodicBoundaryConditions, ElectrostaticParameters,
                    GravitationalParameters, GravitationalSystem, InfiniteBox,
                    LangevinThermostat, LennardJonesParameters, MagneticParticle,
                    MagnetostaticParameters, MassBody, NBodySimulation, NoseHooverThermostat,
                    NullThermostat, PeriodicBoundaryConditions, PotentialNBodySystem,
                    PotentialParameters, ReactionFieldParameters, SPCFwParameters,
                    WaterMolecule, WaterSPCFw, get_masses, kb_SI)
from .generators import generate_bodies_in_cell_nodes
from .pdb import extract_from_pdb, load_water_molecules_from_pdb, save_to_pdb, write_pdb_data

__all__ = [n for n in dir() if not n.startswith("_")]
