"""Seeded initial conditions for the benchmark systems (numpy, host side).

- fcc_lattice / lj_liquid: 4-site FCC basis on an m x m x m lattice at reduced density
  rho, Maxwell-Boltzmann velocities with the centre-of-mass momentum removed and the
  temperature rescaled to T (md_temperature with 3N degrees of freedom).
- plummer: Plummer sphere by Aarseth, Henon & Wielen (A&A 37, 183, 1974), G = M = a = 1.
- water_box: SPC/Fw molecules on a jittered cubic O lattice with uniformly random
  orientations (nm / kJ mol^-1 / ps / e units, as the reference's water tests use).
- lj_batch: replica sizes, temperatures and thermostats of the batched LJ workload.

The reference's own generator, generate_bodies_in_cell_nodes (src/bodies.jl:139-159),
draws Julia MersenneTwister normals; `generate_bodies_in_cell_nodes` below keeps its
simple-cubic placement but uses numpy's generator, so its velocities differ.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .types import MassBody, WaterMolecule

# SPC/Fw (Wu, Tepper & Voth, J. Chem. Phys. 124, 024503 (2006)) in nm, kJ/mol, ps, e
KCAL = 4.184
SPCFW = dict(
    mO=15.999, mH=1.00794, qO=-0.82, qH=0.41,
    sigma=0.3165492, eps=0.1554253 * KCAL,
    rOH=0.1012, kbond=1059.162 * KCAL * 100.0,
    aHOH=113.24 * math.pi / 180.0, kangle=75.90 * KCAL,
    k_el=138.935458, kb=8.3144598e-3,
)
DA_PER_NM3_PER_G_CM3 = 602.214076  # 1 g/cm^3 in Da/nm^3


def fcc_lattice(cells: int, L: float) -> np.ndarray:
    a = L / cells
    basis = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
    g = np.arange(cells)
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    corners = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], axis=1).astype(np.float64)
    pos = (corners[:, None, :] + basis[None, :, :]).reshape(-1, 3) * a
    return np.ascontiguousarray(pos + 0.25 * a)


def maxwell_boltzmann(n: int, T: float, masses, kb: float = 1.0, seed: int = 0,
                      rescale: bool = True) -> np.ndarray:
    rng = np.random.default_rng(seed)
    m = np.broadcast_to(np.asarray(masses, dtype=np.float64), (n,))
    v = rng.standard_normal((n, 3)) * np.sqrt(kb * T / m)[:, None]
    p = (m[:, None] * v).sum(axis=0) / m.sum()
    v -= p[None, :]
    if rescale and n > 1 and T > 0:
        t_now = float((m[:, None] * v * v).sum()) / (kb * 3 * n)
        v *= math.sqrt(T / t_now)
    return np.ascontiguousarray(v)


@dataclass
class LJSystem:
    u: np.ndarray       # (N, 3) positions
    v: np.ndarray       # (N, 3) velocities
    L: float
    masses: np.ndarray


def lj_liquid(cells: int, rho: float = 0.8442, T: float = 0.72, seed: int = 0,
              jitter: float = 0.0) -> LJSystem:
    n = 4 * cells ** 3
    L = (n / rho) ** (1.0 / 3.0)
    u = fcc_lattice(cells, L)
    rng = np.random.default_rng(seed + 7919)
    if jitter > 0:
        u = u + jitter * rng.standard_normal(u.shape)
    v = maxwell_boltzmann(n, T, 1.0, 1.0, seed)
    return LJSystem(np.ascontiguousarray(u), v, L, np.ones(n))


def plummer(n: int, seed: int = 0, mass_cut: float = 0.999):
    """Equal-mass Plummer sphere (Aarseth, Henon & Wielen 1974), G = M = a = 1, in the
    centre-of-mass frame.  Returns (u, v, masses)."""
    rng = np.random.default_rng(seed)
    X1 = rng.uniform(0.0, mass_cut, n)
    X1 = np.maximum(X1, 1e-12)
    r = 1.0 / np.sqrt(X1 ** (-2.0 / 3.0) - 1.0)

    def iso(k):
        z = rng.uniform(-1.0, 1.0, k)
        ph = rng.uniform(0.0, 2.0 * math.pi, k)
        s = np.sqrt(1.0 - z * z)
        return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)

    u = r[:, None] * iso(n)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        X4 = rng.uniform(0.0, 1.0, todo.size)
        X5 = rng.uniform(0.0, 1.0, todo.size)
        ok = 0.1 * X5 < X4 * X4 * (1.0 - X4 * X4) ** 3.5
        q[todo[ok]] = X4[ok]
        todo = todo[~ok]
    ve = math.sqrt(2.0) * (1.0 + r * r) ** (-0.25)
    v = (q * ve)[:, None] * iso(n)
    u -= u.mean(axis=0)
    v -= v.mean(axis=0)
    m = np.full(n, 1.0 / n)
    return np.ascontiguousarray(u), np.ascontiguousarray(v), m


def random_rotations(k: int, rng) -> np.ndarray:
    """k uniformly random rotation matrices (unit quaternions, Shoemake 1992)."""
    u1, u2, u3 = rng.uniform(size=(3, k))
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], 1)
    x, y, z, w = q.T
    R = np.empty((k, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


@dataclass
class WaterBox:
    u: np.ndarray        # (3 nmol, 3) atoms O, H1, H2 per molecule
    v: np.ndarray
    L: float
    masses: np.ndarray
    charges: np.ndarray
    molecules: list      # WaterMolecule bodies


def water_box(per_side: int = 16, density_g_cm3: float = 0.997, T: float = 300.0,
              seed: int = 0, jitter: float = 0.01) -> WaterBox:
    p = SPCFW
    nmol = per_side ** 3
    m_h2o = p["mO"] + 2 * p["mH"]
    rho = density_g_cm3 * DA_PER_NM3_PER_G_CM3
    L = (nmol * m_h2o / rho) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    a = L / per_side
    g = (np.arange(per_side) + 0.5) * a
    cx, cy, cz = np.meshgrid(g, g, g, indexing="ij")
    O = np.stack([cx.ravel(), cy.ravel(), cz.ravel()], 1) + jitter * rng.standard_normal((nmol, 3))
    R = random_rotations(nmol, rng)
    r0, th = p["rOH"], p["aHOH"]
    h1 = np.array([r0, 0.0, 0.0])
    h2 = np.array([r0 * math.cos(th), r0 * math.sin(th), 0.0])
    H1 = O + R @ h1
    H2 = O + R @ h2
    u = np.empty((3 * nmol, 3))
    u[0::3], u[1::3], u[2::3] = O, H1, H2
    ms = np.empty(3 * nmol)
    ms[0::3] = p["mO"]; ms[1::3] = p["mH"]; ms[2::3] = p["mH"]
    qs = np.empty(3 * nmol)
    qs[0::3] = p["qO"]; qs[1::3] = p["qH"]; qs[2::3] = p["qH"]
    v = maxwell_boltzmann(3 * nmol, T, ms, p["kb"], seed + 1, rescale=False)
    mols = [WaterMolecule(MassBody(u[3 * i], v[3 * i], p["mO"]),
                          MassBody(u[3 * i + 1], v[3 * i + 1], p["mH"]),
                          MassBody(u[3 * i + 2], v[3 * i + 2], p["mH"])) for i in range(nmol)]
    return WaterBox(np.ascontiguousarray(u), v, L, ms, qs, mols)


@dataclass
class Replica:
    cells: int
    N: int
    L: float
    T: float
    thermostat: str      # "langevin" or "andersen"


def lj_batch(count: int = 512, seed: int = 0, nmin: int = 3, nmax: int = 10):
    """Replica table of the batched workload: N = 4 n^3, n ~ U{nmin..nmax},
    T* ~ U[0.7, 1.5], Langevin (gamma 1.0) or Andersen (nu 0.01) by a fair coin."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        N = 4 * n ** 3
        out.append(Replica(n, N, (N / 0.8442) ** (1.0 / 3.0), float(rng.uniform(0.7, 1.5)),
                           "langevin" if rng.uniform() < 0.5 else "andersen"))
    return out


def generate_bodies_in_cell_nodes(n: int, m: float, v_dev: float, L: float, seed: int = None):
    """Simple-cubic placement of src/bodies.jl:139-159 with numpy normal velocities."""
    rng = np.random.default_rng(n if seed is None else seed)
    vel = v_dev * rng.standard_normal((3, n))
    dL = L / math.ceil(n ** (1.0 / 3.0))
    bodies = []
    count = 0
    xs = np.arange(dL / 2, L, dL)
    for x in xs:
        for y in xs:
            for z in xs:
                if count >= n:
                    return bodies
                bodies.append(MassBody(np.array([x, y, z]), vel[:, count].copy(), m))
                count += 1
    return bodies
