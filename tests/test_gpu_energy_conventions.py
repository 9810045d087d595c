"""GPU energies against closed forms computed here (numpy), independent of the oracle:
the reference's Lennard-Jones energy includes U(R) for every pair beyond the cutoff
(src/nbody_simulation_result.jl:293-319), the reaction-field pair energy is shifted to
zero at the cutoff, and truncated Coulomb adds q_j/R beyond it (:321-351)."""
import math

import numpy as np
import pytest

from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def _U(r2):
    q = 1.0 / r2
    return 4.0 * (q ** 6 - q ** 3)


def _shifted_sum(u, L, R):
    d = u[:, None, :] - u[None, :, :]
    d -= L * np.round(d / L)
    r2 = np.einsum("ijk,ijk->ij", d, d)[np.triu_indices(u.shape[0], 1)]
    r2 = r2[r2 < R * R]
    return float(np.sum(_U(r2) - _U(R * R)))


@pytest.mark.parametrize("prec", [_lib.PREC_FP64_REF, _lib.PREC_MIXED])
def test_lj_energy_is_shifted_sum_plus_tail(nbx, prec):
    print(f"[case] LJ N=864 energy conventions prec={prec}", flush=True)
    s = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=21, jitter=0.02)
    n, R = s.u.shape[0], 2.5
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, s.masses)
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, R)
    ctx.set_precision(prec)
    ctx.set_state(s.u, s.v)
    es = _shifted_sum(s.u, s.L, R)
    tail = n * (n - 1) / 2 * _U(R * R)
    ke, pe, _ = ctx.energies()
    assert pe == pytest.approx(es + tail, rel=1e-10)
    assert ke == pytest.approx(0.5 * float(np.sum(s.masses[:, None] * s.v ** 2)), rel=1e-13)
    if prec == _lib.PREC_MIXED:
        ke2, pe_shift = ctx.learn_energy()   # list-based, shifted definition
        assert pe_shift == pytest.approx(es, rel=1e-10)
        assert ke2 == pytest.approx(ke, rel=1e-13)


@pytest.mark.parametrize("eps_rf", [78.5, math.inf])
def test_reaction_field_and_coulomb_pair_energies(nbx, eps_rf):
    print(f"[case] RF / Coulomb pair energies eps_rf={eps_rf}", flush=True)
    k, R, q = 138.935458, 0.9, np.array([-0.82, 0.41])
    krf = 1.0 / (2 * R ** 3) if math.isinf(eps_rf) else (eps_rf - 1) / ((2 * eps_rf + 1) * R ** 3)
    crf = 1.0 / R + krf * R * R
    for kind in ("rf", "coulomb"):
        ctx = _lib.Context(2, _lib.SYS_POTENTIAL, np.array([15.999, 1.008]), q)
        ctx.set_boundary(_lib.BC_CUBIC, [5.0, 0, 0, 0, 0, 0])
        if kind == "rf":
            ctx.add_reaction_field(k, R, eps_rf)
        else:
            ctx.add_electrostatic(k, R)
        for r in (0.3, 0.85, 1.4):
            u = np.array([[1.0, 1.0, 1.0], [1.0, 1.0 + r, 1.0]])
            ctx.set_state(u, np.zeros_like(u))
            pe = ctx.energies()[1]
            if kind == "rf":
                expect = k * q[0] * q[1] * (1 / r + krf * r * r - crf) if r < R else 0.0
            else:
                expect = k * q[0] * q[1] / min(r, R)
            assert pe == pytest.approx(expect, rel=1e-13, abs=1e-15)
