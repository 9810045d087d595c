"""The Julia binding's call sequences (nbodysimulator.jl_amd/julia/NBodySimulatorMI355X),
replayed call for call through the C ABI with ctypes, against the FP64 oracle.

* The per-atom closure of `MI355XPotential(p)`: one context per wrapped potential (create
  -> set_boundary -> add_<p> -> set_kb -> set_precision), then, on the call with i == 1 of
  an RHS evaluation, nbx_accelerations(ctx, 0, u, 0, n, cache) into a zeroed cache, and
  `dv[:, i] += cache[:, i]` for every i, closure after closure, as soode_system! calls
  them (src/nbody_to_ode.jl:474-488).
* `run_simulation(sim, MI355XVerlet())` with a Nose-Hoover thermostat: frame by frame
  (nbx_vv_run of `every` steps, nbx_get_state, nbx_get_thermostat_state), with zeta and
  its rate in column n + 1 (the reference's u[3n+1], v[3n+1], src/nbody_to_ode.jl:1-20).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def _ctx_for(n, ms, qs, L, potential):
    lib = _lib.load()
    h = C.c_void_p()
    qp = _lib.dptr(qs) if qs is not None else None
    assert lib.nbx_create(C.byref(h), n, _lib.SYS_POTENTIAL, _lib.dptr(ms), qp) == 0
    box = np.array([L, 0, 0, 0, 0, 0], dtype=np.float64)
    assert lib.nbx_set_boundary(h, _lib.BC_CUBIC, _lib.dptr(box)) == 0
    kind, p = potential
    if kind == "lj":
        assert lib.nbx_add_lennard_jones(h, *p) == 0
    else:
        assert lib.nbx_add_electrostatic(h, *p) == 0
    assert lib.nbx_set_kb(h, 1.0) == 0
    assert lib.nbx_set_precision(h, _lib.PREC_FP64_REF) == 0
    return h


def test_wrapped_closures_add_into_dv_like_soode_system(nbx):
    print("[case] Julia closure protocol: LJ + Coulomb, N=500", flush=True)
    s = gen.lj_liquid(5, rho=0.8442, T=0.72, seed=31, jitter=0.03)
    n = s.u.shape[0]
    rng = np.random.default_rng(31)
    qs = rng.choice([-0.5, 0.5], size=n).astype(np.float64)
    ms = np.ascontiguousarray(s.masses, dtype=np.float64)
    pots = [("lj", (1.0, 1.0, 2.5)), ("el", (1.3, 2.2))]
    lib = _lib.load()
    ctxs = [_ctx_for(n, ms, qs, s.L, p) for p in pots]
    try:
        u = np.ascontiguousarray(s.u)          # column-major 3 x n == row-major n x 3
        for rhs_eval in range(2):
            uu = u + 0.01 * rhs_eval
            dv = np.zeros((n, 3))              # soode_system! zeroes dv[:, i] first
            caches = []
            for h in ctxs:                     # the i == 1 call of each closure
                cache = np.zeros((n, 3))
                assert lib.nbx_accelerations(h, 0, _lib.dptr(uu), 0, n, _lib.dptr(cache)) == 0
                caches.append(cache)
            for i in range(n):                 # closures in order, atom after atom
                for cache in caches:
                    dv[i] += cache[i]
            orc = O.System(ms, charges=qs, bc=O.make_bc(O.BC_CUBIC, s.L))
            orc.add(O.TERM_LJ, *pots[0][1]).add(O.TERM_ELEC, *pots[1][1])
            ref = orc.rhs(uu)
            assert np.array_equal(dv, ref)
        # a cache that starts from another closure's contribution adds onto it (the shim
        # zeroes its own cache, but nbx_accelerations itself adds, like `+=`)
        pre = rng.normal(size=(n, 3))
        acc = pre.copy()
        assert lib.nbx_accelerations(ctxs[0], 0, _lib.dptr(u), 0, n, _lib.dptr(acc)) == 0
        only_lj = O.System(ms, bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, *pots[0][1])
        # the term is rounded first and then added (no contraction): bit for bit pre + ref
        assert np.array_equal(acc, pre + only_lj.rhs(u))
    finally:
        for h in ctxs:
            lib.nbx_destroy(h)


def test_nosehoover_frames_carry_zeta_in_column_n_plus_1(nbx):
    print("[case] Julia run_simulation protocol: Nose-Hoover LJ N=256", flush=True)
    s = gen.lj_liquid(4, rho=0.8442, T=0.72, seed=7, jitter=0.02)
    n = s.u.shape[0]
    T0, tau, dt, nsteps, every = 0.5, 0.1, 0.005, 60, 20
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, np.ones(n))
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(_lib.PREC_FP64_REF)
    ctx.set_thermostat(_lib.THERM_NOSEHOOVER, T0, tau)
    ctx.set_state(s.u, s.v)
    frames_u, frames_v = [], []
    for k in range(1 + nsteps // every):
        if k > 0:
            ctx.vv_run(dt, every)
        u, v = ctx.get_state()
        z, xi = ctx.thermostat_state()
        uf = np.zeros((n + 1, 3)); vf = np.zeros((n + 1, 3))
        uf[:n], vf[:n] = u, v
        uf[n, 0], vf[n, 0] = z, xi
        frames_u.append(uf); frames_v.append(vf)
    assert frames_u[0][n, 0] == 0.0 and frames_v[0][n, 0] == 0.0
    orc = O.System(np.ones(n), bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5)
    orc.thermostat(O.THERM_NOSEHOOVER, T0, tau)
    for k in range(1, 1 + nsteps // every):
        uo, vo, (zo, xo) = orc.vv_run_nh_state(s.u, s.v, dt, k * every)
        assert np.array_equal(frames_u[k][:n], uo) and np.array_equal(frames_v[k][:n], vo)
        assert frames_u[k][n, 0] == zo and frames_v[k][n, 0] == xo
    assert frames_u[-1][n, 0] != 0.0
