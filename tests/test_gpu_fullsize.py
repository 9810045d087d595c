"""Parity at the benchmark sizes: gravity N=65,536 (Plummer), LJ N=32,000 (FCC 20^3,
rho* 0.8442), SPC/Fw water 4,096 molecules (reaction field 0.9 nm), 512 batched LJ
replicas.  The FP64 oracle is evaluated on sampled target atoms (its per-atom loops are
O(N) per target); the size-independent properties are Newton's third law (the
momentum sum of the accelerations) and the per-target error bars of the small cases.
Lattices are jittered (0.02 sigma): a perfect lattice has zero net force.

Bars: FP64 path bit for bit against the oracle's per-target accelerations (the same
summation order); mixed path max |a - a_ref| / max |a_ref| <= 5e-5 for gravity
(unsoftened 1/r^2 on a 65,536-body sphere: 1.5e-5 measured) and <= 2e-6 for LJ and
water (9.2e-7 and 4.4e-7 measured); momentum sums below 1e-5 of the sum of |m a|; batched replicas
<= 2e-5 (FP32 LDS positions)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def mixed_accel(ctx, u):
    """a(u) of the mixed path: one velocity-Verlet step of dt = 2^-60 from rest."""
    ctx.set_state(u, np.zeros_like(u))
    dt = 2.0 ** -60
    _, vf = ctx.vv_run(dt, 1, save_every=1)
    return (vf[1] - vf[0]) / dt


def relerr(a, ref):
    return float(np.abs(a - ref).max() / np.abs(ref).max())


def momentum_ratio(a, m):
    p = (a * m[:, None]).sum(axis=0)
    return float(np.abs(p).max() / np.abs(a * m[:, None]).sum(axis=0).max())


def test_gravity_65536(nbx):
    print("[case] full size gravity N=65536", flush=True)
    n = 65536
    u, v, m = gen.plummer(n, seed=2024)
    targets = np.random.default_rng(3).choice(n, 256, replace=False)
    lib = O.load()
    ref = np.zeros((len(targets), 3))
    uu, mm = np.ascontiguousarray(u), np.ascontiguousarray(m)
    for k, i in enumerate(targets):
        lib.orc_grav_accel(O.dp(ref[k]), O.dp(uu), int(i), n, O.dp(mm), 1.0)
    f64 = _lib.Context(n, _lib.SYS_POTENTIAL, m)
    f64.add_gravitational(1.0)
    a64 = f64.rhs(u)
    assert np.array_equal(a64[targets], ref)
    mx = _lib.Context(n, _lib.SYS_POTENTIAL, m)
    mx.add_gravitational(1.0)
    mx.set_precision(_lib.PREC_MIXED)
    am = mixed_accel(mx, u)
    assert relerr(am[targets], ref) <= 5e-5
    assert momentum_ratio(am, m) <= 1e-5


def test_lj_32000(nbx):
    print("[case] full size LJ N=32000", flush=True)
    # jittered lattice: a perfect one has zero net force, and a relative error against
    # it would measure rounding noise only
    s = gen.lj_liquid(20, rho=0.8442, T=0.9, seed=5, jitter=0.02)
    n = s.u.shape[0]
    assert n == 32000
    targets = np.random.default_rng(4).choice(n, 256, replace=False)
    lib = O.load()
    ref = np.zeros((len(targets), 3))
    uu = np.ascontiguousarray(s.u)
    idx = np.arange(n, dtype=np.int64)
    bc = O.make_bc(O.BC_CUBIC, s.L)
    for k, i in enumerate(targets):
        lib.orc_lj_accel(O.dp(ref[k]), O.dp(uu), int(i), idx.ctypes.data_as(C.POINTER(C.c_int64)),
                         n, O.dp(np.ones(n)), 1.0, 1.0, 2.5, C.byref(bc))
    ctxs = []
    for prec in (_lib.PREC_FP64_REF, _lib.PREC_MIXED):
        c = _lib.Context(n, _lib.SYS_POTENTIAL, s.masses)
        c.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
        c.add_lennard_jones(1.0, 1.0, 2.5)
        c.set_kb(1.0)
        c.set_precision(prec)
        ctxs.append(c)
    a64 = ctxs[0].rhs(s.u)
    assert np.array_equal(a64[targets], ref)
    am = mixed_accel(ctxs[1], s.u)
    print(f"[case] LJ N=32000 mixed force error {relerr(am[targets], ref):.2e}", flush=True)
    assert relerr(am[targets], ref) <= 2e-6
    assert momentum_ratio(am, s.masses) <= 1e-5


def test_lj_list_entries_of_both_widths(nbx):
    """Verlet lists hold 16-bit atom indices up to 65,536 atoms and 32-bit ones above: an
    FCC liquid of 4 x 26^3 = 70,304 atoms (32-bit) and the same density at 4 x 25^3 =
    62,500 atoms (16-bit) against the oracle on sampled targets, with a short run whose
    energy drift stays small (the lists of both widths through rebuilds)."""
    lib = O.load()
    for cells in (25, 26):
        s = gen.lj_liquid(cells, rho=0.8442, T=0.9, seed=6, jitter=0.02)
        n = s.u.shape[0]
        print(f"[case] mixed LJ N={n} list entries {'32' if n > 65536 else '16'}-bit", flush=True)
        targets = np.random.default_rng(cells).choice(n, 128, replace=False)
        ref = np.zeros((len(targets), 3))
        uu = np.ascontiguousarray(s.u)
        idx = np.arange(n, dtype=np.int64)
        bc = O.make_bc(O.BC_CUBIC, s.L)
        for k, i in enumerate(targets):
            lib.orc_lj_accel(O.dp(ref[k]), O.dp(uu), int(i), idx.ctypes.data_as(C.POINTER(C.c_int64)),
                             n, O.dp(np.ones(n)), 1.0, 1.0, 2.5, C.byref(bc))
        c = _lib.Context(n, _lib.SYS_POTENTIAL, s.masses)
        c.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
        c.add_lennard_jones(1.0, 1.0, 2.5)
        c.set_kb(1.0)
        c.set_precision(_lib.PREC_MIXED)
        am = mixed_accel(c, s.u)
        print(f"[case] force error {relerr(am[targets], ref):.2e}", flush=True)
        assert relerr(am[targets], ref) <= 2e-6
        c.set_state(s.u, s.v)
        e0 = sum(c.energies())
        c.vv_run(0.005, 400)
        e1 = sum(c.energies())
        assert c.stats()["list_builds"] >= 2
        assert abs(e1 - e0) <= 1e-4 * abs(e0)


def test_water_4096(nbx):
    print("[case] full size water 4096 molecules", flush=True)
    w = gen.water_box(per_side=16, seed=7)
    n = len(w.masses)
    Rc = 0.9
    P = gen.SPCFW
    orc = O.System(w.masses, charges=w.charges, water=True, bc=O.make_bc(O.BC_CUBIC, w.L),
                   kb=P["kb"])
    orc.water_params((P["eps"], P["sigma"], Rc), O.TERM_RF, (P["k_el"], Rc, math.inf),
                     P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    ref = orc.rhs(w.u)
    out = []
    for prec in (_lib.PREC_FP64_REF, _lib.PREC_MIXED):
        c = _lib.Context(n, _lib.SYS_WATER, w.masses, w.charges)
        c.set_boundary(_lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
        c.set_kb(P["kb"])
        c.add_lennard_jones(P["eps"], P["sigma"], Rc)
        c.add_reaction_field(P["k_el"], Rc, math.inf)
        c.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
        c.set_precision(prec)
        out.append(c)
    assert relerr(out[0].rhs(w.u), ref) <= 1e-12
    am = mixed_accel(out[1], w.u)
    print(f"[case] water mixed force error {relerr(am, ref):.2e}", flush=True)
    assert relerr(am, ref) <= 2e-6
    assert momentum_ratio(am, w.masses) <= 1e-5
    # list builds on the bins of the previous build (half-size cells of at least
    # (list radius + 1.5 skin) / 2, 5 x 5 x 5 stencil): a small skin rebuilds often, and
    # the trajectory must follow the FP64 path's (a pair missed by a stale bin would move
    # atoms by ~1e-3 nm)
    skin = 0.06
    for c in out:
        c.set_state(w.u, w.v)
    out[1].set_skin(skin)
    out[1].vv_run(0.0005, 200)
    out[0].vv_run(0.0005, 200)
    st = out[1].stats()
    um, _ = out[1].get_state()
    ur, _ = out[0].get_state()
    dev = float(np.abs(um - ur).max())
    print(f"[case] water rebuilds: {st['list_builds']} builds, {st['cells_per_side']} cells per "
          f"side, max |du| vs FP64 {dev:.2e} nm", flush=True)
    assert st["cells_per_side"] == int(w.L // (0.5 * (Rc + skin + 1.5 * skin)))
    assert st["list_builds"] >= 5
    assert dev < 1e-5


def test_batch_512_replicas(nbx):
    print("[case] full size batch 512 replicas", flush=True)
    reps = gen.lj_batch(512, seed=0)
    systems = [gen.lj_liquid(r.cells, rho=0.8442, T=r.T, seed=k, jitter=0.02)
               for k, r in enumerate(reps)]
    natoms = [x.u.shape[0] for x in systems]
    b = _lib.Batch(natoms, [x.L for x in systems], np.concatenate([x.masses for x in systems]),
                   1.0, 1.0, 2.5, 1.0)
    U = np.concatenate([x.u for x in systems])
    b.set_state(U, np.zeros_like(U))
    dt = 2.0 ** -60
    b.run(dt, 1)
    _, V = b.get_state()
    A = V / dt
    offs = np.concatenate([[0], np.cumsum(natoms)])
    for k in np.random.default_rng(5).choice(512, 8, replace=False):
        x = systems[k]
        ref = O.System(x.masses, bc=O.make_bc(O.BC_CUBIC, x.L)).add(O.TERM_LJ, 1, 1, 2.5).rhs(x.u)
        assert relerr(A[offs[k]:offs[k + 1]], ref) <= 2e-5, (k, natoms[k])
