"""GPU parity: libnbx (through its C ABI) against the FP64 oracle on identical seeded inputs.

FP64 reference-order path: identical arithmetic, so the bar is bit equality except
where a libm function is involved (acos in the angle term; exp/log/cos/sin in the
stochastic thermostats), which must agree to 1e-12 relative.
Mixed path (FP32 pair forces, Kahan sums, FP64 state): norm-wise force error
max_i |a_i - a_ref_i| / max_i |a_ref_i| <= 2e-5 for LJ / Coulomb / RF / gravity.
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu

MIXED_TOL = 2e-5


def relerr(a, ref):
    return float(np.abs(a - ref).max() / max(np.abs(ref).max(), 1e-300))


def lj_pair(n, L, prec=_lib.PREC_FP64_REF, bc_kind=_lib.BC_CUBIC, R=2.5, m=1.0, therm=None):
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, np.full(n, m))
    box = [L, 0, 0, 0, 0, 0] if bc_kind == _lib.BC_CUBIC else [0, L, 0, L, 0, L]
    ctx.set_boundary(bc_kind, box if bc_kind != _lib.BC_INFINITE else None)
    ctx.add_lennard_jones(1.0, 1.0, R)
    ctx.set_kb(1.0)
    ctx.set_precision(prec)
    obc = O.make_bc(bc_kind, L if bc_kind == _lib.BC_CUBIC else [0, L, 0, L, 0, L]) \
        if bc_kind != _lib.BC_INFINITE else O.make_bc()
    orc = O.System(np.full(n, m), bc=obc).add(O.TERM_LJ, 1.0, 1.0, R)
    if therm:
        ctx.set_thermostat(*therm)
        orc.thermostat(*therm)
    return ctx, orc


@pytest.fixture(scope="module")
def c1(nbx):
    s = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=11, jitter=0.02)
    return s


# ---------------------------------------------------------------------------
# FP64 reference-order forces
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("bc_kind", [_lib.BC_CUBIC, _lib.BC_PERIODIC, _lib.BC_INFINITE])
def test_fp64_lj_forces_bitwise(c1, bc_kind):
    print(f"[case] fp64 LJ N=864 bc={bc_kind}", flush=True)
    ctx, orc = lj_pair(864, c1.L, bc_kind=bc_kind)
    a = ctx.rhs(c1.u)
    ref = orc.rhs(c1.u)
    assert np.array_equal(a, ref), relerr(a, ref)


def test_fp64_lj_unwrapped_positions_bitwise(c1):
    # unwrapped coordinates several boxes away exercise the while-loop wraps
    shift = np.random.default_rng(2).integers(-3, 4, c1.u.shape) * c1.L
    u = c1.u + shift
    ctx, orc = lj_pair(864, c1.L)
    assert np.array_equal(ctx.rhs(u), orc.rhs(u))


def test_fp64_coulomb_rf_gravity_bitwise():
    rng = np.random.default_rng(4)
    n, L = 300, 9.0
    u = rng.uniform(0, L, (n, 3))
    q = rng.choice([-1.0, 1.0], n)
    m = rng.uniform(1, 2, n)
    for kind, args in ((O.TERM_ELEC, (2.0, 3.0)), (O.TERM_ELEC, (2.0, math.inf)),
                       (O.TERM_RF, (2.0, 3.0, 78.5))):
        print(f"[case] fp64 electrostatics kind={kind} args={args}", flush=True)
        ctx = _lib.Context(n, _lib.SYS_POTENTIAL, m, q)
        ctx.set_boundary(_lib.BC_CUBIC, [L, 0, 0, 0, 0, 0])
        (ctx.add_electrostatic if kind == O.TERM_ELEC else ctx.add_reaction_field)(*args)
        orc = O.System(m, charges=q, bc=O.make_bc(O.BC_CUBIC, L)).add(kind, *args)
        a, ref = ctx.rhs(u), orc.rhs(u)
        assert np.array_equal(a, ref), relerr(a, ref)
    u, v, mm = gen.plummer(2048, seed=1)
    ctx = _lib.Context(2048, _lib.SYS_POTENTIAL, mm)
    ctx.add_gravitational(1.0)
    orc = O.System(mm).add(O.TERM_GRAV, 1.0)
    a, ref = ctx.rhs(u), orc.rhs(u)
    assert np.array_equal(a, ref), relerr(a, ref)


def water_ctx(nmol_side=3, elec_kind=O.TERM_ELEC, prec=_lib.PREC_FP64_REF, therm=None, seed=3):
    w = gen.water_box(per_side=nmol_side, seed=seed)
    P = gen.SPCFW
    n = len(w.masses)
    Rc = min(0.9, 0.45 * w.L)
    ctx = _lib.Context(n, _lib.SYS_WATER, w.masses, w.charges)
    ctx.set_boundary(_lib.BC_CUBIC, [w.L, 0, 0, 0, 0, 0])
    ctx.set_kb(P["kb"])
    ctx.add_lennard_jones(P["eps"], P["sigma"], Rc)
    elec = (P["k_el"], Rc) if elec_kind == O.TERM_ELEC else (P["k_el"], Rc, math.inf)
    (ctx.add_electrostatic if elec_kind == O.TERM_ELEC else ctx.add_reaction_field)(*elec)
    ctx.set_spcfw(P["rOH"], P["aHOH"], P["kbond"], P["kangle"])
    ctx.set_precision(prec)
    orc = O.System(w.masses, charges=w.charges, water=True, bc=O.make_bc(O.BC_CUBIC, w.L),
                   kb=P["kb"])
    orc.water_params((P["eps"], P["sigma"], Rc), elec_kind, elec, P["rOH"], P["aHOH"],
                     P["kbond"], P["kangle"])
    if therm:
        ctx.set_thermostat(*therm)
        orc.thermostat(*therm)
    return ctx, orc, w


@pytest.mark.parametrize("elec_kind", [O.TERM_ELEC, O.TERM_RF])
def test_fp64_water_rhs(elec_kind):
    print(f"[case] fp64 water rhs elec={elec_kind}", flush=True)
    ctx, orc, w = water_ctx(4, elec_kind, therm=(_lib.THERM_BERENDSEN, 300.0, 0.1))
    a, ref = ctx.rhs(w.u, w.v), orc.rhs(w.u, w.v)
    # the angle term goes through acos (OCML vs glibc): 1e-12, everything else exact
    assert relerr(a, ref) <= 1e-12


def test_fp64_closure_protocol_per_term_and_range(c1):
    ctx, orc = lj_pair(864, c1.L)
    dv = np.full((864, 3), 0.5)
    ctx.accelerations(0, c1.u, 100, 300, dv)
    ref = orc.rhs(c1.u)
    exp = np.full((864, 3), 0.5)
    exp[100:300] += ref[100:300]
    assert np.array_equal(dv, exp)
    with pytest.raises(IndexError):
        ctx.accelerations(0, c1.u, 0, 865, dv)
    with pytest.raises(IndexError):
        ctx.accelerations(1, c1.u, 0, 10, dv)


def test_water_closure_terms_sum_to_rhs():
    ctx, orc, w = water_ctx(3)
    n = len(w.masses)
    dv = np.zeros((n, 3))
    for term in range(4):
        ctx.accelerations(term, w.u, 0, n, dv)
    assert relerr(dv, orc.rhs(w.u)) <= 1e-12


# ---------------------------------------------------------------------------
# FP64 velocity Verlet, thermostats, energies, rdf
# ---------------------------------------------------------------------------

def test_fp64_vv_trajectory_bitwise(c1):
    print("[case] fp64 VV 100 steps N=864", flush=True)
    ctx, orc = lj_pair(864, c1.L)
    ctx.set_state(c1.u, c1.v)
    ctx.vv_run(0.005, 100)
    u, v = ctx.get_state()
    uo, vo, _ = orc.vv_run(c1.u, c1.v, 0.005, 100)
    assert np.array_equal(u, uo) and np.array_equal(v, vo)
    assert ctx.steps_taken() == 100


def test_fp64_vv_split_run_and_frames(c1):
    ctx, orc = lj_pair(864, c1.L)
    ctx.set_state(c1.u, c1.v)
    uf, vf = ctx.vv_run(0.005, 20, save_every=5)
    ctx.vv_run(0.005, 30)
    u, v = ctx.get_state()
    uo, vo, (ufo, vfo) = orc.vv_run(c1.u, c1.v, 0.005, 50, save_every=5)
    assert np.array_equal(u, uo) and np.array_equal(v, vo)
    assert uf.shape == (5, 864, 3)
    assert np.array_equal(uf, ufo[:5]) and np.array_equal(vf, vfo[:5])


@pytest.mark.parametrize("therm", [(_lib.THERM_BERENDSEN, 1.2, 0.05),
                                   (_lib.THERM_ANDERSEN, 1.1, 5.0),
                                   (_lib.THERM_LANGEVIN, 1.3, 1.0)])
def test_fp64_thermostats_match_oracle(c1, therm):
    print(f"[case] fp64 thermostat {therm}", flush=True)
    ctx, orc = lj_pair(864, c1.L, therm=therm)
    ctx.set_state(c1.u, c1.v)
    ctx.vv_run(0.005, 40, seed=1234)
    u, v = ctx.get_state()
    uo, vo, _ = orc.vv_run(c1.u, c1.v, 0.005, 40, seed=1234)
    assert relerr(u, uo) <= 1e-12 and relerr(v, vo) <= 1e-11


def test_fp64_water_vv_with_berendsen():
    ctx, orc, w = water_ctx(3, therm=(_lib.THERM_BERENDSEN, 300.0, 0.1))
    ctx.set_state(w.u, w.v)
    ctx.vv_run(0.0005, 50)
    u, v = ctx.get_state()
    uo, vo, _ = orc.vv_run(w.u, w.v, 0.0005, 50)
    assert relerr(u, uo) <= 1e-12 and relerr(v, vo) <= 1e-10


def test_energies_and_temperature(c1):
    ctx, orc = lj_pair(864, c1.L)
    ctx.set_state(c1.u, c1.v)
    ke, pe, pg = ctx.energies()
    assert ke == orc.kinetic_energy(c1.v)
    # The reference adds all 372,816 pair terms into one running sum; the GPU sums per
    # atom, then over atoms.  Both orders carry rounding error of ~1e-12 relative here
    # (the terms cancel), so the bar is 1e-11; the forces above are bit-identical.
    assert pe == pytest.approx(orc.potential_energy(c1.u), rel=1e-11)
    assert pg == 0.0
    lib = O.load()
    T = lib.orc_md_temperature(O.dp(c1.v), O.dp(np.ones(864)), 1.0, 864, 0)
    assert ctx.temperature() == T
    cw, ow, w = water_ctx(3)
    cw.set_state(w.u, w.v)
    assert cw.energies()[1] == pytest.approx(ow.potential_energy(w.u), rel=1e-12)


def test_gravity_potential_energy():
    u, v, m = gen.plummer(1000, seed=2)
    ctx = _lib.Context(1000, _lib.SYS_POTENTIAL, m)
    ctx.add_gravitational(1.0)
    ctx.set_state(u, v)
    ke, pe, pg = ctx.energies()
    assert pe == 0.0
    assert pg == O.grav_potential(u, m, 1.0)


def test_rdf_matches_oracle(c1):
    ctx, orc = lj_pair(864, c1.L)
    frames = np.stack([c1.u, c1.u + 0.01, c1.u - 0.02])
    rs, gr = ctx.rdf(frames)
    rso, gro = O.rdf(frames, c1.L)
    assert np.array_equal(rs, rso) and np.array_equal(gr, gro)


# ---------------------------------------------------------------------------
# Mixed precision path
# ---------------------------------------------------------------------------

def _mixed_accel(ctx, u, v):
    """a = f(u) of the mixed path, read back from one VV step of dt = 2^-60 from rest:
    v' = dt (a(u) + a(u'))/2 with u' = u + dt^2 a/2 == u in FP64, so v'/dt = a(u)."""
    ctx.set_state(u, v)
    dt = 2.0 ** -60
    uf, vf = ctx.vv_run(dt, 1, save_every=1)
    return (vf[1] - vf[0]) / dt


def test_mixed_force_error_lj_water_gravity(c1):
    ctx, orc = lj_pair(864, c1.L, prec=_lib.PREC_MIXED)
    zero = np.zeros_like(c1.u)
    a = _mixed_accel(ctx, c1.u, zero)
    assert relerr(a, orc.rhs(c1.u)) <= MIXED_TOL
    u, v, m = gen.plummer(4096, seed=5)
    cg = _lib.Context(4096, _lib.SYS_POTENTIAL, m)
    cg.add_gravitational(1.0)
    cg.set_precision(_lib.PREC_MIXED)
    ag = _mixed_accel(cg, u, np.zeros_like(u))
    assert relerr(ag, O.System(m).add(O.TERM_GRAV, 1.0).rhs(u)) <= MIXED_TOL
    for elec_kind in (O.TERM_ELEC, O.TERM_RF):
        cw, ow, w = water_ctx(5, elec_kind, prec=_lib.PREC_MIXED)
        aw = _mixed_accel(cw, w.u, np.zeros_like(w.u))
        assert relerr(aw, ow.rhs(w.u)) <= MIXED_TOL


def test_mixed_gravity_equal_and_unequal_masses():
    # with one mass for all bodies the tiles sum r_ij / r^3 and the finish applies G m in
    # FP64; otherwise m_j enters every pair term.  Both against the FP64 oracle, with a
    # common mass that is not a power of two and with masses spread over a factor 3.
    u, _, m = gen.plummer(4096, seed=5)
    rng = np.random.default_rng(11)
    for masses, G in ((m, 1.0), (np.full_like(m, 0.3), 0.7),
                      (m * rng.uniform(0.5, 1.5, size=m.shape), 1.0)):
        cg = _lib.Context(4096, _lib.SYS_POTENTIAL, masses)
        cg.add_gravitational(G)
        cg.set_precision(_lib.PREC_MIXED)
        ag = _mixed_accel(cg, u, np.zeros_like(u))
        ref = O.System(masses).add(O.TERM_GRAV, G).rhs(u)
        assert relerr(ag, ref) <= MIXED_TOL, (masses[:2], G)


def test_mixed_gravity_si_masses_stay_finite():
    # SI units: G = 6.67e-11, planetary to stellar masses (1e24 .. 2e30 kg) over 1e11 m.
    # The FP32 tiles carry m / m_max, so the self term's guarded 1/r^3 times the mass stays
    # finite (with m itself in FP32 it overflowed to inf, and 0 * inf gave NaN).
    print("[case] mixed gravity SI masses N=2048", flush=True)
    rng = np.random.default_rng(17)
    n = 2048
    masses = 10.0 ** rng.uniform(24.0, np.log10(2e30), size=n)
    d = rng.normal(size=(n, 3))
    u = 1e11 * d / np.linalg.norm(d, axis=1)[:, None] * rng.uniform(0.05, 1.0, size=(n, 1)) ** (1 / 3)
    G = 6.67e-11
    cg = _lib.Context(n, _lib.SYS_POTENTIAL, masses)
    cg.add_gravitational(G)
    cg.set_precision(_lib.PREC_MIXED)
    ag = _mixed_accel(cg, u, np.zeros_like(u))
    assert np.all(np.isfinite(ag))
    ref = O.System(masses).add(O.TERM_GRAV, G).rhs(u)
    assert relerr(ag, ref) <= MIXED_TOL


def test_mixed_lj_periodic_images_of_the_fixed_point_positions(c1):
    """The Lennard-Jones list kernel takes displacements from 32-bit box fractions (the
    integer difference is the minimum image): atoms moved by whole boxes, up to 3 L either
    way, give the same forces to the rounding of their fractions, and both match the oracle.
    A dropped wrap or a sign error in the fraction would move a shifted atom's image."""
    print("[case] mixed LJ N=864, atoms shifted by whole boxes", flush=True)
    rng = np.random.default_rng(17)
    shifted = c1.u + c1.L * rng.integers(-3, 4, size=c1.u.shape)
    ctx, orc = lj_pair(864, c1.L, prec=_lib.PREC_MIXED)
    zero = np.zeros_like(c1.u)
    a0 = _mixed_accel(ctx, c1.u, zero)
    a1 = _mixed_accel(ctx, shifted, zero)
    assert relerr(a1, a0) <= 1e-6
    assert relerr(a1, orc.rhs(c1.u)) <= MIXED_TOL


def test_mixed_all_pairs_mode_matches(c1):
    # periodic lo/hi wrapping is not minimum image: the mixed path uses all pairs
    ctx, orc = lj_pair(864, c1.L, prec=_lib.PREC_MIXED, bc_kind=_lib.BC_PERIODIC)
    a = _mixed_accel(ctx, c1.u, np.zeros_like(c1.u))
    assert relerr(a, orc.rhs(c1.u)) <= MIXED_TOL


def test_mixed_nve_energy_drift_and_determinism(c1):
    print("[case] mixed NVE 2000 steps N=864", flush=True)
    s = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=3)
    ctx, orc = lj_pair(864, s.L, prec=_lib.PREC_MIXED)
    ctx.set_state(s.u, s.v)
    ke0, pe0, _ = ctx.energies()
    ctx.vv_run(0.005, 1000)
    ctx.vv_run(0.005, 1000)
    ke1, pe1, _ = ctx.energies()
    u1, v1 = ctx.get_state()
    e0, e1 = ke0 + pe0, ke1 + pe1
    assert abs(e1 - e0) / abs(e0) < 5e-4
    ctx.set_state(s.u, s.v)
    ctx.vv_run(0.005, 2000)
    u2, v2 = ctx.get_state()
    assert np.array_equal(u1, u2) and np.array_equal(v1, v2)


def test_mixed_thermostats_follow_oracle_statistically(c1):
    for therm in ((_lib.THERM_LANGEVIN, 1.2, 1.0), (_lib.THERM_ANDERSEN, 1.2, 20.0),
                  (_lib.THERM_BERENDSEN, 1.2, 0.05)):
        ctx, _ = lj_pair(864, c1.L, prec=_lib.PREC_MIXED, therm=therm)
        ctx.set_state(c1.u, c1.v)
        ctx.vv_run(0.005, 1500, seed=9)
        assert ctx.temperature() == pytest.approx(1.2, rel=0.1), therm


def test_mixed_count_pairs_and_kernel_timer(c1):
    ctx, _ = lj_pair(864, c1.L, prec=_lib.PREC_MIXED)
    ctx.set_state(c1.u, c1.v)
    pairs = ctx.count_pairs()
    d = c1.u[:, None, :] - c1.u[None, :, :]
    d -= c1.L * np.round(d / c1.L)
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, np.inf)
    assert pairs == int((r2 < 6.25).sum())
    ms, name = ctx.time_force_kernel(5)
    assert ms > 0 and name == "k_nb_list"
    # issue-rate calibration: an independent v_rsq_f32 stream issues slower than v_fmac_f32
    assert ctx.valu_probe(0, 2) > ctx.valu_probe(2, 2) > 0


def test_every_lane_count_gives_the_same_physics(c1, monkeypatch):
    # the Verlet list is interleaved for G lanes per atom; every G the autotuner can pick
    # (and the learned stride's descriptor and energy kernels over the same list) must
    # read it correctly: forces against the FP64 oracle, the same pair count, and the
    # descriptor of the learned stride against its oracle
    s = gen.lj_liquid(5, rho=0.8442, T=0.9, seed=3, jitter=0.03)
    orc = O.System(np.ones(500), bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5)
    a_ref = orc.rhs(s.u)
    fo, _ = O.lr_features(s.u, a_ref, s.v, s.L)
    counts, pes = set(), []
    for lanes in (1, 2, 4, 8, 16, 32):
        print(f"[case] lanes {lanes}", flush=True)
        monkeypatch.setenv("NBX_LANES", str(lanes))
        ctx, _ = lj_pair(500, s.L, prec=_lib.PREC_MIXED)
        assert ctx.stats()["lanes"] == lanes
        assert relerr(_mixed_accel(ctx, s.u, np.zeros_like(s.u)), a_ref) <= MIXED_TOL, lanes
        ctx.set_state(s.u, s.v)
        counts.add(ctx.count_pairs())
        f, _ = ctx.features()
        assert np.abs(f[:, :32] - fo[:, :32]).max() <= 1e-5 * np.abs(fo[:, :32]).max(), lanes
        pes.append(ctx.learn_energy()[1])
        cw, ow, w = water_ctx(4, O.TERM_RF, prec=_lib.PREC_MIXED)
        assert cw.stats()["lanes"] == lanes
        aw = _mixed_accel(cw, w.u, np.zeros_like(w.u))
        assert relerr(aw, ow.rhs(w.u)) <= MIXED_TOL, lanes
    assert len(counts) == 1
    assert max(abs(p - pes[0]) for p in pes) <= 1e-6 * abs(pes[0])


def test_list_overflow_replays_the_chunk():
    # an imploding gas: list capacity is sized for the initial density and is exceeded
    # inside a run; the chunk is restored from its device backup and replayed with larger
    # lists.  The run must follow the FP64 path (FP32 pair forces aside) and the lists
    # must have grown on the way.
    s = gen.lj_liquid(6, rho=0.25, T=0.1, seed=5, jitter=0.05)
    centre = 0.5 * s.L
    v = s.v - 2.0 * (s.u - centre) / centre
    mixed, _ = lj_pair(864, s.L, prec=_lib.PREC_MIXED)
    mixed.set_state(s.u, v)
    mixed.vv_run(0.002, 1)
    slots0 = mixed.stats()["list_slots"]
    mixed.set_state(s.u, v)
    mixed.vv_run(0.002, 500)
    slots1 = mixed.stats()["list_slots"]
    ref, _ = lj_pair(864, s.L)
    ref.set_state(s.u, v)
    ref.vv_run(0.002, 500)
    um, _ = mixed.get_state()
    ur, _ = ref.get_state()
    dev = float(np.abs(um - ur).max())
    print(f"[case] implosion: list slots {slots0} -> {slots1}, max |du| vs FP64 {dev:.2e}",
          flush=True)
    assert slots1 > slots0
    assert dev < 1e-5


def test_kernel_timing_over_a_run(c1):
    # one timed launch per step (graphs off while timing), and the same trajectory bit for
    # bit as the untimed (graph-replayed) run
    u, v, m = gen.plummer(2048, seed=4)
    cases = []
    g = _lib.Context(2048, _lib.SYS_POTENTIAL, m)
    g.add_gravitational(1.0)
    cases.append((g, u, v, 120, "k_grav_tiled"))
    lj, _ = lj_pair(864, c1.L, prec=_lib.PREC_MIXED)
    cases.append((lj, c1.u, c1.v, 120, "k_nb_list"))
    for ctx, u0, v0, steps, kname in cases:
        ctx.set_precision(_lib.PREC_MIXED)
        ctx.set_state(u0, v0)
        ctx.vv_run(0.001, steps)
        ref = ctx.get_state()
        ctx.set_state(u0, v0)
        ctx.set_kernel_timing(True)
        ctx.vv_run(0.001, steps)
        ms, launches, name = ctx.kernel_timing()
        ctx.set_kernel_timing(False)
        got = ctx.get_state()
        assert name == kname and launches == steps and ms > 0, (name, launches, ms)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), kname
        assert ctx.kernel_timing()[1] == 0


# ---------------------------------------------------------------------------
# Edge cases
# ---------------------------------------------------------------------------

def test_edge_cases():
    ctx = _lib.Context(0, _lib.SYS_POTENTIAL, np.zeros(0))
    ctx.add_lennard_jones(1, 1, 2.5)
    ctx.set_state(np.zeros((0, 3)))
    ctx.vv_run(0.01, 10)
    one = _lib.Context(1, _lib.SYS_POTENTIAL, [1.0])
    one.add_lennard_jones(1, 1, 2.5)
    one.set_boundary(_lib.BC_CUBIC, [5.0, 0, 0, 0, 0, 0])
    assert np.array_equal(one.rhs(np.array([[1.0, 2.0, 3.0]])), np.zeros((1, 3)))
    with pytest.raises(FloatingPointError):
        one.set_state(np.array([[np.nan, 0, 0]]))
    with pytest.raises(ValueError):
        _lib.Context(4, _lib.SYS_WATER, np.ones(4))
    with pytest.raises(ValueError):
        one.set_thermostat(_lib.THERM_ANDERSEN, 1.0, -1.0)
    two = _lib.Context(2, _lib.SYS_POTENTIAL, [1.0, 1.0])
    two.add_gravitational(1.0)
    two.add_lennard_jones(1.0, 1.0, 2.5)
    two.set_precision(_lib.PREC_MIXED)
    two.set_state(np.array([[0, 0, 0], [1.5, 0, 0.0]]))
    with pytest.raises(NotImplementedError):
        two.vv_run(0.01, 1)
