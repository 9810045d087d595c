"""Spatially sorted state inside mixed-path runs (NBX_SORT_EVERY).

Inside nbx_vv_run the state of a Lennard-Jones Verlet-list run is kept sorted by cell and
re-sorted at absolute steps that are multiples of the interval; outside a run it is in the
user's order.  The bar: a sorted run follows an unsorted run from the same state to the
rounding of the FP32 pair sums (their order changes with the slots), thermostat noise is
drawn for the user's atom index (a wrong index would move every atom by ~1e-3 after 100
Langevin steps), frames come back in the user's order, split runs reproduce an unsplit run
bit for bit, and entry points used between runs see the user's order.  The inputs are
shuffled so that the sorts change the order (`nbx_get_stats` out[7] counts them).
"""
import numpy as np
import pytest

from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def shuffled_liquid(seed, T):
    """A 4,000-atom LJ liquid with its atoms in a random order: a state in generation order
    is already nearly in cell order, and the library keeps such an order."""
    s = gen.lj_liquid(10, rho=0.8442, T=T, seed=seed, jitter=0.02)
    order = np.random.default_rng(seed).permutation(s.u.shape[0])
    s.u, s.v = s.u[order].copy(), s.v[order].copy()
    return s


def lj_ctx(s, therm=True):
    ctx = _lib.Context(s.u.shape[0], _lib.SYS_POTENTIAL, s.masses)
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(_lib.PREC_MIXED)
    if therm:
        ctx.set_thermostat(_lib.THERM_LANGEVIN, 1.2, 1.0)
    ctx.set_state(s.u, s.v)
    return ctx


def test_sorted_run_follows_unsorted_run(nbx, monkeypatch):
    s = shuffled_liquid(21, 1.2)   # N = 4,000
    out = {}
    for every in ("0", "40"):
        print(f"[case] N=4000 Langevin, sort every {every}", flush=True)
        monkeypatch.setenv("NBX_SORT_EVERY", every)
        ctx = lj_ctx(s)
        uf, vf = ctx.vv_run(0.005, 100, seed=9, save_every=25)
        u, v = ctx.get_state()
        out[every] = (u, v, uf, vf)
        # the shuffled state is sorted at step 0 (then at 40 and 80 if it has mixed enough)
        r = ctx.stats()["resorts"]
        assert r == 0 if every == "0" else 1 <= r <= 3
    du = float(np.abs(out["0"][0] - out["40"][0]).max())
    dv = float(np.abs(out["0"][1] - out["40"][1]).max())
    dfr = float(np.abs(out["0"][2] - out["40"][2]).max())
    print(f"[case] sorted vs unsorted after 100 steps: max |du| {du:.2e}, |dv| {dv:.2e}, "
          f"frames {dfr:.2e}", flush=True)
    assert du < 1e-5 and dv < 1e-4 and dfr < 1e-5
    # the first frame is the initial state, in the user's order
    assert np.array_equal(out["40"][2][0], s.u)


def test_sorted_split_runs_are_bitwise_equal(nbx, monkeypatch):
    monkeypatch.setenv("NBX_SORT_EVERY", "50")
    s = shuffled_liquid(22, 1.2)
    print("[case] sorted runs 137 + 163 against 300 steps", flush=True)
    a = lj_ctx(s)
    a.vv_run(0.005, 137, seed=3)
    a.vv_run(0.005, 163, seed=3)
    ua, va = a.get_state()
    b = lj_ctx(s)
    b.vv_run(0.005, 300, seed=3)
    ub, vb = b.get_state()
    assert a.stats()["resorts"] > 0 and a.stats()["resorts"] == b.stats()["resorts"]
    assert np.array_equal(ua, ub) and np.array_equal(va, vb)


def test_entry_points_between_sorted_runs_use_user_order(nbx, monkeypatch):
    s = shuffled_liquid(23, 0.9)
    res = {}
    for every in ("0", "30"):
        print(f"[case] NVE with calls between runs, sort every {every}", flush=True)
        monkeypatch.setenv("NBX_SORT_EVERY", every)
        ctx = lj_ctx(s, therm=False)
        ctx.vv_run(0.005, 60)
        pairs = ctx.count_pairs()
        ke, pe = ctx.learn_energy()
        ctx.vv_run(0.005, 60)
        u, _ = ctx.get_state()
        res[every] = (pairs, ke + pe, u)
        assert (ctx.stats()["resorts"] > 0) == (every != "0")
    print(f"[case] pairs {res['0'][0]} / {res['30'][0]}, energy {res['0'][1]:.10g} / "
          f"{res['30'][1]:.10g}", flush=True)
    assert abs(res["0"][0] - res["30"][0]) <= 2     # a pair within rounding of the cutoff
    assert abs(res["0"][1] - res["30"][1]) <= 1e-7 * abs(res["0"][1])
    assert float(np.abs(res["0"][2] - res["30"][2]).max()) < 1e-5
