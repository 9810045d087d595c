"""GPU tests of the autotuner: it picks a candidate launch shape, restores the state and
step count, and the tuned context integrates the same physics."""
import numpy as np
import pytest

from nbodysimulator import _lib
from nbodysimulator import generators as gen

pytestmark = pytest.mark.gpu


def test_autotune_list_plan_restores_state(nbx):
    print("[case] autotune LJ N=864", flush=True)
    s = gen.lj_liquid(6, rho=0.8442, T=0.72, seed=21)
    ctx = _lib.Context(864, _lib.SYS_POTENTIAL, np.ones(864))
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(_lib.PREC_MIXED)
    ctx.set_state(s.u, s.v)
    ke0, pe0, _ = ctx.energies()
    r = ctx.autotune(0.005, 60)
    assert r["shape0"] in (2, 4, 8, 16, 32)
    assert any(abs(r["shape1"] - f * 2.5) < 1e-12
               for f in (0.06, 0.09, 0.12, 0.18, 0.24, 0.32, 0.40))
    assert r["best_ms_per_step"] <= r["entry_ms_per_step"] and r["candidates"] >= 20
    u, v = ctx.get_state()
    assert np.array_equal(u, s.u) and np.array_equal(v, s.v)
    assert ctx.steps_taken() == 0
    assert ctx.stats()["lanes"] == int(r["shape0"])
    ctx.vv_run(0.005, 500)
    ke1, pe1, _ = ctx.energies()
    assert abs((ke1 + pe1) - (ke0 + pe0)) <= 1e-4 * abs(ke0 + pe0)


def test_autotune_gravity(nbx):
    print("[case] autotune gravity N=4096", flush=True)
    u, v, m = gen.plummer(4096, seed=9)
    ctx = _lib.Context(4096, _lib.SYS_POTENTIAL, m)
    ctx.add_gravitational(1.0)
    ctx.set_precision(_lib.PREC_MIXED)
    ctx.set_state(u, v)
    ke0, _, pg0 = ctx.energies()
    r = ctx.autotune(1e-3, 20)
    assert r["shape0"] in (1, 2, 4) and r["shape1"] in (1, 2, 4, 8, 16)
    u1, _ = ctx.get_state()
    assert np.array_equal(u1, u)
    ctx.vv_run(1e-3, 100)
    ke1, _, pg1 = ctx.energies()
    assert abs((ke1 + pg1) - (ke0 + pg0)) <= 1e-5 * abs(ke0 + pg0)


def _lj500(seed):
    s = gen.lj_liquid(5, rho=0.8442, T=0.8, seed=seed)
    ctx = _lib.Context(500, _lib.SYS_POTENTIAL, np.ones(500))
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(_lib.PREC_MIXED)
    ctx.set_state(s.u, s.v)
    return ctx


def test_autotune_results_are_kept_in_process_and_file(nbx, tmp_path, monkeypatch):
    print("[case] autotune cache LJ N=500", flush=True)
    first = tmp_path / "tune.tsv"
    monkeypatch.setenv("NBX_TUNE_CACHE", str(first))
    r1 = _lj500(31).autotune(0.005, 37)
    assert r1["candidates"] >= 20
    lines = first.read_text().splitlines()
    assert len(lines) == 1 and lines[0].startswith("list n=500 ")
    # the same configuration again (another state): the kept shape, no trials
    c2 = _lj500(32)
    r2 = c2.autotune(0.005, 37)
    assert r2["candidates"] == 0
    assert (r2["shape0"], r2["shape1"]) == (r1["shape0"], r1["shape1"])
    assert c2.stats()["lanes"] == int(r1["shape0"])
    # a cache file written elsewhere is read: its line overrides the kept result
    key = lines[0].split("\t")[0]
    other_lanes = 4 if int(r1["shape0"]) != 4 else 16
    second = tmp_path / "other.tsv"
    second.write_text(f"{key}\t{other_lanes}\t0.75\t0.01\t0.02\n")
    monkeypatch.setenv("NBX_TUNE_CACHE", str(second))
    c3 = _lj500(33)
    r3 = c3.autotune(0.005, 37)
    assert r3["candidates"] == 0 and r3["shape0"] == other_lanes and r3["shape1"] == 0.75
    c3.vv_run(0.005, 50)
    assert c3.stats()["lanes"] == other_lanes
