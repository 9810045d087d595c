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
