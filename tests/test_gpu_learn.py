"""GPU tests of the learned stride through the C ABI: the BF16 MFMA forward and gradient
against the numpy oracle, training, the descriptor against the C oracle, the training
targets, the corrected integrator with a zero network (= plain velocity Verlet at the
stride), the energy-drift fallback, and what a trained network gains over the
uncorrected strided step on held-out trajectories.

Tolerances: the kernels round inputs, weights, activations and gradients to BF16 at the
same points as the oracle but accumulate in FP32 in another order, so an activation can
land on the other side of a BF16 rounding boundary (2^-8 relative); outputs are compared
at 2% of their RMS, gradients at 3% (norm-wise) with a cosine above 0.999.
"""
import numpy as np
import pytest

import oracle_lib as O
from nbodysimulator import _lib
from nbodysimulator import generators as gen
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def random_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    for k, off in M.OFFSETS.items():
        fan = M.SHAPES[k][-1] if k.startswith("W") else M.SHAPES["W" + k[1]][1]
        size = int(np.prod(M.SHAPES[k]))
        p[off:off + size] = scale * rng.uniform(-1, 1, size) * np.sqrt(3.0 / fan)
    norm = np.concatenate([rng.normal(0, 0.5, 35), rng.uniform(0.5, 2.0, 35),
                           rng.uniform(0.5, 2.0, 6)])
    return p, norm


def lj_ctx(cells, T, seed, jitter=0.0):
    s = gen.lj_liquid(cells, rho=0.8442, T=T, seed=seed, jitter=jitter)
    n = s.u.shape[0]
    ctx = _lib.Context(n, _lib.SYS_POTENTIAL, np.ones(n))
    ctx.set_boundary(_lib.BC_CUBIC, [s.L, 0, 0, 0, 0, 0])
    ctx.add_lennard_jones(1.0, 1.0, 2.5)
    ctx.set_kb(1.0)
    ctx.set_precision(_lib.PREC_MIXED)
    ctx.set_state(s.u, s.v)
    return ctx, s


def test_mlp_forward_matches_oracle(nbx):
    print("[case] MLP forward n=300", flush=True)
    p, norm = random_params(1)
    m = _lib.Model(0)
    m.set(p, norm)
    feats = np.random.default_rng(2).normal(1.0, 1.5, size=(300, 35))
    y = m.forward(feats)
    yo, _ = M.forward(p, norm, feats, rounding=True)
    rms = np.sqrt(np.mean(yo * yo))
    assert np.abs(y - yo).max() <= 0.02 * rms, (np.abs(y - yo).max(), rms)
    # the unrounded network is further away than the rounding-aware oracle
    y64, _ = M.forward(p, norm, feats, rounding=False)
    assert np.linalg.norm(y - yo) <= np.linalg.norm(y - y64)


def test_mlp_gradient_matches_oracle(nbx):
    print("[case] MLP gradient n=512", flush=True)
    p, norm = random_params(3)
    m = _lib.Model(0)
    m.set(p, norm)
    rng = np.random.default_rng(4)
    feats = rng.normal(1.0, 1.5, size=(512, 35))
    targets = rng.normal(size=(512, 6)) * norm[70:76]
    g, loss = m.gradient(feats, targets)
    lo, go = M.loss_and_gradient(p, norm, feats, targets, rounding=True)
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    cos = np.dot(g, go) / (np.linalg.norm(g) * np.linalg.norm(go))
    assert rel <= 0.03 and cos >= 0.999, (rel, cos)
    # every block of the flat layout (weights and biases of the four layers) agrees
    for k, off in M.OFFSETS.items():
        size = int(np.prod(M.SHAPES[k]))
        a, b = g[off:off + size], go[off:off + size]
        assert np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)) >= 0.995, k


def test_parameters_roundtrip_and_init(nbx):
    m = _lib.Model(7)
    p, nm = m.get()
    assert p.shape == (M.NPARAM,)
    # U(-1/sqrt(fan_in), 1/sqrt(fan_in)) per layer
    assert np.abs(p[M.OFFSETS["W1"]:M.OFFSETS["W2"]]).max() <= 1 / np.sqrt(35) + 1e-7
    assert np.abs(p[M.OFFSETS["W2"]:]).max() <= 1 / np.sqrt(128) + 1e-7
    assert np.array_equal(nm[35:70], np.ones(35)) and np.array_equal(nm[70:], np.ones(6))
    q, norm = random_params(8)
    m.set(q, norm)
    p2, nm2 = m.get()
    assert np.array_equal(p2, q.astype(np.float32).astype(np.float64))
    assert np.array_equal(nm2, norm)
    with pytest.raises(FloatingPointError):
        m.set(np.full(M.NPARAM, np.nan))


def test_training_fits_a_smooth_map(nbx):
    print("[case] MLP training 400 iterations", flush=True)
    rng = np.random.default_rng(5)
    A = rng.normal(size=(35, 6)) / np.sqrt(35)
    feats = rng.normal(2.0, 3.0, size=(32768, 35))
    targets = np.tanh(((feats - 2.0) / 3.0) @ A) * 0.01
    m = _lib.Model(11)
    m.add_samples(feats, targets)
    assert m.num_samples == 32768
    m.fit_normalization()
    _, nm = m.get()
    assert np.allclose(nm[:35], feats.mean(0), rtol=1e-5, atol=1e-5)
    assert np.allclose(nm[70:], np.sqrt((targets ** 2).mean(0)), rtol=1e-5)
    mse0, base0 = m.evaluate(0, 32768)
    losses = m.train(400, batch=1024, lr=2e-3, seed=1)
    assert np.all(np.isfinite(losses))
    mse1, base1 = m.evaluate(0, 32768)
    assert np.allclose(base0, base1)
    assert mse1.mean() < 0.1 * mse0.mean(), (mse0, mse1)
    assert losses[-50:].mean() < 0.2 * losses[:10].mean()


def test_features_match_oracle(nbx):
    print("[case] descriptor N=500", flush=True)
    ctx, s = lj_ctx(5, 0.9, seed=3, jitter=0.03)
    f, fr = ctx.features()
    orc = O.System(np.ones(500), bc=O.make_bc(O.BC_CUBIC, s.L)).add(O.TERM_LJ, 1.0, 1.0, 2.5)
    a = orc.rhs(s.u)
    fo, fro = O.lr_features(s.u, a, s.v, s.L)
    assert np.abs(f[:, :32] - fo[:, :32]).max() <= 1e-5 * np.abs(fo[:, :32]).max()
    assert np.allclose(f[:, 32:], fo[:, 32:], rtol=1e-5, atol=1e-5)
    # the frame follows the direction of a (mixed precision here, FP64 in the oracle):
    # its error scales like the force error over |a_i| (and over |v_perp| for e2)
    an = np.linalg.norm(a, axis=1)
    ferr = 5e-6 * an.max()
    tol_e1 = 1e-6 + 4 * ferr / an
    tol_e2 = tol_e1 * (1.0 + np.linalg.norm(s.v, axis=1) / fo[:, 33])
    assert np.all(np.abs(fr[:, :3] - fro[:, :3]).max(1) <= tol_e1)
    assert np.all(np.abs(fr[:, 3:] - fro[:, 3:]).max(1) <= tol_e2)
    assert np.mean(np.abs(fr - fro).max(1) <= 1e-5) > 0.9


def test_collected_targets(nbx):
    print("[case] collect stride 1 and 8", flush=True)
    ctx, s = lj_ctx(5, 0.9, seed=4)
    m = _lib.Model(0)
    # one plain step: the strided step is the step itself, the targets vanish
    ctx.learn_collect(m, 0.005, 1, 4, 100, seed=1)
    f, t = m.get_samples()
    assert f.shape == (400, 35) and np.all(np.isfinite(f))
    assert np.abs(t).max() <= 1e-12
    m.clear_samples()
    ctx.learn_collect(m, 0.005, 8, 20, 200, seed=2)
    f, t = m.get_samples()
    D = 0.04
    rms = np.sqrt((t ** 2).mean(0))
    # third-order local error of the strided step: |dx| ~ D^3 |da/dt| / 6, |dv| ~ D^2
    assert 1e-5 < rms[0] < 3e-2 and 1e-4 < rms[3] < 3e-1, rms
    # the e3 components have no preferred sign (reflection symmetry of the frame)
    assert abs(t[:, 2].mean()) < 0.2 * rms[2] and abs(t[:, 5].mean()) < 0.2 * rms[5]
    assert ctx.steps_taken() == 4 + 160


def test_zero_network_is_plain_vv_at_the_stride(nbx):
    print("[case] learned run with a zero network", flush=True)
    ctx, s = lj_ctx(5, 0.9, seed=6)
    ref, _ = lj_ctx(5, 0.9, seed=6)
    m = _lib.Model(0)
    nm = np.concatenate([np.zeros(35), np.ones(35), np.ones(6)])
    m.set(np.zeros(M.NPARAM), nm)
    rep = ctx.learned_run(m, 0.005, 4, 20, drift_tol=1.0, check_every=5)
    assert rep["strided_steps"] == 20 and not rep["tripped"] and rep["plain_steps"] == 0
    ref.vv_run(0.02, 20)
    u1, v1 = ctx.get_state()
    u2, v2 = ref.get_state()
    # same arithmetic; the Verlet lists are rebuilt at other moments (the strided run
    # rebuilds at every monitor block), which reorders the FP32 pair sums (~1e-7 of a)
    assert np.abs(u1 - u2).max() <= 1e-6
    assert np.abs(v1 - v2).max() <= 1e-5
    ke, pe = ctx.learn_energy()
    assert abs((ke + pe) - rep["E"]) <= 1e-9 * abs(rep["E"])


def test_drift_monitor_falls_back_to_plain_vv(nbx):
    print("[case] learned run fallback", flush=True)
    ctx, s = lj_ctx(5, 0.9, seed=7)
    ref, _ = lj_ctx(5, 0.9, seed=7)
    p, _ = random_params(9)
    m = _lib.Model(0)
    # small random corrections (1e-3 sigma per step): no overlaps, but a clear energy drift
    nm = np.concatenate([np.zeros(35), np.ones(35), np.full(6, 1e-3)])
    m.set(p, nm)
    rep = ctx.learned_run(m, 0.005, 8, 30, drift_tol=1e-4, check_every=10)
    assert rep["tripped"] and rep["strided_before_trip"] == 0
    assert rep["strided_steps"] == 0 and rep["plain_steps"] == 30 * 8
    ref.vv_run(0.005, 240)
    u1, _ = ctx.get_state()
    u2, _ = ref.get_state()
    assert np.abs(u1 - u2).max() <= 1e-6
    assert ctx.steps_taken() == 240


def test_learned_stride_at_the_benchmark_size(nbx):
    """The learned part at LJ N=32,000 (FCC 20^3, rho* 0.8442): collection (targets of the
    stride-1 step vanish), the zero network as plain velocity Verlet at the stride, and the
    fallback to plain steps, each against plain runs of the same context type."""
    print("[case] learned part at N=32000: collect", flush=True)
    ctx, s = lj_ctx(20, 0.9, seed=21, jitter=0.02)
    m = _lib.Model(0)
    ctx.learn_collect(m, 0.005, 1, 3, 256, seed=4)
    f, t = m.get_samples()
    assert f.shape == (768, 35) and np.all(np.isfinite(f))
    assert np.abs(t).max() <= 1e-12
    m.clear_samples()
    ctx.learn_collect(m, 0.005, 8, 4, 256, seed=5)
    f, t = m.get_samples()
    assert f.shape == (1024, 35) and np.all(np.isfinite(t))
    assert ctx.steps_taken() == 3 + 32
    print("[case] learned part at N=32000: zero network", flush=True)
    ctx, s = lj_ctx(20, 0.9, seed=22, jitter=0.02)
    ref, _ = lj_ctx(20, 0.9, seed=22, jitter=0.02)
    nm = np.concatenate([np.zeros(35), np.ones(35), np.ones(6)])
    m.set(np.zeros(M.NPARAM), nm)
    rep = ctx.learned_run(m, 0.005, 4, 20, drift_tol=1.0, check_every=5)
    assert rep["strided_steps"] == 20 and not rep["tripped"] and rep["plain_steps"] == 0
    ref.vv_run(0.02, 20)
    u1, v1 = ctx.get_state()
    u2, v2 = ref.get_state()
    du, dv = np.abs(u1 - u2).max(), np.abs(v1 - v2).max()
    print(f"[case] zero network vs plain VV at 4 dt: |du| {du:.2e} |dv| {dv:.2e}", flush=True)
    # the lists are rebuilt at other moments than in the plain run, which reorders the FP32
    # pair sums (~1e-7 of a); 20 steps of the fluid grow that to well under these bars
    assert du <= 1e-5 and dv <= 1e-4
    ke, pe = ctx.learn_energy()
    assert abs((ke + pe) - rep["E"]) <= 1e-9 * abs(rep["E"])
    print("[case] learned part at N=32000: fallback", flush=True)
    ctx, s = lj_ctx(20, 0.9, seed=23, jitter=0.02)
    ref, _ = lj_ctx(20, 0.9, seed=23, jitter=0.02)
    p, _ = random_params(9)
    m.set(p, np.concatenate([np.zeros(35), np.ones(35), np.full(6, 1e-3)]))
    rep = ctx.learned_run(m, 0.005, 8, 10, drift_tol=1e-4, check_every=10)
    assert rep["tripped"] and rep["strided_before_trip"] == 0
    assert rep["strided_steps"] == 0 and rep["plain_steps"] == 80
    ref.vv_run(0.005, 80)
    u1, _ = ctx.get_state()
    u2, _ = ref.get_state()
    assert np.abs(u1 - u2).max() <= 1e-5
    assert ctx.steps_taken() == 80


def test_trained_network_beats_the_uncorrected_stride(nbx):
    """The network's gain: on held-out trajectories its corrected 8-dt step is closer to
    eight plain steps (the yardstick) than the uncorrected 8-dt step, along the force."""
    print("[case] train on N=864 trajectories, evaluate held out", flush=True)
    m = _lib.Model(3)
    for seed in range(6):
        ctx, _ = lj_ctx(6, 0.9, seed=100 + seed)
        ctx.learn_collect(m, 0.005, 8, 60, 256, seed=seed)
    ntot = m.num_samples
    held = 60 * 256
    assert ntot == 6 * held
    m.fit_normalization()
    # train on the first five trajectories only (rows are in trajectory order)
    feats, targets = m.get_samples(0, ntot)
    train = _lib.Model(3)
    train.add_samples(feats[:-held], targets[:-held])
    train.set(None, m.get()[1])
    train.train(1500, batch=2048, lr=1e-3, seed=5)
    test = _lib.Model(3)
    test.add_samples(feats[-held:], targets[-held:])
    test.set(*train.get())
    mse, base = test.evaluate(0, held)
    print("held-out mse / uncorrected:", np.round(mse / base, 3), flush=True)
    assert mse[0] < 0.85 * base[0] and mse[3] < 0.85 * base[3], (mse, base)
    assert mse.sum() < 0.95 * base.sum()
