"""Pins of the learned-stride oracle (CPU): the numpy MLP against PyTorch autograd and
torch's bfloat16 rounding and Adam, the C descriptor against closed forms and
invariances."""
import math

import numpy as np
import pytest
import torch

from oracle import mlp_oracle as M

# ---------------------------------------------------------------------------------
# MLP
# ---------------------------------------------------------------------------------


def random_model(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.NPARAM)
    for k, off in M.OFFSETS.items():
        fan = M.SHAPES[k][-1] if k.startswith("W") else M.SHAPES["W" + k[1]][1]
        size = int(np.prod(M.SHAPES[k]))
        p[off:off + size] = scale * rng.uniform(-1, 1, size) / math.sqrt(fan)
    norm = np.concatenate([rng.normal(0, 0.5, 35), rng.uniform(0.5, 2.0, 35),
                           rng.uniform(0.5, 2.0, 6)])
    return p, norm


def torch_net(p):
    d = M.unpack(p)
    net = torch.nn.Sequential(torch.nn.Linear(35, 128), torch.nn.SiLU(), torch.nn.Linear(128, 128),
                              torch.nn.SiLU(), torch.nn.Linear(128, 128), torch.nn.SiLU(),
                              torch.nn.Linear(128, 6)).double()
    with torch.no_grad():
        for li, k in zip((0, 2, 4, 6), ("1", "2", "3", "4")):
            net[li].weight.copy_(torch.tensor(d["W" + k]))
            net[li].bias.copy_(torch.tensor(d["b" + k]))
    return net


def test_parameter_layout_counts():
    assert M.OFFSETS["b4"] + 6 == M.NPARAM == 35 * 128 + 128 + 2 * (128 * 128 + 128) + 6 * 128 + 6
    p = np.arange(M.NPARAM, dtype=np.float64)
    assert np.array_equal(M.pack(M.unpack(p)), p)


def test_forward_matches_torch_fp64():
    p, norm = random_model(1)
    feats = np.random.default_rng(2).normal(size=(64, 35))
    y, _ = M.forward(p, norm, feats, rounding=False)
    x = (feats - norm[:35]) * norm[35:70]
    yt = torch_net(p)(torch.tensor(x)).detach().numpy()
    assert np.allclose(y, yt, rtol=1e-12, atol=1e-13)


def test_gradient_matches_torch_autograd():
    p, norm = random_model(3)
    rng = np.random.default_rng(4)
    feats = rng.normal(size=(96, 35))
    targets = rng.normal(size=(96, 6))
    loss, g = M.loss_and_gradient(p, norm, feats, targets, rounding=False)
    net = torch_net(p)
    x = torch.tensor((feats - norm[:35]) * norm[35:70])
    t = torch.tensor(targets / norm[70:76])
    lt = ((net(x) - t) ** 2).mean()
    lt.backward()
    gt = {}
    for li, k in zip((0, 2, 4, 6), ("1", "2", "3", "4")):
        gt["W" + k] = net[li].weight.grad.numpy()
        gt["b" + k] = net[li].bias.grad.numpy()
    assert abs(loss - lt.item()) < 1e-12 * max(1.0, abs(loss))
    assert np.allclose(g, M.pack(gt), rtol=1e-9, atol=1e-14)


def test_gradient_matches_finite_differences():
    p, norm = random_model(5)
    rng = np.random.default_rng(6)
    feats = rng.normal(size=(16, 35))
    targets = rng.normal(size=(16, 6))
    _, g = M.loss_and_gradient(p, norm, feats, targets, rounding=False)
    for idx in (7, M.OFFSETS["b1"] + 3, M.OFFSETS["W2"] + 999, M.OFFSETS["W3"] + 12345,
                M.OFFSETS["W4"] + 500, M.OFFSETS["b4"] + 5):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        lp, _ = M.loss_and_gradient(pp, norm, feats, targets, rounding=False)
        lm, _ = M.loss_and_gradient(pm, norm, feats, targets, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1e-3, abs(g[idx])), (idx, fd, g[idx])


def test_bf16_rounding_matches_torch():
    x = np.random.default_rng(7).normal(size=20000) * np.exp(np.random.default_rng(8).normal(size=20000) * 5)
    x = np.concatenate([x, [1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -0.0, 0.0, 65504.0]])
    ours = M.bf16(x)
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)
    assert M.bf16(1.0 + 2 ** -8) == 1.0              # tie to even
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6  # tie to even (upwards)


def test_rounded_gradient_close_to_exact():
    p, norm = random_model(9)
    rng = np.random.default_rng(10)
    feats = rng.normal(size=(256, 35))
    targets = rng.normal(size=(256, 6))
    _, g = M.loss_and_gradient(p, norm, feats, targets, rounding=False)
    _, gq = M.loss_and_gradient(p, norm, feats, targets, rounding=True)
    rel = np.linalg.norm(gq - g) / np.linalg.norm(g)
    assert rel < 0.05, rel
    # a dropped term (e.g. no SiLU derivative) is far outside that band
    assert np.dot(gq, g) / (np.linalg.norm(gq) * np.linalg.norm(g)) > 0.998


def test_adam_matches_torch():
    rng = np.random.default_rng(11)
    p = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    w = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([w], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(50)
    v = np.zeros(50)
    q = p.copy()
    for t, g in enumerate(grads, start=1):
        opt.zero_grad()
        w.grad = torch.tensor(g)
        opt.step()
        q, m, v = M.adam_step(q, g, m, v, t, 1e-2)
    assert np.allclose(q, w.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_fit_normalization_and_rotation():
    rng = np.random.default_rng(12)
    f = rng.normal(3.0, 2.0, size=(5000, 35))
    t = rng.normal(0.0, 0.5, size=(5000, 6))
    nm = M.fit_normalization(f, t)
    assert np.allclose(nm[:35], f.mean(0))
    assert np.allclose(1.0 / nm[35:70], f.std(0))
    assert np.allclose(nm[70:76], np.sqrt((t * t).mean(0)))
    # rotation back to the lab frame preserves lengths and uses a right-handed frame
    e1 = np.array([[0.0, 0.0, 1.0]])
    e2 = np.array([[1.0, 0.0, 0.0]])
    dx, dv = M.rotate_out(np.array([[1.0, 2.0, 3.0, 0.0, 0.0, 1.0]]), np.hstack([e1, e2]),
                          np.ones(6))
    assert np.allclose(dx, [[2.0, 3.0, 1.0]])   # e3 = e1 x e2 = y axis
    assert np.allclose(dv, [[0.0, 1.0, 0.0]])


# ---------------------------------------------------------------------------------
# Descriptor
# ---------------------------------------------------------------------------------

def centres():
    return 0.8 + 1.7 * np.arange(32) / 31.0


def switch(r):
    return 1.0 if r < 2.0 else (0.5 * (1 + math.cos(math.pi * (r - 2.0) / 0.5)) if r < 2.5 else 0.0)


@pytest.mark.parametrize("r", [0.95, 1.5, 2.2, 2.49, 2.6])
def test_descriptor_two_atoms_closed_form(oracle, r):
    L = 10.0
    u = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]])
    a = np.array([[0.0, 2.0, 0.0], [0.0, -2.0, 0.0]])
    v = np.array([[1.0, 1.0, 0.0], [0.0, 0.0, 3.0]])
    f, fr = oracle.lr_features(u, a, v, L)
    g = np.exp(-0.5 * ((r - centres()) / 0.1) ** 2) * switch(r)
    assert np.allclose(f[0, :32], g, rtol=1e-13, atol=1e-300)
    assert np.allclose(f[1, :32], g, rtol=1e-13, atol=1e-300)
    # atom 0: e1 = +y, v.e1 = 1, normal part (1, 0, 0), |a| = 2
    assert np.allclose(f[0, 32:], [1.0, 1.0, 2.0])
    assert np.allclose(fr[0], [0, 1, 0, 1, 0, 0])
    # atom 1: e1 = -y, v along z entirely normal
    assert np.allclose(f[1, 32:], [0.0, 3.0, 2.0])
    assert np.allclose(fr[1], [0, -1, 0, 0, 0, 1])


def test_descriptor_minimum_image_and_invariances(oracle):
    rng = np.random.default_rng(13)
    L = 6.0
    u = rng.uniform(0, L, size=(40, 3))
    a = rng.normal(size=(40, 3))
    v = rng.normal(size=(40, 3))
    f, fr = oracle.lr_features(u, a, v, L)
    # translation by a lattice vector and a shift (periodic images) change nothing
    f2, fr2 = oracle.lr_features(u + np.array([L, -2 * L, 0.3]), a, v, L)
    assert np.allclose(f, f2, rtol=1e-10, atol=1e-12)
    # a rotation of a, v (with positions rotated about the box centre by a cubic symmetry)
    Rm = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    c = np.array([L / 2] * 3)
    f3, fr3 = oracle.lr_features((u - c) @ Rm.T + c, a @ Rm.T, v @ Rm.T, L)
    assert np.allclose(f, f3, rtol=1e-10, atol=1e-12)
    assert np.allclose(fr3[:, :3], fr[:, :3] @ Rm.T, atol=1e-12)
    # frames are orthonormal
    e1, e2 = fr[:, :3], fr[:, 3:]
    assert np.allclose((e1 * e1).sum(1), 1.0)
    assert np.allclose((e2 * e2).sum(1), 1.0)
    assert np.allclose((e1 * e2).sum(1), 0.0, atol=1e-12)
    # velocity reconstructs from its frame components
    vr = f[:, 32:33] * e1 + f[:, 33:34] * e2
    assert np.allclose(vr, v, atol=1e-12)


def test_descriptor_degenerate_frames(oracle):
    u = np.array([[0.0, 0.0, 0.0], [5.0, 5.0, 5.0]])
    a = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    v = np.array([[0.0, 2.0, 0.0], [0.0, 0.0, -4.0]])
    f, fr = oracle.lr_features(u, a, v, 20.0)
    assert np.allclose(fr[0], [1, 0, 0, 0, 1, 0])      # a = 0: e1 = x
    assert np.allclose(f[1, 32:], [-4.0, 0.0, 1.0])   # v parallel to e1
    assert np.allclose(fr[1], [0, 0, 1, 1, 0, 0])      # e2 from the least aligned axis
