"""CPU study of a conservative learned correction for the strided velocity-Verlet step.

Not part of the test suite (pytest does not collect this file): it backs the learned-
stride design decision with numbers.  Run from the repository root:

    python tests/study_learned_correction.py [strides, default 3,4,6,8] [iters, default 800]

System: the Lennard-Jones benchmark fluid at N = 864 (FCC 6^3, rho* = 0.8442, T* = 0.9
start, cutoff 2.5, truncated force and shifted energy as the reference's
src/basic_potentials.jl:240-272 and src/nbody_simulation_result.jl:293-319), FP64,
minimum image, all pairs (PyTorch on the CPU).

Correction: a per-atom energy E_theta(d_i) of the 32 radial density bins d_i (the
descriptor of the GPU learned stride: Gaussian bins of width 0.1 at 0.8 .. 2.5 with a
cosine switch over 2.0 .. 2.5), MLP 32-128-128-128-1 with SiLU.  The strided step is plain
velocity Verlet of length D = stride * dt on U + sum_i E_theta(d_i): a conservative force
keeps the step symplectic.  Training is force matching (double backpropagation through
the descriptor) against the part of the one-step residual that is even in the velocity,
    T_i = [r_i(x, v) + r_i(x, -v)] / 2,   r_i = 2 (x_i(t + D) - x_i - D v_i - D^2 a_i / 2) / D^2,
with x(t + D) from `stride` plain steps at dt: the odd part, (D / 3) (grad F) . v at
leading order, averages to zero at fixed positions, so no position-only correction can
learn it.

Output per stride: the largest relative deviation of the total (shifted) energy over 10k
dt of simulated time from an equilibrated state, plain and corrected, and the held-out
mean squared error of the correction over that of predicting zero.
"""
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "nbodysimulator.jl_amd"))
from nbodysimulator import generators as gen  # noqa: E402

DT = 0.005
RC = 2.5
R2 = RC * RC
UC = 4 * ((1 / R2) ** 6 - (1 / R2) ** 3)
CENTRES = torch.linspace(0.8, 2.5, 32, dtype=torch.float64)
WIDTH = 0.1


def main():
    strides = [int(a) for a in (sys.argv[1] if len(sys.argv) > 1 else "3,4,6,8").split(",")]
    iters = int(sys.argv[2]) if len(sys.argv) > 2 else 800
    torch.manual_seed(0)
    np.random.seed(0)
    s = gen.lj_liquid(6, rho=0.8442, T=0.9, seed=5)
    L, n = s.L, s.u.shape[0]

    def pairs(x):
        d = x[:, None, :] - x[None, :, :]
        d = d - L * torch.round(d / L)
        r2 = (d * d).sum(-1)
        m = (r2 < R2) & (r2 > 0)
        i, j = m.nonzero(as_tuple=True)
        return i, j, d[i, j], r2[i, j]

    def lj(x):
        i, _, d, r2 = pairs(x)
        s6 = (1 / r2) ** 3
        f = (24 * (2 * s6 * s6 - s6) / r2)[:, None] * d
        return torch.zeros_like(x).index_add_(0, i, f), 0.5 * (4 * (s6 * s6 - s6) - UC).sum()

    def descriptor(i, rij):
        r = rij.norm(dim=-1)
        fc = torch.where(r < 2.0, torch.ones_like(r),
                         0.5 * (1 + torch.cos(math.pi * (r - 2.0) / 0.5)))
        g = torch.exp(-(r[:, None] - CENTRES[None, :]) ** 2 / (2 * WIDTH * WIDTH)) * fc[:, None]
        return torch.zeros(n, 32, dtype=torch.float64).index_add_(0, i, g)

    class Net(torch.nn.Module):
        def __init__(self):
            super().__init__()
            self.l = torch.nn.Sequential(
                torch.nn.Linear(32, 128), torch.nn.SiLU(), torch.nn.Linear(128, 128),
                torch.nn.SiLU(), torch.nn.Linear(128, 128), torch.nn.SiLU(),
                torch.nn.Linear(128, 1)).double()
            self.mu = torch.zeros(32, dtype=torch.float64)
            self.sd = torch.ones(32, dtype=torch.float64)

        def forward(self, d):
            return self.l((d - self.mu) / self.sd)[:, 0]

    def corr_force(net, x, create_graph=False):
        xr = x.detach().requires_grad_(True)
        i, j, _, _ = pairs(x.detach())
        d = xr[i] - xr[j]
        d = d - L * torch.round(d.detach() / L)
        e = net(descriptor(i, d)).sum()
        g, = torch.autograd.grad(e, xr, create_graph=create_graph)
        return -g

    def vv(x, v, F, h, steps, fn):
        for _ in range(steps):
            x = x + h * v + 0.5 * h * h * F
            Fn = fn(x)
            v = v + 0.5 * h * (F + Fn)
            F = Fn
        return x, v, F

    plain = lambda y: lj(y)[0]  # noqa: E731
    x, v = torch.tensor(s.u), torch.tensor(s.v)
    F = plain(x)
    x, v, F = vv(x, v, F, DT, 2000, plain)
    print(f"equilibrated: T = {(v * v).sum().item() / (3 * n):.3f}", flush=True)
    nframes = 200
    X, V, FF = [x], [v], [F]
    for _ in range(nframes + 8):
        x, v, F = vv(x, v, F, DT, 10, plain)
        X.append(x); V.append(v); FF.append(F)
    X, V, FF = torch.stack(X), torch.stack(V), torch.stack(FF)

    def targets(stride, frames):
        D = stride * DT
        out = []
        for f in frames:
            acc = 0
            for sg in (1.0, -1.0):
                x1, _, _ = vv(X[f], sg * V[f], FF[f], DT, stride, plain)
                dx = x1 - (X[f] + D * sg * V[f] + 0.5 * D * D * FF[f])
                dx = dx - L * torch.round(dx / L)
                acc = acc + (2 * dx / (D * D)) / 2
            out.append(acc)
        return torch.stack(out)

    def energy(x, v):
        return 0.5 * (v * v).sum().item() + lj(x)[1].item()

    def run(net, stride, x0, v0, tol=1e-4):
        D = stride * DT
        fn = plain if net is None else (lambda y: plain(y) + corr_force(net, y).detach())
        x, v = x0, v0
        F = fn(x)
        e0 = energy(x, v)
        maxdev, done, total = 0.0, 0, int(round(10000 * DT / D))
        ck = max(1, 200 // stride)
        while done < total:
            m = min(ck, total - done)
            x, v, F = vv(x, v, F, D, m, fn)
            done += m
            dev = abs(energy(x, v) - e0) / abs(e0)
            maxdev = max(maxdev, dev if math.isfinite(dev) else math.inf)
            if not maxdev < 100 * tol:
                break
        return done, total, maxdev

    xs, vs = X[-1], V[-1]
    for stride in strides:
        done, total, dev = run(None, stride, xs, vs)
        print(f"stride {stride} plain: {done}/{total} steps, max rel. deviation {dev:.2e}",
              flush=True)
        frames = list(range(0, nframes, 2))
        T = targets(stride, frames)
        net = Net()
        with torch.no_grad():
            ds = torch.cat([descriptor(pairs(X[f])[0], pairs(X[f])[2]) for f in frames[:20]])
            net.mu, net.sd = ds.mean(0), ds.std(0) + 1e-3
        opt = torch.optim.Adam(net.parameters(), lr=1e-3)
        sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, iters)
        t0 = time.time()
        for _ in range(iters):
            loss = 0
            for q in np.random.randint(0, len(frames), 4):
                loss = loss + (corr_force(net, X[frames[q]], True) - T[q]).pow(2).mean() / 4
            opt.zero_grad()
            loss.backward()
            opt.step()
            sched.step()
        held = list(range(1, nframes, 50))
        Th = targets(stride, held)
        mse = sum((corr_force(net, X[f]).detach() - Th[k]).pow(2).mean().item()
                  for k, f in enumerate(held)) / len(held)
        done, total, dev = run(net, stride, xs, vs)
        print(f"stride {stride} corrected: {done}/{total} steps, max rel. deviation {dev:.2e}; "
              f"held-out mse / zero {mse / Th.pow(2).mean().item():.3f} "
              f"(training {time.time() - t0:.0f} s)", flush=True)


if __name__ == "__main__":
    main()
