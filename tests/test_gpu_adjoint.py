"""GPU parity of the differentiation form (nbx_force_vjp, nns_amd.adjoint).

Bars:
  vector-Jacobian product on the forward tree: against the oracle's FP64 VJP
  of the same interaction lists (orc_tree_vjp, mac_mode 1), error norm
  <= 2e-5 of the gradient norm for dL/dx and dL/dm;
  theta = 0 (every node opened: the direct sum), through BarnesHutAccel and
  through a 16-step Rollout with checkpoints every 8 steps: against torch
  autograd of the same FP64 direct-sum map on the CPU, error norm <= 1e-4 of
  the gradient norm (FP32 trajectory against FP64).
"""
import numpy as np
import pytest

import nns_amd
from oracle import oracle as O

from test_gpu_parity import make_case

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("kind,n,theta", [("plummer", 4096, 0.5), ("star_disk", 20000, 0.5),
                                          ("binary_disk", 30000, 0.7), ("cluster", 9000, 0.5),
                                          ("coincident", 5000, 0.5), ("plummer", 3, 0.5)])
def test_vjp_matches_oracle(kind, n, theta):
    torch = _torch()
    print(f"case vjp {kind} n={n} theta={theta}", flush=True)
    x, y, z, vx, vy, vz, m = make_case(kind, n)
    p = nns_amd.default_params(theta=theta)
    m_nodes = 24 * n + 12000 if kind == "coincident" else None
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=p, m=m_nodes)
    bh.rebuild()
    u = np.random.default_rng(7).normal(size=(n, 3)).astype(np.float32)
    ut = [torch.as_tensor(u[:, k].copy(), device=bh.dev) for k in range(3)]
    gx, gy, gz, gm = [t.cpu().numpy() for t in bh.vjp(*ut)]
    bh.ctx.check()
    cap = 23 * n + 12000 if kind == "coincident" else 0
    t = O.KeyTree(x, y, z, m, cap=cap)
    ogx, ogm = t.vjp(theta, float(p.softening) ** 2, u.astype(np.float64), mac_mode=1)
    g = np.stack([gx, gy, gz], 1)
    assert rel_err(g, ogx) <= 2e-5, rel_err(g, ogx)
    assert rel_err(gm, ogm) <= 2e-5, rel_err(gm, ogm)


@pytest.mark.parametrize("kind,n", [("plummer", 4096), ("coincident", 3000)])
def test_vjp_unsoftened_matches_oracle(kind, n):
    """eps = 0: the walk compiled without softening, where a body at the same
    point as the lane's own is left out (|d|^2 = 0), as in the forward walk."""
    torch = _torch()
    print(f"case vjp unsoftened {kind} n={n}", flush=True)
    x, y, z, vx, vy, vz, m = make_case(kind, n, seed=3)
    p = nns_amd.default_params(theta=0.5, softening=0.0)
    m_nodes = 24 * n + 12000 if kind == "coincident" else None
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=p, m=m_nodes)
    bh.rebuild()
    u = np.random.default_rng(8).normal(size=(n, 3)).astype(np.float32)
    ut = [torch.as_tensor(u[:, k].copy(), device=bh.dev) for k in range(3)]
    gx, gy, gz, gm = [t.cpu().numpy() for t in bh.vjp(*ut)]
    bh.ctx.check()
    cap = 23 * n + 12000 if kind == "coincident" else 0
    t = O.KeyTree(x, y, z, m, cap=cap)
    ogx, ogm = t.vjp(0.5, 0.0, u.astype(np.float64), mac_mode=1)
    g = np.stack([gx, gy, gz], 1)
    assert np.isfinite(g).all() and np.isfinite(gm).all()
    assert rel_err(g, ogx) <= 2e-5, rel_err(g, ogx)
    assert rel_err(gm, ogm) <= 2e-5, rel_err(gm, ogm)


def _direct_accel_t(torch, x, y, z, m, eps2):
    dx = x[None, :] - x[:, None]
    dy = y[None, :] - y[:, None]
    dz = z[None, :] - z[:, None]
    r2 = dx * dx + dy * dy + dz * dz
    s = torch.where(r2 > 0, (r2 + eps2) ** -1.5, torch.zeros_like(r2))
    w = m[None, :] * s
    return (w * dx).sum(1), (w * dy).sum(1), (w * dz).sum(1)


def test_accel_function_theta0_matches_direct_autograd():
    torch = _torch()
    from nns_amd.adjoint import BarnesHutAccel, Slots
    n = 512
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("plummer", n, 3)
    p = nns_amd.default_params(theta=0.0)
    slots = Slots(n, 1, p)
    dev = slots.bh[0].dev
    xs = [torch.tensor(a, device=dev, requires_grad=True) for a in (x, y, z, m)]
    a = BarnesHutAccel.apply(*xs, slots)
    w = torch.as_tensor(np.random.default_rng(2).normal(size=(3, n)).astype(np.float32), device=dev)
    (a[0] * w[0] + a[1] * w[1] + a[2] * w[2]).sum().backward()
    slots.check()
    ref = [torch.tensor(np.asarray(v, np.float64), requires_grad=True) for v in (x, y, z, m)]
    ar = _direct_accel_t(torch, *ref, float(p.softening) ** 2)
    wr = w.cpu().double()
    (ar[0] * wr[0] + ar[1] * wr[1] + ar[2] * wr[2]).sum().backward()
    for got, want in zip(xs, ref):
        assert rel_err(got.grad.cpu().numpy(), want.grad.numpy()) <= 1e-4


def test_rollout_theta0_matches_direct_autograd():
    torch = _torch()
    from nns_amd.adjoint import Slots, rollout
    n, steps, dt, every = 256, 16, 1e-3, 8
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("plummer", n, 5)
    p = nns_amd.default_params(theta=0.0)
    slots = Slots(n, every, p)
    dev = slots.bh[0].dev
    tgt = np.random.default_rng(4).normal(size=(3, n))
    ins = [torch.tensor(a, device=dev, requires_grad=True) for a in (x, y, z, vx, vy, vz, m)]
    out = rollout(slots, *ins, steps=steps, dt=dt, every=every)
    tg = torch.as_tensor(tgt.astype(np.float32), device=dev)
    loss = ((out[0] - tg[0]) ** 2 + (out[1] - tg[1]) ** 2 + (out[2] - tg[2]) ** 2).mean()
    loss.backward()
    slots.check()

    ref = [torch.tensor(np.asarray(a, np.float64), requires_grad=True) for a in (x, y, z, vx, vy, vz, m)]
    X, Y, Z, VX, VY, VZ, M = ref
    eps2 = float(p.softening) ** 2
    for _ in range(steps):
        ax, ay, az = _direct_accel_t(torch, X, Y, Z, M, eps2)
        VX, VY, VZ = VX + dt * ax, VY + dt * ay, VZ + dt * az
        X, Y, Z = X + dt * VX, Y + dt * VY, Z + dt * VZ
    tr = torch.as_tensor(tgt)
    lr = ((X - tr[0]) ** 2 + (Y - tr[1]) ** 2 + (Z - tr[2]) ** 2).mean()
    lr.backward()
    assert abs(loss.item() - lr.item()) <= 1e-5 * abs(lr.item())
    for got, want in zip(ins, ref):
        assert rel_err(got.grad.cpu().numpy(), want.grad.numpy()) <= 1e-4, rel_err(got.grad.cpu().numpy(),
                                                                                    want.grad.numpy())
