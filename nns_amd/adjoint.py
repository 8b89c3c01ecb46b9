"""Differentiable Barnes-Hut over the C-ABI.

* ``BarnesHutAccel`` -- torch.autograd.Function: accelerations of bodies at
  (x, y, z) with masses m; its backward is ``nbx_force_vjp`` on the tree of the
  same positions (rebuilt if the context has walked another tree since).
* ``Rollout`` -- K steps of the reference's integrator (v += dt a; x += d dt v,
  3D/CUDA/CUDA_NBody/src/Kernels.cu:581-600, here with d = 1) as one autograd
  node.  Its backward is the discrete adjoint of that map,
      lv' = lv + dt lx,   lx <- lx + dt (dA/dx)^T lv',   lm <- lm + dt (dA/dm)^T lv',
  taken step by step from the last state to the first.  States are
  checkpointed every ``every`` steps; the backward re-runs one segment from its
  checkpoint with one context per step, so that every step's tree is still
  held when its vector-Jacobian product is taken (no second build).  The last
  step of a re-run segment only builds its tree: the state after it is the
  next checkpoint (or the final state), already known.  The forward's last
  segment is not re-run at all: its trees are still in the contexts (each
  context remembers which rollout step its tree belongs to).

No counterpart in the reference; the force and update arithmetic is that of
the plain method.
"""
from __future__ import annotations

import numpy as np
import torch

from . import BarnesHut, Params, default_params


class Slots:
    """`count` independent BarnesHut contexts for the same n (one per step of a
    checkpoint segment)."""

    def __init__(self, n: int, count: int, params: Params | None = None, device: int = 0):
        self.n = int(n)
        self.params = params if params is not None else default_params()
        z = np.zeros(self.n, np.float32)
        self.bh = [BarnesHut(z, z, z, z, z, z, z, params=self.params, device=device) for _ in range(count)]
        self.tag = [None] * count  # what slot k's tree belongs to (None: unknown)

    def tree(self, k: int, x, y, z, m, tag=None):
        """Builds slot k's tree for (x, y, z, m), remembered under `tag`."""
        bh, n = self.bh[k], self.n
        for key, t in (("x", x), ("y", y), ("z", z), ("mass", m)):
            bh.arr[key][:n].copy_(t)
        bh.rebuild()
        self.tag[k] = tag

    def accel(self, k: int, x, y, z, m, tag=None):
        """Builds slot k's tree for (x, y, z, m) and returns its accelerations (new tensors)."""
        bh, n = self.bh[k], self.n
        self.tree(k, x, y, z, m, tag)
        bh.forces()
        return bh.arr["ax"][:n].clone(), bh.arr["ay"][:n].clone(), bh.arr["az"][:n].clone()

    def vjp(self, k: int, ux, uy, uz):
        """(dL/dx, dL/dy, dL/dz, dL/dm) over slot k's current tree."""
        return self.bh[k].vjp(ux, uy, uz, g=self.params.gravity)

    def check(self):
        for bh in self.bh:
            bh.ctx.check()


class BarnesHutAccel(torch.autograd.Function):
    """a(x, y, z, m) through slot 0 of `slots`; backward through nbx_force_vjp."""

    @staticmethod
    def forward(ctx, x, y, z, m, slots: Slots):
        ax, ay, az = slots.accel(0, x, y, z, m)
        ctx.slots = slots
        ctx.save_for_backward(x, y, z, m)
        return ax, ay, az

    @staticmethod
    def backward(ctx, gax, gay, gaz):
        x, y, z, m = ctx.saved_tensors
        slots = ctx.slots
        zero = torch.zeros_like(x)
        u = [g.contiguous() if g is not None else zero for g in (gax, gay, gaz)]
        slots.tree(0, x, y, z, m)  # the tree of these positions
        gx, gy, gz, gm = slots.vjp(0, *u)
        return gx, gy, gz, gm, None


def _step(slots: Slots, k: int, s, m, dt: float, tag=None):
    x, y, z, vx, vy, vz = s
    ax, ay, az = slots.accel(k, x, y, z, m, tag)
    vx = torch.add(vx, ax, alpha=dt)
    vy = torch.add(vy, ay, alpha=dt)
    vz = torch.add(vz, az, alpha=dt)
    return (torch.add(x, vx, alpha=dt), torch.add(y, vy, alpha=dt), torch.add(z, vz, alpha=dt), vx, vy, vz)


class Rollout(torch.autograd.Function):
    """Final state (x, y, z, vx, vy, vz) after `steps` steps from the initial
    state and masses; gradients w.r.t. all seven inputs."""

    @staticmethod
    def forward(ctx, x, y, z, vx, vy, vz, m, slots: Slots, steps: int, dt: float, every: int):
        s = (x, y, z, vx, vy, vz)
        ckpt = []
        run = object()  # names this rollout's trees in the slots
        for k in range(steps):
            if k % every == 0:
                ckpt.append(tuple(t.clone() for t in s))
            s = _step(slots, k % every, s, m, dt, (run, k))
        ctx.slots, ctx.steps, ctx.dt, ctx.every, ctx.run = slots, steps, dt, every, run
        ctx.ckpt = ckpt
        ctx.save_for_backward(m)
        return tuple(t.clone() for t in s)

    @staticmethod
    def backward(ctx, gx, gy, gz, gvx, gvy, gvz):
        (m,) = ctx.saved_tensors
        slots, steps, dt, every = ctx.slots, ctx.steps, ctx.dt, ctx.every
        zero = torch.zeros_like(m)
        lx, ly, lz, lvx, lvy, lvz = [g if g is not None else zero for g in (gx, gy, gz, gvx, gvy, gvz)]
        lm = torch.zeros_like(m)
        nseg = (steps + every - 1) // every
        for seg in reversed(range(nseg)):
            k0 = seg * every
            k1 = min(steps, k0 + every)
            if any(slots.tag[k - k0] != (ctx.run, k) for k in range(k0, k1)):
                # re-run the segment: slot j keeps the tree of state k0 + j; the
                # last state's tree is all its step needs (the state after it is known)
                s = ctx.ckpt[seg]
                for k in range(k0, k1 - 1):
                    s = _step(slots, k - k0, s, m, dt, (ctx.run, k))
                slots.tree(k1 - 1 - k0, s[0], s[1], s[2], m, (ctx.run, k1 - 1))
            for k in reversed(range(k0, k1)):
                # x_{k+1} = x_k + dt v_{k+1}, v_{k+1} = v_k + dt a(x_k)
                lvx = torch.add(lvx, lx, alpha=dt)
                lvy = torch.add(lvy, ly, alpha=dt)
                lvz = torch.add(lvz, lz, alpha=dt)
                ax, ay, az, am = slots.vjp(k - k0, lvx, lvy, lvz)
                lx = torch.add(lx, ax, alpha=dt)
                ly = torch.add(ly, ay, alpha=dt)
                lz = torch.add(lz, az, alpha=dt)
                lm = torch.add(lm, am, alpha=dt)
        return lx, ly, lz, lvx, lvy, lvz, lm, None, None, None, None


def rollout(slots: Slots, x, y, z, vx, vy, vz, m, steps: int, dt: float, every: int = 8):
    return Rollout.apply(x, y, z, vx, vy, vz, m, slots, steps, dt, every)
