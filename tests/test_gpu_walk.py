"""The force walk's own work counters against the oracle's count of the same
walk: nbx_walk_stats and nbx_batch_walk_stats report the child steps (one step
tests one child for the 64 lanes of a group) and frames (opened cells) that
the wavefront walk performed, summed over its groups; orc_group_walk counts the
union of the group's per-body opened sets on the oracle's key tree.  Both must
agree exactly, and the interactions must be the per-body counts."""
import numpy as np
import pytest

import nns_amd

pytestmark = pytest.mark.gpu

CASES = [("plummer", 65536, 0.5), ("binary_disk", 40000, 0.5), ("plummer", 30000, 0.8),
         ("star_disk", 20000, 0.3), ("plummer", 100, 0.5)]


@pytest.mark.parametrize("kind,n,theta", CASES)
def test_walk_stats_match_oracle(oracle, kind, n, theta):
    import torch
    assert torch.cuda.is_available()
    print(f"case walk {kind} n={n} theta={theta}", flush=True)
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions(kind, n, 1)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=nns_amd.default_params(theta=theta))
    bh.rebuild()
    w = bh.walk_stats()
    bh.ctx.check()
    t = oracle.KeyTree(x, y, z, m)
    eps2 = float(np.float32(np.float32(bh.ctx.params.softening) ** 2))
    th = float(np.float32(theta))
    steps, frames = t.group_walk(th, eps2)
    _, cnt = t.accel(th, eps2, mac_mode=1)
    assert w["groups"] == len(steps) == (n + 63) // 64
    assert w["child_steps"] == int(steps.sum())
    assert w["frames"] == int(frames.sum())
    assert w["interactions"] == int(cnt.sum())
    assert 0.0 < w["useful_lane_fraction"] <= 1.0


def test_batch_walk_stats_match_oracle(oracle):
    """The ensemble's groups never straddle two systems, so its counters are
    the sums of the per-system walks, each with its own theta."""
    import torch
    assert torch.cuda.is_available()
    sizes = [1024, 3000, 65, 2, 777, 5000]
    thetas = [0.3, 0.5, 0.7, 0.3, 0.7, 0.5]
    print(f"case batch walk sizes={sizes}", flush=True)
    parts = [nns_amd.initial_conditions("plummer", n, seed=200 + s) for s, n in enumerate(sizes)]
    x, y, z, vx, vy, vz, m = [np.concatenate([p[k] for p in parts]) for k in range(7)]
    b = nns_amd.Batch(x, y, z, vx, vy, vz, m, sizes, thetas)
    b.step()
    w = b.walk_stats()
    b.check()
    eps2 = float(np.float32(np.float32(b.params.softening) ** 2))
    steps = frames = inter = groups = 0
    for s, n in enumerate(sizes):
        lo, hi = b.off[s], b.off[s + 1]
        t = oracle.KeyTree(x[lo:hi], y[lo:hi], z[lo:hi], m[lo:hi])
        st, fr = t.group_walk(float(np.float32(thetas[s])), eps2)
        steps += int(st.sum())
        frames += int(fr.sum())
        groups += len(st)
        inter += int(t.accel(float(np.float32(thetas[s])), eps2, mac_mode=1)[1].sum())
    assert (w["groups"], w["child_steps"], w["frames"], w["interactions"]) == (groups, steps, frames, inter)
