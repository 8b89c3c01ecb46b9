"""GPU parity of the batched form: an ensemble of independent systems of
uneven size, each with its own theta, stepped by one launch per stage; every
system is checked against the oracle run on that system alone.  The graph
replay and the autotuned launch shape must reproduce the eager step bit for bit.
"""
import numpy as np
import pytest

import nns_amd

pytestmark = pytest.mark.gpu


def ensemble(sizes, seed0=100):
    parts = [nns_amd.initial_conditions("plummer", int(n), seed=seed0 + s) for s, n in enumerate(sizes)]
    return [np.concatenate([p[k] for p in parts]) for k in range(7)]


SIZES = [1024, 3000, 65, 1, 2, 777, 5000, 1500]
THETAS = [0.3, 0.5, 0.7, 0.5, 0.3, 0.7, 0.5, 0.3]


def test_ensemble_step_matches_per_system_oracle(oracle):
    import torch
    assert torch.cuda.is_available()
    x, y, z, vx, vy, vz, m = ensemble(SIZES)
    b = nns_amd.Batch(x, y, z, vx, vy, vz, m, SIZES, THETAS)
    b.step()
    b.check()
    torch.cuda.synchronize()
    counts = b.count_interactions().cpu().numpy()
    b.check()
    ax, ay, az = b.host("ax"), b.host("ay"), b.host("az")
    keys = b.copy_buffer(nns_amd.BUF_KEYS, b.N, np.uint64)
    perm = b.copy_buffer(nns_amd.BUF_PERM, b.N, np.int32)
    ext = b.copy_buffer(nns_amd.BUF_EXTENT, 4 * b.S, np.float64).reshape(b.S, 4)
    eps2 = float(np.float32(np.float32(b.params.softening) ** 2))
    for s, n in enumerate(SIZES):
        lo, hi = b.off[s], b.off[s + 1]
        xs, ys, zs, ms = x[lo:hi], y[lo:hi], z[lo:hi], m[lo:hi]
        _, e = oracle.bbox(xs, ys, zs)
        assert np.array_equal(ext[s], e), s
        assert np.array_equal(keys[lo:hi], oracle.keys_morton(xs, ys, zs)), s
        if n < 2:
            assert np.array_equal(perm[lo:hi], [lo])
            assert ax[lo] == 0 and counts[lo] == 0
            continue
        t = oracle.KeyTree(xs, ys, zs, ms)
        assert np.array_equal(perm[lo:hi] - lo, t.perm), s
        acc, cnt = t.accel(float(np.float32(THETAS[s])), eps2, mac_mode=1)
        assert np.array_equal(counts[lo:hi], cnt), s
        ag = np.stack([ax[lo:hi], ay[lo:hi], az[lo:hi]], 1).astype(np.float64)
        rel = np.linalg.norm(ag - acc, axis=1) / np.linalg.norm(acc, axis=1)
        assert np.sqrt((rel ** 2).mean()) < 1e-5 and rel.max() < 2e-3, (s, rel.max())
    # the update of every system is the symplectic Euler of its own accelerations
    ex = oracle.update(x, y, z, vx, vy, vz, ax, ay, az, b.params.timestep, b.params.dampening)
    for k, e in zip(("x", "y", "z", "vx", "vy", "vz"), ex):
        assert np.array_equal(b.host(k), e), k


def test_graph_replay_and_autotune_match_eager_steps():
    import torch
    sizes = [2048, 4096, 1024, 8192, 3000]
    x, y, z, vx, vy, vz, m = ensemble(sizes, seed0=7)
    eager = nns_amd.Batch(x, y, z, vx, vy, vz, m, sizes, [0.5] * 5)
    graph = nns_amd.Batch(x, y, z, vx, vy, vz, m, sizes, [0.5] * 5)
    stream = torch.cuda.Stream()
    for k in range(4):
        eager.step()
        with torch.cuda.stream(stream):
            graph.graph_step(stream=stream.cuda_stream)
        stream.synchronize()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        for key in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            assert np.array_equal(eager.host(key), graph.host(key)), (k, key)
    eager.check()
    graph.check()
    before = {k: eager.host(k) for k in ("x", "vx", "ax")}
    w = eager.autotune()
    assert w in (2, 4, 8)
    for k, v in before.items():  # the tuner leaves the bound arrays alone
        assert np.array_equal(eager.host(k), v), k
    eager.step()
    graph.step()
    for key in ("x", "vx", "ax"):
        assert np.array_equal(eager.host(key), graph.host(key)), key


def test_single_system_autotune_keeps_results(oracle):
    import torch
    n = 30000
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("plummer", n, seed=9)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    bh.update()
    bh.ctx.check()
    x1, y1, z1 = bh.host("x"), bh.host("y"), bh.host("z")
    bh.stages()  # leaves the tree of (x1, y1, z1) and its forces
    a_before = bh.host("ax")
    w = nns_amd.autotune(bh.ctx)
    assert w in (2, 4, 8)
    a = bh.arr
    bh.kw.computeForces(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"], a["mass"],
                        a["sorted"], a["child"], a["minX"], a["maxX"], n, 1.0)
    torch.cuda.synchronize()
    assert np.array_equal(bh.host("ax"), a_before)
