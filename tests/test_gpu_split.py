"""The force launch splits its groups over the XCDs by the work the previous
launch over the same groups measured.  Every group is walked by exactly one
wave whatever the split, so repeated launches on one tree -- the first with an
even split, the next ones with measured ones, then an even split again after a
launch over a different number of groups -- give the same accelerations bit for
bit, for the sizes where the split is degenerate (fewer groups than XCDs, a
count not a multiple of the 4-group chunks) as for large ones."""
import numpy as np
import pytest

import nns_amd

pytestmark = pytest.mark.gpu


def _acc(bh):
    return np.stack([bh.host("ax"), bh.host("ay"), bh.host("az")])


@pytest.mark.parametrize("kind,n", [("plummer", 100), ("plummer", 64 * 13 + 5), ("cluster_like", 70001),
                                    ("plummer", 300000)])
def test_split_does_not_change_accelerations(kind, n):
    import torch
    assert torch.cuda.is_available()
    print(f"case split {kind} n={n}", flush=True)
    if kind == "cluster_like":  # a dense core and a sparse halo: uneven groups
        r = np.random.default_rng(5)
        p = np.vstack([r.normal(0, 1e-3, (n // 2, 3)), r.normal(0, 1.0, (n - n // 2, 3))]).astype(np.float32)
        x, y, z = p.T.copy()
        vx = vy = vz = np.zeros(n, np.float32)
        m = np.full(n, 1.0 / n, np.float32)
    else:
        x, y, z, vx, vy, vz, m = nns_amd.initial_conditions(kind, n, 3)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    bh.rebuild()
    bh.forces()
    bh.ctx.check()
    first = _acc(bh)
    for _ in range(3):
        bh.forces()
        bh.ctx.check()
        assert np.array_equal(_acc(bh), first)
    counts = bh.count_interactions().cpu().numpy()
    bh.forces()
    bh.ctx.check()
    assert np.array_equal(_acc(bh), first)
    assert np.array_equal(bh.count_interactions().cpu().numpy(), counts)
    # a launch over another number of groups in the same context resets the split
    k = max(2, n // 3)
    small = nns_amd.BarnesHut(x[:k], y[:k], z[:k], vx[:k], vy[:k], vz[:k], m[:k], ctx=bh.ctx, m=bh.m)
    small.rebuild()
    small.forces()
    small.ctx.check()
    ref_small = nns_amd.BarnesHut(x[:k], y[:k], z[:k], vx[:k], vy[:k], vz[:k], m[:k])
    ref_small.rebuild()
    ref_small.forces()
    ref_small.ctx.check()
    assert np.array_equal(_acc(small), _acc(ref_small))
    bh.rebuild()
    bh.forces()
    bh.ctx.check()
    assert np.array_equal(_acc(bh), first)
