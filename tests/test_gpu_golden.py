"""GPU results against the reference's own outputs (tests/golden/, written by
the reference's OpenMP code compiled in oracle/_ref; see tests/golden/make_golden.py).

Bars:
  keys: bit-exact with Body::discretizePosition + Body::getKey
  (3D/OpenMP/OpenMP_NBody/src/Body.cpp:13-32);
  sorted order: the GPU's stable order, restricted to the last body of each
  key, is the reference's std::map<key, Body*> order -- the order of its
  `hashed` traversal (Interaction.cpp:156-177,190-197);
  accelerations (unsoftened cases): the reference's Tree::insert +
  treeInteraction increments, converted from its SI units to G = 1
  (dv = TIME_STEP G_SI / TO_METERS^2 a).  The GPU takes the opening test in
  FP32 (crit_L < |d|^2), the reference in FP64 (l/D < theta), so a body with
  a cell within rounding of the boundary can meet a different list; such
  bodies are found by comparing the GPU's interaction counts with the
  oracle's walk under the reference's predicate (equal to the reference's
  output to 1e-14).  Bars: at most 1 % of the bodies change lists; on the
  others RMS relative error <= 1e-5; every body within 2e-3.
"""
import numpy as np
import pytest

import nns_amd
from conftest import golden
from oracle import oracle as O

from test_gpu_parity import run_to_forces

pytestmark = pytest.mark.gpu

KEY_CASES = ["cloud", "disk", "dups", "lattice"]
TREE_CASES = ["blob_t05", "disk_t05"]  # the unsoftened ones: the reference's softened kernel is not Plummer


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


@pytest.mark.parametrize("case", KEY_CASES)
def test_keys_and_hashed_order_match_reference(case):
    _torch()
    print(f"case golden keys {case}", flush=True)
    g = golden(f"keys_{case}.npz")
    n = len(g["x"])
    z0 = np.zeros(n, np.float32)
    bh = run_to_forces(g["x"], g["y"], g["z"], z0, z0, z0, np.full(n, 1.0 / n, np.float32))
    keys = bh.ctx.copy_buffer(nns_amd.BUF_KEYS, n, np.uint64)
    assert np.array_equal(keys, g["keys"])
    perm = bh.ctx.copy_buffer(nns_amd.BUF_PERM, n, np.int32)
    sk = keys[perm]
    assert np.all(sk[1:] >= sk[:-1])
    last = np.ones(n, bool)
    last[:-1] = sk[1:] != sk[:-1]
    assert np.array_equal(perm[last], g["map_order"])


@pytest.mark.parametrize("case", TREE_CASES)
def test_accelerations_match_reference_tree_walk(case):
    _torch()
    print(f"case golden tree {case}", flush=True)
    g = golden(f"tree_{case}.npz")
    assert float(g["softening_au"]) == 0.0
    n = len(g["x"])
    z0 = np.zeros(n, np.float32)
    p = nns_amd.default_params(theta=float(g["theta"]), softening=0.0, gravity=1.0)
    bh = run_to_forces(g["x"], g["y"], g["z"], z0, z0, z0, g["m"], params=p)
    ag = np.stack([bh.host("ax"), bh.host("ay"), bh.host("az")], 1).astype(np.float64)
    cnt_gpu = bh.count_interactions().cpu().numpy()
    _, cnt_ref = O.KeyTree(g["x"], g["y"], g["z"], g["m"]).accel(float(g["theta"]), 0.0, mac_mode=0)
    a_ref = g["dv"] / (O.G_SI / O.TO_METERS ** 2)
    norm = np.linalg.norm(a_ref, axis=1)
    assert (norm > 0).all()
    rel = np.linalg.norm(ag - a_ref, axis=1) / norm
    same = cnt_gpu == cnt_ref
    print(f"bodies with another list: {(~same).sum()} of {n}", flush=True)
    assert (~same).sum() <= 0.01 * n
    assert np.sqrt((rel[same] ** 2).mean()) <= 1e-5, np.sqrt((rel[same] ** 2).mean())
    assert rel.max() <= 2e-3, rel.max()
