"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Runs oracle/_ref/ref_harness (the reference's 3D/OpenMP/OpenMP_NBody sources
compiled by oracle/Makefile) on small seeded inputs and stores inputs and
outputs as .npz:

  keys_*.npz   positions -> Body::getKey keys and the std::map key order
               (3D/OpenMP/OpenMP_NBody/src/Body.cpp:13-32, Interaction.cpp:156-177)
  tree_*.npz   positions, masses, theta, softening, root octant -> the velocity
               increments of Tree::insert + Interaction::treeInteraction in SI
               units (Tree.cpp:71-278, Interaction.cpp:17-98) and the node dump
  stars_*.npz  suns and bodies -> the velocity increments of the star loop
               (Interaction.cpp:111-124) and the state after
               Interaction::updateBodies (:219-255)

Usage: make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def key_inputs():
    r = np.random.default_rng(11)
    cases = {}
    # generic cloud plus the corners of its bounding box and exact copies of extremes
    x = r.normal(0, 1, 1500).astype(np.float32)
    y = r.normal(0, 2, 1500).astype(np.float32)
    z = r.normal(0, 0.5, 1500).astype(np.float32)
    ex = np.array([x.min(), x.max(), x.min(), x.max(), 0.0], np.float32)
    ey = np.array([y.min(), y.max(), y.max(), y.min(), 0.0], np.float32)
    ez = np.array([z.min(), z.max(), z.min(), z.max(), 0.0], np.float32)
    cases["cloud"] = (np.concatenate([x, ex]), np.concatenate([y, ey]), np.concatenate([z, ez]))
    # thin disk: x and y dominate the span, z quantises onto few bins
    rad = r.uniform(0.3, 2.0, 1000); ph = r.uniform(0, 2 * np.pi, 1000)
    cases["disk"] = ((rad * np.cos(ph)).astype(np.float32), (rad * np.sin(ph)).astype(np.float32),
                     r.normal(0, 0.02, 1000).astype(np.float32))
    # duplicates: the map keeps the last body of each key
    bx = r.uniform(-1, 1, 200).astype(np.float32)
    by = r.uniform(-1, 1, 200).astype(np.float32)
    bz = r.uniform(-1, 1, 200).astype(np.float32)
    idx = r.integers(0, 200, 100)
    cases["dups"] = (np.concatenate([bx, bx[idx]]), np.concatenate([by, by[idx]]), np.concatenate([bz, bz[idx]]))
    # lattice: coordinates on exact binary fractions (bin boundaries)
    g = np.arange(8, dtype=np.float32) / 8.0
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    cases["lattice"] = (gx.ravel().copy(), gy.ravel().copy(), gz.ravel().copy())
    return cases


def tree_inputs():
    r = np.random.default_rng(12)
    cases = {}
    n = 1200
    x = r.normal(0, 1, n); y = r.normal(0, 1, n); z = r.normal(0, 0.4, n)
    m = r.uniform(0.5, 2.0, n)
    cases["blob_t05"] = (x, y, z, m, 0.5, 0.0)
    cases["blob_t07_soft"] = (x, y, z, m, 0.7, 0.05)
    n = 1500
    rad = r.uniform(0.3, 2.0, n); ph = r.uniform(0, 2 * np.pi, n)
    x = rad * np.cos(ph); y = rad * np.sin(ph); z = r.normal(0, 0.02, n)
    m = np.full(n, 1e-3)
    m[:2] = 0.5
    x[:2] = [0.1, -0.1]; y[:2] = 0.0; z[:2] = 0.0
    cases["disk_t05"] = (x, y, z, m, 0.5, 0.0)
    return cases


def star_inputs():
    """Suns (m, x, y, z, vx, vy, vz) and bodies for the star loop, with
    positions in AU and velocities in m/s as the reference keeps them; one body
    sits on a sun (zero distance, left out)."""
    r = np.random.default_rng(13)
    cases = {}

    def disk(n, m):
        rad = r.uniform(0.1, 1.0, n); ph = r.uniform(0, 2 * np.pi, n)
        x = rad * np.cos(ph); y = rad * np.sin(ph); z = r.normal(0, 0.01, n)
        v = r.normal(0, 3e4, (3, n))
        return [np.full(n, m) * r.uniform(0.5, 1.5, n), x, y, z, v[0], v[1], v[2]]

    one = [np.array([1.0]), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1), np.zeros(1)]
    b = disk(400, 1e-6)
    for k in range(1, 7):
        b[k][17] = one[k][0] if k < 4 else b[k][17]
    cases["stars_one"] = (one, b, 0.0, 3600)
    two = [np.array([0.5, 0.5]), np.array([0.1, -0.1]), np.zeros(2), np.zeros(2),
           np.zeros(2), np.array([-2.0e4, 2.0e4]), np.zeros(2)]
    b = disk(300, 1e-7)
    cases["stars_two"] = (two, b, 0.0, 3600)
    cases["stars_two_soft"] = (two, b, 0.01 * O.TO_METERS, 3600)
    return cases


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/ref_harness missing: run `make -C oracle ref` first")
    for name, (x, y, z) in key_inputs().items():
        keys, order = O.ref_keys(x, y, z)
        np.savez_compressed(os.path.join(HERE, f"keys_{name}.npz"), x=x, y=y, z=z, keys=keys, map_order=order)
        print(f"keys_{name}: n={len(x)} map={len(order)}")
    for name, (x, y, z, m, theta, soft_au) in tree_inputs().items():
        # float32 positions, as the engine stores them, promoted exactly to double
        x32, y32, z32 = (np.asarray(a, np.float32) for a in (x, y, z))
        m32 = np.asarray(m, np.float32)
        mm, ext = O.bbox(x32, y32, z32)
        lroot = np.ldexp(ext[3] / 2097151.0, 21)  # side of the key grid's root cell
        center = ext[:3] + 0.5 * lroot
        soft_m = soft_au * O.TO_METERS
        dv, nodes = O.ref_tree(x32.astype(np.float64), y32.astype(np.float64), z32.astype(np.float64),
                               m32.astype(np.float64), theta, soft_m, center, lroot)
        np.savez_compressed(os.path.join(HERE, f"tree_{name}.npz"), x=x32, y=y32, z=z32, m=m32,
                            theta=theta, softening_m=soft_m, softening_au=soft_au, center=center,
                            length=lroot, dv=dv, node_level=nodes["level"], node_external=nodes["external"],
                            node_d=nodes["d"])
        print(f"tree_{name}: n={len(x)} nodes={len(nodes)}")
    for name, (suns, bods, soft_m, ts) in star_inputs().items():
        out = O.ref_stars(suns, bods, soft_m, ts)
        keys = ("m", "x", "y", "z", "vx", "vy", "vz")
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), softening_m=soft_m, time_step=ts,
                            **{f"s_{k}": a for k, a in zip(keys, suns)}, **{f"b_{k}": a for k, a in zip(keys, bods)},
                            **out)
        print(f"{name}: suns={len(suns[0])} bodies={len(bods[0])}")


if __name__ == "__main__":
    main()
