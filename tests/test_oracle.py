"""Pins the CPU oracle (oracle/nbx_oracle.c) before it is trusted as the GPU checker.

Every function of the oracle is checked against at least one of: the golden
vectors written by the reference's own compiled code (tests/golden/), a
closed form, an invariant, or brute force on a tiny input.
"""
import numpy as np
import pytest

from conftest import golden

KEY_CASES = ["cloud", "disk", "dups", "lattice"]
TREE_CASES = ["blob_t05", "blob_t07_soft", "disk_t05"]


# ------------------------------------------------------------------ keys
@pytest.mark.parametrize("case", KEY_CASES)
def test_keys_match_reference(oracle, case):
    g = golden(f"keys_{case}.npz")
    k = oracle.keys_morton(g["x"], g["y"], g["z"])
    assert np.array_equal(k, g["keys"])


@pytest.mark.parametrize("case", KEY_CASES)
def test_stable_order_matches_reference_map(oracle, case):
    """std::map<key, Body*> keeps one body per key, the last one inserted
    (Interaction.cpp:166); the stable order restricted to the last body of each
    key must equal the reference's map order."""
    g = golden(f"keys_{case}.npz")
    perm, sk = oracle.sort_stable(g["keys"])
    assert np.all(sk[1:] >= sk[:-1])
    last = np.ones(len(sk), bool)
    last[:-1] = sk[1:] != sk[:-1]
    assert np.array_equal(perm[last], g["map_order"])


def test_key_corners_and_axes(oracle):
    ext = np.array([0.0, 0.0, 0.0, 1.0])
    k = lambda x, y, z: int(oracle.lib().orc_morton_key(x, y, z, ext))
    assert k(0, 0, 0) == 1 << 63
    assert k(1, 1, 1) == (1 << 64) - 1
    # one axis at the maximum: exactly the bits of that axis (x -> 3j, y -> 3j+1, z -> 3j+2)
    xs = sum(1 << (3 * j) for j in range(21))
    assert k(1, 0, 0) == (1 << 63) | xs
    assert k(0, 1, 0) == (1 << 63) | (xs << 1)
    assert k(0, 0, 1) == (1 << 63) | (xs << 2)
    # upper half of the root on x sets digit 0 (bits 60..62) to 1
    assert (k(0.75, 0.1, 0.1) >> 60) & 7 == 1
    assert (k(0.1, 0.75, 0.1) >> 60) & 7 == 2
    assert (k(0.1, 0.1, 0.75) >> 60) & 7 == 4


def test_keys_monotone_along_axis(oracle):
    x = np.linspace(0, 1, 1001).astype(np.float32)
    z0 = np.zeros_like(x)
    k = oracle.keys_morton(x, z0, z0)
    assert np.all(np.diff(k.astype(object)) >= 0)


def _lebesgue(ix, iy, iz, L):
    key = 0
    for j in range(L):
        key |= (((ix >> j) & 1) | (((iy >> j) & 1) << 1) | (((iz >> j) & 1) << 2)) << (3 * j)
    return key << (3 * (21 - L))


@pytest.mark.parametrize("L", [1, 2, 3])
def test_hilbert_is_a_face_adjacent_bijection(oracle, L):
    """Lebesgue2Hilbert (SFC/src/Tree.cpp:82-104) restricted to level L must be
    a bijection onto 0..8^L-1 whose consecutive indices are face neighbours."""
    side = 1 << L
    cells = {}
    for ix in range(side):
        for iy in range(side):
            for iz in range(side):
                h = oracle.lebesgue2hilbert(_lebesgue(ix, iy, iz, L), L) >> (3 * (21 - L))
                cells[h] = (ix, iy, iz)
    assert sorted(cells) == list(range(side ** 3))
    for h in range(side ** 3 - 1):
        a, b = np.array(cells[h]), np.array(cells[h + 1])
        assert np.abs(a - b).sum() == 1, (h, cells[h], cells[h + 1])


def test_hilbert_full_keys_order_consistent(oracle):
    """Hilbert keys of full-depth Morton keys are distinct for distinct cells and
    keep whole octants contiguous (a level-1 prefix maps to a level-1 prefix)."""
    r = np.random.default_rng(3)
    x, y, z = (r.uniform(0, 1, 4000).astype(np.float32) for _ in range(3))
    mk = oracle.keys_morton(x, y, z)
    hk = oracle.keys_hilbert(mk)
    assert len(np.unique(hk)) == len(np.unique(mk))
    top_m = (mk >> np.uint64(60)) & np.uint64(7)
    top_h = hk >> np.uint64(60)
    for d in range(8):
        assert len(np.unique(top_h[top_m == d])) <= 1


def test_sort_stable_equals_numpy(oracle):
    r = np.random.default_rng(5)
    keys = (r.integers(0, 50, 20000).astype(np.uint64) << np.uint64(40)) | np.uint64(1 << 63)
    perm, sk = oracle.sort_stable(keys)
    ref = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, ref.astype(np.int32))
    assert np.array_equal(sk, keys[ref])


# ------------------------------------------------ reference tree restatement
@pytest.mark.parametrize("case", TREE_CASES)
def test_geo_tree_matches_reference_bit_for_bit(oracle, case):
    g = golden(f"tree_{case}.npz")
    t = oracle.GeoTree(g["x"], g["y"], g["z"], g["m"], g["center"], float(g["length"]))
    assert t.err == 0
    dv = t.accel_reference_si(float(g["theta"]), float(g["softening_m"]))
    assert np.array_equal(dv, g["dv"])
    # node sets: level, geometry, mass and centre of mass, exactly
    ours = np.concatenate([t.ints[:, :1].astype(np.float64), t.dbl], axis=1)
    ref = np.concatenate([g["node_level"][:, None].astype(np.float64), g["node_d"]], axis=1)
    o1 = np.lexsort(ours[:, 4::-1].T)
    o2 = np.lexsort(ref[:, 4::-1].T)
    assert np.array_equal(ours[o1], ref[o2])


def _key_cells_geometry(t):
    """Level and geometric centre of every key-tree cell from its key prefix."""
    lvl = t.level
    key = t.skeys[t.start]
    l = t.lcell0 / (2.0 ** lvl)
    out = np.zeros((t.ncell, 3))
    for c in range(t.ncell):
        L = int(lvl[c])
        pre = int(key[c]) >> (63 - 3 * L) if L else 0
        ix = iy = iz = 0
        for j in range(L):
            d = (pre >> (3 * j)) & 7
            ix |= (d & 1) << j; iy |= ((d >> 1) & 1) << j; iz |= ((d >> 2) & 1) << j
        out[c] = t.ext[:3] + (np.array([ix, iy, iz]) + 0.5) * l[c]
    return out


@pytest.mark.parametrize("case", TREE_CASES)
def test_key_tree_is_the_reference_tree(oracle, case):
    """With the root octant placed on the key grid's root cell, the reference's
    midpoint octree and the key-prefix tree have the same cells, masses and
    centres of mass, and the same Barnes-Hut accelerations."""
    g = golden(f"tree_{case}.npz")
    kt = oracle.KeyTree(g["x"], g["y"], g["z"], g["m"])
    gt = oracle.GeoTree(g["x"], g["y"], g["z"], g["m"], g["center"], float(g["length"]))
    assert kt.err == 0 and gt.err == 0
    internal = gt.ints[:, 1] == 0
    assert internal.sum() == kt.ncell
    geo = np.concatenate([gt.ints[internal, :1], gt.dbl[internal][:, [0, 1, 2, 4, 5, 6, 7]]], axis=1)
    cen = _key_cells_geometry(kt)
    key = np.concatenate([kt.level[:, None], cen, kt.mass[:, None], kt.com], axis=1)
    o1 = np.lexsort(np.round(geo[:, 3::-1].T, 9))
    o2 = np.lexsort(np.round(key[:, 3::-1].T, 9))
    assert np.array_equal(geo[o1, 0], key[o2, 0])
    np.testing.assert_allclose(key[o2, 1:4], geo[o1, 1:4], rtol=0, atol=1e-12 * float(g["length"]))
    np.testing.assert_allclose(key[o2, 4], geo[o1, 4], rtol=1e-13)
    np.testing.assert_allclose(key[o2, 5:], geo[o1, 5:], rtol=0, atol=1e-13 * float(g["length"]))
    eps2 = float(g["softening_au"]) ** 2
    a_key, _ = kt.accel(float(g["theta"]), eps2, mac_mode=0)
    a_geo = gt.accel_plummer(float(g["theta"]), eps2)
    rel = np.linalg.norm(a_key - a_geo, axis=1) / np.linalg.norm(a_geo, axis=1)
    assert rel.max() < 1e-11


@pytest.mark.parametrize("case", TREE_CASES)
def test_reference_si_units_equal_plummer_g1(oracle, case):
    """With zero softening the reference's SI increment is TIME_STEP*G/TO_METERS^2
    times the G = 1 acceleration of the same tree walk (Interaction.cpp:17-55)."""
    g = golden(f"tree_{case}.npz")
    if float(g["softening_au"]) != 0.0:
        pytest.skip("softened reference kernel is d(d^2+eps^2), not Plummer")
    gt = oracle.GeoTree(g["x"], g["y"], g["z"], g["m"], g["center"], float(g["length"]))
    a = gt.accel_plummer(float(g["theta"]), 0.0)
    scale = oracle.G_SI / oracle.TO_METERS ** 2
    np.testing.assert_allclose(g["dv"], a * scale, rtol=1e-12, atol=0)


STAR_CASES = ["stars_one", "stars_two", "stars_two_soft"]


def _star_case(g):
    keys = ("m", "x", "y", "z", "vx", "vy", "vz")
    return [g[f"s_{k}"] for k in keys], [g[f"b_{k}"] for k in keys]


@pytest.mark.parametrize("case", STAR_CASES)
def test_star_forces_match_reference_bit_for_bit(oracle, case):
    """The suns (bodies outside the tree, Interaction.cpp:111-124) and the
    update (:219-255) restated: the reference's own velocity increments and
    state after Interaction::updateBodies, bit for bit (one sun with a body on
    it, a binary, a softened binary)."""
    g = golden(f"{case}.npz")
    suns, bods = _star_case(g)
    ts, soft = float(g["time_step"]), float(g["softening_m"])
    sa, ba = oracle.star_forces(suns[:4], bods[:4], kernel=0, softening_m=soft, time_step=ts)
    assert np.array_equal(sa, g["sun_dv"]) and np.array_equal(ba, g["body_dv"])
    sx = oracle.update_reference(*suns[1:], sa, ts)
    bx = oracle.update_reference(*bods[1:], ba, ts)
    assert np.array_equal(np.stack(sx[:3], 1), g["sun_x"]) and np.array_equal(np.stack(sx[3:], 1), g["sun_v"])
    assert np.array_equal(np.stack(bx[:3], 1), g["body_x"]) and np.array_equal(np.stack(bx[3:], 1), g["body_v"])


@pytest.mark.parametrize("case", ["stars_one", "stars_two"])
def test_star_si_units_equal_plummer_g1(oracle, case):
    """With zero softening the reference's SI increments of the star loop are
    TIME_STEP * G / TO_METERS^2 times the G = 1 Plummer accelerations."""
    g = golden(f"{case}.npz")
    suns, bods = _star_case(g)
    sa, ba = oracle.star_forces(suns[:4], bods[:4], kernel=1, eps2=0.0)
    scale = float(g["time_step"]) * oracle.G_SI / oracle.TO_METERS ** 2
    np.testing.assert_allclose(g["sun_dv"], sa * scale, rtol=1e-12, atol=0)
    np.testing.assert_allclose(g["body_dv"], ba * scale, rtol=1e-12, atol=0)


def test_star_forces_closed_forms(oracle):
    """One sun and one body: a_b = m_s d (d^2 + eps^2)^-3/2 towards the sun and
    the opposite pull on the sun; for many suns and bodies the pair forces
    cancel (sum of m a is zero) and a body on a sun feels nothing from it."""
    eps2 = 0.01
    a_s, a_b = oracle.star_forces(([2.0], [0.0], [0.0], [0.0]), ([3.0], [1.0], [0.0], [0.0]), eps2=eps2)
    f = 1.0 / (1.0 + eps2) ** 1.5
    assert a_b[0] == pytest.approx([-2.0 * f, 0.0, 0.0], rel=1e-15)
    assert a_s[0] == pytest.approx([3.0 * f, 0.0, 0.0], rel=1e-15)
    r = np.random.default_rng(44)
    sm, bm = r.uniform(0.5, 1.0, 3), r.uniform(1e-6, 1e-5, 500)
    sp, bp = r.normal(0, 0.2, (3, 3)), r.normal(0, 1, (3, 500))
    bp[:, 5] = sp[:, 1]
    a_s, a_b = oracle.star_forces((sm, *sp), (bm, *bp), eps2=eps2)
    mom = (sm[:, None] * a_s).sum(0) + (bm[:, None] * a_b).sum(0)
    assert np.abs(mom).max() <= 1e-14 * (sm[:, None] * np.abs(a_s)).sum()
    d = sp[:, [0, 2]] - bp[:, [5]]  # body 5 sits on sun 1: only suns 0 and 2 pull it
    w = sm[[0, 2]] / ((d * d).sum(0) + eps2) ** 1.5
    np.testing.assert_allclose(a_b[5], (d * w).sum(1), rtol=1e-13)


def test_key_tree_invariants_brute_force(oracle):
    r = np.random.default_rng(7)
    n = 700
    x, y, z = (r.normal(0, 1, n).astype(np.float32) for _ in range(3))
    m = r.uniform(0.1, 1.0, n).astype(np.float32)
    t = oracle.KeyTree(x, y, z, m)
    assert t.err == 0
    assert t.count[0] == n and t.start[0] == 0 and t.level[0] == 0
    np.testing.assert_allclose(t.mass[0], m.astype(np.float64).sum(), rtol=1e-14)
    com = np.array([(m.astype(np.float64) * a.astype(np.float64)).sum() for a in (x, y, z)]) / t.mass[0]
    np.testing.assert_allclose(t.com[0], com, rtol=1e-13, atol=1e-15)
    keys = oracle.keys_morton(x, y, z)
    for c in range(t.ncell):
        L = int(t.level[c])
        members = t.perm[t.start[c]:t.start[c] + t.count[c]]
        pre = keys[members] >> np.uint64(63 - 3 * L) if L else np.zeros(len(members), np.uint64)
        assert len(np.unique(pre)) == 1                      # one prefix per cell
        full = keys >> np.uint64(63 - 3 * L) if L else np.zeros(n, np.uint64)
        assert (full == pre[0]).sum() == t.count[c]           # and every body with it
        mb = m[members].astype(np.float64)
        np.testing.assert_allclose(t.mass[c], mb.sum(), rtol=1e-14)
        for ch in t.child[c]:
            if ch <= -2:
                assert t.level[-2 - ch] == L + 1 and t.parent[-2 - ch] == c


def test_theta_zero_is_direct_sum(oracle):
    r = np.random.default_rng(8)
    n = 600
    x, y, z = (r.normal(0, 1, n).astype(np.float32) for _ in range(3))
    m = r.uniform(0.1, 1.0, n).astype(np.float32)
    t = oracle.KeyTree(x, y, z, m)
    a, cnt = t.accel(1e-12, 1e-4, mac_mode=0)
    d = oracle.direct_accel(x, y, z, m, 1e-4)
    np.testing.assert_allclose(a, d, rtol=1e-12, atol=1e-12 * np.abs(d).max())
    assert np.all(cnt == n - 1)


def test_device_predicate_mode_agrees(oracle):
    """The FP32 opening predicate of the GPU changes the interaction list only at
    ties; accelerations of the two modes agree to FP32 resolution."""
    r = np.random.default_rng(9)
    n = 3000
    x, y, z = (r.normal(0, 1, n).astype(np.float32) for _ in range(3))
    m = np.full(n, 1.0 / n, np.float32)
    t = oracle.KeyTree(x, y, z, m)
    a0, c0 = t.accel(0.5, 1e-6, mac_mode=0)
    a1, c1 = t.accel(0.5, 1e-6, mac_mode=1)
    rel = np.linalg.norm(a1 - a0, axis=1) / np.linalg.norm(a0, axis=1)
    assert rel.max() < 1e-5
    assert (c0 != c1).mean() < 1e-2


# ------------------------------------------------------- group walk work
def _group_case(oracle, seed, n):
    r = np.random.default_rng(seed)
    x, y, z = (r.normal(0, 1, n).astype(np.float32) for _ in range(3))
    x[: n // 4] *= 0.05  # a dense clump: deep cells and one-child chains
    m = r.uniform(0.1, 1.0, n).astype(np.float32)
    return oracle.KeyTree(x, y, z, m)


@pytest.mark.parametrize("theta", [0.3, 0.5])
def test_group_walk_of_one_body_is_its_own_walk(oracle, theta):
    """A group of one body tests exactly the children of its own walk, each of
    them accepted (an interaction, or the body itself) or opened (a frame):
    steps = interactions + frames + 1, the 1 being the body's own leaf, which it
    always reaches for theta < 1/sqrt(3) (a cell holding the body is within
    sqrt(3) of its side of the body, so it is never accepted)."""
    t = _group_case(oracle, 11, 2500)
    eps2 = float(np.float32(1e-3) ** 2)
    order = t.group_order()
    steps, frames = t.group_walk(theta, eps2, order=order, group=1)
    _, cnt = t.accel(theta, eps2, mac_mode=1, targets=order)
    assert (frames > 0).all()
    assert np.array_equal(steps, cnt + frames + 1)


def test_group_walk_wide_theta_and_unions(oracle):
    """At theta = 1.2 a body may accept a cell that holds it and never reach its
    own leaf (steps = interactions + frames + {0, 1}).  A group's frames are the
    union of its bodies' opened cells: no fewer than any one body's, no more
    than their sum; a group of one body repeated is that body's walk."""
    t = _group_case(oracle, 12, 3000)
    eps2 = float(np.float32(1e-3) ** 2)
    order = t.group_order()
    s1, f1 = t.group_walk(1.2, eps2, order=order, group=1)
    _, cnt = t.accel(1.2, eps2, mac_mode=1, targets=order)
    d = s1 - cnt - f1
    assert set(np.unique(d)) <= {0, 1} and (d == 0).any()
    s64, f64 = t.group_walk(1.2, eps2, order=order, group=64)
    for g in range(len(f64)):
        fg = f1[64 * g:64 * g + 64]
        assert fg.max() <= f64[g] <= fg.sum()
        assert s1[64 * g:64 * g + 64].max() <= s64[g]
    assert (f64 < np.add.reduceat(f1, np.arange(0, len(f1), 64))).any()  # shared frames counted once
    rep = np.repeat(order[:40], 64)
    sr, fr = t.group_walk(1.2, eps2, order=rep, group=64)
    assert np.array_equal(sr, s1[:40]) and np.array_equal(fr, f1[:40])
    # the default group order is key order re-sorted by the top ten Hilbert digits
    top = oracle.keys_hilbert(t.skeys & np.uint64((1 << 63) - 1)) >> np.uint64(33)
    assert np.all(np.diff(top[np.argsort(t.perm)][order].astype(np.int64)) >= 0)
    with pytest.raises(ValueError):
        oracle.KeyTree(*(np.zeros(1, np.float32) for _ in range(4))).group_walk(0.5, eps2)


# ------------------------------------------------------------ closed forms
def test_direct_two_body_and_cube(oracle):
    eps = 0.1
    x = np.array([0.0, 2.0], np.float32); y = np.zeros(2, np.float32); z = np.zeros(2, np.float32)
    m = np.array([1.0, 3.0], np.float32)
    a = oracle.direct_accel(x, y, z, m, eps * eps)
    assert a[0, 0] == pytest.approx(3.0 * 2.0 / (4.0 + eps * eps) ** 1.5, rel=1e-14)
    assert a[1, 0] == pytest.approx(-1.0 * 2.0 / (4.0 + eps * eps) ** 1.5, rel=1e-14)
    # body at the centre of a cube of equal masses feels no force
    c = np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)] + [[0, 0, 0]], np.float32)
    mm = np.ones(9, np.float32)
    a = oracle.direct_accel(c[:, 0], c[:, 1], c[:, 2], mm, 0.0)
    assert np.abs(a[8]).max() < 1e-15


def _abs_term_sum(x, y, z, m, eps2):
    """sum_j |m_j d_ij| (|d_ij|^2 + eps^2)^-3/2 for every body i, in numpy FP64:
    the size of the direct sum without cancellation, the scale that the rounding
    of any summation order is bounded by."""
    p = np.stack([x, y, z], 1).astype(np.float64)
    mm = m.astype(np.float64)
    out = np.zeros(len(x))
    for lo in range(0, len(x), 500):
        d = p[None, :, :] - p[lo:lo + 500, None, :]
        d2 = (d * d).sum(-1)
        w = np.where(d2 > 0, mm[None, :] / np.where(d2 > 0, (d2 + eps2) ** 1.5, 1.0), 0.0)
        out[lo:lo + 500] = (w * np.sqrt(d2)).sum(1)
    return out


@pytest.mark.parametrize("eps2", [0.0, 1e-6])
def test_direct_fast_equals_direct(oracle, eps2):
    """orc_direct_accel_fast, the FP64 reference of the bench's accuracy figures,
    sums the closed form of orc_direct_accel, possibly in another order (it may
    be vectorised over sources).  With coincident bodies (left out at zero distance, also when
    eps = 0) the two agree to the rounding of the summation:
    |a_fast - a| <= 1e-13 * sum_j |term_j| for every body; a target list gives
    the same values bit for bit as the full call."""
    r = np.random.default_rng(31)
    n = 4000
    p = r.normal(0, 1, (n, 3))
    p[100:141] = p[0]      # 42 bodies on one point
    p[200:203] = p[7]      # and 4 on another
    x, y, z = (p[:, k].astype(np.float32) for k in range(3))
    m = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    a = oracle.direct_accel(x, y, z, m, eps2)
    af = oracle.direct_accel_fast(x, y, z, m, eps2)
    assert np.isfinite(af).all()
    scale = _abs_term_sum(x, y, z, m, eps2)
    err = np.linalg.norm(af - a, axis=1)
    assert (err <= 1e-13 * scale).all(), (err / scale).max()
    tg = np.sort(np.concatenate([[0, 7, 100, 140, 200], r.choice(n, 300, replace=False)])).astype(np.int32)
    assert np.array_equal(oracle.direct_accel_fast(x, y, z, m, eps2, tg), af[tg])


def test_direct_fast_closed_forms(oracle):
    """Two bodies and the centre of a cube, as for the scalar direct sum."""
    eps = 0.1
    x = np.array([0.0, 2.0], np.float32); y = np.zeros(2, np.float32); z = np.zeros(2, np.float32)
    m = np.array([1.0, 3.0], np.float32)
    a = oracle.direct_accel_fast(x, y, z, m, eps * eps)
    assert a[0, 0] == pytest.approx(3.0 * 2.0 / (4.0 + eps * eps) ** 1.5, rel=1e-14)
    assert a[1, 0] == pytest.approx(-1.0 * 2.0 / (4.0 + eps * eps) ** 1.5, rel=1e-14)
    assert a[0, 1] == 0.0 and a[1, 2] == 0.0
    c = np.array([[i, j, k] for i in (-1, 1) for j in (-1, 1) for k in (-1, 1)] + [[0, 0, 0]], np.float32)
    a = oracle.direct_accel_fast(c[:, 0], c[:, 1], c[:, 2], np.ones(9, np.float32), 0.0)
    assert np.abs(a[8]).max() < 1e-15


def _span_reference(xf, yf, zf):
    """Interaction.cpp:143-147 on the float inputs promoted to double:
    xSpan = |xMax - xMin| (and y, z); maxSpan = xSpan > ySpan ? xSpan : ySpan;
    maxSpan = maxSpan < zSpan ? zSpan : maxSpan."""
    lo = [float(v.min()) for v in (xf, yf, zf)]
    hi = [float(v.max()) for v in (xf, yf, zf)]
    xs, ys, zs = (abs(h - l) for h, l in zip(hi, lo))
    ms = xs if xs > ys else ys
    ms = zs if ms < zs else ms
    return lo, ms


@pytest.mark.parametrize("case", KEY_CASES + ["long_x", "long_y", "long_z"])
def test_bbox_is_min_max_and_span(oracle, case):
    """orc_bbox: the float min/max of each axis exactly, and the FP64 extent
    (origin = the minima, maxSpan by the reference's formula) exactly.  Clouds
    stretched along each axis in turn make each axis's span the one chosen."""
    if case.startswith("long_"):
        r = np.random.default_rng(41)
        p = r.normal(0, 1, (3000, 3)) + 5.0
        p[:, "xyz".index(case[-1])] *= 7.0
        x, y, z = (p[:, k].astype(np.float32) for k in range(3))
    else:
        g = golden(f"keys_{case}.npz")
        x, y, z = g["x"], g["y"], g["z"]
    mm, ext = oracle.bbox(x, y, z)
    assert np.array_equal(mm, np.array([x.min(), x.max(), y.min(), y.max(), z.min(), z.max()], np.float32))
    lo, span = _span_reference(x, y, z)
    assert ext[0] == lo[0] and ext[1] == lo[1] and ext[2] == lo[2]
    assert ext[3] == span
    if case.startswith("long_"):
        k = "xyz".index(case[-1])
        assert span == float(mm[2 * k + 1]) - float(mm[2 * k])


def test_bh_two_cluster_monopole(oracle):
    """A tight cluster far away is accepted as one monopole: the force is the
    cluster's total mass at its centre of mass."""
    r = np.random.default_rng(10)
    cl = r.normal(0, 1e-3, (50, 3)) + np.array([100.0, 0, 0])
    pts = np.vstack([[[0.0, 0.0, 0.0]], cl]).astype(np.float32)
    m = np.concatenate([[1.0], r.uniform(0.5, 1.5, 50)]).astype(np.float32)
    t = oracle.KeyTree(pts[:, 0], pts[:, 1], pts[:, 2], m)
    a, cnt = t.accel(0.5, 0.0, mac_mode=0, targets=[0])
    mc = m[1:].astype(np.float64)
    com = (mc[:, None] * pts[1:].astype(np.float64)).sum(0) / mc.sum()
    expect = mc.sum() * com / np.linalg.norm(com) ** 3
    np.testing.assert_allclose(a[0], expect, rtol=1e-12)
    assert cnt[0] == 1


def test_update_is_symplectic_euler(oracle):
    r = np.random.default_rng(4)
    v = [r.normal(0, 1, 100).astype(np.float32) for _ in range(9)]
    x, y, z, vx, vy, vz = oracle.update(*v, 1e-3, 0.5)
    for pos, vel, v0, p0, a in ((x, vx, v[3], v[0], v[6]), (y, vy, v[4], v[1], v[7]), (z, vz, v[5], v[2], v[8])):
        ev = v0.astype(np.float64) + 1e-3 * a.astype(np.float64)
        np.testing.assert_allclose(vel, ev, rtol=2e-7, atol=1e-9)
        ex = p0.astype(np.float64) + np.float32(0.5 * np.float32(1e-3)) * vel.astype(np.float64)
        np.testing.assert_allclose(pos, ex, rtol=2e-7, atol=1e-9)


# ------------------------------------------------------------------ adjoint
def _direct_sum_autograd(x, y, z, m, eps2, u):
    import torch
    X = torch.tensor(np.stack([x, y, z], 1).astype(np.float64), requires_grad=True)
    M = torch.tensor(m.astype(np.float64), requires_grad=True)
    d = X[None, :, :] - X[:, None, :]
    d2 = (d ** 2).sum(-1)
    w = torch.where(d2 > 0, M[None, :] / (d2 + eps2) ** 1.5, torch.zeros_like(d2))
    a = (w[:, :, None] * d).sum(1)
    (torch.tensor(u) * a).sum().backward()
    return X.grad.numpy(), M.grad.numpy()


def test_vjp_theta_zero_is_direct_sum_autograd(oracle):
    """With every cell opened the walk is the direct sum, and its VJP must be
    torch's FP64 autograd of the direct sum."""
    r = np.random.default_rng(21)
    n = 300
    x, y, z = (r.normal(0, 1, n).astype(np.float32) for _ in range(3))
    m = r.uniform(0.5, 1.5, n).astype(np.float32)
    u = r.normal(0, 1, (n, 3))
    t = oracle.KeyTree(x, y, z, m)
    gx, gm = t.vjp(1e-12, 1e-2, u, mac_mode=0)
    ex, em = _direct_sum_autograd(x, y, z, m, 1e-2, u)
    np.testing.assert_allclose(gx, ex, rtol=0, atol=1e-12 * np.abs(ex).max())
    np.testing.assert_allclose(gm, em, rtol=0, atol=1e-12 * np.abs(em).max())


def test_vjp_two_clusters_central_differences(oracle):
    """Two clusters far apart see each other as one accepted cell, so the VJP
    goes through the centre-of-mass chain; central differences of
    L = sum u.a (tree rebuilt for each perturbed input) must agree."""
    r = np.random.default_rng(22)
    a = r.normal(0, 0.5, (12, 3)) + [-50.0, 0, 0]
    b = r.normal(0, 0.5, (12, 3)) + [50.0, 3.0, -2.0]
    p = np.vstack([a, b]).astype(np.float32)
    m = r.uniform(0.5, 2.0, 24).astype(np.float32)
    u = r.normal(0, 1, (24, 3))
    eps2, theta = 0.1, 0.5
    t = oracle.KeyTree(p[:, 0], p[:, 1], p[:, 2], m)
    gx, gm = t.vjp(theta, eps2, u, mac_mode=0)
    _, cnt0 = t.accel(theta, eps2, mac_mode=0)
    assert np.all(cnt0 < 23)  # some cells accepted

    def loss(pp, mm):
        tt = oracle.KeyTree(pp[:, 0], pp[:, 1], pp[:, 2], mm)
        acc, cnt = tt.accel(theta, eps2, mac_mode=0)
        return float((u * acc).sum()), np.array_equal(cnt, cnt0)

    checked = 0
    for k in range(24):
        for ax in range(3):
            pp, pm = p.copy(), p.copy()
            pp[k, ax] += np.float32(2e-3)
            pm[k, ax] -= np.float32(2e-3)
            delta = float(pp[k, ax]) - float(pm[k, ax])
            (lp, okp), (lm, okm) = loss(pp, m), loss(pm, m)
            if not (okp and okm):
                continue  # an interaction list changed: not differentiable here
            checked += 1
            assert (lp - lm) / delta == pytest.approx(gx[k, ax], rel=1e-3, abs=1e-6 * np.abs(gx).max()), (k, ax)
        mp, mm_ = m.copy(), m.copy()
        mp[k] *= np.float32(1.001)
        mm_[k] *= np.float32(0.999)
        (lp, _), (lm, _) = loss(p, mp), loss(p, mm_)
        assert (lp - lm) / (float(mp[k]) - float(mm_[k])) == pytest.approx(gm[k], rel=1e-4,
                                                                            abs=1e-8 * np.abs(gm).max()), k
    assert checked >= 54
