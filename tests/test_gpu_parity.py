"""GPU parity: every stage of the HIP engine against the CPU oracle on the same
seeded inputs, through the C ABI (nns_amd.KernelsWrapper).

Bars:
  bounding box, keys, sorted order, tree cells and child slots, cell counts and
  start positions, FP64 cell mass / centre of mass, interaction counts,
  update: bit-exact;
  accelerations: FP32 against the oracle's FP64 evaluation of the same
  interaction lists, RMS relative error <= 1e-5 and max <= 2e-3 (the max is
  set by FP32 differences of nearby coordinates, |x| eps_32 / |d|).
"""
import numpy as np
import pytest

import nns_amd

pytestmark = pytest.mark.gpu

THETA = np.float32(0.5)


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def make_case(kind, n, seed=1):
    if kind in ("star_disk", "plummer", "binary_disk"):
        return nns_amd.initial_conditions(kind, n, seed)
    r = np.random.default_rng(seed)
    if kind == "cluster":  # two tight clusters and a halo: deep trees and chains
        a = r.normal(0, 1e-4, (n // 3, 3)) + 0.3
        b = r.normal(0, 1e-3, (n // 3, 3)) - 0.2
        c = r.normal(0, 1.0, (n - 2 * (n // 3), 3))
        p = np.vstack([a, b, c]).astype(np.float32)
    elif kind == "coincident":  # groups of 1..6 bodies at identical positions
        base = r.normal(0, 1, (n, 3)).astype(np.float32)
        idx = np.repeat(np.arange(n), r.integers(1, 7, n))[:n]
        p = base[idx]
    elif kind == "line":  # collinear bodies: one key axis only, long one-child chains
        t = np.sort(r.uniform(-1, 1, n))
        p = np.stack([t, 0.5 * t, -0.25 * t], 1).astype(np.float32)
    elif kind == "far_clumps":  # two clumps 1e3 apart: empty space, long chains down to each clump
        a = r.normal(0, 0.1, (n // 2, 3))
        b = r.normal(0, 0.1, (n - n // 2, 3)) + 1e3
        p = np.vstack([a, b]).astype(np.float32)
    elif kind == "offset":  # a unit cloud at 1e4 from the origin: the extent, not the origin, sets the keys
        p = (r.normal(0, 1, (n, 3)) + 1e4).astype(np.float32)
    else:
        raise ValueError(kind)
    z0 = np.zeros(n, np.float32)
    m = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), z0, z0.copy(), z0.copy(), m


def run_to_forces(x, y, z, vx, vy, vz, m, params=None, m_nodes=None):
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=params, m=m_nodes)
    a, kw, n, mm = bh.arr, bh.kw, bh.n, bh.m
    kw.resetArrays(a["mutex"], a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["sorted"],
                   a["child"], a["index"], a["minX"], a["maxX"], a["minY"], a["maxY"], a["minZ"], a["maxZ"], n, mm)
    kw.computeBoundingBox(a["mutex"], a["x"], a["y"], a["z"], a["minX"], a["maxX"], a["minY"], a["maxY"],
                          a["minZ"], a["maxZ"], n)
    kw.buildTree(a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["child"], a["index"], a["minX"],
                 a["maxX"], a["minY"], a["maxY"], a["minZ"], a["maxZ"], n, mm)
    kw.centreOfMass(a["x"], a["y"], a["z"], a["mass"], a["index"], n)
    kw.sort(a["count"], a["start"], a["sorted"], a["child"], a["index"], n)
    kw.computeForces(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"], a["mass"],
                     a["sorted"], a["child"], a["minX"], a["maxX"], n, bh.ctx.params.gravity)
    bh.ctx.check()
    return bh


def gpu_cells(bh):
    n = bh.n
    index = int(bh.arr["index"].item())
    nc = index - n
    lvl = bh.ctx.copy_buffer(nns_amd.BUF_CELL_LEVEL, index, np.int32)[n:]
    m64 = bh.ctx.copy_buffer(nns_amd.BUF_CELL_MASS64, index, np.float64)[n:]
    com = bh.ctx.copy_buffer(nns_amd.BUF_CELL_COM64, 3 * index, np.float64).reshape(index, 3)[n:]
    start = bh.host("start", index)[n:]
    count = bh.host("count", index)[n:]
    child = bh.host("child", 8 * index).reshape(index, 8)
    root = bh.ctx.copy_buffer(nns_amd.BUF_ROOT, 4, np.float64)
    return dict(nc=nc, level=lvl, mass=m64, com=com, start=start, count=count, child=child, root=root,
                fx=bh.host("x", index)[n:], fy=bh.host("y", index)[n:], fz=bh.host("z", index)[n:],
                fm=bh.host("mass", index)[n:])


def check_tree(bh, t):
    """Cells, children, FP64 summaries and counts against the oracle's key tree."""
    n = bh.n
    g = gpu_cells(bh)
    # oracle cell 0 is the root; GPU cells are [n, index).  A cell of the key
    # tree is named by (level, first sorted position): match the two sets.
    assert g["nc"] == t.ncell - 1
    gk = g["level"].astype(np.int64) << 32 | g["start"].astype(np.int64)
    ok = t.level[1:].astype(np.int64) << 32 | t.start[1:].astype(np.int64)
    gi, oi = np.argsort(gk, kind="stable"), np.argsort(ok, kind="stable")
    assert np.array_equal(gk[gi], ok[oi])
    assert np.unique(gk).size == gk.size
    oc = np.empty(g["nc"], np.int64)
    oc[gi] = oi + 1                                        # oracle cell of each GPU cell
    assert np.array_equal(g["count"], t.count[oc])
    assert np.array_equal(g["mass"], t.mass[oc])          # FP64, bit for bit
    assert np.array_equal(g["com"], t.com[oc])
    assert np.array_equal(g["fx"], t.c32[oc, 0]) and np.array_equal(g["fm"], t.c32[oc, 3])
    assert g["root"][0] == t.mass[0] and np.array_equal(g["root"][1:], t.com[0])

    def enc(v):  # GPU child values -> oracle encoding (-1 empty, body id, -2 - cell)
        v = v.astype(np.int64)
        assert ((v >= -1) & (v < n + g["nc"])).all()
        out = v.copy()
        cell = v >= n
        out[cell] = -2 - oc[v[cell] - n]
        return out
    assert np.array_equal(enc(g["child"][0]), t.child[0])
    ch = enc(g["child"][n:])
    bad = np.nonzero((ch != t.child[oc]).any(axis=1))[0]
    assert bad.size == 0, (bad[:4], ch[bad[:4]], t.child[oc[bad[:4]]])


def check_forces(bh, t, oracle, eps2, tol_rms=1e-5, tol_max=2e-3):
    a_orc, cnt_orc = t.accel(bh.ctx.params.theta, eps2, mac_mode=1)
    ag = np.stack([bh.host("ax"), bh.host("ay"), bh.host("az")], 1).astype(np.float64)
    G = bh.ctx.params.gravity
    norm = np.linalg.norm(a_orc, axis=1)
    ok = norm > 0
    rel = np.linalg.norm(ag[ok] - G * a_orc[ok], axis=1) / (G * norm[ok])
    assert np.sqrt((rel ** 2).mean()) <= tol_rms, np.sqrt((rel ** 2).mean())
    assert rel.max() <= tol_max, rel.max()
    cnt_gpu = bh.count_interactions().cpu().numpy()
    assert np.array_equal(cnt_gpu, cnt_orc)
    return rel, cnt_orc


CASES = [("star_disk", 4096), ("star_disk", 32768), ("plummer", 20000), ("binary_disk", 50000),
         ("cluster", 9000), ("coincident", 5000), ("plummer", 2), ("plummer", 3), ("plummer", 130),
         ("line", 6000), ("far_clumps", 12000), ("offset", 8000)]


@pytest.mark.parametrize("kind,n", CASES)
def test_pipeline_matches_oracle(oracle, kind, n):
    _torch()
    print(f"case {kind} n={n}", flush=True)
    x, y, z, vx, vy, vz, m = make_case(kind, n)
    params = nns_amd.default_params(theta=float(THETA))
    # coincident bodies hang off chains of cells down to level 21: more cells than 2n + 12000
    m_nodes = 24 * n + 12000 if kind == "coincident" else None
    bh = run_to_forces(x, y, z, vx, vy, vz, m, params, m_nodes=m_nodes)
    # bounding box and extent
    mm, ext = oracle.bbox(x, y, z)
    got = np.array([bh.arr[k].item() for k in ("minX", "maxX", "minY", "maxY", "minZ", "maxZ")], np.float32)
    assert np.array_equal(got, mm)
    assert np.array_equal(bh.ctx.copy_buffer(nns_amd.BUF_EXTENT, 4, np.float64), ext)
    # keys and sorted order
    t = oracle.KeyTree(x, y, z, m, cap=(m_nodes - n) if m_nodes else 0)
    assert t.err == 0
    keys = bh.ctx.copy_buffer(nns_amd.BUF_KEYS, n, np.uint64)
    assert np.array_equal(keys, oracle.keys_morton(x, y, z))
    assert np.array_equal(bh.ctx.copy_buffer(nns_amd.BUF_SORTED_KEYS, n, np.uint64), t.skeys)
    assert np.array_equal(bh.host("sorted"), t.perm)
    check_tree(bh, t)
    check_forces(bh, t, oracle, float(np.float32(np.float32(params.softening) ** 2)))


def test_centre_of_mass_twice_resums_every_cell():
    """A second centreOfMass on the same tree writes every cell again, subtree
    cells included (their job counter is set back after each summary): the
    cells' masses and centres, zeroed in between, come back bit for bit."""
    _torch()
    n = 200_000
    print(f"case centre of mass twice binary_disk n={n}", flush=True)
    bh = run_to_forces(*make_case("binary_disk", n))
    a, kw = bh.arr, bh.kw
    index = int(a["index"].item())
    assert index - n > 1000  # subtree cells as well as top cells
    first = {k: a[k][n:index].clone() for k in ("x", "y", "z", "mass")}
    m64 = bh.ctx.copy_buffer(nns_amd.BUF_CELL_MASS64, index, np.float64)
    for k in first:
        a[k][n:index] = 0.0
    kw.centreOfMass(a["x"], a["y"], a["z"], a["mass"], a["index"], n)
    bh.ctx.check()
    for k, v in first.items():
        assert bool((a[k][n:index] == v).all()), k
    assert np.array_equal(bh.ctx.copy_buffer(nns_amd.BUF_CELL_MASS64, index, np.float64), m64)


@pytest.mark.parametrize("n,m", [(1000, 14000), (4097, 2 * 4097 + 12000), (70000, 70001)])
def test_reset_arrays_post_state(n, m):
    """nbx_reset_arrays leaves exactly what resetArraysKernel (Kernels.cu:21-64)
    leaves, read back from arrays filled with garbage first: every child slot
    -1; count 1 for bodies and 0 for cells; cell x, y, z, mass 0 and the
    bodies' untouched; start -1; sorted 0; *index = n; *mutex and the six
    bounding-box scalars 0."""
    torch = _torch()
    print(f"case reset n={n} m={m}", flush=True)
    dev = torch.device("cuda")
    r = np.random.default_rng(n)
    f = {k: torch.as_tensor(r.normal(size=m).astype(np.float32), device=dev) for k in ("x", "y", "z", "mass")}
    keep = {k: v[:n].cpu().numpy().copy() for k, v in f.items()}
    ints = {k: torch.as_tensor(r.integers(-9, 9, size=m).astype(np.int32), device=dev)
            for k in ("count", "start", "sorted")}
    child = torch.as_tensor(r.integers(-9, 9, size=8 * m).astype(np.int32), device=dev)
    sc = {k: torch.full((1,), 7, dtype=torch.int32, device=dev) for k in ("index", "mutex")}
    bb = {k: torch.full((1,), 3.5, device=dev) for k in ("minX", "maxX", "minY", "maxY", "minZ", "maxZ")}
    ctx = nns_amd.Context(n, m)
    kw = nns_amd.KernelsWrapper(ctx)
    kw.resetArrays(sc["mutex"], f["x"], f["y"], f["z"], f["mass"], ints["count"], ints["start"], ints["sorted"],
                   child, sc["index"], bb["minX"], bb["maxX"], bb["minY"], bb["maxY"], bb["minZ"], bb["maxZ"], n, m)
    torch.cuda.synchronize()
    assert (child.cpu().numpy() == -1).all()
    cnt = ints["count"].cpu().numpy()
    assert (cnt[:n] == 1).all() and (cnt[n:] == 0).all()
    for k in ("x", "y", "z", "mass"):
        v = f[k].cpu().numpy()
        assert np.array_equal(v[:n], keep[k]) and (v[n:] == 0).all(), k
    assert (ints["start"].cpu().numpy() == -1).all()
    assert (ints["sorted"].cpu().numpy() == 0).all()
    assert sc["index"].item() == n and sc["mutex"].item() == 0
    assert all(v.item() == 0.0 for v in bb.values())
    ctx.check()


def _cluster_case(core, n, seed):
    """`core` bodies in one tight cluster (sigma 1e-4 of an extent of about 2,
    so the cell holding them sits near level 12 and every cell above it holds
    all of them) and n - core spread bodies."""
    r = np.random.default_rng(seed)
    p = np.vstack([r.normal(0.3, 1e-4, (core, 3)), r.uniform(-1, 1, (n - core, 3))]).astype(np.float32)
    z0 = np.zeros(n, np.float32)
    m = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), z0, z0.copy(), z0.copy(), m


@pytest.mark.parametrize("core,n", [(1024, 1100), (1025, 1100), (0, 1024), (0, 1025), (0, 2049), (2049, 2100)])
def test_subtree_size_boundaries(oracle, core, n):
    """Cells of exactly kSubtreeMax = 1024 bodies (one subtree) and of 1025 (a
    top cell), alone or as a chain of top cells above a tight cluster, and body
    counts just past a multiple of 1024 (the top-cell samples)."""
    _torch()
    print(f"case boundary core={core} n={n}", flush=True)
    x, y, z, vx, vy, vz, m = _cluster_case(core, n, 11) if core else make_case("plummer", n)
    params = nns_amd.default_params(theta=float(THETA))
    bh = run_to_forces(x, y, z, vx, vy, vz, m, params)
    t = oracle.KeyTree(x, y, z, m)
    assert t.err == 0
    assert np.array_equal(bh.ctx.copy_buffer(nns_amd.BUF_SORTED_KEYS, n, np.uint64), t.skeys)
    assert np.array_equal(bh.host("sorted"), t.perm)
    check_tree(bh, t)
    check_forces(bh, t, oracle, float(np.float32(np.float32(params.softening) ** 2)))


@pytest.mark.parametrize("kind,n", [("star_disk", 32768), ("coincident", 5000), ("plummer", 130)])
def test_locked_build_matches_oracle(oracle, kind, n):
    """The lock-based subtree insertion (build_locked = 1) builds the same tree."""
    _torch()
    print(f"case locked {kind} n={n}", flush=True)
    x, y, z, vx, vy, vz, m = make_case(kind, n)
    params = nns_amd.default_params(theta=float(THETA), build_locked=1)
    m_nodes = 24 * n + 12000 if kind == "coincident" else None
    bh = run_to_forces(x, y, z, vx, vy, vz, m, params, m_nodes=m_nodes)
    t = oracle.KeyTree(x, y, z, m, cap=(m_nodes - n) if m_nodes else 0)
    assert np.array_equal(bh.host("sorted"), t.perm)
    check_tree(bh, t)
    check_forces(bh, t, oracle, float(np.float32(np.float32(params.softening) ** 2)))


def test_full_size_binary_disk(oracle):
    """The headline workload at its full size (binary + disk, N = 4,194,304,
    theta 0.5, eps 1e-3, Kahan on): keys, sorted order, every cell, child slot
    and FP64 summary bit for bit; interaction counts and accelerations on a
    seeded sample of 4096 bodies against the oracle's walk of the same tree."""
    _torch()
    n = 4194304
    print(f"case binary_disk n={n}", flush=True)
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("binary_disk", n, 1)
    params = nns_amd.default_params(theta=float(THETA))
    bh = run_to_forces(x, y, z, vx, vy, vz, m, params)
    t = oracle.KeyTree(x, y, z, m)
    assert t.err == 0
    assert np.array_equal(bh.ctx.copy_buffer(nns_amd.BUF_SORTED_KEYS, n, np.uint64), t.skeys)
    assert np.array_equal(bh.host("sorted"), t.perm)
    check_tree(bh, t)
    eps2 = float(np.float32(np.float32(params.softening) ** 2))
    tg = np.sort(np.random.default_rng(9).choice(n, 4096, replace=False)).astype(np.int32)
    a_orc, cnt_orc = t.accel(params.theta, eps2, mac_mode=1, targets=tg)
    ag = np.stack([bh.host("ax")[tg], bh.host("ay")[tg], bh.host("az")[tg]], 1).astype(np.float64)
    G = params.gravity
    rel = np.linalg.norm(ag - G * a_orc, axis=1) / (G * np.linalg.norm(a_orc, axis=1))
    assert np.sqrt((rel ** 2).mean()) <= 1e-5 and rel.max() <= 2e-3, (np.sqrt((rel ** 2).mean()), rel.max())
    assert np.array_equal(bh.count_interactions().cpu().numpy()[tg], cnt_orc)


def test_maximum_size(oracle):
    """The largest system a context takes, N = 48,805,888 (47,662 samples of
    1024 bodies fill the tree's 2^20-entry top table): a Gaussian sphere at
    theta 0.5.  Sorted keys and order bit for bit, through the properties that
    fix a stable sort (keys in order and equal to the oracle's keys of the
    listed bodies, a permutation, equal keys in index order); the FP64 root mass
    and centre of mass against FP64 sums; accelerations of 256 seeded bodies
    against an FP64 direct sum, to the Barnes-Hut error at theta 0.5 (the
    oracle's walk of a 200K Gaussian sphere: RMS 1.5e-3, max 7.5e-3)."""
    torch = _torch()
    n = 48805888
    print(f"case gaussian n={n}", flush=True)
    r = np.random.default_rng(11)
    p = r.normal(size=(n, 3)).astype(np.float32)
    x, y, z = (p[:, k].copy() for k in range(3))
    del p
    m = np.full(n, 1.0 / n, np.float32)
    z0 = np.zeros(n, np.float32)
    params = nns_amd.default_params(theta=0.5)
    bh = run_to_forces(x, y, z, z0, z0, z0, m, params)
    keys = oracle.keys_morton(x, y, z)
    sk = bh.ctx.copy_buffer(nns_amd.BUF_SORTED_KEYS, n, np.uint64)
    srt = bh.host("sorted").astype(np.int64)
    assert np.array_equal(sk, keys[srt])
    seen = np.zeros(n, bool)
    seen[srt] = True
    assert seen.all()
    up, eq = sk[1:] > sk[:-1], sk[1:] == sk[:-1]
    assert (up | (eq & (srt[1:] > srt[:-1]))).all()
    del keys, sk, srt, seen, up, eq
    m64 = m.astype(np.float64)
    M = m64.sum()
    root = bh.ctx.copy_buffer(nns_amd.BUF_ROOT, 4, np.float64)
    assert abs(root[0] - M) <= 1e-12 * M
    for k, v in enumerate((x, y, z)):
        assert abs(root[1 + k] - np.dot(m64, v.astype(np.float64)) / M) <= 1e-10
    tg = np.sort(r.choice(n, 256, replace=False))
    dev = bh.dev
    P = torch.stack([torch.as_tensor(v, device=dev).double() for v in (x, y, z)], 1)
    W = torch.as_tensor(m64, device=dev)
    T = P[torch.as_tensor(tg, device=dev)]
    eps2 = float(np.float32(np.float32(params.softening) ** 2))
    acc = torch.zeros(256, 3, dtype=torch.float64, device=dev)
    for lo in range(0, n, 1 << 18):
        d = P[lo:lo + (1 << 18)][None, :, :] - T[:, None, :]
        w = W[lo:lo + (1 << 18)][None, :] * ((d * d).sum(-1) + eps2).pow(-1.5)
        acc += (w[..., None] * d).sum(1)
    ref = params.gravity * acc.cpu().numpy()
    ag = np.stack([bh.host("ax")[tg], bh.host("ay")[tg], bh.host("az")[tg]], 1).astype(np.float64)
    rel = np.linalg.norm(ag - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"direct-sum error rms {np.sqrt((rel ** 2).mean()):.3e} max {rel.max():.3e}", flush=True)
    assert np.sqrt((rel ** 2).mean()) <= 3e-3 and rel.max() <= 2e-2


def test_step_captured_in_graph_matches_eager():
    """nbx_step makes no host synchronisation, copy, memset or allocation when
    not timed, so a caller can capture it on its stream into a graph.  Each of
    three replays of one captured step, with eager work and synchronisations
    between them, matches the eager step of a twin system bit for bit."""
    torch = _torch()
    n = 30000
    print(f"case graph capture plummer n={n}", flush=True)
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("plummer", n, 4)
    eager = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    graph = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=side):
        graph.update(stream=side.cuda_stream)
    torch.cuda.synchronize()
    for k in range(3):
        g.replay()
        torch.cuda.synchronize()
        eager.update()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        for key in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
            assert np.array_equal(eager.host(key), graph.host(key)), (k, key)
    eager.ctx.check()
    graph.ctx.check()


def test_unsoftened_and_no_kahan(oracle):
    _torch()
    x, y, z, vx, vy, vz, m = make_case("coincident", 3000, seed=5)
    params = nns_amd.default_params(theta=0.6, softening=0.0, kahan=0)
    bh = run_to_forces(x, y, z, vx, vy, vz, m, params, m_nodes=24 * 3000 + 12000)
    t = oracle.KeyTree(x, y, z, m, cap=23 * 3000 + 12000)
    check_tree(bh, t)
    check_forces(bh, t, oracle, 0.0)


def test_keys_hilbert_and_sort_extension(oracle):
    torch = _torch()
    x, y, z, *_ = make_case("plummer", 70000, seed=2)
    ctx = nns_amd.Context(70000)
    kw = nns_amd.KernelsWrapper(ctx)
    dev = torch.device("cuda")
    tx, ty, tz = (torch.as_tensor(a, device=dev) for a in (x, y, z))
    mins = [torch.zeros(1, device=dev) for _ in range(6)]
    mutex = torch.zeros(1, dtype=torch.int32, device=dev)
    kw.computeBoundingBox(mutex, tx, ty, tz, *mins, 70000)
    keys = torch.zeros(70000, dtype=torch.int64, device=dev)
    kw.computeKeys(tx, ty, tz, 70000, keys, nns_amd.KEY_HILBERT)
    hk = keys.cpu().numpy().view(np.uint64)
    assert np.array_equal(hk, oracle.keys_hilbert(oracle.keys_morton(x, y, z)))
    sk = torch.zeros_like(keys)
    perm = torch.zeros(70000, dtype=torch.int32, device=dev)
    kw.sortKeys(keys, 70000, sk, perm)
    p_orc, sk_orc = oracle.sort_stable(hk)
    assert np.array_equal(perm.cpu().numpy(), p_orc)
    assert np.array_equal(sk.cpu().numpy().view(np.uint64), sk_orc)


def test_update_bit_exact_and_step_matches_stages(oracle):
    torch = _torch()
    x, y, z, vx, vy, vz, m = make_case("star_disk", 20000, seed=3)
    p = nns_amd.default_params()
    bh = run_to_forces(x, y, z, vx, vy, vz, m, p)
    ax, ay, az = bh.host("ax"), bh.host("ay"), bh.host("az")
    a = bh.arr
    bh.kw.update(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"], bh.n,
                 p.timestep, p.dampening)
    torch.cuda.synchronize()
    ex = oracle.update(x, y, z, vx, vy, vz, ax, ay, az, p.timestep, p.dampening)
    for k, e in zip(("x", "y", "z", "vx", "vy", "vz"), ex):
        assert np.array_equal(bh.host(k), e), k
    # nbx_step on a fresh copy gives the same state as the seven stage calls
    bh2 = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=p)
    ms = bh2.update(stage_times=True)
    assert len(ms) == 7 and all(v >= 0 for v in ms)
    for k in ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"):
        assert np.array_equal(bh2.host(k), bh.host(k)), k


def test_ten_step_run_tracks_oracle(oracle):
    """The plain-method CPU case (central star + disk, N = 32,768, 10 steps) with
    the star inside the tree as an ordinary body: the GPU trajectory and the
    oracle's FP64-force trajectory stay within 1e-5.  (The reference's own code
    keeps the star outside the tree: tests/test_gpu_stars.py.)"""
    _torch()
    n = 32768
    x, y, z, vx, vy, vz, m = make_case("star_disk", n, seed=4)
    p = nns_amd.default_params()
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=p)
    st = [a.copy() for a in (x, y, z, vx, vy, vz)]
    eps2 = float(np.float32(np.float32(p.softening) ** 2))
    for _ in range(10):
        bh.update()
        t = oracle.KeyTree(st[0], st[1], st[2], m)
        acc, _ = t.accel(0.5, eps2, mac_mode=1)
        st = oracle.update(*st, acc[:, 0].astype(np.float32), acc[:, 1].astype(np.float32),
                           acc[:, 2].astype(np.float32), p.timestep, p.dampening)
    bh.ctx.check()
    pos = np.stack([bh.host("x"), bh.host("y"), bh.host("z")], 1)
    ref = np.stack(st[:3], 1)
    assert np.abs(pos - ref).max() < 1e-5


def test_error_paths_do_not_fault(oracle):
    torch = _torch()
    # capacity: m barely above n
    x, y, z, vx, vy, vz, m = make_case("plummer", 5000, seed=6)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, m=5000 + 50)
    with pytest.raises(nns_amd.NbxError) as e:
        bh.update()
        bh.ctx.check()
    assert e.value.code == -2
    # bucket overflow: nine bodies on one point
    x, y, z, vx, vy, vz, m = make_case("plummer", 1000, seed=6)
    x[:9], y[:9], z[:9] = x[0], y[0], z[0]
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    with pytest.raises(nns_amd.NbxError) as e:
        bh.update()
        bh.ctx.check()
    assert e.value.code == -3
    # non-finite input
    x, y, z, vx, vy, vz, m = make_case("plummer", 1000, seed=6)
    x[17] = np.nan
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    with pytest.raises(nns_amd.NbxError) as e:
        bh.update()
        bh.ctx.check()
    assert e.value.code == -5
    # stage order
    x, y, z, vx, vy, vz, m = make_case("plummer", 1000, seed=6)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    a = bh.arr
    with pytest.raises(nns_amd.NbxError) as e:
        bh.kw.computeForces(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"],
                            a["mass"], a["sorted"], a["child"], a["minX"], a["maxX"], bh.n, 1.0)
    assert e.value.code == -7
    # the context recovers after a reset
    bh.update()
    bh.ctx.check()
    torch.cuda.synchronize()
    # the tree is built over the bodies of the last bounding box only
    a = bh.arr
    kw, n, mm = bh.kw, bh.n, bh.m
    kw.resetArrays(a["mutex"], a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["sorted"],
                   a["child"], a["index"], a["minX"], a["maxX"], a["minY"], a["maxY"], a["minZ"], a["maxZ"], n, mm)
    kw.computeBoundingBox(a["mutex"], a["x"], a["y"], a["z"], a["minX"], a["maxX"], a["minY"], a["maxY"],
                          a["minZ"], a["maxZ"], n - 1)
    with pytest.raises(nns_amd.NbxError) as e:
        kw.buildTree(a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["child"], a["index"], a["minX"],
                     a["maxX"], a["minY"], a["maxY"], a["minZ"], a["maxZ"], n, mm)
    assert e.value.code == -1 and "bounding_box" in str(e.value)


def _half_massless(m, x):
    """Masses with every body of the half x < 0 at zero: each cell inside that
    half has zero total mass, so many subtree jobs raise the latch."""
    out = m.copy()
    out[x < 0.0] = 0.0
    return out


@pytest.mark.parametrize("locked", [0, 1])
def test_errors_raised_inside_job_queue_kernels(oracle, locked):
    """Errors latched from inside the persistent job-queue kernels while other
    workgroups of the same launch are still taking jobs: zero-mass cells in the
    subtree summary (NBX_ERR_NONFINITE), and for the lock-based build a
    capacity overflow inside the subtree insertion (NBX_ERR_CAPACITY).  Each
    workgroup leaves as a whole; the check returns the code, once, and the
    next step on valid input recovers."""
    torch = _torch()
    n = 60000
    print(f"case latch inside job-queue kernels locked={locked} n={n}", flush=True)
    params = nns_amd.default_params(build_locked=locked)
    x, y, z, vx, vy, vz, m = make_case("plummer", n, seed=13)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, _half_massless(m, x), params=params)
    with pytest.raises(nns_amd.NbxError) as e:
        bh.update()
        bh.ctx.check()
    assert e.value.code == -5
    bh.arr["mass"][:n] = torch.as_tensor(m, device=bh.dev)
    bh.update()
    bh.ctx.check()
    if not locked:
        return
    # room for the top cells (more than 1024 bodies, wired before the subtrees)
    # and 100 cells more: the insertion of the subtrees runs out
    t = oracle.KeyTree(x, y, z, m)
    ntop = int((t.count[1:] > 1024).sum())
    assert t.ncell - 1 > ntop + 100
    print(f"case locked capacity n={n} top cells={ntop}", flush=True)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m, params=params, m=n + ntop + 100)
    with pytest.raises(nns_amd.NbxError) as e:
        bh.update()
        bh.ctx.check()
    assert e.value.code == -2
    # still latched (no reset since): reported again, and the state is untouched
    with pytest.raises(nns_amd.NbxError) as e:
        bh.ctx.check()
    assert e.value.code == -2


def test_error_of_an_earlier_step_is_reported_later():
    """A step whose summary latches an error leaves the bodies where they were
    (the update stage does nothing); later steps clear the latch but keep the
    error, which the next nbx_check reports once."""
    torch = _torch()
    n = 20000
    print(f"case sticky error n={n}", flush=True)
    x, y, z, vx, vy, vz, m = make_case("plummer", n, seed=15)
    bh = nns_amd.BarnesHut(x, y, z, vx, vy, vz, m)
    bh.update()
    bh.update()
    torch.cuda.synchronize()
    keys = ("x", "y", "z", "vx", "vy", "vz")
    before = {k: bh.host(k).copy() for k in keys}
    bh.arr["mass"][:n] = torch.as_tensor(_half_massless(m, before["x"]), device=bh.dev)
    bh.update()  # latches NONFINITE
    torch.cuda.synchronize()
    for k in keys:
        assert np.array_equal(bh.host(k), before[k]), k
    bh.arr["mass"][:n] = torch.as_tensor(m, device=bh.dev)
    bh.update()
    bh.update()
    torch.cuda.synchronize()
    assert not np.array_equal(bh.host("x"), before["x"])  # the later steps moved the bodies
    with pytest.raises(nns_amd.NbxError) as e:
        bh.ctx.check()
    assert e.value.code == -5
    bh.ctx.check()  # reported once


def test_force_waves_parameter_reaches_the_context(oracle):
    """nbx_params.force_waves sets the force kernel's waves per block (0 keeps
    the current choice in nbx_set_params); groups are independent of the block
    shape, so the accelerations are the same bit for bit."""
    _torch()
    ctx = nns_amd.Context(1000, params=nns_amd.default_params(force_waves=8))
    assert ctx.get_params().force_waves == 8
    ctx.set_params(nns_amd.default_params(force_waves=0))
    assert ctx.get_params().force_waves == 8
    ctx.set_params(nns_amd.default_params(force_waves=2))
    assert ctx.get_params().force_waves == 2
    ctx.close()
    x, y, z, vx, vy, vz, m = make_case("plummer", 30000, seed=16)
    acc = []
    for w in (2, 4, 8):
        bh = run_to_forces(x, y, z, vx, vy, vz, m, nns_amd.default_params(force_waves=w))
        assert bh.ctx.get_params().force_waves == w
        acc.append(np.stack([bh.host("ax"), bh.host("ay"), bh.host("az")]))
    assert np.array_equal(acc[0], acc[1]) and np.array_equal(acc[1], acc[2])


def test_valu_probe_chains_converge():
    """The FP32 VALU probe behind the bench's measured issue rates runs both of
    its chains to their fixed point (x -> 0.999 x + 1e-3 and x -> x^-1/2 both
    end at 1, so each lane's sum of 8 registers is 8) and reports a time."""
    _torch()
    print("case valu probe", flush=True)
    for kind, iters, tol in (("fma", 16384, 1e-3), ("rsq", 256, 1e-4)):
        rate, ms, out = nns_amd.probe_valu(kind, blocks=512, iters=iters)
        assert rate > 0 and ms > 0
        v = out.cpu().numpy()
        assert v.size == 512 * 256 and np.abs(v - 8.0).max() <= tol, (kind, np.abs(v - 8.0).max())
