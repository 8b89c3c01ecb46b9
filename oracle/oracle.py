"""ctypes front end of the CPU checker (oracle/nbx_oracle.c) and of the reference
harness (oracle/_ref/ref_harness).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The engine under nns_amd/ never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

# Reference physical constants used by Interaction::singleInteraction
# (3D/OpenMP/OpenMP_NBody/include/Constants.h:7-9).
TO_METERS = 1.496e11
G_SI = 6.67408e-11

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")

_lib = None


def build():
    """Compile liboracle.so (and the reference harness when the reference is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if os.path.isdir("/root/reference/3D/OpenMP/OpenMP_NBody"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    L.orc_version.restype = C.c_int
    L.orc_bbox.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p, _f64p]
    L.orc_morton_key.restype = C.c_uint64
    L.orc_morton_key.argtypes = [C.c_float, C.c_float, C.c_float, _f64p]
    L.orc_keys_morton.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f64p, _u64p]
    L.orc_lebesgue2hilbert.restype = C.c_uint64
    L.orc_lebesgue2hilbert.argtypes = [C.c_uint64, C.c_int]
    L.orc_keys_hilbert.argtypes = [C.c_int64, _u64p, _u64p]
    L.orc_sort_stable.argtypes = [C.c_int64, _u64p, _i32p, _u64p]
    L.orc_tree_build.restype = C.c_void_p
    L.orc_tree_build.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p, C.c_int64, C.POINTER(C.c_int)]
    L.orc_tree_ncells.restype = C.c_int64
    L.orc_tree_ncells.argtypes = [C.c_void_p]
    L.orc_tree_lcell0.restype = C.c_double
    L.orc_tree_lcell0.argtypes = [C.c_void_p]
    L.orc_tree_export.argtypes = [C.c_void_p] + [C.c_void_p] * 11
    L.orc_tree_free.argtypes = [C.c_void_p]
    L.orc_crit0.restype = C.c_float
    L.orc_crit0.argtypes = [C.c_double, C.c_double]
    L.orc_tree_accel.restype = C.c_int
    L.orc_tree_accel.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int64,
                                 C.c_void_p, _f64p, C.c_void_p]
    L.orc_direct_accel.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p, C.c_double, C.c_int64,
                                   C.c_void_p, _f64p]
    L.orc_tree_vjp.restype = C.c_int
    L.orc_group_walk.restype = C.c_int
    L.orc_group_walk.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int64, _i32p, C.c_void_p,
                                 _i32p, C.c_void_p, C.c_void_p]
    L.orc_tree_vjp.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, _f64p, _f64p, _f64p]
    L.orc_direct_accel_fast.restype = C.c_int
    L.orc_direct_accel_fast.argtypes = [C.c_int64, _f32p, _f32p, _f32p, _f32p, C.c_double, C.c_int64,
                                        C.c_void_p, _f64p]
    L.orc_update.argtypes = [C.c_int64] + [_f32p] * 9 + [C.c_float, C.c_float]
    L.orc_geo_build.restype = C.c_void_p
    L.orc_geo_build.argtypes = [C.c_int64, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.POINTER(C.c_int)]
    L.orc_geo_nnodes.restype = C.c_int64
    L.orc_geo_nnodes.argtypes = [C.c_void_p]
    L.orc_geo_export.argtypes = [C.c_void_p, _i32p, _f64p, C.c_void_p]
    L.orc_geo_accel.argtypes = [C.c_void_p, C.c_int64, _f64p, _f64p, _f64p, _f64p, C.c_double,
                                C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.c_double, _f64p, C.c_void_p]
    L.orc_geo_free.argtypes = [C.c_void_p]
    L.orc_star_forces.argtypes = [C.c_int64] + [_f64p] * 4 + [C.c_int64] + [_f64p] * 4 + \
        [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _f64p, _f64p]
    L.orc_update_reference.argtypes = [C.c_int64] + [_f64p] * 7 + [C.c_double, C.c_double]
    _lib = L
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def bbox(x, y, z):
    mm = np.zeros(6, np.float32)
    ext = np.zeros(4, np.float64)
    lib().orc_bbox(len(x), _f32(x), _f32(y), _f32(z), mm, ext)
    return mm, ext


def keys_morton(x, y, z, ext=None):
    if ext is None:
        ext = bbox(x, y, z)[1]
    out = np.zeros(len(x), np.uint64)
    lib().orc_keys_morton(len(x), _f32(x), _f32(y), _f32(z), _f64(ext), out)
    return out


def lebesgue2hilbert(lebesgue, level=21):
    return int(lib().orc_lebesgue2hilbert(int(lebesgue), int(level)))


def keys_hilbert(morton):
    morton = np.ascontiguousarray(morton, dtype=np.uint64)
    out = np.zeros(len(morton), np.uint64)
    lib().orc_keys_hilbert(len(morton), morton, out)
    return out


def sort_stable(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    perm = np.zeros(len(keys), np.int32)
    sk = np.zeros(len(keys), np.uint64)
    lib().orc_sort_stable(len(keys), keys, perm, sk)
    return perm, sk


class KeyTree:
    """Key-prefix octree with FP64 summaries (see nbx_oracle.c)."""

    def __init__(self, x, y, z, m, cap=0):
        self.x, self.y, self.z, self.m = _f32(x), _f32(y), _f32(z), _f32(m)
        self.n = len(self.x)
        err = C.c_int(0)
        self._h = lib().orc_tree_build(self.n, self.x, self.y, self.z, self.m, int(cap), C.byref(err))
        self.err = err.value
        nc = lib().orc_tree_ncells(self._h)
        self.ncell = nc
        self.level = np.zeros(nc, np.int32)
        self.start = np.zeros(nc, np.int32)
        self.count = np.zeros(nc, np.int32)
        self.parent = np.zeros(nc, np.int32)
        self.child = np.zeros((nc, 8), np.int32)
        self.mass = np.zeros(nc, np.float64)
        self.com = np.zeros((nc, 3), np.float64)
        self.c32 = np.zeros((nc, 4), np.float32)
        self.perm = np.zeros(self.n, np.int32)
        self.skeys = np.zeros(self.n, np.uint64)
        self.ext = np.zeros(4, np.float64)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        lib().orc_tree_export(self._h, p(self.level), p(self.start), p(self.count), p(self.parent),
                              p(self.child), p(self.mass), p(self.com), p(self.c32), p(self.perm),
                              p(self.skeys), p(self.ext))
        self.lcell0 = lib().orc_tree_lcell0(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_tree_free(h)
            self._h = None

    def crit0(self, theta):
        return float(lib().orc_crit0(self.lcell0, float(theta)))

    def accel(self, theta, eps2, mac_mode=0, targets=None):
        """FP64 Barnes-Hut accelerations; returns (acc[k,3], ninter[k])."""
        if targets is None:
            k, tp = self.n, None
        else:
            targets = np.ascontiguousarray(targets, dtype=np.int32)
            k, tp = len(targets), targets.ctypes.data_as(C.c_void_p)
        acc = np.zeros(3 * k, np.float64)
        cnt = np.zeros(k, np.int64)
        rc = lib().orc_tree_accel(self._h, float(theta), float(eps2), int(mac_mode), k, tp, acc,
                                  cnt.ctypes.data_as(C.c_void_p))
        if rc != 0:
            raise RuntimeError("orc_tree_accel failed")
        return acc.reshape(k, 3), cnt

    def group_order(self, levels=10):
        """Body ids in the wavefront walk's group order: key order, re-sorted
        stably by the top `levels` Hilbert digits of the key."""
        top = keys_hilbert(self.skeys & np.uint64((1 << 63) - 1)) >> np.uint64(3 * (21 - levels))
        return self.perm[np.argsort(top, kind="stable")].astype(np.int32)

    def group_walk(self, theta, eps2, order=None, group=64):
        """Child steps and frames of the group walk (orc_group_walk) over groups of
        `group` consecutive bodies of `order`; returns (steps[g], frames[g])."""
        order = self.group_order() if order is None else np.ascontiguousarray(order, dtype=np.int32)
        first = np.arange(0, len(order), group, dtype=np.int64)
        cnt = np.minimum(group, len(order) - first).astype(np.int32)
        steps = np.zeros(len(first), np.int64)
        frames = np.zeros(len(first), np.int64)
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        rc = lib().orc_group_walk(self._h, float(theta), float(eps2), len(first), order, p(first), cnt,
                                  p(steps), p(frames))
        if rc == -2:
            raise ValueError("orc_group_walk: a tree of fewer than two bodies has no cell")
        if rc != 0:
            raise MemoryError("orc_group_walk: out of memory")
        return steps, frames

    def vjp(self, theta, eps2, u, mac_mode=0):
        """dL/dx (n,3) and dL/dm (n) for L = sum_i u_i . a_i with the lists held fixed."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        gx = np.zeros(3 * self.n, np.float64)
        gm = np.zeros(self.n, np.float64)
        lib().orc_tree_vjp(self._h, float(theta), float(eps2), int(mac_mode), u, gx, gm)
        return gx.reshape(self.n, 3), gm


def direct_accel(x, y, z, m, eps2, targets=None):
    x, y, z, m = _f32(x), _f32(y), _f32(z), _f32(m)
    if targets is None:
        k, tp = len(x), None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        k, tp = len(targets), targets.ctypes.data_as(C.c_void_p)
    acc = np.zeros(3 * k, np.float64)
    lib().orc_direct_accel(len(x), x, y, z, m, float(eps2), k, tp, acc)
    return acc.reshape(k, 3)


def direct_accel_fast(x, y, z, m, eps2, targets=None):
    """FP64 direct sum, vectorised (agrees with direct_accel to FP64 rounding)."""
    x, y, z, m = _f32(x), _f32(y), _f32(z), _f32(m)
    if targets is None:
        k, tp = len(x), None
    else:
        targets = np.ascontiguousarray(targets, dtype=np.int32)
        k, tp = len(targets), targets.ctypes.data_as(C.c_void_p)
    acc = np.zeros(3 * k, np.float64)
    if lib().orc_direct_accel_fast(len(x), x, y, z, m, float(eps2), k, tp, acc) != 0:
        raise MemoryError("orc_direct_accel_fast: allocation of the FP64 source copies failed")
    return acc.reshape(k, 3)


def update(x, y, z, vx, vy, vz, ax, ay, az, dt, d=1.0):
    arrs = [np.array(a, dtype=np.float32, copy=True) for a in (x, y, z, vx, vy, vz)]
    lib().orc_update(len(arrs[0]), *arrs, _f32(ax), _f32(ay), _f32(az), float(dt), float(d))
    return arrs


class GeoTree:
    """The reference's pointer octree (Tree.cpp insert version 3) restated."""

    def __init__(self, x, y, z, m, center, length):
        self.x, self.y, self.z, self.m = _f64(x), _f64(y), _f64(z), _f64(m)
        self.n = len(self.x)
        err = C.c_int(0)
        cx, cy, cz = (float(c) for c in center)
        self._h = lib().orc_geo_build(self.n, self.x, self.y, self.z, self.m, cx, cy, cz,
                                      float(length), C.byref(err))
        self.err = err.value
        nn = lib().orc_geo_nnodes(self._h)
        self.ints = np.zeros(2 * nn, np.int32)
        self.dbl = np.zeros(8 * nn, np.float64)
        lib().orc_geo_export(self._h, self.ints, self.dbl, None)
        self.ints = self.ints.reshape(nn, 2)
        self.dbl = self.dbl.reshape(nn, 8)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.orc_geo_free(h)
            self._h = None

    def accel_reference_si(self, theta, softening_m, time_step=1.0):
        acc = np.zeros(3 * self.n, np.float64)
        lib().orc_geo_accel(self._h, self.n, self.x, self.y, self.z, self.m, float(theta), 0,
                            TO_METERS, G_SI, float(time_step), float(softening_m), 0.0, acc, None)
        return acc.reshape(self.n, 3)

    def accel_plummer(self, theta, eps2):
        acc = np.zeros(3 * self.n, np.float64)
        lib().orc_geo_accel(self._h, self.n, self.x, self.y, self.z, self.m, float(theta), 1,
                            TO_METERS, G_SI, 1.0, 0.0, float(eps2), acc, None)
        return acc.reshape(self.n, 3)


# ------------------------------------------------------------ reference harness
def have_ref():
    return os.access(REF_HARNESS, os.X_OK)


def ref_keys(x, y, z):
    """Keys and std::map order from the reference's Body::getKey (needs oracle/_ref)."""
    x, y, z = _f32(x), _f32(y), _f32(z)
    n = len(x)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<q", n))
            f.write(x.tobytes()); f.write(y.tobytes()); f.write(z.tobytes())
        subprocess.run([REF_HARNESS, "keys", fin, fout], check=True)
        buf = open(fout, "rb").read()
    n2 = struct.unpack_from("<q", buf, 0)[0]
    keys = np.frombuffer(buf, np.uint64, n2, 8).copy()
    off = 8 + 8 * n2
    k = struct.unpack_from("<q", buf, off)[0]
    order = np.frombuffer(buf, np.int32, k, off + 8).copy()
    return keys, order


def ref_tree(x, y, z, m, theta, softening_m, center, length):
    """Velocity increments and node dump from the reference Tree/Interaction code."""
    x, y, z, m = _f64(x), _f64(y), _f64(z), _f64(m)
    n = len(x)
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<q", n))
            for a in (x, y, z, m):
                f.write(a.tobytes())
        args = [REF_HARNESS, "tree", fin, fout, repr(float(theta)), repr(float(softening_m))]
        args += [repr(float(c)) for c in center] + [repr(float(length))]
        subprocess.run(args, check=True)
        buf = open(fout, "rb").read()
    n2 = struct.unpack_from("<q", buf, 0)[0]
    dv = np.frombuffer(buf, np.float64, 3 * n2, 8).reshape(n2, 3).copy()
    off = 8 + 24 * n2
    nn = struct.unpack_from("<q", buf, off)[0]
    off += 8
    rec = np.dtype([("level", "<i4"), ("external", "<i4"), ("d", "<f8", (8,))])
    nodes = np.frombuffer(buf, rec, nn, off).copy()
    return dv, nodes


# ------------------------------------------------------------ stars ("suns")
def star_forces(suns, bodies, kernel=1, eps2=0.0, softening_m=0.0, time_step=1.0):
    """Direct, symmetric forces between the suns and the bodies (suns are not in
    the tree; Interaction.cpp:111-124).  suns, bodies: (m, x, y, z) arrays.
    kernel 1: Plummer G = 1 accelerations; kernel 0: the reference's SI
    velocity increments.  Returns (sun_acc [ns, 3], body_acc [nb, 3])."""
    sm, sx, sy, sz = (_f64(a) for a in suns)
    bm, bx, by, bz = (_f64(a) for a in bodies)
    sa = np.zeros(3 * len(sm))
    ba = np.zeros(3 * len(bm))
    lib().orc_star_forces(len(sm), sm, sx, sy, sz, len(bm), bm, bx, by, bz, int(kernel), TO_METERS, G_SI,
                          float(time_step), float(softening_m), float(eps2), sa, ba)
    return sa.reshape(-1, 3), ba.reshape(-1, 3)


def update_reference(x, y, z, vx, vy, vz, dv, time_step):
    """Interaction::updateBodies (Interaction.cpp:219-255): v += dv; x += TIME_STEP v / TO_METERS."""
    a = [_f64(v).copy() for v in (x, y, z, vx, vy, vz)]
    acc = _f64(dv).reshape(-1).copy()
    lib().orc_update_reference(len(a[0]), *a, acc, float(time_step), TO_METERS)
    return a


def ref_stars(suns, bodies, softening_m, time_step):
    """The reference's star loop and update (needs oracle/_ref).  suns, bodies:
    (m, x, y, z, vx, vy, vz) arrays.  Returns dict of sun_dv, body_dv, and the
    positions and velocities after Interaction::updateBodies."""
    s = [_f64(a) for a in suns]
    b = [_f64(a) for a in bodies]
    ns, nb = len(s[0]), len(b[0])
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<qq", ns, nb))
            for a in s + b:
                f.write(a.tobytes())
        subprocess.run([REF_HARNESS, "stars", fin, fout, repr(float(softening_m)), str(int(time_step))], check=True)
        buf = open(fout, "rb").read()
    ns2, nb2 = struct.unpack_from("<qq", buf, 0)
    off = 16
    out = {}
    for key, k in (("sun_dv", ns2), ("body_dv", nb2), ("sun_x", ns2), ("sun_v", ns2), ("body_x", nb2),
                   ("body_v", nb2)):
        out[key] = np.frombuffer(buf, np.float64, 3 * k, off).reshape(k, 3).copy()
        off += 24 * k
    return out
