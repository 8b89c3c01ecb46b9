"""nns_amd -- Python front end of the nbx Barnes-Hut engine (libnbx.so, HIP for gfx950).

The engine is the C ABI in include/nbx.h; this module binds it with ctypes.
PyTorch is used only to own device memory and to name the stream.  There is no
CPU fallback: if libnbx.so is missing or no GPU is visible, the calls raise.

Classes:
  Params          nbx_params
  Context         one nbx_context (owns the engine's scratch)
  KernelsWrapper  the reference's launch-wrapper interface (3D/CUDA/CUDA_NBody/
                  include/KernelsWrapper.cuh:15-50): resetArrays ... update, each
                  returning the elapsed milliseconds when timing=True
  BarnesHut       the reference's simulation object (include/BarnesHut.cuh:18-127)
                  over device arrays of the reference layout
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnbx.so")

NBX_OK = 0
ERRORS = {
    -1: "NBX_ERR_INVALID", -2: "NBX_ERR_CAPACITY", -3: "NBX_ERR_BUCKET", -4: "NBX_ERR_STACK",
    -5: "NBX_ERR_NONFINITE", -6: "NBX_ERR_HIP", -7: "NBX_ERR_ORDER",
}
KEY_MORTON, KEY_HILBERT = 0, 1
BUF_KEYS, BUF_SORTED_KEYS, BUF_PERM, BUF_CELL_LEVEL = 1, 2, 3, 4
BUF_CELL_MASS64, BUF_CELL_COM64, BUF_ROOT, BUF_EXTENT = 5, 6, 7, 8

# Exported symbols of include/nbx.h (checked by tests/test_abi.py).
EXPORTS = [
    "nbx_default_params", "nbx_create", "nbx_destroy", "nbx_set_params", "nbx_get_params",
    "nbx_last_error", "nbx_check", "nbx_reset_arrays", "nbx_compute_bounding_box", "nbx_build_tree",
    "nbx_centre_of_mass", "nbx_sort", "nbx_compute_forces", "nbx_update", "nbx_compute_keys",
    "nbx_sort_keys", "nbx_count_interactions", "nbx_step", "nbx_copy_buffer", "nbx_ic_star_disk",
    "nbx_ic_plummer", "nbx_ic_binary_disk", "nbx_autotune", "nbx_batch_create", "nbx_batch_destroy",
    "nbx_batch_last_error", "nbx_batch_bind", "nbx_batch_step", "nbx_batch_graph_step", "nbx_batch_check",
    "nbx_batch_count_interactions", "nbx_batch_autotune", "nbx_batch_copy_buffer", "nbx_force_vjp", "nbx_probe_valu",
    "nbx_star_forces", "nbx_walk_stats", "nbx_batch_walk_stats", "nbx_sort_keys_static",
]


def _walk_dict(v):
    inter, steps, frames, groups = (int(x) for x in v)
    return {"interactions": inter, "child_steps": steps, "frames": frames, "groups": groups,
            "useful_lane_fraction": inter / (64.0 * steps) if steps else 0.0,
            "work_inflation": 64.0 * steps / inter if inter else 0.0}


class NbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(C.Structure):
    _fields_ = [("theta", C.c_float), ("softening", C.c_float), ("gravity", C.c_float),
                ("timestep", C.c_float), ("dampening", C.c_float), ("kahan", C.c_int),
                ("block_size", C.c_int), ("force_waves", C.c_int), ("build_locked", C.c_int)]


STAR_KEYS = ("sx", "sy", "sz", "smass", "svx", "svy", "svz", "sax", "say", "saz")
MAX_STARS = 8  # NBX_MAX_STARS


class Arrays(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("x", "y", "z", "mass", "vx", "vy", "vz", "ax", "ay", "az",
                                          "count", "start", "sorted", "child", "index", "mutex",
                                          "minX", "maxX", "minY", "maxY", "minZ", "maxZ")] + \
               [("n", C.c_int), ("m", C.c_int)] + [(k, C.c_void_p) for k in STAR_KEYS] + [("nstars", C.c_int)]


_lib = None


def lib():
    """Loads libnbx.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m nns_amd.build` (no CPU fallback)")
    # torch first: libnbx.so then binds to the HIP runtime torch has loaded.
    # Loaded the other way round, the process maps two HIP and two HSA
    # runtimes, and the second one to start finds no GPU.
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    vp, i, f, fp = C.c_void_p, C.c_int, C.c_float, C.POINTER(C.c_float)
    L.nbx_default_params.argtypes = [C.POINTER(Params)]
    L.nbx_create.argtypes = [C.POINTER(Params), i, i, i, C.POINTER(vp)]
    L.nbx_destroy.argtypes = [vp]
    L.nbx_set_params.argtypes = [vp, C.POINTER(Params)]
    L.nbx_get_params.argtypes = [vp, C.POINTER(Params)]
    L.nbx_last_error.restype = C.c_char_p
    L.nbx_last_error.argtypes = [vp]
    L.nbx_check.argtypes = [vp, vp]
    L.nbx_reset_arrays.argtypes = [vp] + [vp] * 16 + [i, i, vp, fp]
    L.nbx_compute_bounding_box.argtypes = [vp] + [vp] * 10 + [i, vp, fp]
    L.nbx_build_tree.argtypes = [vp] + [vp] * 14 + [i, i, vp, fp]
    L.nbx_centre_of_mass.argtypes = [vp] + [vp] * 5 + [i, vp, fp]
    L.nbx_sort.argtypes = [vp] + [vp] * 5 + [i, vp, fp]
    L.nbx_compute_forces.argtypes = [vp] + [vp] * 14 + [i, f, vp, fp]
    L.nbx_update.argtypes = [vp] + [vp] * 9 + [i, f, f, vp, fp]
    L.nbx_compute_keys.argtypes = [vp, vp, vp, vp, i, i, vp, vp, fp]
    L.nbx_sort_keys.argtypes = [vp, vp, i, vp, vp, vp, fp]
    L.nbx_count_interactions.argtypes = [vp, i, vp, vp]
    L.nbx_star_forces.argtypes = [vp] + [vp] * 7 + [i] + [vp] * 7 + [i, f, vp, fp]
    L.nbx_walk_stats.argtypes = [vp, vp, vp]
    L.nbx_sort_keys_static.argtypes = [vp, i, vp, vp, vp]
    L.nbx_batch_walk_stats.argtypes = [vp, vp, vp]
    L.nbx_force_vjp.argtypes = [vp, vp, vp, vp, f, vp, vp, vp, vp, vp, fp]
    L.nbx_step.argtypes = [vp, C.POINTER(Arrays), vp, fp]
    L.nbx_copy_buffer.argtypes = [vp, i, vp, C.c_int64, vp]
    for name in ("nbx_ic_star_disk", "nbx_ic_plummer", "nbx_ic_binary_disk"):
        getattr(L, name).argtypes = [i, C.c_uint64] + [vp] * 7
    ip = C.POINTER(C.c_int)
    L.nbx_autotune.argtypes = [vp, vp, ip]
    L.nbx_probe_valu.argtypes = [i, i, i, vp, C.c_int64, vp, fp]
    L.nbx_batch_create.argtypes = [C.POINTER(Params), i, vp, vp, i, C.POINTER(vp)]
    L.nbx_batch_destroy.argtypes = [vp]
    L.nbx_batch_last_error.restype = C.c_char_p
    L.nbx_batch_last_error.argtypes = [vp]
    L.nbx_batch_bind.argtypes = [vp] + [vp] * 10
    L.nbx_batch_step.argtypes = [vp, vp, fp]
    L.nbx_batch_graph_step.argtypes = [vp, vp]
    L.nbx_batch_check.argtypes = [vp, vp]
    L.nbx_batch_count_interactions.argtypes = [vp, vp, vp]
    L.nbx_batch_autotune.argtypes = [vp, vp, ip]
    L.nbx_batch_copy_buffer.argtypes = [vp, i, vp, C.c_int64, vp]
    _lib = L
    return L


def default_params(**kw) -> Params:
    p = Params()
    lib().nbx_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _ptr(t):
    if t is None:
        return None
    if isinstance(t, int):
        return t
    return t.data_ptr()


def _stream(stream):
    if stream is None:
        import torch
        return torch.cuda.current_stream().cuda_stream
    return stream


# ------------------------------------------------------------------ diagnostics
def probe_valu(kind: str, blocks: int = 8192, iters: int = 4096, device: int = 0):
    """Issue-rate probe of the FP32 vector ALU (nbx_probe_valu): kind "fma"
    (v_fma_f32) or "rsq" (v_rsq_f32).  Returns (instructions per second per
    lane-op, kernel ms, the per-lane results as a device tensor)."""
    import torch
    k = {"fma": 0, "rsq": 1}[kind]
    out = torch.empty(256 * blocks, dtype=torch.float32, device=torch.device("cuda", device))
    ms = C.c_float()
    rc = lib().nbx_probe_valu(k, blocks, iters, C.c_void_p(out.data_ptr()), out.numel(), _stream(None), C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"nbx_probe_valu failed ({rc})")
    ops = 256.0 * blocks * iters * 8
    return ops / (ms.value * 1e-3), ms.value, out


# ------------------------------------------------------------------ initial conditions
def initial_conditions(kind: str, n: int, seed: int = 1):
    """Host arrays (x, y, z, vx, vy, vz, mass) of a seeded configuration:
    'star_disk' (1 + n-1 disk bodies), 'plummer', 'binary_disk'."""
    arrs = [np.zeros(n, np.float32) for _ in range(7)]
    fn = {"star_disk": "nbx_ic_star_disk", "plummer": "nbx_ic_plummer",
          "binary_disk": "nbx_ic_binary_disk"}[kind]
    rc = getattr(lib(), fn)(int(n), int(seed), *[a.ctypes.data_as(C.c_void_p) for a in arrs])
    if rc != NBX_OK:
        raise NbxError(rc, f"{fn} failed")
    x, y, z, vx, vy, vz, m = arrs
    return x, y, z, vx, vy, vz, m


# ------------------------------------------------------------------ context
class Context:
    def __init__(self, n_max: int, m_max: int | None = None, params: Params | None = None, device: int = 0):
        self.n_max = int(n_max)
        self.m_max = int(m_max if m_max is not None else 2 * n_max + 12000)
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        rc = lib().nbx_create(C.byref(self.params), self.n_max, self.m_max, int(device), C.byref(h))
        if rc != NBX_OK:
            raise NbxError(rc, lib().nbx_last_error(None).decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nbx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _rc(self, rc):
        if rc != NBX_OK:
            raise NbxError(rc, lib().nbx_last_error(self.h).decode())

    def set_params(self, params: Params):
        self.params = params
        self._rc(lib().nbx_set_params(self.h, C.byref(params)))

    def get_params(self) -> Params:
        """The context's parameters (force_waves: the kernel shape in use)."""
        p = Params()
        self._rc(lib().nbx_get_params(self.h, C.byref(p)))
        return p

    def check(self, stream=None):
        self._rc(lib().nbx_check(self.h, _stream(stream)))

    def copy_buffer(self, which: int, count: int, dtype):
        out = np.zeros(count, dtype)
        self._rc(lib().nbx_copy_buffer(self.h, which, out.ctypes.data_as(C.c_void_p), count, _stream(None)))
        return out


class KernelsWrapper:
    """Mirror of the reference's KernelsWrapper (include/KernelsWrapper.cuh:15-50).

    Arguments are device tensors (or raw device pointers as int) in the
    reference's order; each method returns the elapsed ms when timing=True,
    else 0.0, like the reference.  Errors raise NbxError."""

    def __init__(self, ctx: Context, stream=None):
        self.ctx = ctx
        self.stream = stream

    def _call(self, fn, *args, timing=False):
        ms = C.c_float(0.0)
        rc = fn(self.ctx.h, *args, _stream(self.stream), C.byref(ms) if timing else None)
        self.ctx._rc(rc)
        return ms.value if timing else 0.0

    def resetArrays(self, mutex, x, y, z, mass, count, start, sorted, child, index, minX, maxX, minY,
                    maxY, minZ, maxZ, n, m, timing=False):
        a = [_ptr(t) for t in (mutex, x, y, z, mass, count, start, sorted, child, index, minX, maxX, minY,
                               maxY, minZ, maxZ)]
        return self._call(lib().nbx_reset_arrays, *a, int(n), int(m), timing=timing)

    def computeBoundingBox(self, mutex, x, y, z, minX, maxX, minY, maxY, minZ, maxZ, n, timing=False):
        a = [_ptr(t) for t in (mutex, x, y, z, minX, maxX, minY, maxY, minZ, maxZ)]
        return self._call(lib().nbx_compute_bounding_box, *a, int(n), timing=timing)

    def buildTree(self, x, y, z, mass, count, start, child, index, minX, maxX, minY, maxY, minZ, maxZ, n, m,
                  timing=False):
        a = [_ptr(t) for t in (x, y, z, mass, count, start, child, index, minX, maxX, minY, maxY, minZ, maxZ)]
        return self._call(lib().nbx_build_tree, *a, int(n), int(m), timing=timing)

    def centreOfMass(self, x, y, z, mass, index, n, timing=False):
        a = [_ptr(t) for t in (x, y, z, mass, index)]
        return self._call(lib().nbx_centre_of_mass, *a, int(n), timing=timing)

    def sort(self, count, start, sorted, child, index, n, timing=False):
        a = [_ptr(t) for t in (count, start, sorted, child, index)]
        return self._call(lib().nbx_sort, *a, int(n), timing=timing)

    def computeForces(self, x, y, z, vx, vy, vz, ax, ay, az, mass, sorted, child, minX, maxX, n, g,
                      timing=False):
        a = [_ptr(t) for t in (x, y, z, vx, vy, vz, ax, ay, az, mass, sorted, child, minX, maxX)]
        return self._call(lib().nbx_compute_forces, *a, int(n), float(g), timing=timing)

    def update(self, x, y, z, vx, vy, vz, ax, ay, az, n, dt, d, timing=False):
        a = [_ptr(t) for t in (x, y, z, vx, vy, vz, ax, ay, az)]
        return self._call(lib().nbx_update, *a, int(n), float(dt), float(d), timing=timing)

    # extensions
    def computeKeys(self, x, y, z, n, keys, key_type=KEY_MORTON, timing=False):
        a = [_ptr(t) for t in (x, y, z)]
        return self._call(lib().nbx_compute_keys, *a, int(n), int(key_type), _ptr(keys), timing=timing)

    def sortKeys(self, keys, n, sorted_keys, perm, timing=False):
        return self._call(lib().nbx_sort_keys, _ptr(keys), int(n), _ptr(sorted_keys), _ptr(perm), timing=timing)

    def starForces(self, sx, sy, sz, smass, sax, say, saz, ns, x, y, z, mass, ax, ay, az, n, g, timing=False):
        """Stars outside the tree (nbx_star_forces; the reference's suns,
        3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:111-124)."""
        a = [_ptr(t) for t in (sx, sy, sz, smass, sax, say, saz)]
        b = [_ptr(t) for t in (x, y, z, mass, ax, ay, az)]
        return self._call(lib().nbx_star_forces, *a, int(ns), *b, int(n), float(g), timing=timing)


STAGES = ("reset", "bbox", "build", "com", "sort", "force", "update")


class BarnesHut:
    """Device arrays in the reference layout (src/BarnesHut.cu:62-102) and the
    step of BarnesHut::update (src/BarnesHut.cu:195-301), without the per-step
    device-to-host copy.  numNodes m = 2N + 12000 unless given."""

    def __init__(self, x, y, z, vx, vy, vz, mass, params: Params | None = None, m=None, device=0, ctx=None,
                 stars=None):
        """stars: optional (x, y, z, vx, vy, vz, mass) host arrays of up to
        MAX_STARS bodies kept outside the tree, the reference's suns
        (3D/OpenMP/OpenMP_NBody/src/Interaction.cpp:111-124); update() then
        couples them to the bodies directly and moves them with the bodies."""
        import torch
        n = len(x)
        self.n = n
        self.m = int(m if m is not None else 2 * n + 12000)
        self.dev = torch.device("cuda", device)
        self.ctx = ctx if ctx is not None else Context(n, self.m, params, device)
        self.kw = KernelsWrapper(self.ctx)
        f32 = dict(dtype=torch.float32, device=self.dev)
        i32 = dict(dtype=torch.int32, device=self.dev)
        self.arr = {}
        for k in ("x", "y", "z", "mass", "vx", "vy", "vz", "ax", "ay", "az"):
            self.arr[k] = torch.zeros(self.m, **f32)
        for k, v in (("x", x), ("y", y), ("z", z), ("vx", vx), ("vy", vy), ("vz", vz), ("mass", mass)):
            self.arr[k][:n] = torch.as_tensor(np.asarray(v, np.float32), device=self.dev)
        self.arr["child"] = torch.zeros(8 * self.m, **i32)
        for k in ("count", "start", "sorted"):
            self.arr[k] = torch.zeros(self.m, **i32)
        for k in ("index", "mutex"):
            self.arr[k] = torch.zeros(1, **i32)
        for k in ("minX", "maxX", "minY", "maxY", "minZ", "maxZ"):
            self.arr[k] = torch.zeros(1, **f32)
        self.nstars = 0 if stars is None else len(stars[0])
        if self.nstars > MAX_STARS:
            raise ValueError(f"at most {MAX_STARS} stars")
        if self.nstars:
            for k, v in zip(("sx", "sy", "sz", "svx", "svy", "svz", "smass"), stars):
                self.arr[k] = torch.as_tensor(np.asarray(v, np.float32), device=self.dev).contiguous()
            for k in ("sax", "say", "saz"):
                self.arr[k] = torch.zeros(self.nstars, **f32)
        self._arrays = Arrays(**{k: v.data_ptr() for k, v in self.arr.items()}, n=n, m=self.m, nstars=self.nstars)
        self.step_count = 0

    def __getattr__(self, k):
        arr = self.__dict__.get("arr")
        if arr is not None and k in arr:
            return arr[k]
        raise AttributeError(k)

    def update(self, stage_times=False, stream=None):
        """One step; returns the 7 per-stage times (ms) when stage_times."""
        ms = (C.c_float * 7)() if stage_times else None
        rc = lib().nbx_step(self.ctx.h, C.byref(self._arrays), _stream(stream), ms)
        self.ctx._rc(rc)
        self.step_count += 1
        return list(ms) if stage_times else None

    def stages(self, timing=False):
        """The same step through the seven KernelsWrapper calls (reference order)."""
        a, kw, n, m = self.arr, self.kw, self.n, self.m
        p = self.ctx.params
        t = [kw.resetArrays(a["mutex"], a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["sorted"],
                            a["child"], a["index"], a["minX"], a["maxX"], a["minY"], a["maxY"], a["minZ"],
                            a["maxZ"], n, m, timing=timing),
             kw.computeBoundingBox(a["mutex"], a["x"], a["y"], a["z"], a["minX"], a["maxX"], a["minY"],
                                   a["maxY"], a["minZ"], a["maxZ"], n, timing=timing),
             kw.buildTree(a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["child"], a["index"],
                          a["minX"], a["maxX"], a["minY"], a["maxY"], a["minZ"], a["maxZ"], n, m, timing=timing),
             kw.centreOfMass(a["x"], a["y"], a["z"], a["mass"], a["index"], n, timing=timing),
             kw.sort(a["count"], a["start"], a["sorted"], a["child"], a["index"], n, timing=timing),
             kw.computeForces(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"],
                              a["mass"], a["sorted"], a["child"], a["minX"], a["maxX"], n, p.gravity, timing=timing),
             kw.update(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"], n,
                       p.timestep, p.dampening, timing=timing)]
        self.step_count += 1
        return t

    def count_interactions(self):
        import torch
        out = torch.zeros(self.n, dtype=torch.int64, device=self.dev)
        self.ctx._rc(lib().nbx_count_interactions(self.ctx.h, self.n, out.data_ptr(), _stream(None)))
        return out

    def walk_stats(self):
        """Work of the force walk on the last tree (nbx_walk_stats): interactions,
        child steps, frames, groups, and the useful fraction of the lane slots."""
        v = (C.c_int64 * 4)()
        self.ctx._rc(lib().nbx_walk_stats(self.ctx.h, v, _stream(None)))
        return _walk_dict(v)

    def rebuild(self, timing=False):
        """Reset, bounding box, build, centre of mass and sort of the current
        positions (the tree that nbx_compute_forces and nbx_force_vjp walk)."""
        a, kw, n, m = self.arr, self.kw, self.n, self.m
        return [kw.resetArrays(a["mutex"], a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["sorted"],
                               a["child"], a["index"], a["minX"], a["maxX"], a["minY"], a["maxY"], a["minZ"],
                               a["maxZ"], n, m, timing=timing),
                kw.computeBoundingBox(a["mutex"], a["x"], a["y"], a["z"], a["minX"], a["maxX"], a["minY"],
                                      a["maxY"], a["minZ"], a["maxZ"], n, timing=timing),
                kw.buildTree(a["x"], a["y"], a["z"], a["mass"], a["count"], a["start"], a["child"], a["index"],
                             a["minX"], a["maxX"], a["minY"], a["maxY"], a["minZ"], a["maxZ"], n, m,
                             timing=timing),
                kw.centreOfMass(a["x"], a["y"], a["z"], a["mass"], a["index"], n, timing=timing),
                kw.sort(a["count"], a["start"], a["sorted"], a["child"], a["index"], n, timing=timing)]

    def forces(self, timing=False):
        a, p = self.arr, self.ctx.params
        return self.kw.computeForces(a["x"], a["y"], a["z"], a["vx"], a["vy"], a["vz"], a["ax"], a["ay"], a["az"],
                                     a["mass"], a["sorted"], a["child"], a["minX"], a["maxX"], self.n, p.gravity,
                                     timing=timing)

    def vjp(self, ux, uy, uz, g=None, timing=False):
        """dL/dx, dL/dy, dL/dz, dL/dm for u = dL/da over the tree of the last sort
        (device tensors of n floats; nbx_force_vjp)."""
        import torch
        out = [torch.empty(self.n, dtype=torch.float32, device=self.dev) for _ in range(4)]
        us = [u.contiguous() for u in (ux, uy, uz)]
        ms = C.c_float(0.0)
        gv = self.ctx.params.gravity if g is None else g
        self.ctx._rc(lib().nbx_force_vjp(self.ctx.h, *[u.data_ptr() for u in us], C.c_float(gv),
                                         *[o.data_ptr() for o in out], _stream(None),
                                         C.byref(ms) if timing else None))
        return (*out, ms.value) if timing else tuple(out)

    def host(self, k, count=None):
        t = self.arr[k]
        if count is None:
            count = self.nstars if k in STAR_KEYS else self.n
        return t[:count].cpu().numpy()


class Batch:
    """Ensemble of independent systems in concatenated device arrays
    (nbx_batch_*).  sys_n: bodies per system; sys_theta: per-system theta."""

    def __init__(self, x, y, z, vx, vy, vz, mass, sys_n, sys_theta=None, params: Params | None = None, device=0):
        import torch
        self.sys_n = np.ascontiguousarray(sys_n, dtype=np.int32)
        self.S = len(self.sys_n)
        self.N = int(self.sys_n.sum())
        self.off = np.concatenate([[0], np.cumsum(self.sys_n)]).astype(np.int64)
        self.params = params if params is not None else default_params()
        th = None if sys_theta is None else np.ascontiguousarray(sys_theta, dtype=np.float32)
        self.sys_theta = th
        h = C.c_void_p()
        rc = lib().nbx_batch_create(C.byref(self.params), self.S, self.sys_n.ctypes.data_as(C.c_void_p),
                                    None if th is None else th.ctypes.data_as(C.c_void_p), int(device), C.byref(h))
        if rc != NBX_OK:
            raise NbxError(rc, lib().nbx_batch_last_error(None).decode())
        self.h = h
        self.dev = torch.device("cuda", device)
        self.arr = {}
        for k, v in (("x", x), ("y", y), ("z", z), ("vx", vx), ("vy", vy), ("vz", vz), ("mass", mass)):
            self.arr[k] = torch.as_tensor(np.asarray(v, np.float32), device=self.dev).contiguous()
        for k in ("ax", "ay", "az"):
            self.arr[k] = torch.zeros(self.N, dtype=torch.float32, device=self.dev)
        a = self.arr
        self._rc(lib().nbx_batch_bind(self.h, *[a[k].data_ptr() for k in
                                                ("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az", "mass")]))

    def _rc(self, rc):
        if rc != NBX_OK:
            raise NbxError(rc, lib().nbx_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().nbx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, stage_times=False, stream=None):
        ms = (C.c_float * 7)() if stage_times else None
        self._rc(lib().nbx_batch_step(self.h, _stream(stream), ms))
        return list(ms) if stage_times else None

    def graph_step(self, stream=None):
        self._rc(lib().nbx_batch_graph_step(self.h, _stream(stream)))

    def check(self, stream=None):
        self._rc(lib().nbx_batch_check(self.h, _stream(stream)))

    def autotune(self, stream=None):
        w = C.c_int(0)
        self._rc(lib().nbx_batch_autotune(self.h, _stream(stream), C.byref(w)))
        return w.value

    def count_interactions(self):
        import torch
        out = torch.zeros(self.N, dtype=torch.int64, device=self.dev)
        self._rc(lib().nbx_batch_count_interactions(self.h, out.data_ptr(), _stream(None)))
        return out

    def walk_stats(self):
        v = (C.c_int64 * 4)()
        self._rc(lib().nbx_batch_walk_stats(self.h, v, _stream(None)))
        return _walk_dict(v)

    def copy_buffer(self, which, count, dtype):
        out = np.zeros(count, dtype)
        self._rc(lib().nbx_batch_copy_buffer(self.h, which, out.ctypes.data_as(C.c_void_p), count, _stream(None)))
        return out

    def host(self, k):
        return self.arr[k].cpu().numpy()


def autotune(ctx: Context, stream=None) -> int:
    w = C.c_int(0)
    ctx._rc(lib().nbx_autotune(ctx.h, _stream(stream), C.byref(w)))
    return w.value
