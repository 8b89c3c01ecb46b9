"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/nbx.h declares, validates arguments before touching a device, and its
host-only initial-condition generators are seeded and have the right
statistics.  No kernel runs here."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import nns_amd

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "nbx.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbx_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(nns_amd.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(nns_amd.EXPORTS) == declared


def test_one_hip_runtime_whatever_the_import_order():
    """Loading the library before anything else of torch still maps a single
    HIP runtime (torch's): two in one process leave the second without a GPU."""
    code = ("import nns_amd; nns_amd.lib(); import torch; "
            "print(len({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l}))")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300,
                       cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().splitlines()[-1] == "1"


def test_default_params():
    p = nns_amd.default_params()
    assert (p.theta, p.softening, p.gravity, p.timestep, p.dampening, p.kahan) == \
        (pytest.approx(0.5), pytest.approx(1e-3), 1.0, pytest.approx(1e-3), 1.0, 1)


def test_create_validates_before_device():
    L = nns_amd.lib()
    p = nns_amd.default_params()
    h = C.c_void_p()
    assert L.nbx_create(C.byref(p), 0, 10, 0, C.byref(h)) == -1
    assert b"n_max" in L.nbx_last_error(None)
    assert L.nbx_create(C.byref(p), 100, 100, 0, C.byref(h)) == -1
    bad = nns_amd.default_params(theta=-1.0)
    assert L.nbx_create(C.byref(bad), 100, 1000, 0, C.byref(h)) == -1
    for bs in (100, 32, 2048):  # block_size: 0 or a multiple of 64 in [64, 1024]
        bad = nns_amd.default_params(block_size=bs)
        assert L.nbx_create(C.byref(bad), 100, 1000, 0, C.byref(h)) == -1
        assert b"invalid parameters" in L.nbx_last_error(None)
    for fw in (1, 3, 16, -2):  # force_waves: 0, 2, 4 or 8
        bad = nns_amd.default_params(force_waves=fw)
        assert L.nbx_create(C.byref(bad), 100, 1000, 0, C.byref(h)) == -1
        assert b"invalid parameters" in L.nbx_last_error(None)
        sizes = (C.c_int * 1)(100)
        assert L.nbx_batch_create(C.byref(bad), 1, sizes, None, 0, C.byref(h)) == -1
        assert b"force_waves" in L.nbx_batch_last_error(None)
    # the largest system: 47,662 samples of 1024 bodies fill the 2^20-entry top table
    assert L.nbx_create(C.byref(p), 48805889, 2 * 48805889 + 12000, 0, C.byref(h)) == -1
    assert b"maximum of 48,805,888" in L.nbx_last_error(None)
    assert L.nbx_create(C.byref(p), 1000, (1 << 27) + 1, 0, C.byref(h)) == -1
    assert b"maximum of 134,217,728 nodes" in L.nbx_last_error(None)
    assert not h.value


def test_batch_create_validates_before_device():
    L = nns_amd.lib()
    p = nns_amd.default_params()
    h = C.c_void_p()
    ones = (C.c_int * 47663)(*([1] * 47663))  # one sample per system: one over the maximum
    assert L.nbx_batch_create(C.byref(p), 47663, ones, None, 0, C.byref(h)) == -1
    assert b"47,662 samples" in L.nbx_batch_last_error(None)
    sizes = (C.c_int * 2)(5, 0)
    assert L.nbx_batch_create(C.byref(p), 2, sizes, None, 0, C.byref(h)) == -1
    assert b">= 1 body" in L.nbx_batch_last_error(None)
    assert not h.value


def test_valu_probe_validates_before_device():
    L = nns_amd.lib()
    ms = C.c_float()
    buf = C.c_void_p(1)  # never dereferenced: every call below fails its checks first
    assert L.nbx_probe_valu(2, 1, 1, buf, 256, None, C.byref(ms)) == -1      # kind
    assert L.nbx_probe_valu(0, 0, 1, buf, 256, None, C.byref(ms)) == -1      # blocks
    assert L.nbx_probe_valu(0, (1 << 22) + 1, 1, buf, 1 << 31, None, C.byref(ms)) == -1
    assert L.nbx_probe_valu(0, 1, 0, buf, 256, None, C.byref(ms)) == -1      # iters
    assert L.nbx_probe_valu(0, 2, 1, buf, 511, None, C.byref(ms)) == -1      # 256 floats per block
    assert L.nbx_probe_valu(0, 1, 1, None, 256, None, C.byref(ms)) == -1     # out
    assert L.nbx_probe_valu(0, 1, 1, buf, 256, None, None) == -1             # ms


def test_null_context_is_rejected():
    L = nns_amd.lib()
    assert L.nbx_check(None, None) == -1
    assert L.nbx_compute_forces(None, *([None] * 14), 10, 1.0, None, None) == -1


def test_ics_are_seeded():
    a = nns_amd.initial_conditions("plummer", 5000, seed=3)
    b = nns_amd.initial_conditions("plummer", 5000, seed=3)
    c = nns_amd.initial_conditions("plummer", 5000, seed=4)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert not np.array_equal(a[0], c[0])


def test_plummer_statistics():
    n = 6000
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("plummer", n, seed=7)
    m64 = m.astype(np.float64)
    assert m64.sum() == pytest.approx(1.0, rel=1e-6)
    pos = np.stack([x, y, z], 1).astype(np.float64)
    vel = np.stack([vx, vy, vz], 1).astype(np.float64)
    assert np.abs(pos.mean(0)).max() < 1e-6 and np.abs(vel.mean(0)).max() < 1e-6
    r = np.linalg.norm(pos, axis=1)
    # median of U on (0, 0.999] is 0.4995 -> r = (0.4995^(-2/3) - 1)^(-1/2)
    assert np.median(r) == pytest.approx((0.4995 ** (-2 / 3) - 1) ** -0.5, rel=0.04)
    assert r.max() <= (0.999 ** (-2 / 3) - 1) ** -0.5 * 1.001
    # virial ratio T/|W| of an equilibrium sphere is 1/2
    T = 0.5 * (m64 * (vel ** 2).sum(1)).sum()
    W = 0.0
    for i in range(0, n, 1000):
        d = np.linalg.norm(pos[i:i + 1000, None, :] - pos[None, :, :], axis=2)
        np.fill_diagonal(d[:, i:i + 1000], np.inf)
        W -= 0.5 * (m64[i:i + 1000, None] * m64[None, :] / d).sum()
    assert T / abs(W) == pytest.approx(0.5, abs=0.04)
    # isotropy
    assert np.abs((vel ** 2).mean(0) / (vel ** 2).mean() - 1).max() < 0.1


def test_binary_disk_layout():
    n = 20000
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("binary_disk", n, seed=1)
    assert (x[0], x[1], m[0], m[1]) == (np.float32(0.1), np.float32(-0.1), 0.5, 0.5)
    assert vy[0] == pytest.approx(0.5 * np.sqrt(5.0)) and vy[1] == pytest.approx(-0.5 * np.sqrt(5.0))
    r = np.hypot(x[2:], y[2:])
    assert r.min() >= 0.3 * (1 - 1e-6) and r.max() <= 2.0 * (1 + 1e-6)
    assert np.std(z[2:]) == pytest.approx(0.02, rel=0.05)
    assert np.all(m[2:] == np.float32(1e-7))
    v = np.hypot(vx[2:], vy[2:])
    np.testing.assert_allclose(v, 1 / np.sqrt(r), rtol=1e-5)
    # tangential
    assert np.abs(x[2:] * vx[2:] + y[2:] * vy[2:]).max() < 1e-5


def test_star_disk_layout():
    n = 32768
    x, y, z, vx, vy, vz, m = nns_amd.initial_conditions("star_disk", n, seed=1)
    assert m[0] == 1.0 and x[0] == 0 and vx[0] == 0
    r = np.hypot(x[1:], y[1:])
    assert r.min() >= 0.1 * (1 - 1e-6) and r.max() <= 1.0 * (1 + 1e-6)
    assert np.std(z[1:]) == pytest.approx(0.01, rel=0.05)
    assert np.all(m[1:] == np.float32(1e-6))


def test_runner_rejects_bad_options():
    """nns_amd/nbx_runner checks its options before it touches the GPU."""
    import subprocess
    runner = os.path.join(os.path.dirname(nns_amd.__file__), "nbx_runner")
    assert os.access(runner, os.X_OK), "nns_amd/nbx_runner is not built (python -m nns_amd.build --runner)"
    for args in (["-b", "100"], ["-b", "2048"], ["-n", "1"], ["-i", "0"], ["--no-such-option"], ["-n"],
                 ["--suns", "9"], ["--suns", "-1"], ["-n", "3", "--suns", "2"], ["-vx"]):
        r = subprocess.run([runner, *args], capture_output=True, text=True, timeout=60)
        assert r.returncode != 0, args
    r = subprocess.run([runner, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-b blocksize" in r.stdout and "-R interval" in r.stdout
