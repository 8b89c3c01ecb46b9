"""nbx_sort_keys_static, the key sort for callers that cannot allocate (no
context, scratch in a static device array), against nbx_sort_keys through the
C ABI: keys and permutation equal bit for bit, on two calls in a row, the second
with a smaller input; sizes above the maximum rejected with NBX_ERR_INVALID and
a message."""
import ctypes as C

import numpy as np
import pytest

import nns_amd

pytestmark = pytest.mark.gpu

MAX_KEYS = 1 << 20


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible GPU"
    return torch


def _keys(torch, n, seed):
    r = np.random.default_rng(seed)
    k = r.integers(0, 2 ** 63, n, dtype=np.int64).view(np.uint64) | np.uint64(1 << 63)
    k[r.integers(0, n, n // 8)] = k[r.integers(0, n, n // 8)]  # repeated keys: the order must stay stable
    return torch.as_tensor(k.view(np.int64), device="cuda")


def test_static_sort_equals_context_sort():
    torch = _torch()
    L = nns_amd.lib()
    ctx = nns_amd.Context(MAX_KEYS)
    kw = nns_amd.KernelsWrapper(ctx)
    stream = torch.cuda.current_stream().cuda_stream
    for n, seed in ((MAX_KEYS, 1), (300001, 2)):
        print(f"case static sort n={n}", flush=True)
        keys = _keys(torch, n, seed)
        sk_ref = torch.zeros_like(keys)
        p_ref = torch.zeros(n, dtype=torch.int32, device="cuda")
        kw.sortKeys(keys, n, sk_ref, p_ref)
        sk = torch.full_like(keys, -1)
        p = torch.full((n,), -1, dtype=torch.int32, device="cuda")
        rc = L.nbx_sort_keys_static(C.c_void_p(keys.data_ptr()), n, C.c_void_p(sk.data_ptr()),
                                    C.c_void_p(p.data_ptr()), C.c_void_p(stream))
        assert rc == 0, L.nbx_last_error(None)
        torch.cuda.synchronize()
        assert torch.equal(sk, sk_ref) and torch.equal(p, p_ref)
        u = sk.cpu().numpy().view(np.uint64)
        assert (u[1:] >= u[:-1]).all()
    rc = L.nbx_sort_keys_static(C.c_void_p(keys.data_ptr()), MAX_KEYS + 1, C.c_void_p(sk.data_ptr()),
                                C.c_void_p(p.data_ptr()), C.c_void_p(stream))
    assert rc == -1 and b"1,048,576" in L.nbx_last_error(None)
    ctx.close()
