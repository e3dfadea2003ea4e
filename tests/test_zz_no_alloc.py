"""lf_tvnet_forward_noalloc (the allocation-free forward pass, scratch in a
static device array) against lf_tvnet_forward (caller workspace), through the
C ABI: identical outputs on two calls in a row, the second with a smaller
input (the array keeps the first call's contents), and a clear error past the
stated size limit (checked on the host before any launch)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_PIXELS = 655360  # LF_TVNET_NOALLOC_MAX_PIXELS in include/latticeflow_amd.h


def _forward_both(L, params, f, lam):
    import torch
    n, H, W = f.shape
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    need = L.lf_tvnet_work_bytes(n, H, W)
    work = torch.empty(need, dtype=torch.uint8, device="cuda")
    ref = torch.full_like(f, float("nan"))
    out = torch.full_like(f, float("nan"))
    assert L.lf_tvnet_forward(ctypes.c_void_p(params.data_ptr()), n, H, W, ctypes.c_void_p(f.data_ptr()),
                              ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(ref.data_ptr()),
                              ctypes.c_void_p(work.data_ptr()), need, s) == 0
    rc = L.lf_tvnet_forward_noalloc(ctypes.c_void_p(params.data_ptr()), n, H, W, ctypes.c_void_p(f.data_ptr()),
                                    ctypes.c_void_p(lam.data_ptr()), ctypes.c_void_p(out.data_ptr()), s)
    assert rc == 0, L.lf_last_error()
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_forward_matches_workspace_forward():
    import torch
    from latticeflow_amd._lib import lib
    from latticeflow_amd.tvnet import TVNet
    L = lib()
    net = TVNet.load()
    rng = np.random.default_rng(0)
    # a large call first (close to the limit), then a smaller one
    for n, H, W in ((2, 512, 600), (3, 37, 70)):
        print("case", n, H, W, flush=True)
        assert n * H * W <= MAX_PIXELS
        f = torch.from_numpy(rng.uniform(-0.5, 0.5, (n, H, W)).astype(np.float32)).cuda()
        lam = torch.from_numpy(rng.uniform(-3, 1, n).astype(np.float32)).cuda()
        ref, out = _forward_both(L, net.params, f, lam)
        assert np.all(np.isfinite(out))
        assert np.array_equal(out, ref)


def test_noalloc_forward_rejects_oversize():
    import torch
    from latticeflow_amd._lib import lib, LF_ERR_BAD_ARG
    L = lib()
    p = torch.zeros(L.lf_tvnet_param_count(), dtype=torch.float32, device="cuda")
    f = torch.zeros(1, dtype=torch.float32, device="cuda")
    # the size check comes before any device work, so these pointers are never read
    rc = L.lf_tvnet_forward_noalloc(ctypes.c_void_p(p.data_ptr()), 1, 1024, MAX_PIXELS // 1024 + 1,
                                    ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                                    ctypes.c_void_p(f.data_ptr()), None)
    assert rc == LF_ERR_BAD_ARG
    assert b"LF_TVNET_NOALLOC_MAX_PIXELS" in L.lf_last_error()
