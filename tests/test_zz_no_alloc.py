"""gf_bsr_resid_f64_noalloc (the allocation-free variant of gf_bsr_resid_f64)
gives the same results as gf_bsr_resid_f64 through the C ABI, on two calls in
a row, the second on a smaller problem (the static scratch keeps the first
call's partials; the second must not read them), and refuses a mesh beyond
its scratch limit."""
import ctypes

import numpy as np
import pytest

import gpu_fem_amd as G
from gpu_fem_amd import meshes as M

pytestmark = pytest.mark.gpu


def _resid_both(ctx, mesh, seed):
    import torch

    p = G.Problem(ctx, mesh, reorder=1, lev_fill=0)
    p.assemble()
    d = p.device_arrays()
    n = mesh.nnodes
    rng = np.random.default_rng(seed)
    x = torch.from_numpy(rng.standard_normal(6 * n)).cuda()
    b = torch.from_numpy(rng.standard_normal(6 * n)).cuda()
    L = G.lib()
    mx, nrm = ctypes.c_double(), ctypes.c_double()
    ctx.check(L.gf_bsr_resid_f64(ctx.h, n, d["krowp"], d["kcols"], d["kvals"],
                                 G.ptr(x), G.ptr(b), ctypes.byref(mx),
                                 ctypes.byref(nrm)))
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.check(L.gf_bsr_resid_f64_noalloc(ctx.h, n, d["krowp"], d["kcols"],
                                         d["kvals"], G.ptr(x), G.ptr(b),
                                         G.ptr(out)))
    ctx.sync()
    got = out.cpu().numpy()
    p.close()
    return (mx.value, nrm.value), (float(got[0]), float(got[1]))


def test_noalloc_resid_matches_allocating_entry_point(gpu_ctx):
    print("[gpu] no-alloc residual: plate 60x60 then plate 9x7", flush=True)
    big = M.plate(60, 60, bc="ss", load="sine")
    small = M.plate(9, 7, bc="ss", load="sine")
    for mesh, seed in ((big, 1), (small, 2)):
        ref, got = _resid_both(gpu_ctx, mesh, seed)
        # same kernels, same summation order: bitwise equal
        assert got == ref, (mesh.name, ref, got)
        assert np.isfinite(got[1]) and got[1] > 0


def test_noalloc_resid_refuses_oversize(gpu_ctx):
    # rejected on the host before any launch (the pointers are not used)
    import torch

    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    st = G.lib().gf_bsr_resid_f64_noalloc(
        gpu_ctx.h, 40 * (1 << 17) + 1, G.ptr(dummy), G.ptr(dummy),
        G.ptr(dummy), G.ptr(dummy), G.ptr(dummy), G.ptr(dummy))
    assert st == -1
    msg = G.lib().gf_context_last_error(gpu_ctx.h).decode()
    assert "scratch limit" in msg
