"""GPU parity of the HIP kernels against the oracle, called through the C ABI.

Tolerances (written here, per the task's contract):
  - index arrays: bit-exact
  - Jacobian / residual values: 1e-12 relative, measured per entry against
    sqrt(|K_ii| |K_jj|) of the oracle (diagonal scaling; the 6 DOF types
    differ by ~10 orders of magnitude) and against max|K| globally
  - BCs, value copy, permutation: bit-exact
  - SpMV in FP64: 1e-12 relative; FP32: 1e-5 relative
  - ILU factors: the defining property (L U)_ij = A_ij on the pattern to
    1e-10 max|A| for the GPU factors themselves, and agreement with the
    oracle's factors to 1e-9 of each block row's largest entry (the device
    contracts a*b+c into FMA, the host compiler does not, and the
    factorisation amplifies the rounding difference); sweeps 1e-10
"""
import ctypes
import sys

import numpy as np
import pytest

import gpu_fem_amd as G
import oracle
from gpu_fem_amd import meshes as M

from gpu_helpers import block_diag_scale, dev, host, max_rel_scaled, oracle_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def warped_plate(nx, ny, seed=0):
    m = M.plate(nx, ny, 1.0, 0.8, thick=0.004)
    rng = np.random.default_rng(seed)
    m.xpts[:, 2] = rng.uniform(-0.02, 0.02, m.nnodes)
    m.xpts[:, :2] += rng.uniform(-0.01, 0.01, (m.nnodes, 2))
    m.data[:, 2] = rng.uniform(0.002, 0.01, m.nelems)
    return m


MESHES = {
    "plate10": lambda: M.plate(10, 10),
    "plate_ccw": lambda: M.plate(6, 5, ccw=True),
    "warped": lambda: warped_plate(9, 7),
    "cylinder": lambda: M.cylinder(16, 12, seed=3),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_jacobian_and_residual(gpu_ctx, name):
    import torch

    say(f"[gpu] jacobian parity mesh={name}")
    m = MESHES[name]()
    s = oracle_problem(m, "rcm", 0)
    nnzb = int(s["krp"][-1])
    rng = np.random.default_rng(7)
    u = rng.standard_normal(6 * m.nnodes) * 1e-4
    kv_o, res_o = oracle.assemble_jacobian(m.conn, m.xpts, m.data, s["emap"],
                                           s["perm"], nnzb, vars=u)
    d_conn, d_x = dev(m.conn, np.int32), dev(m.xpts, np.float64)
    d_data, d_u = dev(m.data, np.float64), dev(u)
    d_map, d_perm = dev(s["emap"]), dev(s["perm"])
    d_k = torch.zeros(36 * nnzb, dtype=torch.float64, device="cuda")
    d_r = torch.zeros(6 * m.nnodes, dtype=torch.float64, device="cuda")
    gpu_ctx.check(G.lib().gf_assemble_jacobian_mitc4_f64(
        gpu_ctx.h, m.nelems, m.nnodes, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u),
        G.ptr(d_data), G.ptr(d_map), G.ptr(d_perm), G.ptr(d_k), G.ptr(d_r)))
    gpu_ctx.sync()
    kv = host(d_k)
    scale = block_diag_scale(s["krp"], s["kcp"], kv_o)
    assert max_rel_scaled(kv, kv_o, scale) <= 1e-12
    assert np.abs(kv - kv_o).max() <= 1e-12 * np.abs(kv_o).max()
    res = host(d_r)
    assert np.abs(res - res_o).max() <= 1e-12 * np.abs(res_o).max()

    # coloured, atomic-free scatter: K and res start dirty (no zeroing needed)
    import ctypes
    plan = ctypes.c_void_p()
    gpu_ctx.check(G.lib().gf_asm_plan_create(
        gpu_ctx.h, m.nelems, m.nnodes, G.ptr(np.ascontiguousarray(m.conn, np.int32)),
        G.ptr(np.ascontiguousarray(s["emap"], np.int32)), nnzb, ctypes.byref(plan)))
    assert 2 <= G.lib().gf_asm_plan_colors(plan) <= 16
    d_k2 = torch.full_like(d_k, 3.0)
    d_rc = torch.full_like(d_r, -2.0)
    gpu_ctx.check(G.lib().gf_assemble_jacobian_colored_f64(
        gpu_ctx.h, plan, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u), G.ptr(d_data),
        G.ptr(d_map), G.ptr(d_perm), G.ptr(d_k2), G.ptr(d_rc)))
    gpu_ctx.sync()
    G.lib().gf_asm_plan_destroy(plan)
    kv2 = host(d_k2)
    assert max_rel_scaled(kv2, kv_o, scale) <= 1e-12
    assert np.abs(kv2 - kv_o).max() <= 1e-12 * np.abs(kv_o).max()
    assert np.abs(host(d_rc) - res_o).max() <= 1e-12 * np.abs(res_o).max()

    # residual-only kernel
    d_r2 = torch.zeros_like(d_r)
    gpu_ctx.check(G.lib().gf_assemble_residual_mitc4_f64(
        gpu_ctx.h, m.nelems, m.nnodes, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u),
        G.ptr(d_data), G.ptr(d_perm), G.ptr(d_r2)))
    r_o = oracle.assemble_residual(m.conn, m.xpts, u, m.data, s["perm"])
    gpu_ctx.sync()
    assert np.abs(host(d_r2) - r_o).max() <= 1e-12 * np.abs(r_o).max()

    # energy
    d_e = torch.zeros(1, dtype=torch.float64, device="cuda")
    gpu_ctx.check(G.lib().gf_energy_mitc4_f64(
        gpu_ctx.h, m.nelems, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u),
        G.ptr(d_data), G.ptr(d_e)))
    gpu_ctx.sync()
    e_o = sum(oracle.element(m.xpts[c], u.reshape(-1, 6)[c].ravel(),
                             m.data[i])[2] for i, c in enumerate(m.conn))
    assert abs(host(d_e)[0] - e_o) <= 1e-12 * abs(e_o)


def test_bc_copy_permute_spmv(gpu_ctx):
    import torch

    say("[gpu] bc/copy/permute/spmv plate 12x9 ss")
    m = M.plate(12, 9, bc="ss", load="sine")
    s = oracle_problem(m, "rcm", 2)
    n = m.nnodes
    nnzb = int(s["krp"][-1])
    rng = np.random.default_rng(5)
    raw = rng.standard_normal(36 * nnzb)
    ref = oracle.apply_bcs_mat(m.bcs, s["krp"], s["kcp"], s["tr"], s["tc"],
                               s["tm"], s["perm"], raw)
    d = {k: dev(s[k]) for k in ("krp", "kcp", "tr", "tc", "tm", "perm", "mrp",
                                "mcp")}
    d_bcs = dev(m.bcs, np.int32)
    d_v = dev(raw)
    L = G.lib()
    gpu_ctx.check(L.gf_bc_apply_mat(gpu_ctx.h, len(m.bcs), G.ptr(d_bcs),
                                    G.ptr(d["krp"]), G.ptr(d["kcp"]),
                                    G.ptr(d["tr"]), G.ptr(d["tc"]),
                                    G.ptr(d["tm"]), G.ptr(d["perm"]),
                                    G.ptr(d_v)))
    gpu_ctx.sync()
    np.testing.assert_array_equal(host(d_v), ref)
    # copy into the ILU(2) pattern (fill zeroed even if the target is dirty)
    d_t = torch.full((36 * int(s["mrp"][-1]),), 7.0, dtype=torch.float64,
                     device="cuda")
    gpu_ctx.check(L.gf_bsr_copy_to_pattern(gpu_ctx.h, n, G.ptr(d["krp"]),
                                           G.ptr(d["kcp"]), G.ptr(d_v),
                                           G.ptr(d["mrp"]), G.ptr(d["mcp"]),
                                           G.ptr(d_t)))
    gpu_ctx.sync()
    np.testing.assert_array_equal(
        host(d_t), oracle.copy_values(s["krp"], s["kcp"], ref, s["mrp"], s["mcp"]))
    # and into a complete-fill pattern whose rows (up to 161 blocks) take
    # up to three chunks of 64 target blocks
    m2 = M.plate(40, 40)
    rp2, cp2 = oracle.node_graph(m2.conn, m2.nnodes)
    krp2, kcp2 = oracle.permute_pattern(rp2, cp2, oracle.rcm(rp2, cp2))
    frp, fcp, _ = oracle.iluk(krp2, kcp2, 1 << 29)
    assert np.diff(frp).max() > 64
    v2 = rng.standard_normal(36 * int(krp2[-1]))
    d_f = torch.full((36 * int(frp[-1]),), -3.0, dtype=torch.float64,
                     device="cuda")
    d_krp2, d_kcp2, d_v2 = dev(krp2, np.int32), dev(kcp2, np.int32), dev(v2)
    d_frp, d_fcp = dev(frp, np.int32), dev(fcp, np.int32)
    gpu_ctx.check(L.gf_bsr_copy_to_pattern(gpu_ctx.h, m2.nnodes, G.ptr(d_krp2),
                                           G.ptr(d_kcp2), G.ptr(d_v2),
                                           G.ptr(d_frp), G.ptr(d_fcp),
                                           G.ptr(d_f)))
    gpu_ctx.sync()
    np.testing.assert_array_equal(
        host(d_f), oracle.copy_values(krp2, kcp2, v2, frp, fcp))
    # vector BCs and permutation
    v = rng.standard_normal(6 * n)
    d_vec = dev(v)
    gpu_ctx.check(L.gf_bc_apply_vec(gpu_ctx.h, len(m.bcs), G.ptr(d_bcs), None,
                                    G.ptr(d_vec)))
    d_out = torch.zeros_like(d_vec)
    gpu_ctx.check(L.gf_permute_vec(gpu_ctx.h, n, G.ptr(d["perm"]), G.ptr(d_vec),
                                   G.ptr(d_out)))
    gpu_ctx.sync()
    vb = oracle.apply_bcs_vec(m.bcs, v)
    np.testing.assert_array_equal(host(d_vec), vb)
    np.testing.assert_array_equal(host(d_out), oracle.permute_vec(s["perm"], vb))
    # SpMV FP64 and FP32
    x = rng.standard_normal(6 * n)
    d_x, d_y = dev(x), torch.zeros(6 * n, dtype=torch.float64, device="cuda")
    gpu_ctx.check(L.gf_bsr_spmv_f64(gpu_ctx.h, n, G.ptr(d["krp"]),
                                    G.ptr(d["kcp"]), G.ptr(d_v), G.ptr(d_x),
                                    G.ptr(d_y)))
    gpu_ctx.sync()
    y_o = oracle.bsr_spmv(s["krp"], s["kcp"], ref, x)
    assert np.abs(host(d_y) - y_o).max() <= 1e-12 * np.abs(y_o).max()
    d_v32 = torch.zeros(36 * nnzb, dtype=torch.float32, device="cuda")
    gpu_ctx.check(L.gf_convert_f64_f32(gpu_ctx.h, 36 * nnzb, G.ptr(d_v),
                                       G.ptr(d_v32)))
    d_x32 = d_x.float()
    d_y32 = torch.zeros(6 * n, dtype=torch.float32, device="cuda")
    gpu_ctx.check(L.gf_bsr_spmv_f32(gpu_ctx.h, n, G.ptr(d["krp"]),
                                    G.ptr(d["kcp"]), G.ptr(d_v32),
                                    G.ptr(d_x32), G.ptr(d_y32)))
    gpu_ctx.sync()
    assert np.abs(host(d_y32) - y_o).max() <= 1e-5 * np.abs(y_o).max()


@pytest.mark.parametrize("k,reorder", [(0, "natural"), (1, "rcm"), (3, "rcm")])
def test_ilu_factor_and_apply(gpu_ctx, k, reorder):
    import torch

    say(f"[gpu] ilu k={k} reorder={reorder} plate 14x11 ss")
    m = M.plate(14, 11, bc="ss", load="sine")
    s = oracle_problem(m, reorder, k)
    lu_o, st = oracle.bsr_ilu(s["mrp"], s["mcp"], s["mv"])
    assert st == 0
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    d_m = dev(s["mv"])
    ilu.factor(d_m)
    gpu_ctx.sync()
    lu = host(d_m)
    # per-block-row relative agreement with the oracle
    rp = s["mrp"]
    for i in range(len(rp) - 1):
        a, b = lu[36 * rp[i]:36 * rp[i + 1]], lu_o[36 * rp[i]:36 * rp[i + 1]]
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    # ILU property of the GPU factors
    from test_oracle_numeric import bsr_to_dense
    Dm = bsr_to_dense(s["mrp"], s["mcp"], lu)
    A = bsr_to_dense(s["mrp"], s["mcp"], s["mv"])
    pat = bsr_to_dense(s["mrp"], s["mcp"], np.ones_like(lu)) != 0
    LU = (np.tril(Dm, -1) + np.eye(len(Dm))) @ np.triu(Dm)
    assert np.abs(LU - A)[pat].max() <= 1e-10 * np.abs(A).max()
    b = np.random.default_rng(9).standard_normal(6 * m.nnodes)
    x_o = oracle.bsr_trsv(s["mrp"], s["mcp"], lu_o, b)
    d_b, d_t, d_x = dev(b), torch.zeros(len(b), dtype=torch.float64,
                                        device="cuda"), torch.zeros(
        len(b), dtype=torch.float64, device="cuda")
    ilu.apply(d_m, d_b, d_t, d_x)
    gpu_ctx.sync()
    assert np.abs(host(d_x) - x_o).max() <= 1e-10 * np.abs(x_o).max()
    # FP32 factorisation and sweeps: valid FP32 ILU (property to 1e-5
    # max|A|) and backward-stable sweeps (componentwise backward error of the
    # two triangular solves within 1e-6; forward errors depend on the
    # conditioning of the factors, ~8% for natural-order ILU(0) here)
    d_m32 = dev(s["mv"].astype(np.float32))
    ilu.factor(d_m32)
    gpu_ctx.sync()
    lu32 = host(d_m32).astype(np.float64)
    D32 = bsr_to_dense(s["mrp"], s["mcp"], lu32)
    L32, U32 = np.tril(D32, -1) + np.eye(len(D32)), np.triu(D32)
    assert np.abs(L32 @ U32 - A)[pat].max() <= 1e-5 * np.abs(A).max()
    d_b32 = d_b.float()
    d_t32, d_x32 = torch.zeros_like(d_b32), torch.zeros_like(d_b32)
    ilu.apply(d_m32, d_b32, d_t32, d_x32)
    gpu_ctx.sync()
    x32 = host(d_x32).astype(np.float64)
    b32 = b.astype(np.float32).astype(np.float64)
    bwd = np.linalg.norm(L32 @ (U32 @ x32) - b32) / np.linalg.norm(
        np.abs(L32) @ (np.abs(U32) @ np.abs(x32)))
    assert bwd <= 1e-6
    ilu.close()


@pytest.mark.parametrize("reorder,k", [("rcm", 3), ("amd", 12)])
def test_ilu_lds_rows_match_global_rows(gpu_ctx, reorder, k, monkeypatch):
    """Rows factored in LDS (ilu_row_lds_kernel) and in global memory
    (ilu_row_kernel, GPUFEM_ILU_GLOBAL) perform the same operations: the
    factors agree bit for bit, FP64 and FP32."""
    say(f"[gpu] ilu lds vs global rows {reorder} ILU({k}) plate 24x24")
    m = M.plate(24, 24, bc="ss", load="sine")
    s = oracle_problem(m, reorder, k)
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    for dt in (np.float64, np.float32):
        d_a = dev(s["mv"].astype(dt))
        ilu.factor(d_a)
        gpu_ctx.sync()
        monkeypatch.setenv("GPUFEM_ILU_GLOBAL", "1")
        d_b = dev(s["mv"].astype(dt))
        ilu.factor(d_b)
        gpu_ctx.sync()
        monkeypatch.delenv("GPUFEM_ILU_GLOBAL")
        a, b = host(d_a), host(d_b)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    ilu.close()


@pytest.mark.parametrize("nbatch", [1, 3, 8])
def test_ilu_batched_pivots_match_per_pivot_rows(gpu_ctx, nbatch, monkeypatch):
    """Rows in global memory with the pivots applied in batches
    (ilu_row_batch_kernel) and one pivot at a time (ilu_row_kernel): every
    entry sees the same operations in the same order, so the factors agree
    bit for bit, FP64 and FP32, on both schedules.  Complete LU of a small
    wingbox: rows of up to 250 blocks, up to ~200 pivots a row."""
    say(f"[gpu] ilu batched vs per-pivot global rows, batch {nbatch}")
    m = M.wingbox(nc=24, nh=4, nsb=4, nribs=8, seed=5)
    s = oracle_problem(m, "amd", 1 << 29)
    monkeypatch.setenv("GPUFEM_ILU_GLOBAL", "1")
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    for dt in (np.float64, np.float32):
        monkeypatch.setenv("GPUFEM_ILU_GLOBAL_KERNEL", "pivot")
        d_ref = dev(s["mv"].astype(dt))
        ilu.factor(d_ref)
        gpu_ctx.sync()
        ref = host(d_ref)
        monkeypatch.delenv("GPUFEM_ILU_GLOBAL_KERNEL")
        monkeypatch.setenv("GPUFEM_ILU_BATCH_N", str(nbatch))
        for sched in ("flow", "level"):
            if sched == "level":
                monkeypatch.setenv("GPUFEM_ILU_SCHEDULE", "level")
            d_a = dev(s["mv"].astype(dt))
            ilu.factor(d_a)
            gpu_ctx.sync()
            a = host(d_a)
            assert np.array_equal(a.view(np.uint8), ref.view(np.uint8)), \
                f"{np.dtype(dt).name} {sched} batch {nbatch}"
        monkeypatch.delenv("GPUFEM_ILU_SCHEDULE")
        monkeypatch.delenv("GPUFEM_ILU_BATCH_N")
    ilu.close()


ILU_SCHEDULE_CASES = [
    # (name, mesh, ordering, level of fill, LDS budget or None, global rows)
    ("plate24_amd_ilu12", lambda: M.plate(24, 24, bc="ss", load="sine"),
     "amd", 12, None, False),
    ("plate24_rcm_ilu3_global", lambda: M.plate(24, 24, bc="ss", load="sine"),
     "rcm", 3, None, True),
    # complete factorisation (rows up to 250 blocks): several LDS phases
    ("wingbox2176_amd_lu", lambda: M.wingbox(nc=24, nh=4, nsb=4, nribs=8,
                                             seed=5),
     "amd", 1 << 29, None, False),
    # rows over 160 blocks in global memory: LDS and global phases mixed
    ("wingbox2176_amd_lu_mixed", lambda: M.wingbox(nc=24, nh=4, nsb=4,
                                                   nribs=8, seed=5),
     "amd", 1 << 29, 48 * 1024, False),
]


@pytest.mark.parametrize("case", ILU_SCHEDULE_CASES, ids=lambda c: c[0])
def test_ilu_dataflow_matches_level_schedule(gpu_ctx, case, monkeypatch):
    """The dataflow factorisation (rows taken in level order, each waiting
    only for its pivot rows) and the level-by-level factorisation are two
    schedules of the same row arithmetic: the factors agree bit for bit,
    FP64 and FP32, over repeated factorisations (flags reused by epoch)."""
    name, make, reorder, k, cap, glob = case
    say(f"[gpu] ilu dataflow vs level schedule {name}")
    m = make()
    s = oracle_problem(m, reorder, k)
    if glob:
        monkeypatch.setenv("GPUFEM_ILU_GLOBAL", "1")
    if cap is not None:
        monkeypatch.setenv("GPUFEM_ILU_LDS_CAP", str(cap))
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    for dt in (np.float64, np.float32):
        monkeypatch.setenv("GPUFEM_ILU_SCHEDULE", "level")
        d_ref = dev(s["mv"].astype(dt))
        ilu.factor(d_ref)
        gpu_ctx.sync()
        ref = host(d_ref)
        monkeypatch.delenv("GPUFEM_ILU_SCHEDULE")
        for rep in range(2):
            d_a = dev(s["mv"].astype(dt))
            ilu.factor(d_a)
            gpu_ctx.sync()
            a = host(d_a)
            assert np.array_equal(a.view(np.uint8), ref.view(np.uint8)), \
                f"{name} {np.dtype(dt).name} rep {rep}"
    ilu.close()


def test_ilu_dataflow_complete_factors(gpu_ctx):
    """Complete factorisation (every fill kept) on the dataflow schedule:
    the factors match the oracle's block LU row by row, and L U reproduces
    the matrix everywhere (an exact LU, no dropped fill)."""
    say("[gpu] ilu dataflow complete LU wingbox 254 nodes")
    from test_oracle_numeric import bsr_to_dense

    m = M.wingbox(nc=10, nh=3, nsb=2, nribs=4, seed=5)
    s = oracle_problem(m, "amd", 1 << 29)
    lu_o, st = oracle.bsr_ilu(s["mrp"], s["mcp"], s["mv"])
    assert st == 0
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    d_lu = dev(s["mv"].copy())
    ilu.factor(d_lu)
    gpu_ctx.sync()
    lu = host(d_lu)
    ilu.close()
    rp = s["mrp"]
    for i in range(len(rp) - 1):
        a, b = lu[36 * rp[i]:36 * rp[i + 1]], lu_o[36 * rp[i]:36 * rp[i + 1]]
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    Dm = bsr_to_dense(s["mrp"], s["mcp"], lu)
    A = bsr_to_dense(s["krp"], s["kcp"], s["kv"])
    LU = (np.tril(Dm, -1) + np.eye(len(Dm))) @ np.triu(Dm)
    assert np.abs(LU - A).max() <= 1e-10 * np.abs(A).max()


def test_zero_pivot_reported(gpu_ctx):
    say("[gpu] zero pivot detection")
    m = M.plate(3, 3)
    s = oracle_problem(m, "natural", 0)
    mv = s["mv"].copy()
    # zero the first diagonal block: no L part updates it, so the pivot is 0
    i = 0
    for kk in range(s["mrp"][i], s["mrp"][i + 1]):
        if s["mcp"][kk] == i:
            mv[36 * kk:36 * kk + 36] = 0.0
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    with pytest.raises(G.GfError) as e:
        ilu.factor(dev(mv))
    assert e.value.code == -4
    ilu.close()


SWEEP_CASES = [
    # (name, mesh, ordering, level of fill): rows of one gather pass, and
    # rows of a complete factorisation (U rows up to 250 blocks: several
    # passes, taken last to first in the U sweep)
    ("plate30_rcm_ilu3", lambda: M.plate(30, 30, bc="ss", load="sine"), "rcm", 3),
    ("wingbox2176_amd_lu", lambda: M.wingbox(nc=24, nh=4, nsb=4, nribs=8, seed=5),
     "amd", 1 << 29),
]


@pytest.mark.parametrize("case", SWEEP_CASES, ids=lambda c: c[0])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_syncfree_sweeps_match_level_schedule(gpu_ctx, dtype, case, monkeypatch):
    """The persistent sync-free sweeps and the level-by-level sweeps are two
    schedules of the same triangular solves (summed in different orders):
    they agree to rounding, and the FP64 result matches the oracle's
    sweeps on the same factors."""
    import torch

    name, make, reorder, k = case
    say(f"[gpu] sweep schedules dtype={np.dtype(dtype).name} {name}")
    m = make()
    s = oracle_problem(m, reorder, k)
    ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
    d_lu = dev(s["mv"].astype(dtype))
    ilu.factor(d_lu)
    gpu_ctx.sync()
    lu = host(d_lu)
    ilu.close()
    b = np.random.default_rng(4).standard_normal(6 * m.nnodes).astype(dtype)
    outs = []
    for mode in ("level", "syncfree"):
        monkeypatch.setenv("GPUFEM_TRSV", mode)
        ilu = G.Ilu(gpu_ctx, s["mrp"], s["mcp"])
        d_b = dev(b)
        d_t, d_x = torch.zeros_like(d_b), torch.zeros_like(d_b)
        for rep in range(3):  # repeated sweeps reuse flags via epochs
            ilu.apply(d_lu, d_b, d_t, d_x)
        gpu_ctx.sync()
        outs.append(host(d_x))
        ilu.close()
    scale = np.abs(outs[0]).max()
    tol = 1e-13 if dtype == np.float64 else 1e-5
    assert np.abs(outs[0] - outs[1]).max() <= tol * scale
    if dtype == np.float64:
        x_o = oracle.bsr_trsv(s["mrp"], s["mcp"], lu, b.astype(np.float64))
        assert np.abs(outs[1] - x_o).max() <= 1e-10 * np.abs(x_o).max()


def test_residual_diagnostic(gpu_ctx):
    """gf_bsr_resid_f64 (CUSPARSE::get_resid, cusparse.h:853-918): the signed
    entry of A x - b of largest magnitude, first index on ties, and the
    2-norm."""
    import ctypes

    say("[gpu] bsr residual diagnostic plate 12x9 ss")
    m = M.plate(12, 9, bc="ss", load="sine")
    s = oracle_problem(m, "rcm", 0)
    n = m.nnodes
    rng = np.random.default_rng(11)
    x = rng.standard_normal(6 * n)
    b = rng.standard_normal(6 * n)
    d = {k: dev(s[k]) for k in ("krp", "kcp", "kv")}
    L = G.lib()
    mx, nr = ctypes.c_double(), ctypes.c_double()
    for xv, bv in ((x, b), (np.zeros(6 * n), None)):
        if bv is None:  # planted tie: |r| maximal at 5 and 17, opposite signs
            bv = rng.uniform(-1.0, 1.0, 6 * n)
            bv[5], bv[17] = -3.0, 3.0
        d_x, d_b = dev(xv), dev(bv)
        gpu_ctx.check(L.gf_bsr_resid_f64(gpu_ctx.h, n, G.ptr(d["krp"]),
                                         G.ptr(d["kcp"]), G.ptr(d["kv"]),
                                         G.ptr(d_x), G.ptr(d_b),
                                         ctypes.byref(mx), ctypes.byref(nr)))
        r = oracle.bsr_spmv(s["krp"], s["kcp"], s["kv"], xv) - bv
        i = int(np.argmax(np.abs(r)))
        assert abs(mx.value - r[i]) <= 1e-12 * abs(r[i])
        assert abs(nr.value - np.linalg.norm(r)) <= 1e-12 * np.linalg.norm(r)
    assert mx.value == 3.0  # r = -b: first of the tied entries is r[5] = 3


def test_compensated_residual(gpu_ctx):
    """gf_bsr_resid_comp_f64 (true residual of the GMRES outer loop) against
    the exactly rounded residual (exact rational arithmetic), on a system
    where b = K x rounded, so r is pure cancellation and the plain FP64
    residual is mostly rounding error."""
    from fractions import Fraction

    say("[gpu] compensated residual plate 6x5 ss")
    m = M.plate(6, 5, bc="ss", load="sine")
    s = oracle_problem(m, "rcm", 0)
    n = m.nnodes
    rng = np.random.default_rng(3)
    x = rng.standard_normal(6 * n) * 1e-2
    b = oracle.bsr_spmv(s["krp"], s["kcp"], s["kv"], x)
    krp, kcp, kv = s["krp"], s["kcp"], s["kv"]
    exact = np.zeros(6 * n)
    mag = np.zeros(6 * n)
    for i in range(n):
        for r in range(6):
            acc = Fraction(b[6 * i + r])
            sm = 0.0
            for k in range(krp[i], krp[i + 1]):
                j = kcp[k]
                for c in range(6):
                    a = kv[36 * k + 6 * r + c]
                    acc -= Fraction(a) * Fraction(x[6 * j + c])
                    sm += abs(a * x[6 * j + c])
            exact[6 * i + r] = float(acc)
            mag[6 * i + r] = sm
    d = {k: dev(s[k]) for k in ("krp", "kcp", "kv")}
    d_x, d_b = dev(x), dev(b)
    d_r = dev(np.zeros(6 * n))
    gpu_ctx.check(G.lib().gf_bsr_resid_comp_f64(
        gpu_ctx.h, n, G.ptr(d["krp"]), G.ptr(d["kcp"]), G.ptr(d["kv"]),
        G.ptr(d_x), G.ptr(d_b), G.ptr(d_r)))
    gpu_ctx.sync()
    r = host(d_r)
    tol = 2.0 ** -100 * mag + 2 * np.spacing(np.abs(exact))
    assert np.all(np.abs(r - exact) <= tol)
    # the plain FP64 residual is far less accurate here
    plain = b - oracle.bsr_spmv(krp, kcp, kv, x)
    assert np.abs(plain - exact).max() > 1e3 * np.abs(r - exact).max()
