"""Pins for the oracle's assembly, boundary conditions, ILU, triangular
sweeps, SpMV and GMRES (oracle/numeric.cpp).

Parity status:
  - assembly scatter, BCs, value copy: pinned by brute-force dense assembly
    (the dense element matrices scattered by conn) and by the definitions of
    the reference kernels (bsr_mat.cuh:2-140)
  - ILU numeric: cuSPARSE's bsrilu02 is a closed library with no recorded
    outputs, so bit parity is unpinned; pinned by the ILU property
    (L U)_ij = A_ij on the pattern, by ILU on the complete fill pattern
    being the exact LU (residual ~1e-14), and FP32 vs FP64 factors
  - triangular sweeps / SpMV: dense triangular solves / dense products
  - GMRES: the reference's 4x4 known answer
    (examples/cusparse_solve/1_example.cu:34-38) embedded in 6x6 blocks, and
    a plate solve against a dense direct solve at 1e-8
"""
import numpy as np
import pytest

import oracle
from gpu_fem_amd import meshes as M


def bsr_to_dense(rowp, cols, vals):
    n = len(rowp) - 1
    A = np.zeros((6 * n, 6 * n), dtype=vals.dtype)
    for i in range(n):
        for k in range(rowp[i], rowp[i + 1]):
            j = cols[k]
            A[6 * i:6 * i + 6, 6 * j:6 * j + 6] = vals[36 * k:36 * k + 36].reshape(6, 6)
    return A


def setup(mesh, reorder="rcm", k=1):
    n = mesh.nnodes
    rp, cp = oracle.node_graph(mesh.conn, n)
    perm = (oracle.rcm(rp, cp) if reorder == "rcm"
            else np.arange(n, dtype=np.int32))
    krp, kcp = oracle.permute_pattern(rp, cp, perm)
    emap = oracle.elem_ind_map(mesh.conn, krp, kcp, perm)
    kv, _ = oracle.assemble_jacobian(mesh.conn, mesh.xpts, mesh.data, emap,
                                     perm, int(krp[-1]), with_res=False)
    tr, tc, tm = oracle.transpose(krp, kcp)
    kv = oracle.apply_bcs_mat(mesh.bcs, krp, kcp, tr, tc, tm, perm, kv)
    mrp, mcp, _ = oracle.iluk(krp, kcp, k)
    mv = oracle.copy_values(krp, kcp, kv, mrp, mcp)
    return dict(perm=perm, krp=krp, kcp=kcp, kv=kv, mrp=mrp, mcp=mcp, mv=mv,
                emap=emap)


def test_assembly_matches_dense_scatter():
    m = M.plate(3, 2)
    rng = np.random.default_rng(0)
    m.xpts[:, 2] = rng.uniform(-0.05, 0.05, m.nnodes)  # warp it
    s = setup(m)
    perm = s["perm"]
    K = np.zeros((6 * m.nnodes,) * 2)
    for e, c in enumerate(m.conn):
        Ke, _, _ = oracle.element(m.xpts[c], data=m.data[e])
        dofs = np.concatenate([6 * perm[a] + np.arange(6) for a in c])
        K[np.ix_(dofs, dofs)] += Ke
    # apply BCs densely (rows and columns of permuted bc dofs, unit diagonal)
    bc = np.array([6 * perm[b // 6] + b % 6 for b in m.bcs])
    K[bc, :] = 0
    K[:, bc] = 0
    K[bc, bc] = 1
    D = bsr_to_dense(s["krp"], s["kcp"], s["kv"])
    np.testing.assert_allclose(D, K, rtol=0, atol=1e-9 * np.abs(K).max())
    # the copy into the ILU pattern keeps K and zeros the fill
    Dm = bsr_to_dense(s["mrp"], s["mcp"], s["mv"])
    np.testing.assert_array_equal(Dm, D)


def test_residual_paths_agree():
    m = M.plate(3, 3)
    rng = np.random.default_rng(1)
    u = rng.standard_normal(6 * m.nnodes) * 1e-4
    s = setup(m)
    kv, res = oracle.assemble_jacobian(m.conn, m.xpts, m.data, s["emap"],
                                       s["perm"], int(s["krp"][-1]), vars=u)
    r2 = oracle.assemble_residual(m.conn, m.xpts, u, m.data, s["perm"])
    np.testing.assert_allclose(res, r2, rtol=1e-12, atol=1e-12 * np.abs(r2).max())
    up = oracle.permute_vec(s["perm"], u)
    Ku = oracle.bsr_spmv(s["krp"], s["kcp"], kv, up)
    np.testing.assert_allclose(Ku, r2, rtol=0, atol=1e-10 * np.abs(r2).max())


@pytest.mark.parametrize("k", [0, 1, 3])
def test_ilu_property_and_sweeps(k):
    m = M.plate(4, 4, bc="ss", load="sine")
    s = setup(m, k=k)
    lu, st = oracle.bsr_ilu(s["mrp"], s["mcp"], s["mv"])
    assert st == 0
    D = bsr_to_dense(s["mrp"], s["mcp"], lu)
    A = bsr_to_dense(s["mrp"], s["mcp"], s["mv"])
    L = np.tril(D, -1) + np.eye(len(D))
    U = np.triu(D)
    mask = A != 0
    pattern = bsr_to_dense(s["mrp"], s["mcp"], np.ones_like(s["mv"])) != 0
    LU = L @ U
    np.testing.assert_allclose(LU[pattern], A[pattern], rtol=0,
                               atol=1e-9 * np.abs(A).max())
    assert mask.sum() > 0
    b = np.random.default_rng(2).standard_normal(len(D))
    x = oracle.bsr_trsv(s["mrp"], s["mcp"], lu, b)
    np.testing.assert_allclose(L @ (U @ x), b, rtol=1e-9, atol=1e-9 * np.abs(b).max())


def test_full_fill_ilu_is_exact_lu():
    m = M.plate(4, 3)
    s = setup(m)
    frp, fcp = oracle.full_fill(s["krp"], s["kcp"])
    mv = oracle.copy_values(s["krp"], s["kcp"], s["kv"], frp, fcp)
    lu, st = oracle.bsr_ilu(frp, fcp, mv)
    A = bsr_to_dense(s["krp"], s["kcp"], s["kv"])
    b = np.random.default_rng(3).standard_normal(len(A))
    x = oracle.bsr_trsv(frp, fcp, lu, b)
    assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b) * 1e3
    # FP32 factors agree with FP64 to single precision
    lu32, st32 = oracle.bsr_ilu(frp, fcp, mv, dtype=np.float32)
    assert st32 == 0
    rel = np.abs(lu32 - lu).max() / np.abs(lu).max()
    assert rel < 1e-5


def test_gmres_reference_4x4_known_answer():
    # examples/cusparse_solve/1_example.cu:34-38, embedded as a_ij * I6 blocks
    A = np.array([[3, 2, 0, 4], [2, 6, 6, 0], [0, 6, 11, 0], [4, 0, 0, 18.]])
    b = np.array([-1, -2, 3, 4.])
    xtrue = np.array([1.20689655, -2.2183908, 1.48275862, -0.04597701])
    rowp = [0]
    cols, vals = [], []
    for i in range(4):
        for j in range(4):
            if A[i, j] != 0:
                cols.append(j)
                vals.append((A[i, j] * np.eye(6)).ravel())
        rowp.append(len(cols))
    rowp, cols = np.array(rowp, np.int32), np.array(cols, np.int32)
    vals = np.concatenate(vals)
    lu, st = oracle.bsr_ilu(rowp, cols, vals)  # ILU(0)
    bb = np.repeat(b, 6)
    x, st, info = oracle.gmres(rowp, cols, vals, rowp, cols, lu, bb, restart=10,
                               rtol=1e-12)
    assert st == 0
    np.testing.assert_allclose(x.reshape(4, 6), np.repeat(xtrue[:, None], 6, 1),
                               atol=1e-8)


@pytest.mark.parametrize("mixed", [False, True])
def test_gmres_plate_vs_dense(mixed):
    m = M.plate(8, 8, thick=0.01, bc="clamped", load="uniform")
    s = setup(m, k=1)
    lu, st = oracle.bsr_ilu(s["mrp"], s["mcp"], s["mv"])
    b = oracle.permute_vec(s["perm"], m.loads)
    x, st, info = oracle.gmres(s["krp"], s["kcp"], s["kv"], s["mrp"], s["mcp"],
                               lu, b, restart=50, rtol=1e-10, mixed=mixed)
    assert st == 0 and info["rel_resid"] <= 1e-10
    A = bsr_to_dense(s["krp"], s["kcp"], s["kv"])
    xd = oracle.dense_solve(A, b)
    assert np.linalg.norm(x - xd) <= 1e-7 * np.linalg.norm(xd)
