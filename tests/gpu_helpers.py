"""Shared helpers for the GPU parity tests (device upload, oracle set-up,
comparison norms)."""
import numpy as np

import gpu_fem_amd as G
import oracle


def dev(a, dtype=None):
    import torch

    a = np.ascontiguousarray(a if dtype is None else np.asarray(a, dtype))
    return torch.from_numpy(a.copy()).to("cuda")


def host(t):
    return t.detach().cpu().numpy()


def oracle_problem(mesh, reorder="rcm", k=0):
    """Symbolic + assembled K (BCs applied) + M pattern, all on the CPU."""
    n = mesh.nnodes
    rp, cp = oracle.node_graph(mesh.conn, n)
    if reorder == "rcm":
        perm = oracle.rcm(rp, cp)
    elif reorder == "amd":
        # The reference's AMD lives in the absent sparse-utils library, so
        # there is no reference permutation to compare with.  The product's
        # ordering is used on both sides once the oracle has checked it is a
        # minimum-degree ordering: replayed on the explicit elimination graph,
        # every pivot has the least exact degree left at its step (the
        # replay is cubic in the node count: meshes above 2,500 nodes rely on
        # the same check in tests/test_oracle_symbolic.py on smaller meshes)
        perm = G.amd_order(rp, cp)
        if n <= 2500:
            dp, dm, _ = oracle.md_replay(rp, cp, perm)
            assert np.array_equal(dp, dm), "AMD pivot above the minimum degree"
    else:
        perm = np.arange(n, dtype=np.int32)
    krp, kcp = oracle.permute_pattern(rp, cp, perm)
    emap = oracle.elem_ind_map(mesh.conn, krp, kcp, perm)
    kv, _ = oracle.assemble_jacobian(mesh.conn, mesh.xpts, mesh.data, emap,
                                     perm, int(krp[-1]), with_res=False)
    tr, tc, tm = oracle.transpose(krp, kcp)
    kv = oracle.apply_bcs_mat(mesh.bcs, krp, kcp, tr, tc, tm, perm, kv)
    mrp, mcp, _ = oracle.iluk(krp, kcp, k)
    mv = oracle.copy_values(krp, kcp, kv, mrp, mcp)
    return dict(perm=perm, krp=krp, kcp=kcp, kv=kv, mrp=mrp, mcp=mcp, mv=mv,
                emap=emap, tr=tr, tc=tc, tm=tm)


def block_diag_scale(rowp, cols, vals):
    """sqrt(|a_ii| |a_jj|) for every stored entry (scale for relative errors)."""
    n = len(rowp) - 1
    d = np.ones(6 * n)
    for i in range(n):
        for k in range(rowp[i], rowp[i + 1]):
            if cols[k] == i:
                d[6 * i:6 * i + 6] = np.abs(np.diag(vals[36 * k:36 * k + 36].reshape(6, 6)))
    d[d == 0] = 1.0
    s = np.zeros_like(vals)
    for i in range(n):
        for k in range(rowp[i], rowp[i + 1]):
            j = cols[k]
            s[36 * k:36 * k + 36] = np.sqrt(np.outer(d[6 * i:6 * i + 6],
                                                     d[6 * j:6 * j + 6])).ravel()
    return s


def max_rel_scaled(a, b, scale):
    return float(np.max(np.abs(a - b) / scale))
