"""GPU parity of the nonlinear shell path (IsotropicShell<..., isNonlinear =
true>, include/shell/physics.h:6-17) against the oracle, through the C ABI.

Tolerances: tangent K(u) and residual R(u) 1e-12 relative (per entry against
sqrt(|K_ii| |K_jj|) and against max|K|; max|R|), energy 1e-12; the Newton
solution within 1e-8 of the oracle's Newton solution (dense solves), with the
nonlinear solution itself more than 1e-4 away from the linear one.
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
    "plate8": lambda: M.plate(8, 8),
    "warped": lambda: warped_plate(9, 7),
    "cylinder": lambda: M.cylinder(16, 12, seed=3),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_tangent_and_residual_match_oracle(gpu_ctx, name):
    import torch

    say(f"[gpu] nonlinear tangent parity mesh={name}")
    m = MESHES[name]()
    s = oracle_problem(m, "rcm", 0)
    nnzb = int(s["krp"][-1])
    rng = np.random.default_rng(9)
    u = 1e-2 * rng.standard_normal(6 * m.nnodes)
    kv_o, res_o, e_o = oracle.assemble_tangent(m.conn, m.xpts, m.data,
                                               s["emap"], s["perm"], nnzb, u)
    d_conn, d_x = dev(m.conn, np.int32), dev(m.xpts, np.float64)
    d_data, d_u = dev(m.data, np.float64), dev(u)
    d_map, d_perm = dev(s["emap"]), dev(s["perm"])
    d_k = torch.zeros(36 * nnzb, dtype=torch.float64, device="cuda")
    d_r = torch.zeros(6 * m.nnodes, dtype=torch.float64, device="cuda")
    d_e = torch.zeros(1, dtype=torch.float64, device="cuda")
    gpu_ctx.check(G.lib().gf_assemble_tangent_mitc4_f64(
        gpu_ctx.h, m.nelems, m.nnodes, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u),
        G.ptr(d_data), G.ptr(d_map), G.ptr(d_perm), G.ptr(d_k), G.ptr(d_r),
        G.ptr(d_e)))
    gpu_ctx.sync()
    kv = host(d_k)
    scale = block_diag_scale(s["krp"], s["kcp"], kv_o)
    assert max_rel_scaled(kv, kv_o, scale) <= 1e-12
    assert np.abs(kv - kv_o).max() <= 1e-12 * np.abs(kv_o).max()
    assert np.abs(host(d_r) - res_o).max() <= 1e-12 * np.abs(res_o).max()
    assert abs(host(d_e)[0] - e_o) <= 1e-12 * abs(e_o)
    # the nonlinear terms are in play: K(u) is not the linear K
    kv_lin, _ = oracle.assemble_jacobian(m.conn, m.xpts, m.data, s["emap"],
                                         s["perm"], nnzb, with_res=False)
    assert np.abs(kv_o - kv_lin).max() > 1e3 * np.abs(kv - kv_o).max()

    # colouring plan: K and res start dirty (plain stores / read-add-writes)
    plan = ctypes.c_void_p()
    gpu_ctx.check(G.lib().gf_asm_plan_create(
        gpu_ctx.h, m.nelems, m.nnodes,
        G.ptr(np.ascontiguousarray(m.conn, np.int32)),
        G.ptr(np.ascontiguousarray(s["emap"], np.int32)), nnzb,
        ctypes.byref(plan)))
    d_k2 = torch.full_like(d_k, 5.0)
    d_r2 = torch.full_like(d_r, -1.0)
    gpu_ctx.check(G.lib().gf_assemble_tangent_colored_f64(
        gpu_ctx.h, plan, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u), G.ptr(d_data),
        G.ptr(d_map), G.ptr(d_perm), G.ptr(d_k2), G.ptr(d_r2)))
    gpu_ctx.sync()
    G.lib().gf_asm_plan_destroy(plan)
    kv2 = host(d_k2)
    assert max_rel_scaled(kv2, kv_o, scale) <= 1e-12
    assert np.abs(host(d_r2) - res_o).max() <= 1e-12 * np.abs(res_o).max()

    # residual only (no K)
    d_r3 = torch.zeros_like(d_r)
    gpu_ctx.check(G.lib().gf_assemble_tangent_mitc4_f64(
        gpu_ctx.h, m.nelems, m.nnodes, G.ptr(d_conn), G.ptr(d_x), G.ptr(d_u),
        G.ptr(d_data), None, G.ptr(d_perm), None, G.ptr(d_r3), None))
    gpu_ctx.sync()
    assert np.abs(host(d_r3) - res_o).max() <= 1e-12 * np.abs(res_o).max()


def dense(rp, cp, kv, n):
    K = np.zeros((6 * n, 6 * n))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            j = cp[p]
            K[6 * i:6 * i + 6, 6 * j:6 * j + 6] = kv[36 * p:36 * p + 36].reshape(6, 6)
    return K


def oracle_newton(m, tol=1e-13, steps=15):
    """Newton with dense solves on the oracle's tangent; also returns the
    first (linear) iterate."""
    n = m.nnodes
    rp, cp = oracle.node_graph(m.conn, n)
    perm = np.arange(n, dtype=np.int32)
    emap = oracle.elem_ind_map(m.conn, rp, cp, perm)
    f = m.loads.copy()
    f[m.bcs] = 0.0
    fn = np.linalg.norm(f)
    u = np.zeros(6 * n)
    ulin = None
    for it in range(steps):
        kv, R, _ = oracle.assemble_tangent(m.conn, m.xpts, m.data, emap, perm,
                                           int(rp[-1]), u)
        r = f - R
        r[m.bcs] = 0.0
        if np.linalg.norm(r) <= tol * fn:
            break
        K = dense(rp, cp, kv, n)
        K[m.bcs, :] = 0.0
        K[:, m.bcs] = 0.0
        K[m.bcs, m.bcs] = 1.0
        u += np.linalg.solve(K, r)
        if ulin is None:
            ulin = u.copy()
    return u, ulin


def loaded_cylinder():
    m = M.cylinder(12, 8, seed=3)
    m.loads = m.loads * 10.0  # 500 kPa, 100 kN/m: the nonlinear terms matter
    return m


@pytest.mark.parametrize("fill,precision", [(1 << 29, 0), (3, 0), (3, 2)])
def test_newton_matches_oracle_newton(gpu_ctx, fill, precision):
    say(f"[gpu] newton cylinder 12x8 fill={fill} precision={precision}")
    m = loaded_cylinder()
    u_o, u_lin = oracle_newton(m)
    nl = np.linalg.norm(u_o - u_lin) / np.linalg.norm(u_o)
    assert nl > 1e-4
    p = G.Problem(gpu_ctx, m, reorder=2, lev_fill=fill, restart=100,
                  rtol=1e-12, max_iters=2000, precision=precision)
    u, info, hist = p.newton_solve(max_newton=12, newton_rtol=1e-11)
    p.close()
    say(f"[gpu]   newton steps={info.newton_iterations} krylov="
        f"{info.krylov_iterations} hist={['%.1e' % h for h in hist]}")
    assert info.status == 0 and hist[-1] <= 1e-11
    rel = np.linalg.norm(u - u_o) / np.linalg.norm(u_o)
    assert rel <= 1e-8, rel
    if fill > 1000:
        # exact tangent solves: the first (linear) step leaves only the
        # quadratic terms' residual, and Newton's quadratic convergence takes
        # it to the rounding floor in one or two more steps, as the oracle's
        assert hist[1] <= 1e-5
        assert info.newton_iterations <= 3


def test_newton_small_load_is_linear_solve(gpu_ctx):
    """At a load small enough that the quadratic terms stay below the
    tolerance, one Newton step is the linear solve."""
    say("[gpu] newton small load = linear solve")
    m = M.cylinder(12, 8, seed=3)
    m.loads = m.loads * 1e-6
    p = G.Problem(gpu_ctx, m, reorder=2, lev_fill=1 << 29, restart=50,
                  rtol=1e-12)
    u, info, hist = p.newton_solve(max_newton=5, newton_rtol=1e-8)
    p.assemble()
    p.factor()
    x, si = p.solve()
    p.close()
    assert info.status == 0 and info.newton_iterations == 1
    assert np.linalg.norm(u - x) <= 1e-8 * np.linalg.norm(x)
