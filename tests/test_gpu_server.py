"""Serving loop: requests of differing meshes admitted under an HBM budget
and solved concurrently, on worker streams or in the batched GMRES engines.  Each result in the mix equals the
same request solved alone (1e-10 relative; the only order-dependent step is
the FP64 atomic accumulation in the Jacobian scatter), every result meets its
tolerance, and admission respects the budget."""
import sys

import numpy as np
import pytest

import gpu_fem_amd as G
from gpu_fem_amd import meshes as M

pytestmark = pytest.mark.gpu


def requests(seed=0, n=10):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        if rng.random() < 0.5:
            nx = int(rng.integers(8, 28))
            m = M.plate(nx, nx, thick=float(rng.uniform(0.004, 0.01)),
                        bc="clamped" if rng.random() < 0.5 else "ss",
                        load="uniform" if rng.random() < 0.5 else "sine",
                        load_mag=float(rng.uniform(500, 5000)))
        else:
            nc = int(rng.integers(8, 20))
            m = M.cylinder(nc, int(rng.integers(6, 16)),
                           t_range=(0.004, 0.01), seed=int(rng.integers(1 << 30)))
        out.append(m)
    return out


# AMD order with ILU(3) converges on every mesh of this mix (thin plates
# and cylinders); natural or RCM order with ILU(2) stagnates on several.
OPTS = dict(reorder=2, lev_fill=3, restart=80, rtol=1e-9, max_iters=20000)


# 0: each worker solves on its own stream; 1: the batched GMRES engines
# (fused launches over the concurrent solves, replayed as a HIP graph)
ENGINES = pytest.mark.parametrize("batch", [0, 1], ids=["streams", "batched"])


@ENGINES
def test_mix_equals_alone(gpu_ctx, batch, monkeypatch):
    monkeypatch.setenv("GPUFEM_SERVE_BATCH", str(batch))
    print(f"[gpu] server mix vs alone, 10 requests, batch={batch}", flush=True)
    reqs = requests()
    alone = []
    for m in reqs:
        p = G.Problem(gpu_ctx, m, **OPTS)
        p.assemble()
        p.factor()
        x, info = p.solve()
        assert info.status == 0
        alone.append(x)
        p.close()
    srv = G.Server(budget_gb=200.0, slots=4)
    for i, m in enumerate(reqs):
        assert srv.submit(i, m, **OPTS) == 0
    srv.drain()
    got = {}
    while True:
        r = srv.poll()
        if r is None:
            break
        rid, x, info, lat = r
        assert info.status == 0 and info.rel_resid <= 1e-9
        got[rid] = x
    st = srv.stats()
    srv.close()
    assert len(got) == len(reqs) and st["completed"] == len(reqs)
    for i, x in enumerate(alone):
        assert np.linalg.norm(got[i] - x) <= 1e-10 * np.linalg.norm(x)
    assert st["p99_s"] >= st["p50_s"] > 0


def test_budget_admission(gpu_ctx):
    print("[gpu] server admission under a tiny budget", flush=True)
    reqs = requests(seed=1, n=6)
    srv = G.Server(budget_gb=0.05, slots=4)
    for i, m in enumerate(reqs):
        assert srv.submit(i, m, **OPTS) == 0
    srv.drain()
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert 0 < st["peak_bytes"] <= 0.05e9
    # a request larger than the whole budget is refused up front
    big = M.plate(300, 300)
    assert srv.submit(99, big, **OPTS) == -7
    srv.close()


def test_stagnating_request_is_rerun_with_more_fill(gpu_ctx):
    """A request whose ILU(1) solve stagnates (free-edge plate, relative
    residual ~0.9) is re-run by the server with ILU(2), converges, and equals
    the same request solved alone with ILU(2)."""
    print("[gpu] server escalation ILU(1) -> ILU(2)", flush=True)
    m = M.plate(16, 16, 2.0, 1.0, E=70e9, nu=0.3, thick=0.005, bc="ss",
                load="nodal_sine", load_mag=1.0)
    opts = dict(reorder=1, restart=100, rtol=1e-9, max_iters=4000,
                stall_cycles=3)
    p = G.Problem(gpu_ctx, m, lev_fill=2, **opts)
    p.assemble()
    p.factor()
    x2, info2 = p.solve()
    p.close()
    assert info2.status == 0
    srv = G.Server(budget_gb=200.0, slots=2)
    assert srv.submit(0, m, lev_fill=1, **opts) == 0
    srv.drain()
    r = srv.poll()
    st = srv.stats()
    srv.close()
    assert r is not None
    _, x, info, _ = r
    assert info.status == 0 and info.rel_resid <= 1e-9
    assert st["escalated"] == 1 and st["completed"] == 1
    assert np.linalg.norm(x - x2) <= 1e-10 * np.linalg.norm(x2)


def served_mix_sample(seed=5, n=6):
    """The low end of the served request mix (gpu_fem_amd.meshes.
    served_requests): plates n x n with n in [32, 40] and 64 x 128
    cylinders (R = 1 m, L = 3 m), thickness uniform in [1, 10] mm,
    uniform-plus-sinusoidal pressure of magnitude uniform in [0.5, 5] kPa,
    clamped or simply supported with probability 1/2."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        t = float(rng.uniform(1e-3, 10e-3))
        p0 = float(rng.uniform(500.0, 5000.0))
        bc = "clamped" if rng.random() < 0.5 else "ss"
        if i % 2 == 0:
            m = int(rng.integers(32, 41))
            mesh = M.plate(m, m, 1.0, 1.0, thick=t, bc=bc, load="uniform_sine",
                           load_mag=p0)
        else:
            mesh = M.cylinder(64, 128, t_range=(t, t),
                              seed=int(rng.integers(1 << 30)), pressure=p0,
                              axial_line_load=0.0, bc=bc)
        out.append(mesh)
    return out


# the served configuration: AMD + ILU(32), GMRES(200) to 1e-8, FP32 ILU and
# sweeps inside FP64 GMRES, stall window of 4 cycles
SERVED = dict(reorder=2, lev_fill=32, restart=200, rtol=1e-8, max_iters=20000,
              stall_cycles=4, precision=2)


@ENGINES
def test_served_mix_against_fp64_oracle(gpu_ctx, batch, monkeypatch):
    """Requests of the served mix arrive as a Poisson stream (rate 20/s) at
    a server with 4 slots; every served result has a true relative residual
    at or below 1e-8 evaluated by the oracle in compensated FP64, lies close
    to the oracle's FP64 solution (solved to 1e-10 or its floor), and equals
    the same request solved alone."""
    import time

    import oracle
    from gpu_helpers import oracle_problem

    monkeypatch.setenv("GPUFEM_SERVE_BATCH", str(batch))
    reqs = served_mix_sample()
    print(f"[gpu] served mix sample (batch={batch}): {[m.name for m in reqs]}",
          flush=True)
    rng = np.random.default_rng(11)
    arrivals = np.cumsum(rng.exponential(1.0 / 20.0, len(reqs)))
    srv = G.Server(budget_gb=200.0, slots=4)
    t0 = time.perf_counter()
    for i, m in enumerate(reqs):
        while time.perf_counter() - t0 < arrivals[i]:
            time.sleep(0.001)
        assert srv.submit(i, m, **SERVED) == 0
    srv.drain()
    got = {}
    while True:
        r = srv.poll()
        if r is None:
            break
        got[r[0]] = r
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and len(got) == len(reqs)
    for i, m in enumerate(reqs):
        _, x, info, lat = got[i]
        assert info.status == 0 and info.rel_resid <= 1e-8
        # alone, same configuration
        p = G.Problem(gpu_ctx, m, **SERVED)
        p.assemble()
        p.factor()
        xa, ia = p.solve()
        perm = p.perm()
        p.close()
        assert np.linalg.norm(x - xa) <= 1e-10 * np.linalg.norm(xa)
        # FP64 oracle: same ordering (checked as a minimum-degree ordering on
        # the smaller meshes), ILU(32) in FP64, GMRES to 1e-10
        s = oracle_problem(m, "amd", SERVED["lev_fill"])
        np.testing.assert_array_equal(perm, s["perm"])
        lu, ok = oracle.bsr_ilu(s["mrp"], s["mcp"], s["mv"])
        assert ok == 0
        b = oracle.permute_vec(s["perm"],
                               oracle.apply_bcs_vec(m.bcs, m.loads))
        xo, _, oi = oracle.gmres(s["krp"], s["kcp"], s["kv"], s["mrp"],
                                 s["mcp"], lu, b, restart=200, rtol=1e-10,
                                 max_iters=600)
        xp = oracle.permute_vec(s["perm"], x)
        r = oracle.bsr_resid_comp(s["krp"], s["kcp"], s["kv"], xp, b)
        tr = np.linalg.norm(r) / np.linalg.norm(b)
        err = np.linalg.norm(xp - xo) / np.linalg.norm(xo)
        print(f"    {m.name}: latency {lat * 1e3:.0f} ms, iters "
              f"{info.iterations}, residual {info.rel_resid:.2e} (oracle "
              f"{tr:.2e}), |x - x_oracle_fp64| / |x_oracle_fp64| = {err:.2e} "
              f"(oracle residual {oi['rel_resid']:.1e})", flush=True)
        assert tr <= 1.01e-8
        # a residual of 1e-8 bounds the displacement error by the condition
        # number; these shells stay within 1e-6 of the FP64 oracle
        assert err <= 1e-6
