"""Serving loop: requests of differing meshes admitted under an HBM budget
and solved concurrently on worker streams.  Each result in the mix equals the
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


def test_mix_equals_alone(gpu_ctx):
    print("[gpu] server mix vs alone, 10 requests", flush=True)
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
