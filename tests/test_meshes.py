"""Mesh generators (host side): invariants of the seeded cylinder and
wingbox generators, which have no counterpart in the reference."""
import numpy as np
import pytest

from gpu_fem_amd import meshes as M


@pytest.mark.parametrize("dims", [(8, 3, 2, 4), (12, 4, 3, 6)])
def test_wingbox_topology(dims):
    nc, nh, nsb, nribs = dims
    m = M.wingbox(nc=nc, nh=nh, nsb=nsb, nribs=nribs, seed=3)
    Ns = (nribs - 1) * nsb
    # node count: two skins, the spars' interior rows, the ribs' interiors
    assert m.nnodes == 2 * (nc + 1) * (Ns + 1) + 2 * (nh - 1) * (Ns + 1) \
        + nribs * (nc - 1) * (nh - 1)
    assert m.nelems == 2 * nc * Ns + 2 * nh * Ns + nribs * nc * nh
    # shared edges: no two nodes at one point, every node used
    assert len(np.unique(np.round(m.xpts, 9), axis=0)) == m.nnodes
    assert len(np.unique(m.conn)) == m.nnodes
    # no degenerate quads; each element's four nodes are distinct
    X = m.xpts[m.conn]
    area = np.linalg.norm(np.cross(X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]), axis=1)
    assert area.min() > 0
    assert all(len(set(e)) == 4 for e in m.conn.tolist())
    # every edge is shared by 1 (free boundary), 2 (inside a panel) or 3
    # (skin-spar, skin-rib, spar-rib junctions) or 4 elements (where a rib
    # meets a skin-spar corner line)
    edges = {}
    for e in m.conn.tolist():
        for a, b in ((e[0], e[1]), (e[1], e[3]), (e[3], e[2]), (e[2], e[0])):
            k = (min(a, b), max(a, b))
            edges[k] = edges.get(k, 0) + 1
    counts = np.bincount(list(edges.values()))
    assert counts[3] > 0 and len(counts) <= 5
    # root clamped: all 6 DOFs of every node at y = 0
    root = np.flatnonzero(np.abs(m.xpts[:, 1]) < 1e-12)
    assert sorted(m.bcs.tolist()) == sorted((6 * root[:, None] +
                                             np.arange(6)).ravel().tolist())
    # panel thicknesses within [2, 20] mm, constant per panel
    assert m.data[:, 2].min() >= 2e-3 and m.data[:, 2].max() <= 20e-3
    assert len(np.unique(m.data[:, 2])) <= m.meta["panels"]


def test_wingbox_load_resultant():
    """The consistent loads integrate the elliptic pressure over the lower
    skin: sum F_z = int_0^L p_root sqrt(1 - (y/L)^2) c(y) dy (the root row's
    share is removed with its BCs, O(1/Ns))."""
    m = M.wingbox(nc=16, nh=4, nsb=8, nribs=6)
    L, cr, ct, p0 = 30.0, 6.0, 1.5, 20e3
    y = np.linspace(0, L, 200001)
    f = p0 * np.sqrt(1 - (y / L) ** 2) * (cr + (ct - cr) * y / L)
    exact = np.sum((f[1:] + f[:-1]) / 2 * np.diff(y))
    fz = m.loads[2::6].sum()
    Ns = 5 * 8
    assert abs(fz - exact) / exact < 2.0 / Ns
    assert abs(m.loads[0::6]).max() < 1e-6 * abs(m.loads[2::6]).max()


def test_wingbox_seeded():
    a, b = M.wingbox(nc=6, nh=2, nsb=1, nribs=3, seed=1), \
        M.wingbox(nc=6, nh=2, nsb=1, nribs=3, seed=1)
    c = M.wingbox(nc=6, nh=2, nsb=1, nribs=3, seed=2)
    assert np.array_equal(a.data, b.data) and not np.array_equal(a.data, c.data)
