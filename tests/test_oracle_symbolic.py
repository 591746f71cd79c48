"""Pins for the oracle's symbolic phase (oracle/symbolic.cpp) and the
product's host symbolic phase (gpu_fem_amd/csrc/symbolic.cpp, amd.cpp).

Parity status: pinned.
  - golden outputs of the reference's own TacsComputeRCMOrder, computeILUk
    (patterns and levels, k = 0..3), get_elem_ind_map and getBandWidth,
    compiled from the reference sources by oracle/build_ref.sh
    (tests/golden/symbolic_ref.npz, made by tests/golden/make_golden.py);
    when oracle/_ref is built, also live comparisons on more meshes
  - node graph: the 2x2-plate rowPtr/colPtr written out by hand in
    examples/cusparse_solve/4_plate.cu:172-189 (nnzb 49) and the closed form
    nnzb = (3nx-2)(3ny-2)
  - fill-in: 28 -> 32 blocks on the 2-element mesh of
    examples/basic_shell/test_fill_in.cu:23-37
  - ILU(k large) equals the complete symbolic LU fill, checked against a
    dense elimination on tiny patterns
  - AMD (the reference's comes from sparse-utils, absent): parity unpinned;
    invariants only (a permutation; fill not worse than natural order on
    grids; deterministic)
"""
import os

import numpy as np
import pytest

import gpu_fem_amd as G
import oracle
from gpu_fem_amd import meshes as M

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden",
                            "symbolic_ref.npz"))
NAMES = sorted({k.split("/")[0] for k in GOLD.files})


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("impl", ["oracle", "product"])
def test_against_reference_golden(name, impl):
    g = lambda k: GOLD[f"{name}/{k}"]
    conn = g("conn")
    nn = len(g("rowp")) - 1
    if impl == "oracle":
        rp, cp = oracle.node_graph(conn, nn)
        perm = oracle.rcm(rp, cp)
        prp, pcp = oracle.permute_pattern(rp, cp, perm)
    else:
        rp, cp = G.node_graph(conn, nn)
        perm = G.rcm_order(rp, cp)
        prp, pcp = G.pattern_permute(rp, cp, perm)
    np.testing.assert_array_equal(rp, g("rowp"))
    np.testing.assert_array_equal(cp, g("cols"))
    np.testing.assert_array_equal(perm, g("rcm"))
    for k in range(4):
        if impl == "oracle":
            frp, fcp, flv = oracle.iluk(prp, pcp, k)
            np.testing.assert_array_equal(flv, g(f"ilu{k}_levs"))
        else:
            frp, fcp = G.iluk_symbolic(prp, pcp, k)
        np.testing.assert_array_equal(frp, g(f"ilu{k}_rowp"))
        np.testing.assert_array_equal(fcp, g(f"ilu{k}_cols"))
    emap = (oracle.elem_ind_map(conn, prp, pcp, perm) if impl == "oracle"
            else G.elem_ind_map(conn, prp, pcp, perm))
    np.testing.assert_array_equal(emap, g("elem_ind_map"))
    bw = (oracle.bandwidth(prp, pcp) if impl == "oracle"
          else G.bandwidth(prp, pcp))
    assert bw == g("bandwidth")[0]


def test_node_graph_hand_written_2x2():
    # examples/cusparse_solve/4_plate.cu:172-189
    rowp = [0, 4, 10, 14, 20, 29, 35, 39, 45, 49]
    cols = [0, 1, 3, 4, 0, 1, 2, 3, 4, 5, 1, 2, 4, 5, 0, 1, 3, 4, 6, 7,
            0, 1, 2, 3, 4, 5, 6, 7, 8, 1, 2, 4, 5, 7, 8, 3, 4, 6, 7,
            3, 4, 5, 6, 7, 8, 4, 5, 7, 8]
    conn = M.plate_conn(2, 2)
    for f in (oracle.node_graph, G.node_graph):
        rp, cp = f(conn, 9)
        np.testing.assert_array_equal(rp, rowp)
        np.testing.assert_array_equal(cp, cols)


@pytest.mark.parametrize("nx,ny", [(1, 1), (5, 3), (40, 17)])
def test_node_graph_closed_form(nx, ny):
    rp, cp = G.node_graph(M.plate_conn(nx, ny), (nx + 1) * (ny + 1))
    assert rp[-1] == (3 * (nx + 1) - 2) * (3 * (ny + 1) - 2)
    c = M.cylinder(12, 7)
    rp, cp = G.node_graph(c.conn, c.nnodes)
    assert rp[-1] == (3 * 12) * (3 * 8 - 2)


def test_fill_in_kat():
    conn = np.array([[0, 1, 2, 3], [1, 3, 4, 5]], np.int32)
    for ng, ilu in ((oracle.node_graph, lambda r, c, k: oracle.iluk(r, c, k)[:2]),
                    (G.node_graph, G.iluk_symbolic)):
        rp, cp = ng(conn, 6)
        assert rp[-1] == 28
        assert ilu(rp, cp, 0)[0][-1] == 28
        assert ilu(rp, cp, 1)[0][-1] == 32
        assert ilu(rp, cp, 5)[0][-1] == 32


def dense_fill(rowp, cols):
    n = len(rowp) - 1
    A = np.zeros((n, n), bool)
    for i in range(n):
        A[i, cols[rowp[i]:rowp[i + 1]]] = True
        A[i, i] = True
    for k in range(n):
        rows = np.nonzero(A[k + 1:, k])[0] + k + 1
        colsk = np.nonzero(A[k, k + 1:])[0] + k + 1
        A[np.ix_(rows, colsk)] = True
    return A


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_iluk_large_k_is_full_fill(seed):
    m = M.plate(5, 4)
    rp, cp = oracle.node_graph(m.conn, m.nnodes)
    perm = np.random.default_rng(seed).permutation(m.nnodes).astype(np.int32)
    prp, pcp = oracle.permute_pattern(rp, cp, perm)
    A = dense_fill(prp, pcp)
    for frp, fcp in (oracle.full_fill(prp, pcp),
                     G.iluk_symbolic(prp, pcp, 10 ** 6)):
        B = np.zeros_like(A)
        for i in range(len(frp) - 1):
            B[i, fcp[frp[i]:frp[i + 1]]] = True
        np.testing.assert_array_equal(A, B)


def test_live_reference_if_built():
    if oracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for conn, nn in [(M.plate_conn(31, 9), 32 * 10),
                     (M.cylinder(20, 11).conn, 20 * 12)]:
        rp, cp = oracle.ref_node_graph(conn, nn)
        perm = oracle.ref_rcm(rp, cp)
        np.testing.assert_array_equal(G.rcm_order(rp, cp), perm)
        prp, pcp = G.pattern_permute(rp, cp, perm)
        for k in (0, 2, 3):
            a = oracle.ref_iluk(prp, pcp, k)
            b = G.iluk_symbolic(prp, pcp, k)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])


def test_transpose_map_definition():
    m = M.plate(6, 5)
    rp, cp = G.node_graph(m.conn, m.nnodes)
    perm = G.rcm_order(rp, cp)
    prp, pcp = G.pattern_permute(rp, cp, perm)
    for tr, tc, tm in (G.transpose_map(prp, pcp), oracle.transpose(prp, pcp)):
        for c in range(len(tr) - 1):
            for t in range(tr[c], tr[c + 1]):
                r = tc[t]
                k = tm[t]
                assert pcp[k] == c and prp[r] <= k < prp[r + 1]
            assert np.all(np.diff(tc[tr[c]:tr[c + 1]]) > 0)


def test_amd_invariants():
    m = M.plate(20, 20)
    rp, cp = G.node_graph(m.conn, m.nnodes)
    p = G.amd_order(rp, cp)
    assert sorted(p) == list(range(m.nnodes))
    np.testing.assert_array_equal(p, G.amd_order(rp, cp))
    nat = G.iluk_symbolic(rp, cp, 10 ** 6)[0][-1]
    prp, pcp = G.pattern_permute(rp, cp, p)
    amd = G.iluk_symbolic(prp, pcp, 10 ** 6)[0][-1]
    rcm = G.rcm_order(rp, cp)
    rrp, rcp = G.pattern_permute(rp, cp, rcm)
    rcmf = G.iluk_symbolic(rrp, rcp, 10 ** 6)[0][-1]
    assert amd < nat and amd < rcmf
    # on tiny graphs the minimum-degree greedy rule gives no more fill than
    # the oracle's exact minimum degree by more than 25%
    t = M.plate(4, 3)
    rp, cp = G.node_graph(t.conn, t.nnodes)
    f_amd = G.iluk_symbolic(*G.pattern_permute(rp, cp, G.amd_order(rp, cp)),
                            10 ** 6)[0][-1]
    f_md = G.iluk_symbolic(*G.pattern_permute(rp, cp, oracle.min_degree(rp, cp)),
                           10 ** 6)[0][-1]
    assert f_amd <= 1.25 * f_md
