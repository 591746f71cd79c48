"""NASTRAN bulk-data reader (gf_bdf_*, replacing TACSMeshLoader,
include/mesh/TACSMeshLoader.h) -- host only, no GPU.

Parity: on every file the reference loader reads, the product's
getAssemblerCreatorData arrays (connectivity, global BC DOFs, coordinates)
equal the reference's bit for bit; the reference is compiled from its own
header into oracle/_ref/libgfref_bdf.so.  Known-answer tests pin the card
layouts independently of it, including the forms the reference misreads
(standard free field, SPC1, continuation markers).
"""
import numpy as np
import pytest

import gpu_fem_amd as G
import oracle
from gpu_fem_amd import meshes as M


def card(name, *fields, w=8):
    """One fixed-field line: the name in columns 0-7, then fields of width w."""
    s = f"{name:<8s}"
    for f in fields:
        s += f"{str(f):<{w}s}"[:w]
    return s.rstrip()


def write(path, lines):
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return str(path)


# Executive control before BEGIN BULK, comments, unsorted and non-contiguous
# ids, NASTRAN exponents, a GRID* pair, a comma GRID with the name in columns
# 0-7, a large-field CQUAD4* with a continuation line, a CQUADR, SPC and SPC*,
# cards that are not read, and a GRID after ENDDATA.
MIXED = [
    "SOL 101",
    "CEND",
    "$ a comment",
    "BEGIN BULK",
    "$       Shell element data for family    SKIN.1",
    "$       Shell element data for family    SPAR.2",
    card("PSHELL", 1, 1, ".005", 1),
    card("MAT1", 1, "70.+9", "", ".3"),
    card("GRID", 30, 0, "1.0", "0.0", "0.0"),
    card("GRID", 10, 0, "0.0", "0.0", "0.0"),
    card("GRID*", 20, "", "5.-1", "1.5-3", w=16),
    card("*", "-2.+0", w=16),
    card("GRID", 40, 0, "1.0D+0", "1.0", "0.0"),
    "GRID    50,0,0.5,1.0,0.25",
    card("GRID", 60, 0, "2.0", "0.0", "-1.25E-1"),
    card("GRID", 70, 0, "2.0", "1.0", "0.0"),
    card("CQUAD4", 7, 1, 10, 20, 40, 30),
    card("CQUAD4*", 3, 2, 30, 60, w=16),
    card("*", 70, 40, w=16),
    card("CQUADR", 5, 2, 20, 30, 40, 50),
    card("SPC", 1, 10, 123456, "0.0"),
    "SPC*    1               60              135             0.0",
    card("SPC", 1, 70, 246, "1.5"),
    card("FORCE", 1, 70, 0, "1.0", "0.0", "0.0", "-1.0"),
    "ENDDATA",
    card("GRID", 99, 0, "9.0", "9.0", "9.0"),
]

# known answer for MIXED: nodes in id order 10 20 30 40 50 60 70
MIXED_X = np.array([[0.0, 0.0, 0.0], [0.5, 1.5e-3, -2.0], [1.0, 0.0, 0.0],
                    [1.0, 1.0, 0.0], [0.5, 1.0, 0.25], [2.0, 0.0, -0.125],
                    [2.0, 1.0, 0.0]])
# elements in id order 3 5 7; CQUAD4/CQUADR corners a b c d stored a b d c
MIXED_CONN = np.array([[2, 5, 3, 6],    # CQUAD4* 3: 30 60 70 40
                       [1, 2, 4, 3],    # CQUADR 5: 20 30 40 50
                       [0, 1, 2, 3]])   # CQUAD4 7: 10 20 40 30
MIXED_BCS = np.array([0, 1, 2, 3, 4, 5,              # node 10: 123456
                      6 * 5 + 0, 6 * 5 + 2, 6 * 5 + 4,  # node 60: 135
                      6 * 6 + 1, 6 * 6 + 3, 6 * 6 + 5])  # node 70: 246


def read(path):
    ml = G.MeshLoader().scan_bdf_file(path)
    try:
        xpts, eptr, conn, comp = ml.connectivity()
        bn, bp, bv, bx = ml.bcs()
        ec, bcs, x2 = ml.assembler_data(6)
        names = [(ml.component_descript(c), ml.element_descript(c))
                 for c in range(ml.num_components)]
    finally:
        ml.close()
    assert np.array_equal(ec, conn) and np.array_equal(x2, xpts)
    return dict(xpts=xpts, eptr=eptr, conn=conn, comp=comp, bn=bn, bp=bp,
                bv=bv, bx=bx, bcs=bcs, names=names)


def test_mixed_file_known_answer(tmp_path):
    r = read(write(tmp_path / "mixed.bdf", MIXED))
    np.testing.assert_array_equal(r["xpts"], MIXED_X)
    np.testing.assert_array_equal(r["eptr"], [0, 4, 8, 12])
    np.testing.assert_array_equal(r["conn"].reshape(-1, 4), MIXED_CONN)
    np.testing.assert_array_equal(r["comp"], [1, 1, 0])
    np.testing.assert_array_equal(r["bn"], [0, 5, 6])
    np.testing.assert_array_equal(r["bp"], [0, 6, 9, 12])
    np.testing.assert_array_equal(r["bcs"], MIXED_BCS)
    np.testing.assert_array_equal(r["bx"], [0.0] * 9 + [1.5] * 3)
    assert r["names"] == [("SKIN.1", "CQUAD4"), ("SPAR.2", "CQUAD4")]


def ref_available():
    return oracle.ref_bdf_lib() is not None


@pytest.mark.skipif(not ref_available(),
                    reason="reference loader not built (no reference tree)")
@pytest.mark.parametrize("name", ["mixed", "plate_small", "plate_large",
                                  "wingbox_small", "cylinder_large"])
def test_matches_reference_loader(tmp_path, name):
    """Connectivity, global BC DOFs and coordinates equal
    TACSMeshLoader::getAssemblerCreatorData's, bit for bit."""
    p = str(tmp_path / f"{name}.bdf")
    if name == "mixed":
        write(p, MIXED)
    elif name == "plate_small":
        M.to_bdf(M.plate(9, 7, bc="ss", load="sine"), p)
    elif name == "plate_large":
        M.to_bdf(M.plate(9, 7, bc="clamped"), p, large_field=True)
    elif name == "wingbox_small":
        M.to_bdf(M.wingbox(nc=10, nh=3, nsb=2, nribs=4, seed=5), p)
    else:
        M.to_bdf(M.cylinder(12, 10, seed=2), p, large_field=True)
    ref = oracle.ref_bdf(p)
    assert ref is not None
    conn_r, bcs_r, x_r = ref
    r = read(p)
    np.testing.assert_array_equal(r["conn"].reshape(-1, 4), conn_r)
    np.testing.assert_array_equal(r["bcs"], bcs_r)
    assert np.array_equal(r["xpts"].view(np.uint64), x_r.view(np.uint64))


def test_standard_free_field_and_spc1(tmp_path):
    """Forms the reference misreads: comma-separated GRID from column 0,
    SPC1 with THRU and a continuation line, '+' continuation and a marker in
    columns 72-80 of an element line."""
    lines = ["GRID,1,,0.,0.,0.", "GRID,2,0,1.,0.,0.", "GRID,3,,1.,2.5-1,0.",
             "GRID,4,,0.,1.,1.+1",
             card("CQUAD4*", 1, 1, 1, 2, w=16).ljust(72) + "*A",
             card("*A", 3, 4, w=16),
             card("SPC1", 1, 3, 1, "THRU", 2).ljust(72) + "+S",
             card("+S", 4),
             card("CTRIA3", 2, 1, 1, 2),
             "+T      3"]
    r = read(write(tmp_path / "free.bdf", lines))
    np.testing.assert_array_equal(r["xpts"], [[0, 0, 0], [1, 0, 0],
                                              [1, 0.25, 0], [0, 1, 10]])
    np.testing.assert_array_equal(r["eptr"], [0, 4, 7])
    np.testing.assert_array_equal(r["conn"], [0, 1, 3, 2, 0, 1, 2])
    np.testing.assert_array_equal(r["bcs"], [2, 8, 20])


@pytest.mark.parametrize("lines,msg", [
    ([card("GRID", 1, 0, "0.", "0.", "0."), card("GRID", 1, 0, "1.", "0.", "0.")],
     "node id 1 defined twice"),
    ([card("GRID", 1, 0, "0.", "0.", "0."), card("CQUAD4", 1, 1, 1, 2, 3, 9)],
     "undefined node 2"),
    ([card("CQUAD4", 1, 1, 1, 2)], "too few nodes"),
    ([card("CQUAD4", 0, 1, 1, 2, 3, 4)], "no positive element"),
    ([card("GRID", 1, 0, "0.", "0.", "0."), card("SPC", 1, 5, 123, "0.")],
     "SPC on undefined node 5"),
])
def test_bad_files_raise(tmp_path, lines, msg):
    with pytest.raises(G.GfError) as e:
        G.MeshLoader().scan_bdf_file(write(tmp_path / "bad.bdf", lines))
    assert e.value.code == -1 and msg in str(e.value)


def test_missing_file_raises(tmp_path):
    with pytest.raises(G.GfError) as e:
        G.MeshLoader().scan_bdf_file(str(tmp_path / "absent.bdf"))
    assert "cannot open" in str(e.value)


def test_from_bdf_round_trip(tmp_path):
    """to_bdf then from_bdf (createFromBDF with one material): the same
    connectivity and BC DOFs, coordinates to the written precision, element
    data from the one material."""
    m = M.wingbox(nc=12, nh=3, nsb=3, nribs=5, seed=1)
    p = str(tmp_path / "wb.bdf")
    M.to_bdf(m, p, large_field=True)
    r = M.from_bdf(p, E=71e9, nu=0.33, thick=0.004)
    np.testing.assert_array_equal(r.conn, m.conn)
    np.testing.assert_array_equal(np.unique(r.bcs), np.unique(m.bcs))
    assert np.abs(r.xpts - m.xpts).max() <= 1e-9 * np.abs(m.xpts).max()
    assert r.data.shape == (m.nelems, 4)
    np.testing.assert_array_equal(r.data[0], [71e9, 0.33, 0.004, 0.0])
    assert r.loads.shape == (6 * m.nnodes,) and not r.loads.any()


def test_from_bdf_rejects_non_quads(tmp_path):
    lines = [card("GRID", k, 0, f"{k}.", "0.", "0.") for k in (1, 2, 3)]
    lines.append(card("CTRIA3", 1, 1, 1, 2, 3))
    with pytest.raises(ValueError):
        M.from_bdf(write(tmp_path / "tri.bdf", lines))
