"""Regenerate the committed golden fixtures (run in a container that has the
reference tree at /root/reference and a built oracle/_ref).

Fixtures (all numeric data, no reference source text):
  kelem_dense_plate1x0p5.npy  the 24x24 element stiffness values printed in
      examples/cusparse_solve/4_plate_soln.py:29-30 (6 significant digits;
      one MITC4 element, Lx=1, Ly=0.5, E=70e9, nu=0.3, t=5e-3, produced by
      examples/plate_demo/_archive/_one_elem_dense.cu through
      createPlateAssembler's counter-clockwise connectivity)
  symbolic_ref.npz  outputs of the reference's own TacsComputeRCMOrder,
      computeILUk (k = 0..3) and get_elem_ind_map compiled from
      include/linalg/reordering_utils.h and include/linalg/bsr_utils.h by
      oracle/build_ref.sh, on small seeded meshes (plates and a cylinder)
  iluk_highfill_ref.npz  the reference's computeILUk at the levels of fill
      the solver runs (k = 6, 12, 16, 32, 56, 64, 96) on AMD- and RCM-ordered
      patterns of mid-size meshes: input pattern, output row pointers and
      level counts per row, and the SHA-256 of the output column and level
      arrays (int32, little-endian)
"""
import hashlib
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF = "/root/reference"


def kelem():
    src = open(os.path.join(REF, "examples/cusparse_solve/4_plate_soln.py")).read()
    m = re.search(r"kelem_dense = np.array\(\[(.*?)\]\)", src, re.S)
    vals = np.array([float(x) for x in m.group(1).split(",")])
    assert vals.size == 576
    np.save(os.path.join(HERE, "kelem_dense_plate1x0p5.npy"), vals.reshape(24, 24))


def meshes():
    from gpu_fem_amd import meshes as M
    out = {}
    for nx, ny in [(2, 2), (3, 5), (8, 8), (12, 7)]:
        m = M.plate(nx, ny)
        out[f"plate{nx}x{ny}"] = (m.conn, m.nnodes)
    c = M.cylinder(8, 6)
    out["cyl8x6"] = (c.conn, c.nnodes)
    return out


def symbolic():
    import oracle
    assert oracle.ref() is not None, "build oracle/_ref first"
    data = {}
    for name, (conn, nn) in meshes().items():
        rp, cp = oracle.ref_node_graph(conn, nn)
        perm = oracle.ref_rcm(rp, cp)
        # permuted no-fill pattern, built by definition from the node graph
        prp, pcp = oracle.permute_pattern(rp, cp, perm)
        data[f"{name}/conn"] = conn
        data[f"{name}/rowp"] = rp
        data[f"{name}/cols"] = cp
        data[f"{name}/rcm"] = perm
        for k in range(4):
            frp, fcp, flv = oracle.ref_iluk(prp, pcp, k)
            data[f"{name}/ilu{k}_rowp"] = frp
            data[f"{name}/ilu{k}_cols"] = fcp
            data[f"{name}/ilu{k}_levs"] = flv
        data[f"{name}/elem_ind_map"] = oracle.ref_elem_ind_map(conn, nn, prp, pcp, perm)
        data[f"{name}/bandwidth"] = np.array([oracle.ref_bandwidth(prp, pcp)])
    np.savez_compressed(os.path.join(HERE, "symbolic_ref.npz"), **data)


HIGHFILL_K = (6, 12, 16, 32, 56, 64, 96)


def highfill_meshes():
    from gpu_fem_amd import meshes as M
    return {"plate40_amd": (M.plate(40, 40), "amd"),
            "plate100_amd": (M.plate(100, 100), "amd"),
            "plate30_rcm": (M.plate(30, 30), "rcm"),
            "cyl24x16_amd": (M.cylinder(24, 16, seed=11), "amd"),
            "wingbox_amd": (M.wingbox(nc=10, nh=3, nsb=2, nribs=4, seed=5),
                            "amd")}


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, "<i4").tobytes()).hexdigest()


def highfill():
    import gpu_fem_amd as G
    import oracle
    assert oracle.ref() is not None, "build oracle/_ref first"
    data = {}
    for name, (m, order) in highfill_meshes().items():
        rp, cp = oracle.ref_node_graph(m.conn, m.nnodes)
        # the input ordering only has to be fixed: the product's AMD or the
        # reference's RCM, stored with the pattern
        perm = G.amd_order(rp, cp) if order == "amd" else oracle.ref_rcm(rp, cp)
        prp, pcp = oracle.permute_pattern(rp, cp, perm)
        data[f"{name}/rowp"] = prp
        data[f"{name}/cols"] = pcp
        full = None
        for k in HIGHFILL_K + (10 ** 6,):
            frp, fcp, flv = oracle.ref_iluk(prp, pcp, k, fill=100.0)
            if k == 10 ** 6:
                full = frp[-1]
                continue
            data[f"{name}/ilu{k}_rowp"] = frp
            data[f"{name}/ilu{k}_rowlev"] = np.add.reduceat(
                flv, frp[:-1]).astype(np.int64) if len(flv) else flv
            data[f"{name}/ilu{k}_sha_cols"] = np.array(sha(fcp))
            data[f"{name}/ilu{k}_sha_levs"] = np.array(sha(flv))
        data[f"{name}/full_nnzb"] = np.array([full])
    np.savez_compressed(os.path.join(HERE, "iluk_highfill_ref.npz"), **data)


if __name__ == "__main__":
    kelem()
    symbolic()
    highfill()
    print("golden fixtures written to", HERE)
