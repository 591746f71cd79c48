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
"""
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


if __name__ == "__main__":
    kelem()
    symbolic()
    print("golden fixtures written to", HERE)
