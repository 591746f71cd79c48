// C wrapper around the reference's NASTRAN reader, TACSMeshLoader<double>
// (include/mesh/TACSMeshLoader.h, compiled as it lies under the reference
// tree by build_ref.sh into oracle/_ref/libgfref_bdf.so).  Test
// infrastructure: the BDF parity tests compare the product reader with it.
// Only the loader's defined accessors are used: scanBDFFile (:763),
// getNumNodes / getNumElements (:161-162), getAssemblerCreatorData
// (:193-224).
#include <cstring>

#include "mesh/TACSMeshLoader.h"

extern "C" {

struct gfref_bdf {
  int fail, nnodes, nelems, nbcs;
  int *conn, *bcs;
  double* xpts;
};

void* gfref_bdf_scan(const char* path, int vars_per_node) {
  auto* r = new gfref_bdf{};
  TACSMeshLoader<double> ml;
  r->fail = ml.scanBDFFile(path);
  if (r->fail) return r;
  ml.getAssemblerCreatorData(vars_per_node, r->nnodes, r->nelems, r->nbcs,
                             r->conn, r->bcs, r->xpts);
  return r;
}

// fail, nnodes, nelems, nbcs (total constrained DOFs)
void gfref_bdf_sizes(void* h, int* out) {
  auto* r = static_cast<gfref_bdf*>(h);
  out[0] = r->fail;
  out[1] = r->nnodes;
  out[2] = r->nelems;
  out[3] = r->nbcs;
}

// conn[4 * nelems] (4-node elements only), bcs[nbcs], xpts[3 * nnodes]
void gfref_bdf_get(void* h, int* conn, int nconn, int* bcs, double* xpts) {
  auto* r = static_cast<gfref_bdf*>(h);
  if (r->fail) return;
  std::memcpy(conn, r->conn, sizeof(int) * nconn);
  std::memcpy(bcs, r->bcs, sizeof(int) * r->nbcs);
  std::memcpy(xpts, r->xpts, sizeof(double) * 3 * r->nnodes);
}

void gfref_bdf_free(void* h) {
  auto* r = static_cast<gfref_bdf*>(h);
  delete[] r->conn;
  delete[] r->bcs;
  delete[] r->xpts;
  delete r;
}
}
