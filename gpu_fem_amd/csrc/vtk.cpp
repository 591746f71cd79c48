// Legacy-VTK (ASCII) writer for a shell mesh and its displacement field
// (replaces printToVTK, include/mesh/vtk_writer.h:8-73).  The file is the
// one the reference writes: an unstructured grid of quads (cell type 9)
// whose corners go around the element (tensor-product corners 0 1 3 2), the
// displacements u v w of each node as the vector field "disp", reals in the
// default stream format (six significant digits, %g), and the reference's
// type name "double64".
#include <cstdio>

#include "../../include/gpufem_amd.h"

extern "C" int gf_write_vtk(const char* path, int nnodes, const double* xpts,
                            int nelems, const int* conn, const double* vars) {
  if (!path || nnodes < 0 || nelems < 0 || (nnodes && !xpts) ||
      (nelems && !conn))
    return GF_ERR_INVALID;
  for (long e = 0; e < 4L * nelems; e++)
    if (conn[e] < 0 || conn[e] >= nnodes) return GF_ERR_INVALID;
  FILE* f = fopen(path, "w");
  if (!f) return GF_ERR_INVALID;
  fprintf(f, "# vtk DataFile Version 2.0\nTACS GPU shell writer\nASCII\n");
  fprintf(f, "DATASET UNSTRUCTURED_GRID\nPOINTS %d double64\n", nnodes);
  for (long i = 0; i < nnodes; i++)
    fprintf(f, "%g %g %g\n", xpts[3 * i], xpts[3 * i + 1], xpts[3 * i + 2]);
  fprintf(f, "CELLS %d %d\n", nelems, 5 * nelems);
  for (long e = 0; e < nelems; e++) {
    const int* c = conn + 4 * e;
    fprintf(f, "4 %d %d %d %d\n", c[0], c[1], c[3], c[2]);
  }
  fprintf(f, "CELL_TYPES %d\n", nelems);
  for (long e = 0; e < nelems; e++) fputs("9\n", f);
  fprintf(f, "POINT_DATA %d\nVECTORS disp double64\n", nnodes);
  for (long i = 0; i < nnodes; i++) {
    const double* u = vars ? vars + 6 * i : nullptr;
    fprintf(f, "%g %g %g\n", u ? u[0] : 0.0, u ? u[1] : 0.0, u ? u[2] : 0.0);
  }
  const bool ok = !ferror(f);
  return (fclose(f) == 0 && ok) ? GF_OK : GF_ERR_INVALID;
}
