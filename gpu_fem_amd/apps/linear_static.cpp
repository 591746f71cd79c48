// Linear-static plate solve written against the C++ host classes, following
// the flow of examples/iterative/gmres.cu:20-771 (createPlateAssembler ->
// RCM -> permuted pattern -> ILU(k) pattern -> post_reordering_update ->
// add_jacobian -> BCs -> ILUk_Preconditioner -> bsr_pre_solve -> GMRES ->
// bsr_post_solve).  Differences from that program: GMRES is the library's
// restarted right-preconditioned solver with a relative tolerance, and the
// connectivity order is an option (--ccw reproduces the reference generator).
// --bdf <file> reads the mesh from a NASTRAN bulk-data file instead, with one
// material (E = 70 GPa, nu = 0.3, t = 5 mm) and fz = 10 on every node, the
// set-up of examples/uCRM/linear_static.cu:13-80 (nxe is then ignored).
//
// usage: linear_static <nxe> <levFill> <rcm 0|1> <restart> <rtol> <out.bin>
//        [--ccw] [--direct] [--bdf file] [--vtk file]
// --vtk writes the mesh and the displacements as legacy VTK (printToVTK).
// prints one JSON line; writes the unpermuted solution (6*nnodes doubles).
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/gpufem/gpufem.hpp"

using namespace gpufem;

int main(int argc, char** argv) {
  if (argc < 7) {
    std::fprintf(stderr, "usage: %s nxe levFill rcm restart rtol out.bin "
                         "[--ccw] [--direct] [--bdf file] [--vtk file]\n",
                 argv[0]);
    return 2;
  }
  const int nxe = std::atoi(argv[1]), nye = nxe;
  const int levFill = std::atoi(argv[2]);
  const bool perform_rcm = std::atoi(argv[3]) != 0;
  const int restart = std::atoi(argv[4]);
  const double tolerance = std::atof(argv[5]);
  const char* out = argv[6];
  bool ccw = false, direct = false;
  const char* bdf = nullptr;
  const char* vtk = nullptr;
  for (int i = 7; i < argc; i++) {
    if (!std::strcmp(argv[i], "--vtk") && i + 1 < argc) vtk = argv[++i];
    if (!std::strcmp(argv[i], "--ccw")) ccw = true;
    if (!std::strcmp(argv[i], "--direct")) direct = true;
    if (!std::strcmp(argv[i], "--bdf") && i + 1 < argc) bdf = argv[++i];
  }

  using T = double;
  using Quad = QuadLinearQuadrature<T>;
  using Director = LinearizedRotation<T>;
  using Basis = ShellQuadBasis<T, Quad, 2>;
  constexpr bool has_ref_axis = false;
  using Data = ShellIsotropicData<T, has_ref_axis>;
  using Physics = IsotropicShell<T, Data>;
  using ElemGroup = ShellElementGroup<T, Director, Basis, Physics>;
  using Assembler = ElementAssembler<T, ElemGroup, VecType, BsrMat>;

  const double Lx = 2.0, Ly = 1.0, E = 70e9, nu = 0.3, thick = 0.005;
  TACSMeshLoader<T> mesh_loader;
  if (bdf && mesh_loader.scanBDFFile(bdf)) {
    std::fprintf(stderr, "linear_static: %s\n", mesh_loader.error().c_str());
    return 1;
  }
  try {
    auto assembler =
        bdf ? Assembler::createFromBDF(mesh_loader, Data(E, nu, thick))
            : createPlateAssembler<Assembler>(nxe, nye, Lx, Ly, E, nu, thick,
                                              ccw);
    std::vector<T> loads_h;
    if (bdf) {
      loads_h.assign(assembler.get_num_vars(), T(0));
      for (int i = 0; i < assembler.get_num_nodes(); i++) loads_h[6 * i + 2] = 10.0;
    } else {
      auto pl = getPlateLoads<T, Physics>(nxe, nye, Lx, Ly, 1.0);
      loads_h.assign(pl.begin(), pl.end());
    }
    if (direct) {
      // direct path: AMD + complete fill (plate_demo/linear_static.cu:29-67)
      assembler.symbolic_factorization(10.0, false);
      auto d_bsr = assembler.getBsrData().createDeviceBsrData();
      BsrMat<VecType<T>> kmat(d_bsr);
      auto res = assembler.createVarsVec();
      auto soln = assembler.createVarsVec();
      assembler.add_jacobian(res, kmat);
      assembler.apply_bcs(kmat);
      auto loads = assembler.createVarsVec(loads_h.data());
      assembler.apply_bcs(loads);
      solvers::direct_LU_solve(kmat, loads, soln);
      auto h = soln.createHostVec();
      if (vtk) printToVTK(assembler, h, vtk);
      FILE* f = std::fopen(out, "wb");
      std::fwrite(h.getPtr(), sizeof(T), h.getSize(), f);
      std::fclose(f);
      std::printf("{\"path\": \"direct\", \"nnodes\": %d, \"nnzb\": %d}\n",
                  assembler.get_num_nodes(), d_bsr.nnzb);
      return 0;
    }
    auto bsr_data = assembler.getBsrData();
    const int nnodes = bsr_data.nnodes;
    // final_perm: new -> old, final_iperm: old -> new (gmres.cu:93-108, 225-234)
    std::vector<int> final_perm(nnodes), final_iperm(nnodes);
    if (perform_rcm) {
      std::vector<int> rcm_vars(nnodes);
      TacsComputeRCMOrder(nnodes, bsr_data.rowPtr, bsr_data.colPtr,
                          rcm_vars.data(), 0, 1);
      for (int k = 0; k < nnodes; k++) final_perm[rcm_vars[k]] = k;
    } else {
      for (int k = 0; k < nnodes; k++) final_perm[k] = k;
    }
    for (int i = 0; i < nnodes; i++) final_iperm[final_perm[i]] = i;
    // permuted no-fill pattern for K, ILU(k) pattern for the preconditioner
    int *kmat_rowPtr = nullptr, *kmat_colPtr = nullptr;
    check(gf_pattern_permute(nnodes, bsr_data.rowPtr, bsr_data.colPtr,
                             final_iperm.data(), &kmat_rowPtr, &kmat_colPtr),
          "permute");
    const int kmat_nnzb = kmat_rowPtr[nnodes];
    int *rowPtr = kmat_rowPtr, *colPtr = kmat_colPtr;
    computeILUk(nnodes, kmat_nnzb, rowPtr, colPtr, levFill, 10.0);
    const int nnzb = rowPtr[nnodes];
    int* lperm = final_iperm.data();
    int* rperm = final_perm.data();
    assembler.bsr_data.post_reordering_update(kmat_nnzb, kmat_rowPtr,
                                              kmat_colPtr, lperm, rperm);
    auto precond_bsr_data = BsrData(nnodes, 6, nnzb, rowPtr, colPtr, lperm, rperm);
    gf_free(kmat_rowPtr);
    gf_free(kmat_colPtr);
    gf_free(rowPtr);
    gf_free(colPtr);
    auto d_precond_bsr_data = precond_bsr_data.createDeviceBsrData();
    auto d_bsr_data = assembler.getBsrData().createDeviceBsrData();
    auto precond_mat = BsrMat<VecType<T>>(d_precond_bsr_data);
    auto kmat = BsrMat<VecType<T>>(d_bsr_data);

    auto res = assembler.createVarsVec();
    auto soln = assembler.createVarsVec();
    assembler.add_jacobian(res, kmat);
    assembler.apply_bcs(res);
    assembler.apply_bcs(kmat);
    auto loads = assembler.createVarsVec(loads_h.data());
    assembler.apply_bcs(loads);

    auto precond = ILUk_Preconditioner<VecType<T>>(precond_mat, kmat);
    precond.factor();

    auto rhs = bsr_pre_solve<DeviceVec<T>>(kmat, loads, soln);
    auto stats = solvers::gmres(kmat, precond, rhs, soln, restart, 20000,
                                tolerance);
    const T max_resid = solvers::get_resid(kmat, rhs, soln);
    bsr_post_solve<DeviceVec<T>>(kmat, loads, soln);

    auto h = soln.createHostVec();
    if (vtk) printToVTK(assembler, h, vtk);
    FILE* f = std::fopen(out, "wb");
    std::fwrite(h.getPtr(), sizeof(T), h.getSize(), f);
    std::fclose(f);
    std::printf("{\"path\": \"gmres\", \"nnodes\": %d, \"nnzb_K\": %d, "
                "\"nnzb_ILU\": %d, \"iterations\": %d, \"rel_resid\": %.6e, "
                "\"max_resid\": %.6e}\n",
                nnodes, kmat_nnzb, nnzb, stats.iterations, stats.rel_resid,
                max_resid);
  } catch (const Error& e) {
    std::fprintf(stderr, "linear_static: %s (code %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
