// Host symbolic phase of a linear-static problem, separated from its device
// set-up so that a caller (the serving loop) can size the device footprint
// exactly before it admits the problem.
#pragma once
#include <string>
#include <vector>

#include "common.h"
#include "krylov.h"

// A linear-static problem resident on the device (gf_problem_*): mesh,
// patterns, maps, K and preconditioner values, Krylov work space.
struct gf_problem_s {
  gf_context ctx = nullptr;
  gf_options opt{};
  int nnodes = 0, nelems = 0, nbcs = 0;
  int nnzb_k = 0, nnzb_m = 0;
  std::vector<int> h_perm;
  // device arrays
  int *conn = nullptr, *bcs = nullptr, *perm = nullptr, *iperm = nullptr;
  double *xpts = nullptr, *data = nullptr;
  int *krowp = nullptr, *kcols = nullptr, *emap = nullptr;
  int *trowp = nullptr, *tcols = nullptr, *tmap = nullptr;
  double *kvals = nullptr, *mvals = nullptr;
  float *kvals32 = nullptr, *mvals32 = nullptr;
  double *rhs = nullptr, *sol = nullptr, *loads = nullptr;
  double *work64 = nullptr, *work64b = nullptr;
  float *t32a = nullptr, *t32b = nullptr, *t32c = nullptr;
  gf_ilu ilu = nullptr;
  gf_asm_plan plan = nullptr;  // element colouring for the atomic-free scatter
  gf::KrylovWork kw;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double t_assemble = 0, t_factor = 0;
};

namespace gf {

struct ProblemPlan {
  int nnodes = 0, nelems = 0, nbcs = 0;
  gf_options opt{};
  std::vector<int> perm, iperm;     // perm[old] = new
  std::vector<int> krp, kcl;        // permuted K pattern
  std::vector<int> mrp, mcl;        // ILU(k) pattern
  std::vector<int> emap;            // elemIndMap, 16 per element
  std::vector<int> tr, tc, tm;      // transpose pattern and block map
  // bytes of device memory gf_problem_create allocates for this plan
  // (rounded to the block classes of a caching context when caching)
  size_t device_bytes(bool caching) const;
};

// node graph -> reordering -> permuted pattern -> ILU(k) -> maps (host only)
int problem_plan(int nnodes, int nelems, const int* conn, int nbcs,
                 const int* bcs, const gf_options& opt, ProblemPlan& out,
                 std::string* why);
// uploads a plan and allocates the problem's device work space
int problem_create_planned(gf_context ctx, ProblemPlan& plan, const int* conn,
                           const double* xpts, const double* data,
                           const int* bcs, gf_problem* out);
// options check shared by gf_problem_create and the serving loop
bool options_valid(const gf_options& o);

}  // namespace gf
