// Host symbolic phase of a linear-static problem, separated from its device
// set-up so that a caller (the serving loop) can size the device footprint
// exactly before it admits the problem.
#pragma once
#include <string>
#include <vector>

#include "common.h"

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
