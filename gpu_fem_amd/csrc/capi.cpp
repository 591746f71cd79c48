// Context and linear-static problem entry points of the C ABI.
//
// gf_problem_* runs the flow of examples/iterative/gmres.cu:62-771 with the
// reference's data conventions: symbolic work on the host (node graph ->
// reordering -> permuted K pattern -> ILU(k) pattern -> elemIndMap ->
// transpose map), everything uploaded once, then on the device: Jacobian
// assembly, boundary conditions, copy into the preconditioner pattern,
// numeric ILU, and GMRES with the permuted right-hand side
// (solve_utils.h:4-16 bsr_pre_solve / bsr_post_solve).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "common.h"
#include "krylov.h"
#include "problem.h"
#include "symbolic.h"

namespace gf {
// worker = true: a serving worker's context, whose stream does not
// synchronise with the legacy default stream and whose freed device blocks
// are cached (see gf_context_s)
int context_create(int device, bool worker, gf_context* out) {
  if (!out) return GF_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return GF_ERR_DEVICE;
  auto* c = new gf_context_s;
  c->device = device;
  c->caching = worker;
  if (hipSetDevice(device) != hipSuccess ||
      (worker ? hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)
              : hipStreamCreate(&c->stream)) != hipSuccess) {
    delete c;
    return GF_ERR_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess)
    c->num_cus = prop.multiProcessorCount;
  *out = c;
  return GF_OK;
}
}  // namespace gf

extern "C" {

int gf_version(void) { return 1; }

int gf_context_create(int device, gf_context* out) {
  return gf::context_create(device, false, out);
}

int gf_context_destroy(gf_context c) {
  if (!c) return GF_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  gf::release_cache(c);
  if (c->pinned) hipHostFree(c->pinned);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return GF_OK;
}

int gf_context_synchronize(gf_context c) {
  if (!c) return GF_ERR_INVALID;
  return gf::check_hip(c, hipStreamSynchronize(c->stream), "synchronize");
}

const char* gf_context_last_error(gf_context c) {
  return c ? c->last_error.c_str() : "null context";
}

void* gf_context_stream(gf_context c) { return c ? (void*)c->stream : nullptr; }

void gf_options_default(gf_options* o) {
  if (!o) return;
  o->reorder = 1;
  o->lev_fill = 0;
  o->restart = 50;
  o->max_iters = 5000;
  o->rtol = 1e-8;
  o->precision = 0;
  o->stall_cycles = 0;
}

}  // extern "C"


namespace {

template <typename T>
int upload(gf_context ctx, T** dst, const T* src, size_t n) {
  GF_TRY(gf::dmalloc(ctx, (void**)dst, sizeof(T) * (n ? n : 1)));
  if (n && src)
    GF_HIP(ctx, hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice,
                               ctx->stream));
  return GF_OK;
}

template <typename T>
int dalloc(gf_context ctx, T** dst, size_t n) {
  return gf::dmalloc(ctx, (void**)dst, sizeof(T) * (n ? n : 1));
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  hipEventElapsedTime(&ms, a, b);
  return ms;
}

// Uploads a plan and allocates the device work space (every device buffer
// of a problem is allocated here; ProblemPlan::device_bytes mirrors the list)
int problem_setup(gf_problem p, gf::ProblemPlan& pl, const int* conn,
                  const double* xpts, const double* data, const int* bcs) {
  gf_context ctx = p->ctx;
  const int n = p->nnodes, ne = p->nelems;
  p->h_perm = pl.perm;
  p->nnzb_k = pl.krp[n];
  p->nnzb_m = pl.mrp[n];
  GF_TRY(upload(ctx, &p->conn, conn, (size_t)4 * ne));
  GF_TRY(upload(ctx, &p->xpts, xpts, (size_t)3 * n));
  GF_TRY(upload(ctx, &p->data, data, (size_t)4 * ne));
  GF_TRY(upload(ctx, &p->bcs, bcs, (size_t)p->nbcs));
  GF_TRY(upload(ctx, &p->perm, pl.perm.data(), (size_t)n));
  GF_TRY(upload(ctx, &p->iperm, pl.iperm.data(), (size_t)n));
  GF_TRY(upload(ctx, &p->krowp, pl.krp.data(), pl.krp.size()));
  GF_TRY(upload(ctx, &p->kcols, pl.kcl.data(), pl.kcl.size()));
  GF_TRY(upload(ctx, &p->emap, pl.emap.data(), pl.emap.size()));
  GF_TRY(gf_asm_plan_create(ctx, ne, n, conn, pl.emap.data(), p->nnzb_k,
                            &p->plan));
  GF_TRY(upload(ctx, &p->trowp, pl.tr.data(), pl.tr.size()));
  GF_TRY(upload(ctx, &p->tcols, pl.tc.data(), pl.tc.size()));
  GF_TRY(upload(ctx, &p->tmap, pl.tm.data(), pl.tm.size()));
  GF_TRY(gf_ilu_create(ctx, n, pl.mrp.data(), pl.mcl.data(), &p->ilu));
  const size_t N = 6 * (size_t)n;
  GF_TRY(dalloc(ctx, &p->kvals, 36 * (size_t)p->nnzb_k));
  GF_TRY(dalloc(ctx, &p->mvals, 36 * (size_t)p->nnzb_m));
  GF_TRY(dalloc(ctx, &p->rhs, N));
  GF_TRY(dalloc(ctx, &p->sol, N));
  GF_TRY(dalloc(ctx, &p->loads, N));
  GF_TRY(dalloc(ctx, &p->work64, N));
  GF_TRY(dalloc(ctx, &p->work64b, N));
  const bool mixed = p->opt.precision != 0;
  if (mixed) {
    if (p->opt.precision == 1)
      GF_TRY(dalloc(ctx, &p->kvals32, 36 * (size_t)p->nnzb_k));
    GF_TRY(dalloc(ctx, &p->mvals32, 36 * (size_t)p->nnzb_m));
    GF_TRY(dalloc(ctx, &p->t32a, N));
    GF_TRY(dalloc(ctx, &p->t32b, N));
    GF_TRY(dalloc(ctx, &p->t32c, N));
  }
  GF_TRY(p->kw.init(ctx, N, p->opt.restart, mixed, p->opt.precision != 1));
  GF_HIP(ctx, hipEventCreate(&p->ev[0]));
  GF_HIP(ctx, hipEventCreate(&p->ev[1]));
  GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GF_OK;
}

}  // namespace

namespace gf {

bool options_valid(const gf_options& o) {
  return o.restart >= 1 && o.restart <= 512 && o.lev_fill >= 0 &&
         o.reorder >= 0 && o.reorder <= 2 && o.precision >= 0 &&
         o.precision <= 2 && o.max_iters >= 0 && o.rtol > 0.0;
}

int problem_plan(int n, int ne, const int* conn, int nbcs, const int* bcs,
                 const gf_options& opt, ProblemPlan& pl, std::string* why) {
  auto fail = [&](int code, const char* msg) {
    if (why) *why = msg;
    return code;
  };
  for (size_t i = 0; i < 4 * (size_t)ne; i++)
    if (conn[i] < 0 || conn[i] >= n)
      return fail(GF_ERR_INVALID, "element node out of range");
  for (int i = 0; i < nbcs; i++)
    if (bcs[i] < 0 || bcs[i] >= 6 * n)
      return fail(GF_ERR_INVALID, "bc dof out of range");
  pl.nnodes = n;
  pl.nelems = ne;
  pl.nbcs = nbcs;
  pl.opt = opt;
  std::vector<int> rp, cl;
  GF_TRY(node_graph(ne, n, 4, conn, rp, cl));
  pl.perm.resize(n);
  if (opt.reorder == 1) {
    GF_TRY(rcm_order(n, rp.data(), cl.data(), 0, 1, pl.perm.data()));
  } else if (opt.reorder == 2) {
    GF_TRY(amd_order(n, rp.data(), cl.data(), pl.perm.data()));
  } else {
    for (int i = 0; i < n; i++) pl.perm[i] = i;
  }
  GF_TRY(pattern_permute(n, rp.data(), cl.data(), pl.perm.data(), pl.krp,
                         pl.kcl));
  GF_TRY(iluk_symbolic(n, pl.krp.data(), pl.kcl.data(), opt.lev_fill, pl.mrp,
                       pl.mcl, nullptr));
  pl.emap.resize((size_t)16 * ne);
  GF_TRY(elem_ind_map(ne, 4, conn, pl.krp.data(), pl.kcl.data(),
                      pl.perm.data(), pl.emap.data()));
  GF_TRY(transpose_map(n, pl.krp.data(), pl.kcl.data(), pl.tr, pl.tc, pl.tm));
  pl.iperm.resize(n);
  for (int i = 0; i < n; i++) pl.iperm[pl.perm[i]] = i;
  return GF_OK;
}

size_t ProblemPlan::device_bytes(bool caching) const {
  const size_t n = nnodes, ne = nelems, N = 6 * n;
  const size_t nk = krp.empty() ? 0 : krp[n], nm = mrp.empty() ? 0 : mrp[n];
  const size_t m = opt.restart;
  const bool mixed = opt.precision != 0;
  size_t total = 0;
  auto add = [&](size_t bytes) {
    if (bytes == 0) bytes = 1;
    total += caching ? cache_class(bytes) : bytes;
  };
  const size_t I = sizeof(int), D = sizeof(double), F = sizeof(float);
  // mesh, permutation, K pattern, maps, assembly plan, transpose
  add(4 * ne * I); add(3 * n * D); add(4 * ne * D); add(nbcs * I);
  add(n * I); add(n * I); add((n + 1) * I); add(nk * I); add(16 * ne * I);
  add(ne * I); add(ne * I);
  add((n + 1) * I); add(nk * I); add(nk * I);
  // ILU analysis: pattern, diagonal, level orders, chunk offsets (at most
  // one chunk per row), row descriptors, pivot, counter, flags, done
  add((n + 1) * I); add(nm * I); add(n * I); add(n * I); add(n * I);
  add((n + 1) * I); add(4 * n * I); add(4 * n * I); add((n + 1) * I);
  add(I); add(8); add(2 * I); add(n * I);
  // values and vectors
  add(36 * nk * D); add(36 * nm * D);
  for (int i = 0; i < 5; i++) add(N * D);
  if (mixed) {
    if (opt.precision == 1) add(36 * nk * F);
    add(36 * nm * F);
    for (int i = 0; i < 3; i++) add(N * F);
  }
  // Krylov work space (KrylovWork::init): basis and, outside GMRES-IR, its
  // preconditioned images
  add(N * (m + 1) * D);
  if (opt.precision != 1) add(N * (m > 0 ? m : 1) * D);
  for (int i = 0; i < 4; i++) add(N * D);
  add((m + 2) * D); add((m + 2) * D); add(2 * D); add((m + 1) * D);
  add(((m + 1) * m + 3 * m + 2) * D); add(256 * (2 * m + 4) * D);
  add((m + 1) * (m + 1) * D); add((2 * m + 4) * D); add((2 * m + 4) * D);
  add(D); add(4 * D); add(I);
  if (mixed)
    for (int i = 0; i < 3; i++) add(N * F);
  return total;
}

int problem_create_planned(gf_context ctx, ProblemPlan& pl, const int* conn,
                           const double* xpts, const double* data,
                           const int* bcs, gf_problem* out) {
  auto* p = new gf_problem_s;
  p->ctx = ctx;
  p->opt = pl.opt;
  p->nnodes = pl.nnodes;
  p->nelems = pl.nelems;
  p->nbcs = pl.nbcs;
  const int st = problem_setup(p, pl, conn, xpts, data, bcs);
  if (st != GF_OK) {
    gf_problem_destroy(p);
    return st;
  }
  *out = p;
  return GF_OK;
}

}  // namespace gf

extern "C" {

int gf_problem_destroy(gf_problem p) {
  if (!p) return GF_OK;
  void* bufs[] = {p->conn,   p->bcs,   p->perm,    p->iperm,   p->xpts,
                  p->data,   p->krowp, p->kcols,   p->emap,    p->trowp,
                  p->tcols,  p->tmap,  p->kvals,   p->mvals,   p->kvals32,
                  p->mvals32, p->rhs,  p->sol,     p->loads,   p->work64,
                  p->work64b, p->t32a, p->t32b,    p->t32c};
  for (void* b : bufs) gf::dfree(p->ctx, b);
  if (p->ilu) gf_ilu_destroy(p->ilu);
  if (p->plan) gf_asm_plan_destroy(p->plan);
  p->kw.release();
  if (p->ev[0]) hipEventDestroy(p->ev[0]);
  if (p->ev[1]) hipEventDestroy(p->ev[1]);
  delete p;
  return GF_OK;
}

int gf_problem_create(gf_context ctx, int nnodes, int nelems, const int* conn,
                      const double* xpts, const double* data, int nbcs,
                      const int* bcs, const gf_options* opt,
                      gf_problem* out) {
  if (!ctx || nnodes <= 0 || nelems <= 0 || !conn || !xpts || !data ||
      nbcs < 0 || (nbcs > 0 && !bcs) || !out)
    return gf::set_error(ctx, GF_ERR_INVALID, "problem_create: bad args");
  gf_options o;
  if (opt)
    o = *opt;
  else
    gf_options_default(&o);
  if (!gf::options_valid(o))
    return gf::set_error(ctx, GF_ERR_INVALID, "problem_create: bad options");
  gf::ProblemPlan pl;
  std::string why;
  int st = gf::problem_plan(nnodes, nelems, conn, nbcs, bcs, o, pl, &why);
  if (st != GF_OK)
    return gf::set_error(ctx, st, "problem_create: %s",
                         why.empty() ? "symbolic phase failed" : why.c_str());
  return gf::problem_create_planned(ctx, pl, conn, xpts, data, bcs, out);
}

int gf_problem_assemble(gf_problem p) {
  if (!p) return GF_ERR_INVALID;
  gf_context ctx = p->ctx;
  GF_HIP(ctx, hipEventRecord(p->ev[0], ctx->stream));
  GF_TRY(gf_assemble_jacobian_colored_f64(ctx, p->plan, p->conn, p->xpts,
                                          nullptr, p->data, p->emap, p->perm,
                                          p->kvals, nullptr));
  GF_TRY(gf_bc_apply_mat(ctx, p->nbcs, p->bcs, p->krowp, p->kcols, p->trowp,
                         p->tcols, p->tmap, p->perm, p->kvals));
  GF_TRY(gf_bsr_copy_to_pattern(ctx, p->nnodes, p->krowp, p->kcols, p->kvals,
                                gf_ilu_device_rowp(p->ilu),
                                gf_ilu_device_cols(p->ilu), p->mvals));
  GF_HIP(ctx, hipEventRecord(p->ev[1], ctx->stream));
  GF_HIP(ctx, hipEventSynchronize(p->ev[1]));
  p->t_assemble = 1e-3 * elapsed(p->ev[0], p->ev[1]);
  return GF_OK;
}

int gf_problem_factor(gf_problem p) {
  if (!p) return GF_ERR_INVALID;
  gf_context ctx = p->ctx;
  GF_HIP(ctx, hipEventRecord(p->ev[0], ctx->stream));
  int zp = -1;
  if (p->opt.precision != 0) {
    GF_TRY(gf_convert_f64_f32(ctx, 36 * (size_t)p->nnzb_m, p->mvals,
                              p->mvals32));
    if (p->opt.precision == 1)
      GF_TRY(gf_convert_f64_f32(ctx, 36 * (size_t)p->nnzb_k, p->kvals,
                                p->kvals32));
    GF_TRY(gf_bsr_ilu_factor_f32(p->ilu, p->mvals32, &zp));
  } else {
    GF_TRY(gf_bsr_ilu_factor_f64(p->ilu, p->mvals, &zp));
  }
  GF_HIP(ctx, hipEventRecord(p->ev[1], ctx->stream));
  GF_HIP(ctx, hipEventSynchronize(p->ev[1]));
  p->t_factor = 1e-3 * elapsed(p->ev[0], p->ev[1]);
  return GF_OK;
}

int gf_problem_solve(gf_problem p, const double* loads, double* x,
                     int is_device, gf_solve_info* info) {
  if (!p || !loads || !x) return GF_ERR_INVALID;
  gf_context ctx = p->ctx;
  hipStream_t s = ctx->stream;
  const size_t N = 6 * (size_t)p->nnodes;
  GF_HIP(ctx, hipEventRecord(p->ev[0], s));
  GF_HIP(ctx, hipMemcpyAsync(p->loads, loads, sizeof(double) * N,
                             is_device ? hipMemcpyDeviceToDevice
                                       : hipMemcpyHostToDevice,
                             s));
  GF_TRY(gf_bc_apply_vec(ctx, p->nbcs, p->bcs, nullptr, p->loads));
  GF_TRY(gf_permute_vec(ctx, p->nnodes, p->perm, p->loads, p->rhs));
  GF_HIP(ctx, hipMemsetAsync(p->sol, 0, sizeof(double) * N, s));
  gf::Operator op;
  op.ctx = ctx;
  op.n = p->nnodes;
  op.krowp = p->krowp;
  op.kcols = p->kcols;
  op.kvals = p->kvals;
  op.kvals32 = p->kvals32;
  op.ilu = p->ilu;
  op.mlu = p->mvals;
  op.mlu32 = p->mvals32;
  op.work64 = p->work64;
  op.work64b = p->work64b;
  op.t32a = p->t32a;
  op.t32b = p->t32b;
  op.t32c = p->t32c;
  gf::GmresResult gr;
  const int st = gf::gmres_solve(p->kw, op, p->rhs, p->sol, p->opt.max_iters,
                                 p->opt.stall_cycles,
                                 p->opt.rtol, p->opt.precision, &gr);
  // x = inverse permutation of sol: x[old] = sol[perm[old]]
  GF_TRY(gf_permute_vec(ctx, p->nnodes, p->iperm, p->sol, p->work64));
  GF_HIP(ctx, hipMemcpyAsync(x, p->work64, sizeof(double) * N,
                             is_device ? hipMemcpyDeviceToDevice
                                       : hipMemcpyDeviceToHost,
                             s));
  GF_HIP(ctx, hipEventRecord(p->ev[1], s));
  GF_HIP(ctx, hipEventSynchronize(p->ev[1]));
  if (info) {
    info->status = st;
    info->iterations = gr.iterations;
    info->outer = gr.outer;
    info->rel_resid = gr.rel_resid;
    info->t_assemble = p->t_assemble;
    info->t_factor = p->t_factor;
    info->t_solve = 1e-3 * elapsed(p->ev[0], p->ev[1]);
  }
  return st;
}

int gf_problem_assemble_tangent(gf_problem p, const double* u, double* res,
                                int is_device) {
  if (!p || !u) return GF_ERR_INVALID;
  gf_context ctx = p->ctx;
  hipStream_t s = ctx->stream;
  const size_t N = 6 * (size_t)p->nnodes;
  GF_HIP(ctx, hipEventRecord(p->ev[0], s));
  const double* du = u;
  if (!is_device) {
    GF_HIP(ctx, hipMemcpyAsync(p->work64, u, sizeof(double) * N,
                               hipMemcpyHostToDevice, s));
    du = p->work64;
  }
  // K(u) and R(u) (R at 6 perm[node] + k, in work64b)
  GF_TRY(gf_assemble_tangent_colored_f64(ctx, p->plan, p->conn, p->xpts, du,
                                         p->data, p->emap, p->perm, p->kvals,
                                         p->work64b));
  GF_TRY(gf_bc_apply_mat(ctx, p->nbcs, p->bcs, p->krowp, p->kcols, p->trowp,
                         p->tcols, p->tmap, p->perm, p->kvals));
  GF_TRY(gf_bsr_copy_to_pattern(ctx, p->nnodes, p->krowp, p->kcols, p->kvals,
                                gf_ilu_device_rowp(p->ilu),
                                gf_ilu_device_cols(p->ilu), p->mvals));
  if (res) {
    // unpermuted R: r[old] = R[perm[old]], constrained entries zero
    GF_TRY(gf_permute_vec(ctx, p->nnodes, p->iperm, p->work64b, p->rhs));
    GF_TRY(gf_bc_apply_vec(ctx, p->nbcs, p->bcs, nullptr, p->rhs));
    GF_HIP(ctx, hipMemcpyAsync(res, p->rhs, sizeof(double) * N,
                               is_device ? hipMemcpyDeviceToDevice
                                         : hipMemcpyDeviceToHost,
                               s));
  }
  GF_HIP(ctx, hipEventRecord(p->ev[1], s));
  GF_HIP(ctx, hipEventSynchronize(p->ev[1]));
  p->t_assemble = 1e-3 * elapsed(p->ev[0], p->ev[1]);
  return GF_OK;
}

namespace {
// ||x|| (device) through the problem's Krylov reduction (deterministic)
int problem_norm(gf_problem p, const double* x, double* out) {
  gf_context ctx = p->ctx;
  GF_TRY(p->kw.dots(x, p->kw.N, 1, x, p->kw.scal, 2));
  GF_HIP(ctx, hipMemcpyAsync(p->kw.host_scal, p->kw.scal, sizeof(double),
                             hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *out = p->kw.host_scal[0];
  return GF_OK;
}
}  // namespace

int gf_problem_newton_solve(gf_problem p, const double* loads, double* u,
                            int is_device, int max_newton, double newton_rtol,
                            double* history, gf_newton_info* info) {
  if (!p || !loads || !u || max_newton < 0 || !(newton_rtol > 0.0))
    return GF_ERR_INVALID;
  gf_context ctx = p->ctx;
  hipStream_t s = ctx->stream;
  const size_t N = 6 * (size_t)p->nnodes;
  const auto t0 = std::chrono::steady_clock::now();
  double *f = nullptr, *x = nullptr, *r = nullptr, *dx = nullptr;
  int st = gf::dmalloc(ctx, (void**)&f, sizeof(double) * N);
  if (!st) st = gf::dmalloc(ctx, (void**)&x, sizeof(double) * N);
  if (!st) st = gf::dmalloc(ctx, (void**)&r, sizeof(double) * N);
  if (!st) st = gf::dmalloc(ctx, (void**)&dx, sizeof(double) * N);
  gf_newton_info ni{};
  ni.status = st;
  auto run = [&]() -> int {
    const hipMemcpyKind k =
        is_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    GF_HIP(ctx, hipMemcpyAsync(f, loads, sizeof(double) * N, k, s));
    GF_HIP(ctx, hipMemcpyAsync(x, u, sizeof(double) * N, k, s));
    GF_TRY(gf_bc_apply_vec(ctx, p->nbcs, p->bcs, nullptr, f));
    GF_TRY(gf_bc_apply_vec(ctx, p->nbcs, p->bcs, nullptr, x));
    double fn = 0.0;
    GF_TRY(problem_norm(p, f, &fn));
    if (fn == 0.0) fn = 1.0;
    for (int it = 0;; it++) {
      // r = f - R(x) with the tangent K(x) assembled and factored
      GF_TRY(gf_problem_assemble_tangent(p, x, r, 1));
      GF_TRY(gf_vec_axpby_f64(ctx, N, 1.0, f, -1.0, r));
      double rn = 0.0;
      GF_TRY(problem_norm(p, r, &rn));
      ni.rel_resid = rn / fn;
      if (history) history[it] = ni.rel_resid;
      ni.newton_iterations = it;
      if (!(ni.rel_resid > newton_rtol)) return GF_OK;
      if (it == max_newton) return GF_ERR_NO_CONVERGENCE;
      GF_TRY(gf_problem_factor(p));
      gf_solve_info si{};
      const int ss = gf_problem_solve(p, r, dx, 1, &si);
      ni.krylov_iterations += si.iterations;
      if (ss != GF_OK) return ss;
      GF_TRY(gf_vec_axpby_f64(ctx, N, 1.0, dx, 1.0, x));
    }
  };
  if (!st) {
    st = run();
    ni.status = st;
    const int cs = hipMemcpyAsync(u, x, sizeof(double) * N,
                                  is_device ? hipMemcpyDeviceToDevice
                                            : hipMemcpyDeviceToHost,
                                  s) == hipSuccess &&
                           hipStreamSynchronize(s) == hipSuccess
                       ? GF_OK
                       : GF_ERR_DEVICE;
    if (st == GF_OK && cs != GF_OK)
      st = gf::set_error(ctx, cs, "newton_solve: copy-out failed");
  }
  gf::dfree(ctx, f);
  gf::dfree(ctx, x);
  gf::dfree(ctx, r);
  gf::dfree(ctx, dx);
  ni.t_total = std::chrono::duration<double>(std::chrono::steady_clock::now() -
                                             t0)
                   .count();
  if (info) *info = ni;
  return st;
}

int gf_problem_sizes(gf_problem p, int* nnodes, int* nnzb_k, int* nnzb_m) {
  if (!p) return GF_ERR_INVALID;
  if (nnodes) *nnodes = p->nnodes;
  if (nnzb_k) *nnzb_k = p->nnzb_k;
  if (nnzb_m) *nnzb_m = p->nnzb_m;
  return GF_OK;
}

int gf_problem_device_arrays(gf_problem p, const int** krowp,
                             const int** kcols, double** kvals,
                             const int** mrowp, const int** mcols,
                             double** mvals, const int** perm) {
  if (!p) return GF_ERR_INVALID;
  if (krowp) *krowp = p->krowp;
  if (kcols) *kcols = p->kcols;
  if (kvals) *kvals = p->kvals;
  if (mrowp) *mrowp = gf_ilu_device_rowp(p->ilu);
  if (mcols) *mcols = gf_ilu_device_cols(p->ilu);
  if (mvals) *mvals = p->mvals;
  if (perm) *perm = p->perm;
  return GF_OK;
}

gf_ilu gf_problem_ilu(gf_problem p) { return p ? p->ilu : nullptr; }

int gf_problem_host_perm(gf_problem p, int* perm) {
  if (!p || !perm) return GF_ERR_INVALID;
  std::memcpy(perm, p->h_perm.data(), sizeof(int) * p->nnodes);
  return GF_OK;
}

}  // extern "C"

extern "C" int gf_gmres_solve(gf_context ctx, int nnodes, const int* krowp,
                              const int* kcols, const double* kvals,
                              const float* kvals32, gf_ilu ilu,
                              const double* mlu, const float* mlu32,
                              const double* b, double* x, int restart,
                              int max_iters, double rtol, int precision,
                              int* iterations, double* rel_resid) {
  const bool mixed = precision != 0;
  if (!ctx || nnodes <= 0 || !krowp || !kcols || !kvals || !ilu || !b || !x ||
      restart < 1 || restart > 512 || precision < 0 || precision > 2 ||
      (!mixed && !mlu) || (mixed && !mlu32) ||
      (precision == 1 && !kvals32))
    return gf::set_error(ctx, GF_ERR_INVALID, "gmres_solve: bad args");
  // the sweeps launch on the ILU handle's stream and the Krylov kernels on
  // ctx's: both must be the same stream for the solve to be ordered
  if (gf::ilu_context(ilu) != ctx)
    return gf::set_error(ctx, GF_ERR_INVALID,
                         "gmres_solve: the ILU handle belongs to another "
                         "context");
  const size_t N = 6 * (size_t)nnodes;
  gf::KrylovWork kw;
  double *w64 = nullptr, *w64b = nullptr;
  float *a = nullptr, *bb = nullptr, *c = nullptr;
  int st = kw.init(ctx, N, restart, mixed, precision != 1);
  if (!st) st = gf::dmalloc(ctx, (void**)&w64, sizeof(double) * N);
  if (!st) st = gf::dmalloc(ctx, (void**)&w64b, sizeof(double) * N);
  if (!st && mixed) {
    st = gf::dmalloc(ctx, (void**)&a, sizeof(float) * N);
    if (!st) st = gf::dmalloc(ctx, (void**)&bb, sizeof(float) * N);
    if (!st) st = gf::dmalloc(ctx, (void**)&c, sizeof(float) * N);
  }
  gf::GmresResult gr;
  if (!st) {
    gf::Operator op;
    op.ctx = ctx;
    op.n = nnodes;
    op.krowp = krowp;
    op.kcols = kcols;
    op.kvals = kvals;
    op.kvals32 = kvals32;
    op.ilu = ilu;
    op.mlu = mlu;
    op.mlu32 = mlu32;
    op.work64 = w64;
    op.work64b = w64b;
    op.t32a = a;
    op.t32b = bb;
    op.t32c = c;
    st = gf::gmres_solve(kw, op, b, x, max_iters, 0, rtol, precision, &gr);
  }
  hipStreamSynchronize(ctx->stream);
  gf::dfree(ctx, w64);
  gf::dfree(ctx, w64b);
  gf::dfree(ctx, a);
  gf::dfree(ctx, bb);
  gf::dfree(ctx, c);
  kw.release();
  if (iterations) *iterations = gr.iterations;
  if (rel_resid) *rel_resid = gr.rel_resid;
  return st;
}
