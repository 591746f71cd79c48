/*
 * gpufem_amd -- C ABI of the MI355X (gfx950) linear-static MITC4 shell path.
 *
 * The reference (smdogroup/gpu_fem) has no C ABI: its operator surface is a
 * header-only C++ template library.  Each entry point below replaces the
 * reference function or closed-library call cited beside it; the C++ host
 * classes in include/gpufem/ (ElementAssembler, BsrData, BsrMat, ...) keep the
 * reference call surface and are built on these functions.
 *
 * Conventions
 *   - every function returns an int status (GF_OK = 0, negative on error);
 *     gf_context_last_error() describes the last failure on a context
 *   - index arrays are int32; BSR blocks are 6x6, row-major inside the block
 *     (val[36*b + 6*r + c]); perm[old] = new, iperm[new] = old
 *   - "d_" pointers are device memory (hipMalloc'ed by the caller), all other
 *     pointers are host memory; device work is stream-ordered on the context
 *     stream and nothing on the hot path synchronises the device
 *   - host arrays returned through out-pointers are allocated by the library
 *     and released with gf_free()
 *   - a context is not thread-safe: use one context per host thread
 */
#ifndef GPUFEM_AMD_H
#define GPUFEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GF_OK = 0,
  GF_ERR_INVALID = -1,       /* bad argument or shape */
  GF_ERR_NOMEM = -2,         /* device or host allocation failed */
  GF_ERR_DEVICE = -3,        /* HIP runtime error */
  GF_ERR_ZERO_PIVOT = -4,    /* singular diagonal block in ILU */
  GF_ERR_BREAKDOWN = -5,     /* GMRES breakdown (H[j+1,j] == 0 early) */
  GF_ERR_NO_CONVERGENCE = -6,/* max iterations reached */
  GF_ERR_OVER_BUDGET = -7,   /* serving admission over the HBM budget */
  GF_ERR_TIMEOUT = -8        /* a bounded device wait expired */
};

typedef struct gf_context_s* gf_context;
typedef struct gf_ilu_s* gf_ilu;
typedef struct gf_problem_s* gf_problem;
typedef struct gf_server_s* gf_server;
typedef struct gf_bdf_s* gf_bdf;

/* ---- context (replaces the implicit default stream + exit() on error of
 *      include/cuda_utils.h:33-83) ---- */
int gf_context_create(int device, gf_context* out);
int gf_context_destroy(gf_context ctx);
int gf_context_synchronize(gf_context ctx);
const char* gf_context_last_error(gf_context ctx);
void* gf_context_stream(gf_context ctx);
void gf_free(void* host_ptr);
int gf_version(void);

/* ---- host symbolic phase (include/linalg/bsr_utils.h, reordering_utils.h) */
/* node adjacency incl. self, sorted (bsr_utils.h:419-476 get_row_col_ptrs*) */
int gf_node_graph(int nelems, int nnodes, int nodes_per_elem, const int* conn,
                  int** rowp, int** cols, int* nnzb);
/* reverse Cuthill-McKee, rcm_vars[old] = new
 * (reordering_utils.h:272 TacsComputeRCMOrder) */
int gf_rcm_order(int nnodes, const int* rowp, const int* cols, int root,
                 int iters, int* rcm_vars);
/* approximate minimum degree, perm[old] = new (bsr_utils.h:510-567 via
 * sparse-utils BSRMatAMDFactorSymbolicCUDA, which is not available) */
int gf_amd_order(int nnodes, const int* rowp, const int* cols, int* perm);
/* B(perm[i], perm[j]) != 0 iff A(i, j) != 0, sorted (gmres.cu:242-252
 * BSRMatApplyPerm) */
int gf_pattern_permute(int nnodes, const int* rowp, const int* cols,
                       const int* perm, int** prowp, int** pcols);
/* level-of-fill ILU(k) pattern (bsr_utils.h:678 computeILUk) */
int gf_iluk_symbolic(int nnodes, const int* rowp, const int* cols,
                     int lev_fill, int** frowp, int** fcols, int* nnzb);
/* block index of (perm[conn[n]], perm[conn[m]]) per element
 * (bsr_utils.h:809 get_elem_ind_map); map has nelems*npe*npe entries */
int gf_elem_ind_map(int nelems, int nodes_per_elem, const int* conn,
                    const int* rowp, const int* cols, const int* perm,
                    int* map);
/* transpose pattern + transpose_block_map (bsr_utils.h:157-211) */
int gf_transpose_map(int nnodes, const int* rowp, const int* cols, int** trowp,
                     int** tcols, int** tmap);
/* max |i - j| over the pattern (bsr_utils.h:655 getBandWidth) */
int gf_bandwidth(int nnodes, const int* rowp, const int* cols);

/* ---- device kernels on caller-owned device buffers ---- */
/* MITC4 Jacobian + residual (assembler.h:267 add_jacobian ->
 * elem_group.cuh:101 add_jacobian_gpu).  data = {E, nu, thick, tOffset} per
 * element.  kvals are accumulated into (callers zero them first), res is
 * accumulated at 6*perm[node]+k; vars may be NULL (zero displacement), res
 * may be NULL. */
int gf_assemble_jacobian_mitc4_f64(gf_context ctx, int nelems, int nnodes,
                                   const int* d_conn, const double* d_xpts,
                                   const double* d_vars, const double* d_data,
                                   const int* d_elem_ind_map,
                                   const int* d_perm, double* d_kvals,
                                   double* d_res);
/* Conflict-free variant of the Jacobian scatter: the elements are coloured
 * on the host (no two elements of a colour share a node) and each colour is
 * one launch whose contributions are plain stores (first element to touch a
 * block) or plain read-add-writes, so K and res need no zeroing and no
 * atomics are used.  The plan holds the colouring and the first-touch masks
 * for one connectivity + elemIndMap (host arrays; nnzb = K blocks).  Results
 * equal gf_assemble_jacobian_mitc4_f64 on zeroed K/res up to the order of
 * the (at most four per entry for quads) FP64 additions. */
typedef struct gf_asm_plan_s* gf_asm_plan;
int gf_asm_plan_create(gf_context ctx, int nelems, int nnodes, const int* conn,
                       const int* elem_ind_map, int nnzb, gf_asm_plan* out);
int gf_asm_plan_destroy(gf_asm_plan plan);
int gf_asm_plan_colors(gf_asm_plan plan);
int gf_assemble_jacobian_colored_f64(gf_context ctx, gf_asm_plan plan,
                                     const int* d_conn, const double* d_xpts,
                                     const double* d_vars, const double* d_data,
                                     const int* d_elem_ind_map,
                                     const int* d_perm, double* d_kvals,
                                     double* d_res);
/* residual only (assembler.h:232 add_residual) */
int gf_assemble_residual_mitc4_f64(gf_context ctx, int nelems, int nnodes,
                                   const int* d_conn, const double* d_xpts,
                                   const double* d_vars, const double* d_data,
                                   const int* d_perm, double* d_res);
/* strain energy (assembler.h:198 add_energy); writes one double */
int gf_energy_mitc4_f64(gf_context ctx, int nelems, const int* d_conn,
                        const double* d_xpts, const double* d_vars,
                        const double* d_data, double* d_energy);
/* Nonlinear element, IsotropicShell<..., isNonlinear = true>
 * (include/shell/physics.h:6-17, strain a2d/a2dshellstrain.h:72-93): the
 * tangent K(u) (the Hessian of the strain energy at d_vars) added into
 * d_kvals through elemIndMap, the residual R(u) (its gradient) added at
 * 6*perm[node]+k into d_res, the energy (physics.h:18-36) written to
 * *d_energy; replaces add_jacobian / add_residual / add_energy of the
 * nonlinear element group (include/base/elem_group.cuh:101-211,
 * include/base/elem_group.h:40-181).  d_kvals, d_res and d_energy may each
 * be NULL; d_kvals and d_res are not zeroed (FP64 atomic adds). */
int gf_assemble_tangent_mitc4_f64(gf_context ctx, int nelems, int nnodes,
                                  const int* d_conn, const double* d_xpts,
                                  const double* d_vars, const double* d_data,
                                  const int* d_elem_ind_map,
                                  const int* d_perm, double* d_kvals,
                                  double* d_res, double* d_energy);
/* the same on a colouring plan (no atomics, no zeroing; deterministic) */
int gf_assemble_tangent_colored_f64(gf_context ctx, gf_asm_plan plan,
                                    const int* d_conn, const double* d_xpts,
                                    const double* d_vars, const double* d_data,
                                    const int* d_elem_ind_map,
                                    const int* d_perm, double* d_kvals,
                                    double* d_res);
/* zero rows and columns of constrained DOFs, unit diagonal
 * (bsr_mat.cuh:2-104 apply_mat_bcs_rows/cols_kernel) */
int gf_bc_apply_mat(gf_context ctx, int nbcs, const int* d_bcs,
                    const int* d_rowp, const int* d_cols, const int* d_trowp,
                    const int* d_tcols, const int* d_tmap, const int* d_perm,
                    double* d_vals);
/* vec[bc] = 0 (vec.cuh:4 apply_vec_bcs_kernel); d_perm may be NULL for an
 * unpermuted vector */
int gf_bc_apply_vec(gf_context ctx, int nbcs, const int* d_bcs,
                    const int* d_perm, double* d_vec);
/* copy K blocks into a superset pattern, zero elsewhere
 * (bsr_mat.cuh:107 copy_mat_values_kernel) */
int gf_bsr_copy_to_pattern(gf_context ctx, int nnodes, const int* d_rowp,
                           const int* d_cols, const double* d_vals,
                           const int* d_trowp, const int* d_tcols,
                           double* d_tvals);
/* out[6*perm[i]+k] = in[6*i+k] (vec.cuh:16 permute_vec_kernel) */
int gf_permute_vec(gf_context ctx, int nnodes, const int* d_perm,
                   const double* d_in, double* d_out);
/* y = A x (cusparseDbsrmv, gmres.cu:588) */
int gf_bsr_spmv_f64(gf_context ctx, int nnodes, const int* d_rowp,
                    const int* d_cols, const double* d_vals, const double* d_x,
                    double* d_y);
int gf_bsr_spmv_f32(gf_context ctx, int nnodes, const int* d_rowp,
                    const int* d_cols, const float* d_vals, const float* d_x,
                    float* d_y);
int gf_convert_f64_f32(gf_context ctx, size_t n, const double* d_in,
                       float* d_out);
/* y = a x + b y (cublasDaxpy / cublasDscal, gmres.cu:581-666) */
int gf_vec_axpby_f64(gf_context ctx, size_t n, double a, const double* d_x,
                     double b, double* d_y);
/* r = b - A x accumulated with error-free transformations (compensated
 * dot products), as accurate as twice-FP64 arithmetic rounded once; the true
 * residual of the GMRES outer loop.  r may alias b but not x. */
int gf_bsr_resid_comp_f64(gf_context ctx, int nnodes, const int* d_rowp,
                          const int* d_cols, const double* d_vals,
                          const double* d_x, const double* d_b, double* d_r);
/* r = A x - b: the entry of r of largest magnitude (signed, first index on
 * ties; CUSPARSE::get_resid, include/linalg/solvers/cusparse.h:853-918) and
 * ||r||_2.  Synchronises the context stream. */
int gf_bsr_resid_f64(gf_context ctx, int nnodes, const int* d_rowp,
                     const int* d_cols, const double* d_vals,
                     const double* d_x, const double* d_b, double* max_entry,
                     double* norm2);

/* gf_bsr_resid_f64 without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  The partial
 * results live in a static device array of the library (3 MiB per device,
 * held until the process ends): meshes of at most 5,242,880 nodes
 * (GF_ERR_INVALID above).  d_out[0] = the entry of A x - b of largest
 * magnitude (signed, first index on ties), d_out[1] = ||A x - b||_2, in
 * device memory, stream-ordered; nothing synchronises.  Calls that overlap
 * on one device share the array and race. */
int gf_bsr_resid_f64_noalloc(gf_context ctx, int nnodes, const int* d_rowp,
                             const int* d_cols, const double* d_vals,
                             const double* d_x, const double* d_b,
                             double* d_out);

/* ---- ILU analysis, numeric factorisation and triangular sweeps
 *      (cusparseXbsrilu02 / cusparseXbsrsv2, gmres.cu:374-465, 609-625) ---- */
/* host pattern in, uploads the pattern and the level schedules */
int gf_ilu_create(gf_context ctx, int nnodes, const int* rowp, const int* cols,
                  gf_ilu* out);
int gf_ilu_destroy(gf_ilu ilu);
int gf_ilu_info(gf_ilu ilu, int* nnzb, int* nlevels_lower,
                int* nlevels_upper);
const int* gf_ilu_device_rowp(gf_ilu ilu);
const int* gf_ilu_device_cols(gf_ilu ilu);
/* GF_ERR_TIMEOUT if a sweep's bounded dependency wait expired since the last
 * check (synchronises the context stream) */
int gf_ilu_check(gf_ilu ilu);
/* in place on device values laid out on the ILU pattern; on
 * GF_ERR_ZERO_PIVOT *zero_pivot holds the first singular block row.
 * Rows are factored on a dataflow schedule (each row waits for its own pivot
 * rows; GPUFEM_ILU_SCHEDULE=level runs one launch per level instead, with
 * bitwise-identical factors); GF_ERR_TIMEOUT if a bounded wait expired.
 * Synchronises the context stream. */
int gf_bsr_ilu_factor_f64(gf_ilu ilu, double* d_vals, int* zero_pivot);
int gf_bsr_ilu_factor_f32(gf_ilu ilu, float* d_vals, int* zero_pivot);
/* x = U^-1 L^-1 b with the factored values (b and x may not alias) */
int gf_bsr_ilu_apply_f64(gf_ilu ilu, const double* d_lu, const double* d_b,
                         double* d_tmp, double* d_x);
int gf_bsr_ilu_apply_f32(gf_ilu ilu, const float* d_lu, const float* d_b,
                         float* d_tmp, float* d_x);

/* ---- right-preconditioned restarted GMRES on device BSR operands
 *      (examples/iterative/gmres.cu:472-771 with restarts and a relative
 *      tolerance).  precision 0: FP64 operator and ILU factors (d_mlu);
 *      precision 1: GMRES-IR, inner FP32 ILU factors (d_mlu32) and FP32 copy
 *      of the operator (d_kvals32), FP64 outer residual and Arnoldi;
 *      precision 2: FP32 ILU factors (d_mlu32) and FP32 sweeps inside FP64
 *      GMRES (FP64 SpMV, d_kvals32 unused).  x holds the initial guess on
 *      entry. */
int gf_gmres_solve(gf_context ctx, int nnodes, const int* d_krowp,
                   const int* d_kcols, const double* d_kvals,
                   const float* d_kvals32, gf_ilu ilu, const double* d_mlu,
                   const float* d_mlu32, const double* d_b, double* d_x,
                   int restart, int max_iters, double rtol, int precision,
                   int* iterations, double* rel_resid);

/* ---- complete linear-static problem (examples/iterative/gmres.cu flow) ---- */
typedef struct {
  int reorder;       /* 0 natural, 1 RCM (root 0), 2 AMD */
  int lev_fill;      /* ILU(k) level of fill */
  int restart;       /* GMRES(m) restart length */
  int max_iters;     /* total Krylov iterations */
  double rtol;       /* ||b - A x|| <= rtol ||b|| in FP64 */
  int precision;     /* 0 FP64; 1 GMRES-IR (FP32 ILU/trsv/SpMV inside,
                        FP64 outer); 2 FP32 ILU/trsv inside FP64 GMRES */
  int stall_cycles;  /* > 0: stop with GF_ERR_NO_CONVERGENCE once the true
                        residual fell by less than 10% over this many restart
                        cycles; 0 (default): run to max_iters as the
                        reference does */
} gf_options;

typedef struct {
  int status;        /* GF_OK or an error code */
  int iterations;    /* Krylov iterations */
  int outer;         /* restart cycles (FP64) or refinement steps (mixed) */
  double rel_resid;  /* true ||b - A x|| / ||b|| in FP64 */
  double t_assemble, t_factor, t_solve; /* seconds (device, stream-timed) */
} gf_solve_info;

void gf_options_default(gf_options* opt);
/* host mesh in (copied); runs the symbolic phase and uploads everything */
int gf_problem_create(gf_context ctx, int nnodes, int nelems, const int* conn,
                      const double* xpts, const double* data, int nbcs,
                      const int* bcs, const gf_options* opt, gf_problem* out);
int gf_problem_destroy(gf_problem p);
/* K = add_jacobian (+ BCs) and copy into the preconditioner pattern */
int gf_problem_assemble(gf_problem p);
/* numeric ILU in the configured precision */
int gf_problem_factor(gf_problem p);
/* solve K x = loads; loads/x are unpermuted host or device arrays of
 * 6*nnodes doubles (device when is_device != 0) */
int gf_problem_solve(gf_problem p, const double* loads, double* x,
                     int is_device, gf_solve_info* info);
/* Nonlinear (Newton) path, IsotropicShell<..., isNonlinear = true>
 * (include/shell/physics.h:6-17).  gf_problem_assemble_tangent: K(u) with
 * the boundary conditions into the problem's K and preconditioner values
 * (the nonlinear add_jacobian + apply_bcs, assembler.h:267-305), and the
 * residual R(u) into res (unpermuted, constrained entries zero; may be
 * NULL); u and res are host or device arrays of 6*nnodes doubles. */
int gf_problem_assemble_tangent(gf_problem p, const double* u, double* res,
                                int is_device);
typedef struct {
  int status;             /* GF_OK or an error code */
  int newton_iterations;  /* tangent solves taken */
  int krylov_iterations;  /* GMRES iterations over all of them */
  double rel_resid;       /* ||f - R(u)|| / ||f||, constrained entries zero */
  double t_total;         /* seconds (host clock) */
} gf_newton_info;
/* Newton's method for R(u) = f: from the initial u (in/out), each step
 * assembles and factors the tangent K(u) and solves K du = f - R(u) with the
 * problem's preconditioned GMRES (its rtol, fill and precision), until
 * ||f - R(u)|| <= newton_rtol ||f|| or max_newton steps (then
 * GF_ERR_NO_CONVERGENCE).  history (may be NULL) receives the relative
 * residual before each step, max_newton + 1 doubles.  The reference has no
 * Newton driver; this is the loop its nonlinear element group is for. */
int gf_problem_newton_solve(gf_problem p, const double* loads, double* u,
                            int is_device, int max_newton, double newton_rtol,
                            double* history, gf_newton_info* info);
/* sizes and device pointers for tests and benchmarks */
int gf_problem_sizes(gf_problem p, int* nnodes, int* nnzb_k, int* nnzb_m);
int gf_problem_device_arrays(gf_problem p, const int** krowp,
                             const int** kcols, double** kvals,
                             const int** mrowp, const int** mcols,
                             double** mvals, const int** perm);
/* the problem's ILU analysis handle (owned by the problem) */
gf_ilu gf_problem_ilu(gf_problem p);
int gf_problem_host_perm(gf_problem p, int* perm);

/* ---- serving loop (no reference counterpart: the served workload's request
 *      loop).  Worker slots, each with its own stream, admit queued requests
 *      FIFO when their estimated device footprint fits the HBM budget and run
 *      gf_problem_create/assemble/factor/solve; device work of concurrent
 *      requests overlaps.  Thread-safe. ---- */
int gf_server_create(int device, double budget_gb, int slots, gf_server* out);
/* copies the request; GF_ERR_OVER_BUDGET if it can never fit */
int gf_server_submit(gf_server s, long long id, int nnodes, int nelems,
                     const int* conn, const double* xpts, const double* data,
                     int nbcs, const int* bcs, const double* loads,
                     const gf_options* opt);
/* 1 and the result of one finished request, or 0 if none is ready */
int gf_server_poll(gf_server s, long long* id, double* x, size_t x_len,
                   gf_solve_info* info, double* latency_s);
/* stats[9] = [completed, failed, queued, active, p50 s, p99 s, peak bytes,
 * elapsed s, requests re-run with a doubled ILU level of fill] */
int gf_server_stats(gf_server s, double* stats);
int gf_server_drain(gf_server s);
int gf_server_destroy(gf_server s);

/* ---- NASTRAN bulk-data mesh reader (TACSMeshLoader,
 *      include/mesh/TACSMeshLoader.h:145-227, 763-1290; host only).
 *      GRID / GRID* / free-field GRID, SPC / SPC* / SPC1, and the element
 *      cards CBAR CQUADR CQUAD4 CQUAD8 CQUAD9 CQUAD CHEXA27 CHEXA CTRIA3
 *      CTETRA; nodes and elements ordered by file id, CQUAD4/CQUADR corners
 *      stored in tensor-product order (1 2 4 3). ---- */
/* scanBDFFile (:763); GF_ERR_INVALID with gf_bdf_error() on a bad file */
int gf_bdf_read(const char* path, gf_bdf* out);
/* the reason the last gf_bdf_read on this thread failed */
const char* gf_bdf_error(void);
int gf_bdf_destroy(gf_bdf b);
/* getNumNodes / getNumElements / getNumComponents (:158-160); conn_size =
 * total element nodes, nbcs = SPC entries, nbc_vars = constrained DOFs */
int gf_bdf_sizes(gf_bdf b, int* nnodes, int* nelems, int* conn_size,
                 int* nbcs, int* nbc_vars, int* ncomponents);
/* getConnectivity (:184-187): xpts[3*nnodes], elem_ptr[nelems+1],
 * conn[conn_size], elem_comp[nelems] (property id - 1); any may be NULL */
int gf_bdf_connectivity(gf_bdf b, double* xpts, int* elem_ptr, int* conn,
                        int* elem_comp);
/* getBCs (:188-189): bc_nodes[nbcs], bc_ptr[nbcs+1], bc_vars[nbc_vars]
 * (component digit - 1), bc_vals[nbc_vars] */
int gf_bdf_bcs(gf_bdf b, int* bc_nodes, int* bc_ptr, int* bc_vars,
               double* bc_vals);
/* getAssemblerCreatorData (:193-224): elem_conn[conn_size], bcs[nbc_vars]
 * = vars_per_node * node + var, xpts[3*nnodes] */
int gf_bdf_assembler_data(gf_bdf b, int vars_per_node, int* elem_conn,
                          int* bcs, double* xpts);
/* getComponentDescript / getElementDescript (:150-151); NULL if out of range */
const char* gf_bdf_component_descript(gf_bdf b, int comp);
const char* gf_bdf_element_descript(gf_bdf b, int comp);

/* ---- legacy-VTK output (printToVTK, include/mesh/vtk_writer.h:8-73; host
 *      only): nodes, quads (corners written around the element) and the
 *      displacement u v w of vars[6*nnodes] (NULL: zero) as "disp" ---- */
int gf_write_vtk(const char* path, int nnodes, const double* xpts, int nelems,
                 const int* conn, const double* vars);

#ifdef __cplusplus
}
#endif
#endif /* GPUFEM_AMD_H */
