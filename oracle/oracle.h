/*
 * CPU oracle for the linear-static MITC4 shell path (test infrastructure).
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * entry and bench.py's cpu_baseline leg load it.  Every function restates the
 * corresponding reference algorithm (file:line cited beside each one) in plain
 * C++ so that the HIP path can be compared with it on identical inputs.
 *
 * Conventions (same as the reference):
 *   - index arrays are int32, block size 6, BSR values row-major inside a
 *     block: val[36*b + 6*r + c]
 *   - perm[old] = new, iperm[new] = old (BsrData::perm / iperm)
 *   - element data per element: {E, nu, thick, tOffset}
 *
 * Arrays returned through int** / double** out-parameters are malloc'ed and
 * must be released with orc_free().
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

void orc_free(void* p);

/* ---- element (include/shell/, include/base/elem_group.h) ---- */
/* K (24x24 row-major), R (24) and strain energy of one MITC4 element. */
void orc_mitc4_element(const double xpts[12], const double vars[24],
                       const double data[4], double* K, double* R,
                       double* energy);
/* 9 strains and detXd*weight at quadrature point iq for given element vars. */
double orc_mitc4_strain(int iq, const double xpts[12], const double vars[24],
                        double E[9]);
// nonlinear strain (a2d/a2dshellstrain.h:72-93) and the local displacement
// gradients u0x, u1x (row-major 3 x 3) at quadrature point iq
double orc_mitc4_strain_nl(int iq, const double xpts[12],
                           const double vars[24], double E[9], double u0x[9],
                           double u1x[9]);
// nonlinear element: K = Hessian, R = gradient of the energy at vars
void orc_mitc4_element_nl(const double xpts[12], const double vars[24],
                          const double data[4], double* K, double* R,
                          double* energy);
void orc_assemble_tangent(int nelems, const int* conn, const double* xpts,
                          const double* vars, const double* data,
                          const int* elem_ind_map, const int* perm,
                          double* kvals, double* res, double* energy);

/* ---- symbolic (include/linalg/bsr_utils.h, reordering_utils.h) ---- */
int orc_node_graph(int nelems, int nnodes, int nodes_per_elem,
                   const int* conn, int** rowp, int** cols);
int orc_rcm(int nnodes, const int* rowp, const int* cols, int root,
            int* rcm_vars);
int orc_permute_pattern(int nnodes, const int* rowp, const int* cols,
                        const int* perm, int** prowp, int** pcols);
int orc_iluk(int nnodes, const int* rowp, const int* cols, int lev_fill,
             int** frowp, int** fcols, int** levs);
void orc_elem_ind_map(int nelems, int nodes_per_elem, const int* conn,
                      const int* rowp, const int* cols, const int* perm,
                      int* elem_ind_map);
int orc_transpose(int nnodes, const int* rowp, const int* cols, int** trowp,
                  int** tcols, int** tmap);
int orc_bandwidth(int nnodes, const int* rowp, const int* cols);
int orc_full_fill(int nnodes, const int* rowp, const int* cols, int** frowp,
                  int** fcols);
int orc_min_degree(int nnodes, const int* rowp, const int* cols, int* perm);
/* replay of an ordering on the explicit elimination graph: exact degree of
   each pivot and least exact degree at that step; returns nnz(L) blocks */
long long orc_md_replay(int nnodes, const int* rowp, const int* cols,
                        const int* perm, int* dpiv, int* dmin);
/* exact minimum degree, ties to the most recently updated node */
int orc_min_degree_stack(int nnodes, const int* rowp, const int* cols,
                         int* perm);

/* ---- assembly / matrix ops (include/linalg/bsr_mat.h, .cuh, vec.h, .cuh) */
void orc_assemble_jacobian(int nelems, const int* conn, const double* xpts,
                           const double* vars, const double* data,
                           const int* elem_ind_map, const int* perm,
                           double* kvals, double* res, int nthreads);
void orc_assemble_residual(int nelems, const int* conn, const double* xpts,
                           const double* vars, const double* data,
                           const int* perm, double* res);
void orc_apply_bcs_mat(int nbcs, const int* bcs, const int* rowp,
                       const int* cols, const int* trowp, const int* tcols,
                       const int* tmap, const int* perm, double* vals);
void orc_apply_bcs_vec(int nbcs, const int* bcs, const int* perm,
                       double* vec);
void orc_copy_values(int nnodes, const int* rowp, const int* cols,
                     const double* vals, const int* trowp, const int* tcols,
                     double* tvals);
void orc_permute_vec(int nnodes, const int* perm, const double* in,
                     double* out);

/* ---- numeric solve (cusparse bsrilu02 / bsrsv2 / bsrmv semantics) ---- */
int orc_bsr_ilu_f64(int nnodes, const int* rowp, const int* cols,
                    double* vals);
int orc_bsr_ilu_f32(int nnodes, const int* rowp, const int* cols, float* vals);
/* rows [0, row_end) only: a timing sample (a row depends on earlier rows) */
int orc_bsr_ilu_rows_f64(int nnodes, const int* rowp, const int* cols,
                         double* vals, int row_end);
void orc_bsr_spmv_f64(int nnodes, const int* rowp, const int* cols,
                      const double* vals, const double* x, double* y);
/* r = b - A x accumulated with error-free transformations (TwoProduct by
   FMA, TwoSum), rounded once */
void orc_bsr_resid_comp_f64(int nnodes, const int* rowp, const int* cols,
                            const double* vals, const double* x,
                            const double* b, double* r);
void orc_bsr_trsv_f64(int nnodes, const int* rowp, const int* cols,
                      const double* lu, const double* b, double* x);
void orc_bsr_trsv_f32(int nnodes, const int* rowp, const int* cols,
                      const float* lu, const float* b, float* x);
/* right-preconditioned restarted GMRES, FP64 (mixed=0) or GMRES-IR with
   FP32 ILU factors and FP32 inner SpMV (mixed=1).  Returns 0 if converged.
   stats[0]=iterations, stats[1]=final true relative residual,
   stats[2]=outer refinement steps. */
int orc_gmres(int nnodes, const int* krowp, const int* kcols,
              const double* kvals, const int* mrowp, const int* mcols,
              const double* mvals_factored, const double* b, double* x,
              int restart, int max_iters, double rtol, int mixed,
              double* stats);
int orc_dense_solve(int n, const double* A, const double* b, double* x);
/* host threads of the parallel loops (default 1; results independent of it) */
void orc_set_threads(int n);

#ifdef __cplusplus
}
#endif
