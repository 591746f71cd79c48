// Oracle (test infrastructure only): assembly, boundary conditions and the
// ILU-preconditioned GMRES solve, restated in plain loops.
//
//   orc_assemble_jacobian  include/base/elem_group.h:121-181 add_jacobian_cpu
//                          (gather by unpermuted conn), scatter as
//                          bsr_mat.h:196-226 addElementMatrixValues through
//                          elemIndMap; residual into 6*perm[conn]+k as the
//                          device path vec.h:278-293 does
//   orc_assemble_residual  include/base/elem_group.h:40-79 add_residual_cpu
//   orc_assemble_tangent   the same loop for the nonlinear element
//                          (IsotropicShell<..., isNonlinear = true>)
//   orc_apply_bcs_mat      bsr_mat.cuh:2-104 (row and column kernels)
//   orc_apply_bcs_vec      vec.cuh:4-13 (optionally through perm)
//   orc_copy_values        bsr_mat.cuh:107-140 copy_mat_values_kernel
//   orc_permute_vec        vec.cuh:16-34 permute_vec_kernel
//   orc_bsr_ilu_*          cusparseXbsrilu02 semantics (gmres.cu:426-457):
//                          in-place incomplete LU on the given block pattern,
//                          L unit lower, U upper incl. diagonal blocks, which is
//                          scalar ILU on the block-expanded pattern
//   orc_bsr_trsv_*         cusparseXbsrsv2 L (unit) then U (non-unit)
//   orc_bsr_spmv_f64       cusparseDbsrmv, y = A x
//   orc_gmres              restarted, right-preconditioned GMRES(m) with
//                          modified Gram-Schmidt and Givens rotations (textbook;
//                          gmres.cu:472-771 is the flow it replaces)
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "oracle.h"

// Host threads of the parallel loops (element matrices, SpMV, residual,
// Gram-Schmidt vector operations); 1 by default.  Results do not depend on
// the count: element matrices are summed in element order, and reductions
// use a fixed split into kRedChunks partial sums when more than one thread
// runs.  The ILU factorisation and the triangular solves stay sequential.
static int g_threads = 1;
constexpr int kRedChunks = 64;
extern "C" void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

extern "C" void orc_assemble_jacobian(int nelems, const int* conn,
                                      const double* xpts, const double* vars,
                                      const double* data,
                                      const int* elem_ind_map, const int* perm,
                                      double* kvals, double* res,
                                      int nthreads) {
  // Element matrices are computed independently (optionally on several host
  // threads), then summed in element order so the result does not depend on
  // the thread count.
  (void)nthreads;
  if (g_threads > 1) {
    // element matrices of a chunk of elements in parallel, then summed in
    // element order
    constexpr int kChunk = 4096;
    std::vector<double> Kc((size_t)kChunk * 576), Rc((size_t)kChunk * 24);
    for (int e0 = 0; e0 < nelems; e0 += kChunk) {
      const int ne = std::min(kChunk, nelems - e0);
#pragma omp parallel for num_threads(g_threads) schedule(static)
      for (int q = 0; q < ne; q++) {
        const int e = e0 + q;
        double xe[12], ve[24];
        for (int a = 0; a < 4; a++) {
          int nd = conn[4 * e + a];
          for (int k = 0; k < 3; k++) xe[3 * a + k] = xpts[3 * nd + k];
          for (int k = 0; k < 6; k++) ve[6 * a + k] = vars ? vars[6 * nd + k] : 0.0;
        }
        orc_mitc4_element(xe, ve, &data[4 * e], &Kc[576 * (size_t)q],
                          &Rc[24 * (size_t)q], nullptr);
      }
      for (int q = 0; q < ne; q++) {
        const int e = e0 + q;
        const double* Ke = &Kc[576 * (size_t)q];
        for (int a = 0; a < 4; a++)
          for (int b = 0; b < 4; b++) {
            double* blk = &kvals[36 * (size_t)elem_ind_map[16 * e + 4 * a + b]];
            for (int r = 0; r < 6; r++)
              for (int c = 0; c < 6; c++)
                blk[6 * r + c] += Ke[24 * (6 * a + r) + 6 * b + c];
          }
        if (res)
          for (int a = 0; a < 4; a++) {
            int nd = perm ? perm[conn[4 * e + a]] : conn[4 * e + a];
            for (int k = 0; k < 6; k++) res[6 * nd + k] += Rc[24 * (size_t)q + 6 * a + k];
          }
      }
    }
    return;
  }
  std::vector<double> Ke(576), Re(24);
  for (int e = 0; e < nelems; e++) {
    double xe[12], ve[24];
    for (int a = 0; a < 4; a++) {
      int nd = conn[4 * e + a];
      for (int k = 0; k < 3; k++) xe[3 * a + k] = xpts[3 * nd + k];
      for (int k = 0; k < 6; k++) ve[6 * a + k] = vars ? vars[6 * nd + k] : 0.0;
    }
    orc_mitc4_element(xe, ve, &data[4 * e], Ke.data(), Re.data(), nullptr);
    for (int a = 0; a < 4; a++)
      for (int b = 0; b < 4; b++) {
        double* blk = &kvals[36 * (size_t)elem_ind_map[16 * e + 4 * a + b]];
        for (int r = 0; r < 6; r++)
          for (int c = 0; c < 6; c++)
            blk[6 * r + c] += Ke[24 * (6 * a + r) + 6 * b + c];
      }
    if (res)
      for (int a = 0; a < 4; a++) {
        int nd = perm ? perm[conn[4 * e + a]] : conn[4 * e + a];
        for (int k = 0; k < 6; k++) res[6 * nd + k] += Re[6 * a + k];
      }
  }
}

// Nonlinear tangent K(u) and residual R(u) (orc_mitc4_element_nl), gathered
// and scattered as orc_assemble_jacobian; energy summed when requested.
extern "C" void orc_assemble_tangent(int nelems, const int* conn,
                                     const double* xpts, const double* vars,
                                     const double* data,
                                     const int* elem_ind_map, const int* perm,
                                     double* kvals, double* res,
                                     double* energy) {
  std::vector<double> Ke(576), Re(24);
  double U = 0.0;
  for (int e = 0; e < nelems; e++) {
    double xe[12], ve[24], Ue = 0.0;
    for (int a = 0; a < 4; a++) {
      int nd = conn[4 * e + a];
      for (int k = 0; k < 3; k++) xe[3 * a + k] = xpts[3 * nd + k];
      for (int k = 0; k < 6; k++) ve[6 * a + k] = vars[6 * nd + k];
    }
    orc_mitc4_element_nl(xe, ve, &data[4 * e], kvals ? Ke.data() : nullptr,
                         Re.data(), &Ue);
    U += Ue;
    if (kvals)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          double* blk = &kvals[36 * (size_t)elem_ind_map[16 * e + 4 * a + b]];
          for (int r = 0; r < 6; r++)
            for (int c = 0; c < 6; c++)
              blk[6 * r + c] += Ke[24 * (6 * a + r) + 6 * b + c];
        }
    if (res)
      for (int a = 0; a < 4; a++) {
        int nd = perm ? perm[conn[4 * e + a]] : conn[4 * e + a];
        for (int k = 0; k < 6; k++) res[6 * nd + k] += Re[6 * a + k];
      }
  }
  if (energy) *energy = U;
}

extern "C" void orc_assemble_residual(int nelems, const int* conn,
                                      const double* xpts, const double* vars,
                                      const double* data, const int* perm,
                                      double* res) {
  double Re[24];
  for (int e = 0; e < nelems; e++) {
    double xe[12], ve[24];
    for (int a = 0; a < 4; a++) {
      int nd = conn[4 * e + a];
      for (int k = 0; k < 3; k++) xe[3 * a + k] = xpts[3 * nd + k];
      for (int k = 0; k < 6; k++) ve[6 * a + k] = vars[6 * nd + k];
    }
    orc_mitc4_element(xe, ve, &data[4 * e], nullptr, Re, nullptr);
    for (int a = 0; a < 4; a++) {
      int nd = perm ? perm[conn[4 * e + a]] : conn[4 * e + a];
      for (int k = 0; k < 6; k++) res[6 * nd + k] += Re[6 * a + k];
    }
  }
}

extern "C" void orc_apply_bcs_mat(int nbcs, const int* bcs, const int* rowp,
                                  const int* cols, const int* trowp,
                                  const int* tcols, const int* tmap,
                                  const int* perm, double* vals) {
  for (int ib = 0; ib < nbcs; ib++) {  // rows (bsr_mat.cuh:2-45)
    int dof = bcs[ib];
    int brow = perm[dof / 6], ir = dof % 6, grow = 6 * brow + ir;
    for (int k = rowp[brow]; k < rowp[brow + 1]; k++)
      for (int c = 0; c < 6; c++)
        vals[36 * (size_t)k + 6 * ir + c] =
            (6 * cols[k] + c == grow) ? 1.0 : 0.0;
  }
  for (int ib = 0; ib < nbcs; ib++) {  // columns (bsr_mat.cuh:48-104)
    int dof = bcs[ib];
    int bcol = perm[dof / 6], ic = dof % 6, gcol = 6 * bcol + ic;
    for (int t = trowp[bcol]; t < trowp[bcol + 1]; t++) {
      int brow = tcols[t];
      for (int r = 0; r < 6; r++)
        vals[36 * (size_t)tmap[t] + 6 * r + ic] =
            (6 * brow + r == gcol) ? 1.0 : 0.0;
    }
  }
}

extern "C" void orc_apply_bcs_vec(int nbcs, const int* bcs, const int* perm,
                                  double* vec) {
  for (int ib = 0; ib < nbcs; ib++) {
    int dof = bcs[ib];
    int idx = perm ? 6 * perm[dof / 6] + dof % 6 : dof;
    vec[idx] = 0.0;
  }
}

extern "C" void orc_copy_values(int n, const int* rowp, const int* cols,
                                const double* vals, const int* trowp,
                                const int* tcols, double* tvals) {
  for (int i = 0; i < n; i++) {
    int tj = trowp[i];
    for (int jp = rowp[i]; jp < rowp[i + 1]; jp++) {
      while (tcols[tj] < cols[jp]) tj++;
      std::memcpy(&tvals[36 * (size_t)tj], &vals[36 * (size_t)jp],
                  36 * sizeof(double));
    }
  }
}

extern "C" void orc_permute_vec(int n, const int* perm, const double* in,
                                double* out) {
  for (int i = 0; i < n; i++)
    for (int k = 0; k < 6; k++) out[6 * perm[i] + k] = in[6 * i + k];
}

namespace {

template <typename T>
int bsr_ilu(int n, const int* rowp, const int* cols, T* vals, int row_end) {
  // Row-oriented (IKJ) block ILU on the given pattern; rows [0, row_end)
  // only (a row depends on earlier rows alone, so a prefix is exact).
  std::vector<int> diag(n, -1);
  for (int i = 0; i < n; i++)
    for (int k = rowp[i]; k < rowp[i + 1]; k++)
      if (cols[k] == i) diag[i] = k;
  for (int i = 0; i < n; i++)
    if (diag[i] < 0) return -(i + 1);
  for (int i = 0; i < row_end; i++) {
    for (int kk = rowp[i]; kk < diag[i]; kk++) {
      int k = cols[kk];
      T* Aik = &vals[36 * (size_t)kk];
      const T* Ukk = &vals[36 * (size_t)diag[k]];
      // Aik <- Aik * Ukk^-1  (solve X Ukk = Aik, Ukk upper triangular)
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) {
          T s = Aik[6 * r + c];
          for (int m = 0; m < c; m++) s -= Aik[6 * r + m] * Ukk[6 * m + c];
          Aik[6 * r + c] = s / Ukk[6 * c + c];
        }
      // A_ij -= L_ik U_kj for j > k present in both rows
      int p = kk + 1;
      for (int jj = diag[k] + 1; jj < rowp[k + 1]; jj++) {
        int j = cols[jj];
        while (p < rowp[i + 1] && cols[p] < j) p++;
        if (p >= rowp[i + 1]) break;
        if (cols[p] != j) continue;
        T* Aij = &vals[36 * (size_t)p];
        const T* Ukj = &vals[36 * (size_t)jj];
        for (int r = 0; r < 6; r++)
          for (int c = 0; c < 6; c++) {
            T s = 0;
            for (int m = 0; m < 6; m++) s += Aik[6 * r + m] * Ukj[6 * m + c];
            Aij[6 * r + c] -= s;
          }
      }
    }
    // factor the diagonal block in place (no pivoting)
    T* D = &vals[36 * (size_t)diag[i]];
    for (int c = 0; c < 6; c++) {
      if (D[6 * c + c] == T(0)) return -(i + 1);
      for (int r = c + 1; r < 6; r++) {
        T l = D[6 * r + c] / D[6 * c + c];
        D[6 * r + c] = l;
        for (int m = c + 1; m < 6; m++) D[6 * r + m] -= l * D[6 * c + m];
      }
    }
    // U_ij <- L_ii^-1 A_ij for j > i
    for (int jj = diag[i] + 1; jj < rowp[i + 1]; jj++) {
      T* A = &vals[36 * (size_t)jj];
      for (int c = 0; c < 6; c++)
        for (int r = 1; r < 6; r++) {
          T s = A[6 * r + c];
          for (int m = 0; m < r; m++) s -= D[6 * r + m] * A[6 * m + c];
          A[6 * r + c] = s;
        }
    }
  }
  return 0;
}

template <typename T>
void bsr_trsv(int n, const int* rowp, const int* cols, const T* lu, const T* b,
              T* x) {
  std::vector<int> diag(n);
  for (int i = 0; i < n; i++)
    for (int k = rowp[i]; k < rowp[i + 1]; k++)
      if (cols[k] == i) diag[i] = k;
  std::vector<T> y(6 * (size_t)n);
  for (int i = 0; i < n; i++) {  // L y = b, unit lower
    T s[6];
    for (int r = 0; r < 6; r++) s[r] = b[6 * i + r];
    for (int kk = rowp[i]; kk < diag[i]; kk++) {
      const T* L = &lu[36 * (size_t)kk];
      const T* yk = &y[6 * (size_t)cols[kk]];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) s[r] -= L[6 * r + c] * yk[c];
    }
    const T* D = &lu[36 * (size_t)diag[i]];
    for (int r = 0; r < 6; r++) {
      for (int c = 0; c < r; c++) s[r] -= D[6 * r + c] * s[c];
      y[6 * i + r] = s[r];
    }
  }
  for (int i = n - 1; i >= 0; i--) {  // U x = y
    T s[6];
    for (int r = 0; r < 6; r++) s[r] = y[6 * i + r];
    for (int kk = diag[i] + 1; kk < rowp[i + 1]; kk++) {
      const T* U = &lu[36 * (size_t)kk];
      const T* xk = &x[6 * (size_t)cols[kk]];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) s[r] -= U[6 * r + c] * xk[c];
    }
    const T* D = &lu[36 * (size_t)diag[i]];
    for (int r = 5; r >= 0; r--) {
      for (int c = r + 1; c < 6; c++) s[r] -= D[6 * r + c] * s[c];
      s[r] /= D[6 * r + r];
    }
    for (int r = 0; r < 6; r++) x[6 * i + r] = s[r];
  }
}

template <typename T>
void bsr_spmv(int n, const int* rowp, const int* cols, const T* vals,
              const T* x, T* y) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int i = 0; i < n; i++) {
    T s[6] = {0, 0, 0, 0, 0, 0};
    for (int k = rowp[i]; k < rowp[i + 1]; k++) {
      const T* A = &vals[36 * (size_t)k];
      const T* xk = &x[6 * (size_t)cols[k]];
      for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) s[r] += A[6 * r + c] * xk[c];
    }
    for (int r = 0; r < 6; r++) y[6 * i + r] = s[r];
  }
}

// r = b - A x with every product and sum carried as an unevaluated pair
// (Ogita, Rump and Oishi, "Accurate sum and dot product", SIAM J. Sci.
// Comput. 26 (2005): TwoProduct through an FMA, TwoSum), rounded once at the
// end: as accurate as the residual in twice the working precision.
void bsr_resid_comp(int n, const int* rowp, const int* cols,
                    const double* vals, const double* x, const double* b,
                    double* r) {
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
  for (int i = 0; i < n; i++) {
    double s[6], c[6];
    for (int q = 0; q < 6; q++) {
      s[q] = b[6 * (size_t)i + q];
      c[q] = 0.0;
    }
    for (int k = rowp[i]; k < rowp[i + 1]; k++) {
      const double* A = &vals[36 * (size_t)k];
      const double* xk = &x[6 * (size_t)cols[k]];
      for (int q = 0; q < 6; q++)
        for (int t = 0; t < 6; t++) {
          const double p = -A[6 * q + t] * xk[t];
          const double pe = std::fma(-A[6 * q + t], xk[t], -p);
          const double sn = s[q] + p;
          const double bb = sn - s[q];
          const double se = (s[q] - (sn - bb)) + (p - bb);
          s[q] = sn;
          c[q] += se + pe;
        }
    }
    for (int q = 0; q < 6; q++) r[6 * (size_t)i + q] = s[q] + c[q];
  }
}

// sum_t a[t] b[t]: serial with one thread, else kRedChunks partial sums in
// parallel added in chunk order
double pdot(const double* a, const double* b, size_t N) {
  if (g_threads <= 1) {
    double s = 0;
    for (size_t t = 0; t < N; t++) s += a[t] * b[t];
    return s;
  }
  double part[kRedChunks];
  const size_t len = (N + kRedChunks - 1) / kRedChunks;
#pragma omp parallel for num_threads(g_threads) schedule(static)
  for (int c = 0; c < kRedChunks; c++) {
    double s = 0;
    const size_t t1 = std::min(N, (c + 1) * len);
    for (size_t t = c * len; t < t1; t++) s += a[t] * b[t];
    part[c] = s;
  }
  double s = 0;
  for (int c = 0; c < kRedChunks; c++) s += part[c];
  return s;
}

double nrm2(const std::vector<double>& v) {
  return std::sqrt(pdot(v.data(), v.data(), v.size()));
}

}  // namespace

extern "C" int orc_bsr_ilu_f64(int n, const int* rowp, const int* cols,
                               double* vals) {
  return bsr_ilu<double>(n, rowp, cols, vals, n);
}
extern "C" int orc_bsr_ilu_f32(int n, const int* rowp, const int* cols,
                               float* vals) {
  return bsr_ilu<float>(n, rowp, cols, vals, n);
}
extern "C" int orc_bsr_ilu_rows_f64(int n, const int* rowp, const int* cols,
                                    double* vals, int row_end) {
  return bsr_ilu<double>(n, rowp, cols, vals,
                         row_end < 0 ? 0 : (row_end > n ? n : row_end));
}
extern "C" void orc_bsr_spmv_f64(int n, const int* rowp, const int* cols,
                                 const double* vals, const double* x,
                                 double* y) {
  bsr_spmv<double>(n, rowp, cols, vals, x, y);
}
extern "C" void orc_bsr_resid_comp_f64(int n, const int* rowp,
                                       const int* cols, const double* vals,
                                       const double* x, const double* b,
                                       double* r) {
  bsr_resid_comp(n, rowp, cols, vals, x, b, r);
}
extern "C" void orc_bsr_trsv_f64(int n, const int* rowp, const int* cols,
                                 const double* lu, const double* b,
                                 double* x) {
  bsr_trsv<double>(n, rowp, cols, lu, b, x);
}
extern "C" void orc_bsr_trsv_f32(int n, const int* rowp, const int* cols,
                                 const float* lu, const float* b, float* x) {
  bsr_trsv<float>(n, rowp, cols, lu, b, x);
}

// GMRES(m), right preconditioned: A M^-1 u = b, x = M^-1 u.
// mixed = 0: M^-1 from FP64 ILU factors, FP64 SpMV.
// mixed = 1: GMRES-IR; the outer loop forms r = b - A x in FP64 and the
//   correction d solves A d = r with an inner GMRES whose preconditioner
//   (FP32 ILU factors) and operator (FP32 copy of A) run in FP32 while the
//   Arnoldi basis, orthogonalisation and update stay FP64.
extern "C" int orc_gmres(int n, const int* krowp, const int* kcols,
                         const double* kvals, const int* mrowp,
                         const int* mcols, const double* mlu, const double* b,
                         double* x, int restart, int max_iters, double rtol,
                         int mixed, double* stats) {
  const size_t N = 6 * (size_t)n;
  size_t nk = 36 * (size_t)krowp[n], nm = 36 * (size_t)mrowp[n];
  std::vector<float> k32, m32;
  if (mixed) {
    k32.resize(nk);
    m32.resize(nm);
    for (size_t i = 0; i < nk; i++) k32[i] = (float)kvals[i];
    for (size_t i = 0; i < nm; i++) m32[i] = (float)mlu[i];
  }
  std::vector<float> t32a(N), t32b(N);
  auto precond = [&](const double* v, double* z) {
    if (!mixed) {
      bsr_trsv<double>(n, mrowp, mcols, mlu, v, z);
    } else {
      for (size_t i = 0; i < N; i++) t32a[i] = (float)v[i];
      bsr_trsv<float>(n, mrowp, mcols, m32.data(), t32a.data(), t32b.data());
      for (size_t i = 0; i < N; i++) z[i] = (double)t32b[i];
    }
  };
  auto inner_op = [&](const double* z, double* w) {  // w = A z
    if (!mixed) {
      bsr_spmv<double>(n, krowp, kcols, kvals, z, w);
    } else {
      for (size_t i = 0; i < N; i++) t32a[i] = (float)z[i];
      bsr_spmv<float>(n, krowp, kcols, k32.data(), t32a.data(), t32b.data());
      for (size_t i = 0; i < N; i++) w[i] = (double)t32b[i];
    }
  };
  std::vector<double> bv(b, b + N), r(N), w(N), z(N), u(N);
  double bnorm = nrm2(bv);
  if (bnorm == 0.0) bnorm = 1.0;
  int m = restart;
  std::vector<std::vector<double>> V(m + 1, std::vector<double>(N));
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  int total_it = 0, outer = 0, status = 1;
  double true_rel = 0.0;
  // Each outer step: true residual r = b - A x in FP64, then one GMRES(m)
  // cycle on A d = r started from d = 0, then x += d.  In FP64 mode this is
  // plain restarted GMRES(m) with the cycle stopping when its residual
  // estimate reaches 0.3 rtol |b| (aiming below the tolerance: on
  // ill-conditioned preconditioned systems the true residual trails the
  // estimate, and a cycle that stops exactly at rtol |b| leaves the true
  // residual just above it, restarting again and again); in mixed mode the
  // cycle stops at a 1e-5 relative reduction (the FP32 inner operator
  // cannot resolve more) and the FP64 outer loop refines.
  for (;;) {
    // the true residual with compensated accumulation: on thin shells
    // |A| |x| exceeds |b| by many orders of magnitude, and b - A x in plain
    // FP64 stalls at a rounding floor above tight tolerances
    bsr_resid_comp(n, krowp, kcols, kvals, x, bv.data(), r.data());
    double beta = nrm2(r);
    true_rel = beta / bnorm;
    if (true_rel <= rtol) {
      status = 0;
      break;
    }
    if (total_it >= max_iters) break;
    outer++;
    double tol_abs = mixed ? std::max(1e-5 * beta, 0.3 * rtol * bnorm)
                           : 0.3 * rtol * bnorm;
    for (size_t i = 0; i < N; i++) V[0][i] = r[i] / beta;
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int j = 0;
    for (; j < m && total_it < max_iters;) {
      precond(V[j].data(), z.data());
      inner_op(z.data(), w.data());
      total_it++;
      for (int i = 0; i <= j; i++) {  // modified Gram-Schmidt
        const double h = pdot(w.data(), V[i].data(), N);
        H[i * m + j] = h;
        const double* vi = V[i].data();
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
        for (size_t t = 0; t < N; t++) w[t] -= h * vi[t];
      }
      double hn = nrm2(w);
      H[(j + 1) * m + j] = hn;
      double* vn = V[j + 1].data();
#pragma omp parallel for num_threads(g_threads) schedule(static) if (g_threads > 1)
      for (size_t t = 0; t < N; t++) vn[t] = hn > 0 ? w[t] / hn : 0.0;
      for (int i = 0; i < j; i++) {
        double a = H[i * m + j], c = H[(i + 1) * m + j];
        H[i * m + j] = cs[i] * a + sn[i] * c;
        H[(i + 1) * m + j] = -sn[i] * a + cs[i] * c;
      }
      double a = H[j * m + j], c = H[(j + 1) * m + j];
      double den = std::sqrt(a * a + c * c);
      cs[j] = den > 0 ? a / den : 1.0;
      sn[j] = den > 0 ? c / den : 0.0;
      H[j * m + j] = cs[j] * a + sn[j] * c;
      H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      j++;
      if (std::fabs(g[j]) <= tol_abs || hn == 0.0) break;
    }
    for (int i = j - 1; i >= 0; i--) {
      double s = g[i];
      for (int t = i + 1; t < j; t++) s -= H[i * m + t] * y[t];
      y[i] = s / H[i * m + i];
    }
    std::fill(u.begin(), u.end(), 0.0);
    for (int i = 0; i < j; i++)
      for (size_t t = 0; t < N; t++) u[t] += y[i] * V[i][t];
    precond(u.data(), z.data());
    for (size_t t = 0; t < N; t++) x[t] += z[t];
  }
  if (stats) {
    stats[0] = total_it;
    stats[1] = true_rel;
    stats[2] = outer;
  }
  return status;
}

extern "C" int orc_dense_solve(int n, const double* A, const double* b,
                               double* x) {
  // Gaussian elimination with partial pivoting.
  std::vector<double> M(A, A + (size_t)n * n), r(b, b + n);
  for (int c = 0; c < n; c++) {
    int p = c;
    for (int i = c + 1; i < n; i++)
      if (std::fabs(M[(size_t)i * n + c]) > std::fabs(M[(size_t)p * n + c]))
        p = i;
    if (M[(size_t)p * n + c] == 0.0) return -1;
    if (p != c) {
      for (int k = 0; k < n; k++)
        std::swap(M[(size_t)p * n + k], M[(size_t)c * n + k]);
      std::swap(r[p], r[c]);
    }
    for (int i = c + 1; i < n; i++) {
      double l = M[(size_t)i * n + c] / M[(size_t)c * n + c];
      if (l == 0.0) continue;
      for (int k = c; k < n; k++) M[(size_t)i * n + k] -= l * M[(size_t)c * n + k];
      r[i] -= l * r[c];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = r[i];
    for (int k = i + 1; k < n; k++) s -= M[(size_t)i * n + k] * x[k];
    x[i] = s / M[(size_t)i * n + i];
  }
  return 0;
}
