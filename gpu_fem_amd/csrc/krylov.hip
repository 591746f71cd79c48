// Krylov building blocks and the restarted, right-preconditioned GMRES(m)
// driver (FP64) plus its mixed-precision GMRES-IR variant.
//
// Replaces the host-driven loop of examples/iterative/gmres.cu:472-771 and the
// cuBLAS level-1 calls it makes (Dnrm2/Ddot/Daxpy/Dscal/Dtrsv, :540-767).  All
// reductions are two-pass and deterministic; the Hessenberg column, Givens
// rotations and least-squares solve stay on the device, and the host reads one
// scalar (the residual estimate) per Arnoldi step.
//
// Orthogonalisation is classical Gram-Schmidt with one re-orthogonalisation
// pass (CGS2): two multi-dot launches and two updates per step instead of the
// (j+1) serial dot/axpy pairs of modified Gram-Schmidt.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "common.h"
#include "krylov.h"

namespace gf {
namespace {

constexpr int kDotBlocks = 256;  // slices of the index range
constexpr int kDotThreads = 256;
constexpr int kDotChunk = 8;

// partial[b * nvec + i] = sum over slice b of V_i . w.  Grid: x = slice of
// the index range (kDotBlocks slices), y = chunk of kDotChunk vectors, so a
// long basis spreads over kDotBlocks x ceil(nvec / kDotChunk) workgroups.
__global__ void __launch_bounds__(kDotThreads)
    multi_dot_partial(size_t N, int nvec, const double* __restrict__ V,
                      size_t ldv, const double* __restrict__ w,
                      double* __restrict__ partial,
                      const int* __restrict__ gate) {
  __shared__ double red[kDotThreads / 64][kDotChunk];
  if (gate && *gate == 0) return;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int i0 = blockIdx.y * kDotChunk;
  const size_t len = (N + gridDim.x - 1) / gridDim.x;
  const size_t t0 = (size_t)blockIdx.x * len;
  const size_t t1 = t0 + len < N ? t0 + len : N;
  double acc[kDotChunk];
#pragma unroll
  for (int c = 0; c < kDotChunk; c++) acc[c] = 0.0;
  for (size_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const double wt = w[t];
#pragma unroll
    for (int c = 0; c < kDotChunk; c++)
      if (i0 + c < nvec) acc[c] += V[(size_t)(i0 + c) * ldv + t] * wt;
  }
#pragma unroll
  for (int c = 0; c < kDotChunk; c++) {
    double v = acc[c];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x < kDotChunk && i0 + (int)threadIdx.x < nvec) {
    double v = 0.0;
    for (int wv = 0; wv < kDotThreads / 64; wv++) v += red[wv][threadIdx.x];
    partial[(size_t)blockIdx.x * nvec + i0 + threadIdx.x] = v;
  }
}

// out[i] (+)= sum_b partial[b*nvec+i]; mode 0: store, 1: add, 2: store sqrt.
// One workgroup per i; fixed summation tree (deterministic).
__global__ void __launch_bounds__(kDotThreads)
    multi_dot_final(int nblocks, int nvec, const double* __restrict__ partial,
                    double* __restrict__ out, int mode,
                    const int* __restrict__ gate) {
  __shared__ double red[kDotThreads / 64];
  const int i = blockIdx.x;
  if (gate && *gate == 0) {  // pass skipped: a stored result reads zero
    if (threadIdx.x == 0 && mode == 0) out[i] = 0.0;
    return;
  }
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  double s = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kDotThreads)
    s += partial[(size_t)b * nvec + i];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < kDotThreads / 64; w++) v += red[w];
    if (mode == 1)
      out[i] += v;
    else if (mode == 2)
      out[i] = sqrt(v);
    else
      out[i] = v;
  }
}

// w -= sum_i h[i] V_i (skipped when *gate == 0)
__global__ void update_w(size_t N, int nvec, const double* __restrict__ V,
                         size_t ldv, const double* __restrict__ h,
                         double* __restrict__ w, const int* __restrict__ gate) {
  if (gate && *gate == 0) return;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  double s = w[t];
  for (int i = 0; i < nvec; i++) s -= h[i] * V[(size_t)i * ldv + t];
  w[t] = s;
}

// second Gram-Schmidt pass only if the first one cancelled more than
// 1 - 1/sqrt(2) of w (Daniel, Gragg, Kaufman and Stewart, Math. Comp. 30,
// 1976): gate = ||w_after|| < ||w_before|| / sqrt(2)
__global__ void reorth_gate(const double* __restrict__ nrm, int* gate) {
  if (threadIdx.x == 0)
    *gate = nrm[0] < 0.70710678118654752 * nrm[1] ? 1 : 0;
}

// u = sum_i y[i] V_i
__global__ void lincomb(size_t N, int nvec, const double* __restrict__ V,
                        size_t ldv, const double* __restrict__ y,
                        double* __restrict__ u) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  double s = 0.0;
  for (int i = 0; i < nvec; i++) s += y[i] * V[(size_t)i * ldv + t];
  u[t] = s;
}

// v = w / *nrm (0 if nrm == 0)
__global__ void scale_by_inv(size_t N, const double* __restrict__ w,
                             const double* __restrict__ nrm,
                             double* __restrict__ v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  const double d = *nrm;
  v[t] = d > 0.0 ? w[t] / d : 0.0;
}

// r = b - r
__global__ void sub_from(size_t N, const double* __restrict__ b,
                         double* __restrict__ r) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N) r[t] = b[t] - r[t];
}

__global__ void axpy_k(size_t N, double a, const double* __restrict__ x,
                       double* __restrict__ y) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N) y[t] += a * x[t];
}

__global__ void to_f32(size_t N, const double* __restrict__ in,
                       float* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N) out[t] = (float)in[t];
}
__global__ void to_f64(size_t N, const float* __restrict__ in,
                       double* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N) out[t] = (double)in[t];
}

// Arnoldi column j: H[0..j, j] = h, H[j+1, j] = hn; apply the previous
// rotations, form the new one, update g; out = |g[j+1]|, g_j1 stored.
// State layout in `st`: H (m+1)*m | cs m | sn m | g m+1
// Column j of the Hessenberg matrix: store h, apply the previous rotations
// (a sequential recurrence, run by one lane out of LDS after the rotations
// and the column are staged there by the whole workgroup), form rotation j
// and update g.
__global__ void __launch_bounds__(256)
    givens_step(int j, int m, const double* __restrict__ h,
                const double* __restrict__ hn, double* __restrict__ st,
                double* out) {
  extern __shared__ double sh[];  // col[m + 1], cs[m], sn[m]
  double* col = sh;
  double* lcs = sh + (m + 1);
  double* lsn = lcs + m;
  double* H = st;
  double* cs = st + (size_t)(m + 1) * m;
  double* sn = cs + m;
  double* g = sn + m;
  for (int i = threadIdx.x; i <= j; i += blockDim.x) col[i] = h[i];
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    lcs[i] = cs[i];
    lsn[i] = sn[i];
  }
  if (threadIdx.x == 0) col[j + 1] = *hn;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < j; i++) {
      const double a = col[i], c = col[i + 1];
      col[i] = lcs[i] * a + lsn[i] * c;
      col[i + 1] = -lsn[i] * a + lcs[i] * c;
    }
    const double a = col[j], c = col[j + 1];
    const double den = sqrt(a * a + c * c);
    const double cj = den > 0.0 ? a / den : 1.0;
    const double sj = den > 0.0 ? c / den : 0.0;
    cs[j] = cj;
    sn[j] = sj;
    col[j] = cj * a + sj * c;
    col[j + 1] = 0.0;
    g[j + 1] = -sj * g[j];
    g[j] = cj * g[j];
    out[0] = fabs(g[j + 1]);
    out[1] = *hn;
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= j + 1; i += blockDim.x) H[i * m + j] = col[i];
}

__global__ void init_g(int m, const double* __restrict__ beta,
                       double* __restrict__ st) {
  double* g = st + (size_t)(m + 1) * m + 2 * m;
  const int i = threadIdx.x;
  if (i <= m) g[i] = (i == 0) ? *beta : 0.0;
}

// y = H[0:k,0:k]^-1 g[0:k] by column-oriented back substitution: one lane
// forms y_i, then the workgroup removes column i from the right-hand side
__global__ void __launch_bounds__(1024)
    solve_upper(int k, int m, const double* __restrict__ st,
                double* __restrict__ y) {
  extern __shared__ double rhs[];  // k entries
  const double* H = st;
  const double* g = st + (size_t)(m + 1) * m + 2 * m;
  for (int i = threadIdx.x; i < k; i += blockDim.x) rhs[i] = g[i];
  __syncthreads();
  for (int i = k - 1; i >= 0; i--) {
    const double yi = rhs[i] / H[(size_t)i * m + i];
    for (int t = threadIdx.x; t < i; t += blockDim.x)
      rhs[t] -= H[(size_t)t * m + i] * yi;
    if (threadIdx.x == 0) y[i] = yi;
    __syncthreads();
  }
}

// ---- DCGS2 Arnoldi (classical Gram-Schmidt with delayed
// reorthogonalisation: Swirydowicz, Langou, Ananthan, Yang and Thomas,
// Numer. Linear Algebra Appl. 28 (2021); Bielich, Langou, Thomas, Swirydowicz,
// Yamazaki and Boman, Parallel Comput. 112 (2022)).  Step j applies the
// operator to q_j as it came out of one Gram-Schmidt pass, then one pass
// over the basis computes both q_j's second-pass coefficients s = Q^T q_j
// and the projections Q^T w; the corrections that make q_j and column j-1
// of H final, and w's projection onto the corrected basis, follow from
// those dot products and the Arnoldi relation, and a second pass over the
// basis writes the final q_j and the next (once-orthogonalised) vector.
// Two reads of the basis per step instead of four.

// partial[b * nout + o], nout = 2 nvec + 1: o < nvec: V_o . x1; nvec + i:
// V_i . x2; 2 nvec: x2 . x2 (chunk 0 only)
// Slice bx of nbx of the index range, vectors by*kDotChunk.. (one
// workgroup; the body of multi_dot2_partial and its batched form).
__device__ __forceinline__ void dot2_partial_body(
    size_t N, int nvec, const double* __restrict__ V, size_t ldv,
    const double* __restrict__ x1, const double* __restrict__ x2,
    double* __restrict__ partial, int bx, int nbx, int by) {
  __shared__ double red[kDotThreads / 64][2 * kDotChunk + 1];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int i0 = by * kDotChunk;
  const int nout = 2 * nvec + 1;
  const size_t len = (N + nbx - 1) / nbx;
  const size_t t0 = (size_t)bx * len;
  const size_t t1 = t0 + len < N ? t0 + len : N;
  double a1[kDotChunk], a2[kDotChunk], aw = 0.0;
#pragma unroll
  for (int c = 0; c < kDotChunk; c++) a1[c] = a2[c] = 0.0;
  for (size_t t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const double u = x1[t], w = x2[t];
    aw += w * w;
#pragma unroll
    for (int c = 0; c < kDotChunk; c++)
      if (i0 + c < nvec) {
        const double v = V[(size_t)(i0 + c) * ldv + t];
        a1[c] += v * u;
        a2[c] += v * w;
      }
  }
#pragma unroll
  for (int c = 0; c <= 2 * kDotChunk; c++) {
    double v = c < kDotChunk ? a1[c] : (c < 2 * kDotChunk ? a2[c - kDotChunk] : aw);
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (lane == 0) red[wave][c] = v;
  }
  __syncthreads();
  if (threadIdx.x <= 2 * kDotChunk) {
    const int c = threadIdx.x;
    double v = 0.0;
    for (int wv = 0; wv < kDotThreads / 64; wv++) v += red[wv][c];
    int o = -1;
    if (c < kDotChunk && i0 + c < nvec)
      o = i0 + c;
    else if (c >= kDotChunk && c < 2 * kDotChunk && i0 + c - kDotChunk < nvec)
      o = nvec + i0 + c - kDotChunk;
    else if (c == 2 * kDotChunk && by == 0)
      o = 2 * nvec;
    if (o >= 0) partial[(size_t)bx * nout + o] = v;
  }
}

__global__ void __launch_bounds__(kDotThreads)
    multi_dot2_partial(size_t N, int nvec, const double* __restrict__ V,
                       size_t ldv, const double* __restrict__ x1,
                       const double* __restrict__ x2,
                       double* __restrict__ partial) {
  dot2_partial_body(N, nvec, V, ldv, x1, x2, partial, blockIdx.x, gridDim.x,
                    blockIdx.y);
}

// Workgroup sum of v (256 threads), result in every thread.
__device__ double block_sum256(double v, double* red) {
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return red[0] + red[1] + red[2] + red[3];
}

// DCGS2 coefficients of step j (one 256-thread workgroup).  d[0..j]: Q^T q_j
// (d[j] = q_j . q_j); d[j+1 .. 2j+1]: Q^T w; d[2j+2]: w . w (finalize: only
// the first j+1 entries).  Hc: the Hessenberg matrix before rotations,
// column-major with leading dimension m+1; column j-1 holds the provisional
// entries of step j-1 and rho[0] the provisional H[j, j-1].  Makes column
// j-1 final, applies the Givens rotations (st: rotated H (row-major, as
// solve_upper reads it) | cs | sn | g), writes out = {|g_j|, H[j, j-1]},
// and unless finalize computes the provisional column j, rho[0] = its
// subdiagonal, and for update2: co[0..j-1] = s, co[m+1 .. m+1+j] = c,
// co[2m+2] = 1/alpha, co[2m+3] = 1/(alpha rho).
__device__ __forceinline__ void dcgs2_coeffs_body(
    int j, int m, const double* __restrict__ d, double* Hc, double* rho,
    double* st, double* co, double* out, int finalize, double* sh) {
  __shared__ double red[4];  // sh: col[m + 2] | lcs[m] | lsn[m] | s[m]
  double* col = sh;
  double* lcs = sh + (m + 2);
  double* lsn = lcs + m;
  double* ls = lsn + m;
  const int tid = threadIdx.x;
  const size_t ldh = (size_t)m + 1;
  double* R = st;
  double* cs = st + (size_t)(m + 1) * m;
  double* sn = cs + m;
  double* g = sn + m;
  for (int i = tid; i < j; i += 256) ls[i] = d[i];
  double alpha = 1.0;
  // finalize 2: the estimate for j iterations only, nothing is written but
  // out[] (the same column finalised with finalize 1 gives the same value)
  const bool commit = finalize != 2;
  if (j > 0) {
    double ss = 0.0;
    for (int i = tid; i < j; i += 256) ss += d[i] * d[i];
    ss = block_sum256(ss, red);  // (synchronises)
    const double sig = d[j] - ss;
    alpha = sig > 0.0 ? sqrt(sig) : 0.0;
    const double r = rho[0];
    // final column j-1: H[i, j-1] += rho s_i, H[j, j-1] = rho alpha
    double* hc = Hc + (size_t)(j - 1) * ldh;
    for (int i = tid; i < j; i += 256) {
      const double h = hc[i] + r * ls[i];
      if (commit) hc[i] = h;
      col[i] = h;
    }
    for (int i = tid; i < j - 1; i += 256) {
      lcs[i] = cs[i];
      lsn[i] = sn[i];
    }
    if (tid == 0) {
      if (commit) hc[j] = r * alpha;
      col[j] = r * alpha;
    }
    __syncthreads();
    if (tid == 0) {
      const int jc = j - 1;
      for (int i = 0; i < jc; i++) {
        const double a = col[i], c = col[i + 1];
        col[i] = lcs[i] * a + lsn[i] * c;
        col[i + 1] = -lsn[i] * a + lcs[i] * c;
      }
      const double a = col[jc], c = col[jc + 1];
      const double den = sqrt(a * a + c * c);
      const double cj = den > 0.0 ? a / den : 1.0;
      const double sj = den > 0.0 ? c / den : 0.0;
      col[jc] = cj * a + sj * c;
      col[jc + 1] = 0.0;
      const double gn = -sj * g[jc];
      out[0] = fabs(gn);
      out[1] = r * alpha;
      if (commit) {
        cs[jc] = cj;
        sn[jc] = sj;
        g[jc + 1] = gn;
        g[jc] = cj * g[jc];
      }
    }
    __syncthreads();
    if (commit)
      for (int i = tid; i <= j; i += 256) R[(size_t)i * m + (j - 1)] = col[i];
  }
  if (finalize) return;
  const double inva = alpha > 0.0 ? 1.0 / alpha : 0.0;
  const double* t = d + (j + 1);
  // c_j = (t_j - s . t) / alpha; c_i = t_i (i < j)
  double std_ = 0.0;
  for (int i = tid; i < j; i += 256) std_ += ls[i] * t[i];
  std_ = block_sum256(std_, red);  // (synchronises: column j-1 is final)
  const double cj = (t[j] - std_) * inva;
  // (H s)_i over the final columns 0..j-1 (column k has rows 0..k+1)
  double cc = 0.0;
  double* hj = Hc + (size_t)j * ldh;
  for (int i = tid; i <= j; i += 256) {
    // all lanes walk k together so each load is one coalesced row segment;
    // entries below the subdiagonal (k < i - 1) are skipped, not read
    double hs = 0.0;
    for (int k = 0; k < j; k++)
      if (k + 1 >= i) hs += Hc[(size_t)k * ldh + i] * ls[k];
    const double ci = i < j ? t[i] : cj;
    hj[i] = (ci - hs) * inva;
    co[m + 1 + i] = ci;
    if (i < j) co[i] = ls[i];
    cc += ci * ci;
  }
  cc = block_sum256(cc, red);
  if (tid == 0) {
    const double ww = d[2 * j + 2];
    double u2 = ww - cc;
    // the norm only scales the next vector, which the next step corrects;
    // keep it positive when cancellation has eaten the difference
    if (!(u2 > 1e-24 * ww)) u2 = 1e-24 * ww;
    const double r1 = ww > 0.0 ? sqrt(u2) * inva : 0.0;
    rho[0] = r1;
    hj[j + 1] = r1;
    co[2 * m + 2] = inva;
    co[2 * m + 3] = (alpha > 0.0 && r1 > 0.0) ? 1.0 / (alpha * r1) : 0.0;
  }
}

__global__ void __launch_bounds__(256)
    dcgs2_coeffs(int j, int m, const double* __restrict__ d, double* Hc,
                 double* rho, double* st, double* co, double* out,
                 int finalize) {
  extern __shared__ double sh[];
  dcgs2_coeffs_body(j, m, d, Hc, rho, st, co, out, finalize, sh);
}

// One pass over V[0..j-1]: q_j = (V_j - sum s_i V_i) / alpha (final) and
// V_{j+1} = (w - sum_{i<j} c_i V_i - c_j q_j) / (alpha rho)
__device__ __forceinline__ void update2_body(size_t t, int j, int m,
                                             double* __restrict__ V,
                                             size_t ldv,
                                             const double* __restrict__ w,
                                             const double* __restrict__ co) {
  const double* sc = co;
  const double* cc = co + m + 1;
  double as = 0.0, ac = 0.0;
  for (int i = 0; i < j; i++) {
    const double v = V[(size_t)i * ldv + t];
    as += sc[i] * v;
    ac += cc[i] * v;
  }
  const double qf = j > 0 ? (V[(size_t)j * ldv + t] - as) * co[2 * m + 2]
                          : V[t];
  if (j > 0) V[(size_t)j * ldv + t] = qf;
  V[(size_t)(j + 1) * ldv + t] = (w[t] - ac - cc[j] * qf) * co[2 * m + 3];
}

// update2 with the same correction applied to Z_j = M^-1 V_j (the
// preconditioned vectors follow the basis they were computed from)
// zj = M^-1 V_j as computed from the uncorrected V_j
__device__ __forceinline__ void update_z_body(size_t t, int j, int m,
                                              double zj,
                                              double* __restrict__ Z,
                                              size_t ldv,
                                              const double* __restrict__ co) {
  if (j > 0) {
    double az = 0.0;
    for (int i = 0; i < j; i++) az += co[i] * Z[(size_t)i * ldv + t];
    zj = (zj - az) * co[2 * m + 2];
  }
  Z[(size_t)j * ldv + t] = zj;
}

__global__ void update2z(size_t N, int j, int m, double* __restrict__ V,
                         size_t ldv, const double* __restrict__ w,
                         const double* __restrict__ co,
                         double* __restrict__ Z) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  update_z_body(t, j, m, Z[(size_t)j * ldv + t], Z, ldv, co);
  update2_body(t, j, m, V, ldv, w, co);
}

__global__ void update2(size_t N, int j, int m, double* __restrict__ V,
                        size_t ldv, const double* __restrict__ w,
                        const double* __restrict__ co) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  update2_body(t, j, m, V, ldv, w, co);
}

inline dim3 grid1(size_t N) { return dim3(div_up(N, 256)); }

}  // namespace

int KrylovWork::init(gf_context c, size_t N_, int m_, bool mixed,
                     bool keep_z) {
  ctx = c;
  N = N_;
  m = m_;
  auto dm = [&](void** p, size_t bytes) {
    return dmalloc(ctx, p, bytes ? bytes : 8);
  };
  GF_TRY(dm((void**)&V, sizeof(double) * N * (m + 1)));
  if (keep_z) GF_TRY(dm((void**)&Z, sizeof(double) * N * (m > 0 ? m : 1)));
  GF_TRY(dm((void**)&w, sizeof(double) * N));
  GF_TRY(dm((void**)&z, sizeof(double) * N));
  GF_TRY(dm((void**)&r, sizeof(double) * N));
  GF_TRY(dm((void**)&tmp, sizeof(double) * N));
  GF_TRY(dm((void**)&h, sizeof(double) * (m + 2)));
  GF_TRY(dm((void**)&h2, sizeof(double) * (m + 2)));
  GF_TRY(dm((void**)&nrm, sizeof(double) * 2));
  GF_TRY(dm((void**)&y, sizeof(double) * (m + 1)));
  GF_TRY(dm((void**)&state, sizeof(double) * ((m + 1) * m + 3 * m + 2)));
  GF_TRY(dm((void**)&partial, sizeof(double) * kDotBlocks * (2 * m + 4)));
  GF_TRY(dm((void**)&hc, sizeof(double) * (m + 1) * (m + 1)));
  GF_TRY(dm((void**)&dd, sizeof(double) * (2 * m + 4)));
  GF_TRY(dm((void**)&co, sizeof(double) * (2 * m + 4)));
  GF_TRY(dm((void**)&rho, sizeof(double)));
  GF_TRY(dm((void**)&scal, sizeof(double) * 4));
  GF_TRY(dm((void**)&gate, sizeof(int)));
  if (mixed) {
    GF_TRY(dm((void**)&a32, sizeof(float) * N));
    GF_TRY(dm((void**)&b32, sizeof(float) * N));
    GF_TRY(dm((void**)&c32, sizeof(float) * N));
  }
  if (!ctx->pinned)
    GF_TRY(check_hip(ctx, hipHostMalloc((void**)&ctx->pinned, sizeof(double) * 8),
                     "hipHostMalloc"));
  host_scal = ctx->pinned;
  GF_TRY(check_hip(ctx, hipEventCreateWithFlags(&ev[0], hipEventDisableTiming),
                   "event"));
  GF_TRY(check_hip(ctx, hipEventCreateWithFlags(&ev[1], hipEventDisableTiming),
                   "event"));
  return GF_OK;
}

void KrylovWork::release() {
  dfree(ctx, V);
  dfree(ctx, Z);
  dfree(ctx, w);
  dfree(ctx, z);
  dfree(ctx, r);
  dfree(ctx, tmp);
  dfree(ctx, h);
  dfree(ctx, h2);
  dfree(ctx, nrm);
  dfree(ctx, y);
  dfree(ctx, state);
  dfree(ctx, partial);
  dfree(ctx, hc);
  dfree(ctx, dd);
  dfree(ctx, co);
  dfree(ctx, rho);
  dfree(ctx, scal);
  dfree(ctx, gate);
  dfree(ctx, a32);
  dfree(ctx, b32);
  dfree(ctx, c32);
  if (ev[0]) hipEventDestroy(ev[0]);
  if (ev[1]) hipEventDestroy(ev[1]);
  *this = KrylovWork();
}

// out[0..nvec) = V_i . w  (mode: 0 store, 1 add, 2 sqrt)
int KrylovWork::dots(const double* Vb, size_t ldv, int nvec, const double* x,
                     double* out, int mode, const int* gate) {
  hipLaunchKernelGGL(multi_dot_partial,
                     dim3(kDotBlocks, div_up(nvec, kDotChunk)),
                     dim3(kDotThreads), 0, ctx->stream, N, nvec, Vb, ldv, x,
                     partial, gate);
  hipLaunchKernelGGL(multi_dot_final, dim3(nvec), dim3(kDotThreads), 0,
                     ctx->stream, (int)kDotBlocks, nvec, partial, out, mode,
                     gate);
  return check_launch(ctx, "multi_dot");
}

#ifndef GPUFEM_GMRES_DCGS2
#define GPUFEM_GMRES_DCGS2 1  // delayed reorthogonalisation (two basis passes)
#endif
constexpr bool kDcgs2 = GPUFEM_GMRES_DCGS2;
#ifndef GPUFEM_GMRES_CYCLE_TARGET
#define GPUFEM_GMRES_CYCLE_TARGET 0.3
#endif
// each restart cycle runs until its residual estimate falls below this
// fraction of rtol ||b|| (see the comment at its use)
constexpr double kCycleTarget = GPUFEM_GMRES_CYCLE_TARGET;
// sweep depth (L + U dependency levels) below which Arnoldi steps are issued
// one ahead of the convergence test (see gmres_solve)
constexpr int kLookaheadLevels = 128;
// an estimate is formed ahead of step j when the rate so far predicts it at
// or below this multiple of the cycle's target
constexpr double kPeekMargin = 4.0;

int gmres_solve(KrylovWork& W, const Operator& op, const double* b, double* x,
                int max_iters, int stall_cycles, double rtol, int prec,
                GmresResult* res) {
  const bool ir = prec == 1;
  gf_context ctx = W.ctx;
  hipStream_t s = ctx->stream;
  const size_t N = W.N;
  const int m = W.m;
  // the preconditioned basis, when the work space keeps it (FP64 and FP32
  // preconditioner inside FP64 GMRES)
  const bool keep_z = W.Z != nullptr && !ir;
  // Look-ahead: Arnoldi step j+1 is issued before the host reads step j's
  // residual estimate, hiding the host round trip (tens of microseconds)
  // between steps; a step issued past the stopping point is a wasted
  // preconditioner application.  With deep sweeps (a step of milliseconds)
  // the waste costs more than the round trips, so look-ahead is used only
  // for shallow sweeps (GPUFEM_GMRES_LOOKAHEAD=0/1 overrides).
  bool ahead = true;
  {
    int nl = 0, nu = 0;
    if (op.ilu && gf_ilu_info(op.ilu, nullptr, &nl, &nu) == GF_OK)
      ahead = nl + nu < kLookaheadLevels;
    if (const char* e = getenv("GPUFEM_GMRES_LOOKAHEAD")) ahead = atoi(e) != 0;
  }
  bool peek_on = true;
  if (const char* e = getenv("GPUFEM_GMRES_PEEK")) peek_on = atoi(e) != 0;
  // ||b||
  GF_TRY(W.dots(b, N, 1, b, W.scal, 2));
  GF_HIP(ctx, hipMemcpyAsync(W.host_scal, W.scal, sizeof(double),
                             hipMemcpyDeviceToHost, s));
  GF_HIP(ctx, hipStreamSynchronize(s));
  double bnorm = W.host_scal[0];
  if (bnorm == 0.0) bnorm = 1.0;
  // a sweep whose bounded wait expired sets this flag; it is cleared here so
  // that a flag left by an earlier, abandoned solve is not reported now, and
  // read at every cycle start, where the host synchronises anyway
  int* sweep_err = op.ilu ? ilu_sweep_err(op.ilu) : nullptr;
  int* host_err = (int*)(W.host_scal + 6);
  if (sweep_err) GF_HIP(ctx, hipMemsetAsync(sweep_err, 0, sizeof(int), s));
  int it = 0, outer = 0, status = GF_ERR_NO_CONVERGENCE;
  double true_rel = 0.0, last_est = 0.0;
  std::vector<double> hist;  // true relative residual at each cycle start
  bool stalled = false;
  static const bool verbose = getenv("GPUFEM_GMRES_VERBOSE") != nullptr;
  for (;;) {
    // r = b - A x in FP64 with compensated accumulation
    GF_TRY(gf_bsr_resid_comp_f64(ctx, op.n, op.krowp, op.kcols, op.kvals, x, b,
                                 W.r));
    GF_TRY(W.dots(W.r, N, 1, W.r, W.nrm, 2));
    GF_HIP(ctx, hipMemcpyAsync(W.host_scal, W.nrm, sizeof(double),
                               hipMemcpyDeviceToHost, s));
    *host_err = 0;
    if (sweep_err)
      GF_HIP(ctx, hipMemcpyAsync(host_err, sweep_err, sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    GF_HIP(ctx, hipStreamSynchronize(s));
    const double beta = W.host_scal[0];
    true_rel = beta / bnorm;
    if (*host_err) {
      // the preconditioner applications since the last cycle start used
      // stand-in values: the iterate cannot be trusted
      status = set_error(ctx, GF_ERR_TIMEOUT,
                         "triangular sweep: a dependency wait expired");
      GF_HIP(ctx, hipMemsetAsync(sweep_err, 0, sizeof(int), s));
      break;
    }
    if (verbose)
      fprintf(stderr, "gmres: cycle %d  iterations %d  true rel. residual %.3e"
              "  last estimate %.3e\n", outer, it, true_rel, last_est / bnorm);
    if (!std::isfinite(beta)) {
      status = set_error(ctx, GF_ERR_BREAKDOWN, "GMRES: non-finite residual");
      break;
    }
    if (true_rel <= rtol) {
      status = GF_OK;
      break;
    }
    if (it >= max_iters) break;
    // stagnation (when asked for): less than a 10% reduction of the true
    // residual over the last stall_cycles cycles ends the solve early
    hist.push_back(true_rel);
    if (stall_cycles > 0 && (int)hist.size() > stall_cycles &&
        true_rel > 0.9 * hist[hist.size() - 1 - stall_cycles]) {
      stalled = true;
      break;
    }
    outer++;
    // cycles aim at 0.3 rtol |b|: the true residual of an ill-conditioned
    // preconditioned system trails the Arnoldi estimate, and cycles that stop
    // exactly at rtol |b| restart again and again just above it (0.3 took
    // 0-8% fewer iterations than 0.1 over 20 served requests and the 400x400
    // plate; 1.0 and above stagnate there)
    const double tol_abs = ir ? std::max(1e-5 * beta, kCycleTarget * rtol * bnorm)
                              : kCycleTarget * rtol * bnorm;
    hipLaunchKernelGGL(scale_by_inv, grid1(N), dim3(256), 0, s, N, W.r, W.nrm,
                       W.V);
    hipLaunchKernelGGL(init_g, dim3(1), dim3(64 * div_up(m + 1, 64)), 0, s, m,
                       W.nrm, W.state);
    // Arnoldi steps are issued one ahead of the convergence test: the host
    // reads the residual estimate of step j while the device runs step j+1,
    // so there is no host round trip between steps.  A step issued past the
    // stopping point only writes column j of H, V_{j+1} and g[j..j+1], none
    // of which enter the update with j columns.
    int issued = 0, j = 0;
    const int limit = std::min(m, max_iters - it);
    bool stop = false;
    if (kDcgs2) {
      // DCGS2: step js (js < limit) applies the operator to V_js and
      // finalises column js-1; step limit only finalises column limit-1.
      // Step js reports the estimate after js iterations (js >= 1), read
      // once step js+1 is queued.
      const size_t shm = sizeof(double) * (4 * (size_t)m + 2);
      double e1 = 0.0, e2 = 0.0;  // the last two estimates of the cycle
      while (!stop) {
        if (issued <= limit) {
          const int js = issued;
          double* vj = W.V + (size_t)js * N;
          // Without look-ahead, step js is issued once the estimate for js - 1
          // iterations is known.  When the convergence rate so far predicts
          // that js iterations meet the target, the estimate for js
          // iterations is formed first from the dot products with v_js alone
          // (no preconditioner application): if it meets the target, column
          // js - 1 is finalised from the same products and the cycle ends
          // without the sweeps of step js.
          if (peek_on && !ahead && js >= 2 && js < limit && e2 > 0.0 &&
              e1 * (e1 / e2) <= kPeekMargin * tol_abs) {
            GF_TRY(W.dots(W.V, N, js + 1, vj, W.dd, 0));
            hipLaunchKernelGGL(dcgs2_coeffs, dim3(1), dim3(256), shm, s, js,
                               m, W.dd, W.hc, W.rho, W.state, W.co, W.scal, 2);
            GF_HIP(ctx, hipMemcpyAsync(W.host_scal + 4, W.scal,
                                       2 * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
            GF_HIP(ctx, hipStreamSynchronize(s));
            const double est = W.host_scal[4], hn = W.host_scal[5];
            if (!std::isfinite(est) || est <= tol_abs || hn == 0.0) {
              hipLaunchKernelGGL(dcgs2_coeffs, dim3(1), dim3(256), shm, s, js,
                                 m, W.dd, W.hc, W.rho, W.state, W.co, W.scal,
                                 1);
              GF_TRY(check_launch(ctx, "arnoldi finalisation"));
              last_est = est;
              j = js;
              stop = true;
              break;
            }
          }
          if (js < limit) {
            // w = A M^-1 v_j (M^-1 v_j kept in Z_j when Z is kept)
            GF_TRY(op.apply_prec_a(vj, W.w, prec,
                                   keep_z ? W.Z + (size_t)js * N : nullptr));
            const int nv = js + 1, nout = 2 * nv + 1;
            hipLaunchKernelGGL(multi_dot2_partial,
                               dim3(kDotBlocks, div_up(nv, kDotChunk)),
                               dim3(kDotThreads), 0, s, N, nv, W.V, N, vj,
                               W.w, W.partial);
            hipLaunchKernelGGL(multi_dot_final, dim3(nout), dim3(kDotThreads),
                               0, s, (int)kDotBlocks, nout, W.partial, W.dd,
                               0, (const int*)nullptr);
            hipLaunchKernelGGL(dcgs2_coeffs, dim3(1), dim3(256), shm, s, js,
                               m, W.dd, W.hc, W.rho, W.state, W.co, W.scal, 0);
            if (keep_z)
              hipLaunchKernelGGL(update2z, grid1(N), dim3(256), 0, s, N, js,
                                 m, W.V, N, W.w, W.co, W.Z);
            else
              hipLaunchKernelGGL(update2, grid1(N), dim3(256), 0, s, N, js, m,
                                 W.V, N, W.w, W.co);
          } else {
            GF_TRY(W.dots(W.V, N, js + 1, vj, W.dd, 0));
            hipLaunchKernelGGL(dcgs2_coeffs, dim3(1), dim3(256), shm, s, js,
                               m, W.dd, W.hc, W.rho, W.state, W.co, W.scal, 1);
          }
          GF_TRY(check_launch(ctx, "arnoldi step"));
          if (js >= 1) {
            GF_HIP(ctx, hipMemcpyAsync(W.host_scal + 2 * (js % 2), W.scal,
                                       2 * sizeof(double),
                                       hipMemcpyDeviceToHost, s));
            GF_HIP(ctx, hipEventRecord(W.ev[js % 2], s));
          }
          issued++;
        }
        // iteration count j + 1 is reported by step j + 1; it is read once
        // step j + 2 is queued when look-ahead is on, at once otherwise
        if (issued > j + 1 &&
            (!ahead || issued - (j + 1) >= 2 || issued > limit)) {
          const int r = j + 1;
          GF_HIP(ctx, hipEventSynchronize(W.ev[r % 2]));
          const double est = W.host_scal[2 * (r % 2)];
          const double hn = W.host_scal[2 * (r % 2) + 1];
          last_est = est;
          e2 = e1;
          e1 = est;
          j = r;
          if (!std::isfinite(est) || est <= tol_abs || hn == 0.0 || j == limit)
            stop = true;
        }
      }
    }
    while (!stop) {
      if (issued < limit) {
        const int js = issued;
        double* vj = W.V + (size_t)js * N;
        GF_TRY(op.apply_prec_a(vj, W.w, prec));  // w = A M^-1 v_j
        // classical Gram-Schmidt, second pass only when needed (DGKS)
        GF_TRY(W.dots(W.w, N, 1, W.w, W.nrm + 1, 2));  // ||w|| before
        GF_TRY(W.dots(W.V, N, js + 1, W.w, W.h, 0));
        hipLaunchKernelGGL(update_w, grid1(N), dim3(256), 0, s, N, js + 1,
                           W.V, N, W.h, W.w, (const int*)nullptr);
        GF_TRY(W.dots(W.w, N, 1, W.w, W.nrm, 2));
        hipLaunchKernelGGL(reorth_gate, dim3(1), dim3(64), 0, s, W.nrm,
                           W.gate);
        GF_TRY(W.dots(W.V, N, js + 1, W.w, W.h2, 0, W.gate));
        hipLaunchKernelGGL(update_w, grid1(N), dim3(256), 0, s, N, js + 1,
                           W.V, N, W.h2, W.w, (const int*)W.gate);
        hipLaunchKernelGGL(axpy_k, dim3(1), dim3(64 * div_up(js + 1, 64)), 0,
                           s, (size_t)(js + 1), 1.0, W.h2, W.h);
        GF_TRY(W.dots(W.w, N, 1, W.w, W.nrm, 2, W.gate));
        hipLaunchKernelGGL(scale_by_inv, grid1(N), dim3(256), 0, s, N, W.w,
                           W.nrm, W.V + (size_t)(js + 1) * N);
        hipLaunchKernelGGL(givens_step, dim3(1), dim3(256),
                           sizeof(double) * (3 * (size_t)m + 1), s, js, m, W.h,
                           W.nrm, W.state, W.scal);
        GF_TRY(check_launch(ctx, "arnoldi step"));
        GF_HIP(ctx, hipMemcpyAsync(W.host_scal + 2 * (js % 2), W.scal,
                                   2 * sizeof(double), hipMemcpyDeviceToHost,
                                   s));
        GF_HIP(ctx, hipEventRecord(W.ev[js % 2], s));
        issued++;
      }
      // test step j once the next one is queued (or nothing more is issued)
      if (issued - j >= 2 || issued == limit) {
        GF_HIP(ctx, hipEventSynchronize(W.ev[j % 2]));
        const double est = W.host_scal[2 * (j % 2)];
        const double hn = W.host_scal[2 * (j % 2) + 1];
        last_est = est;
        j++;
        if (!std::isfinite(est) || est <= tol_abs || hn == 0.0 || j == limit)
          stop = true;
      }
    }
    it += j;
    GF_HIP(ctx, hipStreamSynchronize(s));
    // x += M^-1 V y
    hipLaunchKernelGGL(solve_upper, dim3(1), dim3(1024),
                       sizeof(double) * (size_t)(j > 0 ? j : 1), s, j, m,
                       W.state, W.y);
    if (keep_z && kDcgs2) {
      // x += Z y: the preconditioned basis is at hand
      hipLaunchKernelGGL(lincomb, grid1(N), dim3(256), 0, s, N, j, W.Z, N,
                         W.y, W.z);
    } else {
      hipLaunchKernelGGL(lincomb, grid1(N), dim3(256), 0, s, N, j, W.V, N,
                         W.y, W.tmp);
      GF_TRY(op.apply_prec(W.tmp, W.z, prec));
    }
    hipLaunchKernelGGL(axpy_k, grid1(N), dim3(256), 0, s, N, 1.0, W.z, x);
    GF_TRY(check_launch(ctx, "gmres update"));
  }
  if (res) {
    res->iterations = it;
    res->outer = outer;
    res->rel_resid = true_rel;
  }
  if (status == GF_ERR_NO_CONVERGENCE)
    set_error(ctx, status,
              "GMRES: %s after %d iterations, rel. residual %.3e > %.1e",
              stalled ? "stagnated" : "iteration limit", it, true_rel, rtol);
  return status;
}

// ---------------------------------------------------------------- operator
int Operator::apply_a64(const double* in, double* out) const {
  return gf_bsr_spmv_f64(ctx, n, krowp, kcols, kvals, in, out);
}

int Operator::apply_prec(const double* in, double* out, int prec) const {
  const size_t N = 6 * (size_t)n;
  if (prec == 0) return gf_bsr_ilu_apply_f64(ilu, mlu, in, work64, out);
  hipLaunchKernelGGL(to_f32, grid1(N), dim3(256), 0, ctx->stream, N, in, t32a);
  GF_TRY(gf_bsr_ilu_apply_f32(ilu, mlu32, t32a, t32b, t32c));
  hipLaunchKernelGGL(to_f64, grid1(N), dim3(256), 0, ctx->stream, N, t32c, out);
  return check_launch(ctx, "apply_prec fp32");
}

int Operator::apply_prec_a(const double* in, double* out, int prec,
                           double* z_out) const {
  const size_t N = 6 * (size_t)n;
  double* zt = z_out ? z_out : work64b;
  if (prec == 0) {
    GF_TRY(gf_bsr_ilu_apply_f64(ilu, mlu, in, work64, zt));
    return gf_bsr_spmv_f64(ctx, n, krowp, kcols, kvals, zt, out);
  }
  hipLaunchKernelGGL(to_f32, grid1(N), dim3(256), 0, ctx->stream, N, in, t32a);
  GF_TRY(gf_bsr_ilu_apply_f32(ilu, mlu32, t32a, t32b, t32c));
  if (prec == 2) {
    hipLaunchKernelGGL(to_f64, grid1(N), dim3(256), 0, ctx->stream, N, t32c,
                       zt);
    GF_TRY(check_launch(ctx, "apply_prec_a fp32 ilu"));
    return gf_bsr_spmv_f64(ctx, n, krowp, kcols, kvals, zt, out);
  }
  GF_TRY(gf_bsr_spmv_f32(ctx, n, krowp, kcols, kvals32, t32c, t32a));
  hipLaunchKernelGGL(to_f64, grid1(N), dim3(256), 0, ctx->stream, N, t32a, out);
  return check_launch(ctx, "apply_prec_a mixed");
}

}  // namespace gf

// ============================================================ batched GMRES
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>

#include "problem.h"

namespace gf {
namespace {

// index slices per solve in the dot products: the single-solve count, so a
// batched solve sums in the same order as gmres_solve
constexpr int kBatchSlices = kDotBlocks;

// Per-solve arrays of the batched step kernels (device copy in a table).
struct KrySlot {
  int n = 0, m = 0;
  long long N = 0;
  const int* krowp = nullptr;
  const int* kcols = nullptr;
  const double* kvals = nullptr;
  double *V = nullptr, *w = nullptr, *hc = nullptr, *rho = nullptr,
         *st = nullptr, *co = nullptr, *dd = nullptr, *partial = nullptr;
  const void* x = nullptr;  // preconditioned vector (sweep output, T)
  void* b = nullptr;        // sweep right-hand side (T)
  void* tmp = nullptr;      // L sweep output (T)
  double* Z = nullptr;      // preconditioned basis (null: not kept)
};

// Arnoldi state of one solve.  mode (set by batch_begin for the step):
// 1 = apply the operator to V_j and orthogonalise, 2 = finalise column
// limit - 1 only, 0 = idle.  stop: iterations of the cycle once it ends
// (estimate at or below tol_abs, breakdown, or the finalising step).
struct SlotState {
  int j, limit, stop, mode;
  double tol_abs, est, hn, pad;
};

__device__ __forceinline__ int slot_of(const long long* off, int ns,
                                       long long g) {
  int lo = 0, hi = ns;  // off[lo] <= g < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ void batch_begin(int ns, SlotState* st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  SlotState& t = st[s];
  t.mode = t.stop ? 0 : (t.j < t.limit ? 1 : (t.j == t.limit ? 2 : 0));
}

__global__ void batch_end(int ns, SlotState* st) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns) return;
  if (st[s].mode > 0) st[s].j += 1;
}

// b = V_j in the sweep precision, tmp = sentinel (solves stepping the
// operator only)
template <typename T>
__global__ void batch_gather(int ns, long long ntot, const long long* off,
                             const KrySlot* ks, const SlotState* st,
                             unsigned long long sentinel) {
  using B = typename std::conditional<sizeof(T) == 8, unsigned long long,
                                      unsigned int>::type;
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  const int s = slot_of(off, ns, g);
  if (st[s].mode != 1) return;
  const KrySlot& k = ks[s];
  const long long t = g - off[s];
  reinterpret_cast<T*>(k.b)[t] = (T)k.V[(size_t)st[s].j * k.N + t];
  reinterpret_cast<B*>(k.tmp)[t] = (B)sentinel;
}

// w = K x, x the sweep output; six lanes per block row, ten rows per wave
template <typename T>
__global__ void __launch_bounds__(256)
    batch_spmv(int ns, long long nrows, const long long* roff,
               const KrySlot* ks, const SlotState* st) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane >= 60) return;
  const long long grow =
      ((long long)blockIdx.x * (blockDim.x / 64) + wave) * 10 + lane / 6;
  const int r = lane % 6;
  if (grow >= nrows) return;
  const int s = slot_of(roff, ns, grow);
  if (st[s].mode != 1) return;
  const KrySlot& k = ks[s];
  const int row = (int)(grow - roff[s]);
  const T* x = reinterpret_cast<const T*>(k.x);
  double acc = 0.0;
  for (int q = k.krowp[row]; q < k.krowp[row + 1]; q++) {
    const double* a = k.kvals + 36 * (size_t)q + 6 * r;
    const T* xb = x + 6 * (size_t)k.kcols[q];
#pragma unroll
    for (int c = 0; c < 6; c++) acc += a[c] * (double)xb[c];
  }
  k.w[6 * (size_t)row + r] = acc;
}

// grid (kBatchSlices, vector chunks, solves)
__global__ void __launch_bounds__(kDotThreads)
    batch_dot2_partial(const KrySlot* ks, const SlotState* st) {
  const int s = blockIdx.z;
  const int mode = st[s].mode;
  if (mode == 0) return;
  const int j = st[s].j, nv = j + 1;
  if ((int)blockIdx.y * kDotChunk >= nv) return;
  const KrySlot& k = ks[s];
  dot2_partial_body((size_t)k.N, nv, k.V, (size_t)k.N,
                    k.V + (size_t)j * k.N, k.w, k.partial, blockIdx.x,
                    kBatchSlices, blockIdx.y);
}

// grid (outputs, solves): dd[i] = sum over slices of partial, in the
// summation order of multi_dot_final
__global__ void __launch_bounds__(kDotThreads)
    batch_dot_final(const KrySlot* ks, const SlotState* st) {
  __shared__ double red[kDotThreads / 64];
  const int s = blockIdx.y;
  if (st[s].mode == 0) return;
  const int nout = 2 * (st[s].j + 1) + 1;
  const int i = blockIdx.x;
  if (i >= nout) return;
  const KrySlot& k = ks[s];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  double v = 0.0;
  for (int b = threadIdx.x; b < kBatchSlices; b += kDotThreads)
    v += k.partial[(size_t)b * nout + i];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if (lane == 0) red[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kDotThreads / 64; w++) t += red[w];
    k.dd[i] = t;
  }
}

// one 256-thread workgroup per solve
__global__ void __launch_bounds__(256)
    batch_coeffs(const KrySlot* ks, SlotState* st) {
  extern __shared__ double sh[];
  __shared__ double out[2];
  const int s = blockIdx.x;
  SlotState& t = st[s];
  const int mode = t.mode;
  if (mode == 0) return;
  const KrySlot& k = ks[s];
  const int j = t.j;
  if (threadIdx.x == 0) {
    out[0] = 0.0;
    out[1] = 1.0;
  }
  __syncthreads();
  dcgs2_coeffs_body(j, k.m, k.dd, k.hc, k.rho, k.st, k.co, out,
                    mode == 2 ? 1 : 0, sh);
  __syncthreads();
  if (threadIdx.x == 0 && j >= 1) {
    const double est = out[0], hn = out[1];
    t.est = est;
    t.hn = hn;
    if (!isfinite(est) || est <= t.tol_abs || hn == 0.0 || mode == 2)
      t.stop = j;
  }
}

// V (and Z, from the sweep output x, when kept) as update2 / update2z
template <typename T>
__global__ void batch_update2(int ns, long long ntot, const long long* off,
                              const KrySlot* ks, const SlotState* st) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  const int s = slot_of(off, ns, g);
  if (st[s].mode != 1) return;
  const KrySlot& k = ks[s];
  const size_t t = (size_t)(g - off[s]);
  const int j = st[s].j;
  if (k.Z)
    update_z_body(t, j, k.m, (double)reinterpret_cast<const T*>(k.x)[t], k.Z,
                  (size_t)k.N, k.co);
  update2_body(t, j, k.m, k.V, (size_t)k.N, k.w, k.co);
}

// one solve: b = (T) u, tmp = sentinel
template <typename T>
__global__ void gather_one(size_t N, const double* __restrict__ u, T* b,
                           T* tmp, unsigned long long sentinel) {
  using B = typename std::conditional<sizeof(T) == 8, unsigned long long,
                                      unsigned int>::type;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
  b[t] = (T)u[t];
  reinterpret_cast<B*>(tmp)[t] = (B)sentinel;
}

// x += (double) z
template <typename T>
__global__ void axpy_from(size_t N, const T* __restrict__ z,
                          double* __restrict__ x) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N) x[t] += (double)z[t];
}

__global__ void fill_one_int(int* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

}  // namespace

// ---------------------------------------------------------------- engine
namespace {
constexpr int kCycleStall = 0;  // (serving passes stall_cycles per request)

struct Solve {
  BatchJob* job = nullptr;
  gf_problem p = nullptr;
  size_t N = 0;
  int n = 0, m = 0;
  IluView iv;
  double bnorm = 1.0, true_rel = 0.0;
  int it = 0, outer = 0;
  std::vector<double> hist;
  SlotState hs{};       // host copy of the state
  int pos = -1;         // position in the tables
  int hidx = -1;        // pinned readback slot (residual, ||b||)
  std::vector<unsigned char> h_solo;  // host copies alive until uploaded
  std::vector<int2> h_sl, h_su;
  // solo sweep of the cycle end: a one-entry table and its chunk lists
  void* d_solo = nullptr;
  int2 *d_sl = nullptr, *d_su = nullptr;
  int nsl = 0, nsu = 0;
  int* d_one = nullptr;
  bool finished = false;
  int status = GF_ERR_NO_CONVERGENCE;
  bool stalled = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
}  // namespace

class BatchEngine {
 public:
  int device = 0, prec = 2, max_slots = 16;
  gf_context ctx = nullptr;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<BatchJob*> incoming;
  bool stopping = false;
  std::vector<Solve*> act;
  // device tables (max_slots entries) and merged chunk lists
  KrySlot* d_kry = nullptr;
  void* d_sweep = nullptr;
  SlotState* d_state = nullptr;
  long long *d_voff = nullptr, *d_roff = nullptr;
  int2 *d_lm = nullptr, *d_um = nullptr;
  size_t lm_cap = 0, um_cap = 0;
  int nl = 0, nu = 0, grid = 1, maxm = 1;
  long long ntot = 0, nrows = 0;
  unsigned long long* d_counters = nullptr;
  int* d_err = nullptr;
  SlotState* h_state = nullptr;  // pinned
  double* h_scal = nullptr;      // pinned
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool dirty = true;
  int steps = 8;
  int waves_per_cu = 3;  // sweep wavefronts per compute unit
  bool use_graph = true;

  bool fp64() const { return prec == 0; }
  size_t tsize() const { return fp64() ? 8 : 4; }

  double* h_norm = nullptr;      // pinned: 2 per pinned slot index
  std::vector<char> hidx_used;
  int start(Solve* v);
  int residual_issue(Solve* v);
  int cycle_decide(Solve* v);
  int cycle_end(Solve* v);
  int finish(Solve* v);
  int rebuild();
  int issue_steps();
  void loop();
  template <typename T> int fill_tables(std::vector<unsigned char>& sw);
};

namespace {
template <typename T>
SweepSlot<T> sweep_entry(const Solve* v, const int* go) {
  SweepSlot<T> e{};
  e.lmeta = v->iv.lmeta;
  e.umeta = v->iv.umeta;
  e.lchunk = v->iv.lchunk;
  e.uchunk = v->iv.uchunk;
  e.cols = v->iv.cols;
  gf_problem p = v->p;
  if (sizeof(T) == 8) {
    e.lu = (const T*)p->mvals;
    e.b = (const T*)p->work64;
    e.tmp = (T*)p->work64b;
    e.x = (T*)p->kw.z;
  } else {
    e.lu = (const T*)p->mvals32;
    e.b = (const T*)p->t32a;
    e.tmp = (T*)p->t32b;
    e.x = (T*)p->t32c;
  }
  e.go = go;
  return e;
}
}  // namespace

template <typename T>
int BatchEngine::fill_tables(std::vector<unsigned char>& sw) {
  sw.resize(sizeof(SweepSlot<T>) * act.size());
  auto* e = reinterpret_cast<SweepSlot<T>*>(sw.data());
  for (size_t i = 0; i < act.size(); i++)
    e[i] = sweep_entry<T>(act[i], &d_state[i].mode);
  return GF_OK;
}

// Tables, merged chunk lists and the step graph for the current solves.
int BatchEngine::rebuild() {
  hipStream_t s = ctx->stream;
  const int ns = (int)act.size();
  std::vector<KrySlot> hk(ns);
  std::vector<long long> voff(ns + 1, 0), roff(ns + 1, 0);
  std::vector<int2> lm, um;
  maxm = 1;
  int gsum = 0;
  for (int i = 0; i < ns; i++) {
    Solve* v = act[i];
    v->pos = i;
    gf_problem p = v->p;
    KrySlot& k = hk[i];
    k.n = v->n;
    k.m = v->m;
    k.N = (long long)v->N;
    k.krowp = p->krowp;
    k.kcols = p->kcols;
    k.kvals = p->kvals;
    k.V = p->kw.V;
    k.w = p->kw.w;
    k.hc = p->kw.hc;
    k.rho = p->kw.rho;
    k.st = p->kw.state;
    k.co = p->kw.co;
    k.dd = p->kw.dd;
    k.partial = p->kw.partial;
    k.Z = p->kw.Z;
    if (fp64()) {
      k.x = p->kw.z;
      k.b = p->work64;
      k.tmp = p->work64b;
    } else {
      k.x = p->t32c;
      k.b = p->t32a;
      k.tmp = p->t32b;
    }
    voff[i + 1] = voff[i] + (long long)v->N;
    roff[i + 1] = roff[i] + v->n;
    maxm = std::max(maxm, v->m);
    gsum += std::max(1, v->iv.grid);
  }
  // chunk lists merged level by level
  auto merge = [&](bool lower, std::vector<int2>& out) {
    out.clear();
    std::vector<size_t> cur(ns, 0);
    for (int lev = 0;; lev++) {
      bool any = false;
      for (int i = 0; i < ns; i++) {
        const std::vector<int>& cl =
            lower ? *act[i]->iv.lchunk_level : *act[i]->iv.uchunk_level;
        while (cur[i] < cl.size() && cl[cur[i]] == lev) {
          out.push_back(make_int2(i, (int)cur[i]));
          cur[i]++;
        }
        any |= cur[i] < cl.size();
      }
      if (!any) break;
    }
  };
  merge(true, lm);
  merge(false, um);
  nl = (int)lm.size();
  nu = (int)um.size();
  ntot = voff[ns];
  nrows = roff[ns];
  grid = std::max(1, std::min(gsum, std::max(1, waves_per_cu * ctx->num_cus / 4)));
  auto ensure = [&](int2** d, size_t& cap, size_t need) -> int {
    if (need <= cap && *d) return GF_OK;
    dfree(ctx, *d);
    *d = nullptr;
    cap = std::max<size_t>(need, 1024);
    return dmalloc(ctx, (void**)d, sizeof(int2) * cap);
  };
  GF_TRY(ensure(&d_lm, lm_cap, lm.size()));
  GF_TRY(ensure(&d_um, um_cap, um.size()));
  std::vector<unsigned char> sw;
  if (fp64())
    fill_tables<double>(sw);
  else
    fill_tables<float>(sw);
  std::vector<SlotState> hs(ns);
  for (int i = 0; i < ns; i++) hs[i] = act[i]->hs;
  if (ns) {
    GF_HIP(ctx, hipMemcpyAsync(d_kry, hk.data(), sizeof(KrySlot) * ns,
                               hipMemcpyHostToDevice, s));
    GF_HIP(ctx, hipMemcpyAsync(d_sweep, sw.data(), sw.size(),
                               hipMemcpyHostToDevice, s));
    GF_HIP(ctx, hipMemcpyAsync(d_state, hs.data(), sizeof(SlotState) * ns,
                               hipMemcpyHostToDevice, s));
    GF_HIP(ctx, hipMemcpyAsync(d_voff, voff.data(), sizeof(long long) * (ns + 1),
                               hipMemcpyHostToDevice, s));
    GF_HIP(ctx, hipMemcpyAsync(d_roff, roff.data(), sizeof(long long) * (ns + 1),
                               hipMemcpyHostToDevice, s));
    GF_HIP(ctx, hipMemcpyAsync(d_lm, lm.data(), sizeof(int2) * lm.size(),
                               hipMemcpyHostToDevice, s));
    GF_HIP(ctx, hipMemcpyAsync(d_um, um.data(), sizeof(int2) * um.size(),
                               hipMemcpyHostToDevice, s));
  }
  GF_HIP(ctx, hipStreamSynchronize(s));
  if (gexec) {
    hipGraphExecDestroy(gexec);
    gexec = nullptr;
  }
  if (graph) {
    hipGraphDestroy(graph);
    graph = nullptr;
  }
  if (ns && use_graph) {
    GF_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int st = issue_steps();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    GF_TRY(st);
    GF_TRY(check_hip(ctx, e, "hipStreamEndCapture"));
    graph = g;
    GF_HIP(ctx, hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
  }
  dirty = false;
  return GF_OK;
}

// `steps` Arnoldi steps of every solve (captured into the graph)
int BatchEngine::issue_steps() {
  hipStream_t s = ctx->stream;
  const int ns = (int)act.size();
  const unsigned long long sent = sweep_sentinel_bits(fp64());
  const dim3 g1(div_up((size_t)ntot, 256)), g1r(div_up((size_t)nrows, 40));
  const dim3 gdot(kBatchSlices, div_up((size_t)maxm + 1, kDotChunk), ns);
  const dim3 gfin(2 * maxm + 3, ns);
  static_assert(kBatchSlices == kDotBlocks, "partials sized for kDotBlocks");
  const size_t shm = sizeof(double) * (4 * (size_t)maxm + 2);
  for (int k = 0; k < steps; k++) {
    hipLaunchKernelGGL(batch_begin, dim3(div_up(ns, 64)), dim3(64), 0, s, ns,
                       d_state);
    if (fp64()) {
      hipLaunchKernelGGL(batch_gather<double>, g1, dim3(256), 0, s, ns, ntot,
                         d_voff, d_kry, d_state, sent);
      GF_TRY(launch_batch_sweeps<double>(
          s, (const SweepSlot<double>*)d_sweep, d_lm, nl, d_um, nu, grid,
          d_counters, d_err));
      hipLaunchKernelGGL(batch_spmv<double>, g1r, dim3(256), 0, s, ns, nrows,
                         d_roff, d_kry, d_state);
    } else {
      hipLaunchKernelGGL(batch_gather<float>, g1, dim3(256), 0, s, ns, ntot,
                         d_voff, d_kry, d_state, sent);
      GF_TRY(launch_batch_sweeps<float>(
          s, (const SweepSlot<float>*)d_sweep, d_lm, nl, d_um, nu, grid,
          d_counters, d_err));
      hipLaunchKernelGGL(batch_spmv<float>, g1r, dim3(256), 0, s, ns, nrows,
                         d_roff, d_kry, d_state);
    }
    hipLaunchKernelGGL(batch_dot2_partial, gdot, dim3(kDotThreads), 0, s,
                       d_kry, d_state);
    hipLaunchKernelGGL(batch_dot_final, gfin, dim3(kDotThreads), 0, s, d_kry,
                       d_state);
    hipLaunchKernelGGL(batch_coeffs, dim3(ns), dim3(256), shm, s, d_kry,
                       d_state);
    if (fp64())
      hipLaunchKernelGGL(batch_update2<double>, g1, dim3(256), 0, s, ns, ntot,
                         d_voff, d_kry, d_state);
    else
      hipLaunchKernelGGL(batch_update2<float>, g1, dim3(256), 0, s, ns, ntot,
                         d_voff, d_kry, d_state);
    hipLaunchKernelGGL(batch_end, dim3(div_up(ns, 64)), dim3(64), 0, s, ns,
                       d_state);
  }
  return check_launch(ctx, "batched Arnoldi step");
}

// rhs = permuted loads with BCs, x = 0, ||b|| into the solve's pinned
// slot; then the first cycle's residual (no synchronisation)
int BatchEngine::start(Solve* v) {
  gf_problem p = v->p;
  hipStream_t s = ctx->stream;
  v->n = p->nnodes;
  v->N = 6 * (size_t)p->nnodes;
  v->m = p->opt.restart;
  GF_TRY(ilu_view(p->ilu, &v->iv));
  GF_HIP(ctx, hipEventCreate(&v->ev0));
  GF_HIP(ctx, hipEventCreate(&v->ev1));
  GF_HIP(ctx, hipEventRecord(v->ev0, s));
  GF_HIP(ctx, hipMemcpyAsync(p->loads, v->job->loads, sizeof(double) * v->N,
                             hipMemcpyHostToDevice, s));
  GF_TRY(gf_bc_apply_vec(ctx, p->nbcs, p->bcs, nullptr, p->loads));
  GF_TRY(gf_permute_vec(ctx, p->nnodes, p->perm, p->loads, p->rhs));
  GF_HIP(ctx, hipMemsetAsync(p->sol, 0, sizeof(double) * v->N, s));
  // solo sweep table and lists for the cycle ends
  GF_TRY(dmalloc(ctx, (void**)&v->d_one, sizeof(int)));
  hipLaunchKernelGGL(fill_one_int, dim3(1), dim3(64), 0, s, v->d_one, 1);
  GF_TRY(dmalloc(ctx, &v->d_solo, fp64() ? sizeof(SweepSlot<double>)
                                         : sizeof(SweepSlot<float>)));
  v->h_solo.resize(fp64() ? sizeof(SweepSlot<double>)
                          : sizeof(SweepSlot<float>));
  if (fp64())
    *reinterpret_cast<SweepSlot<double>*>(v->h_solo.data()) =
        sweep_entry<double>(v, v->d_one);
  else
    *reinterpret_cast<SweepSlot<float>*>(v->h_solo.data()) =
        sweep_entry<float>(v, v->d_one);
  v->nsl = (int)v->iv.lchunk_level->size();
  v->nsu = (int)v->iv.uchunk_level->size();
  v->h_sl.resize(v->nsl);
  v->h_su.resize(v->nsu);
  for (int c = 0; c < v->nsl; c++) v->h_sl[c] = make_int2(0, c);
  for (int c = 0; c < v->nsu; c++) v->h_su[c] = make_int2(0, c);
  GF_TRY(dmalloc(ctx, (void**)&v->d_sl, sizeof(int2) * std::max(1, v->nsl)));
  GF_TRY(dmalloc(ctx, (void**)&v->d_su, sizeof(int2) * std::max(1, v->nsu)));
  GF_HIP(ctx, hipMemcpyAsync(v->d_solo, v->h_solo.data(), v->h_solo.size(),
                             hipMemcpyHostToDevice, s));
  GF_HIP(ctx, hipMemcpyAsync(v->d_sl, v->h_sl.data(), sizeof(int2) * v->nsl,
                             hipMemcpyHostToDevice, s));
  GF_HIP(ctx, hipMemcpyAsync(v->d_su, v->h_su.data(), sizeof(int2) * v->nsu,
                             hipMemcpyHostToDevice, s));
  KrylovWork& W = p->kw;
  hipLaunchKernelGGL(multi_dot_partial, dim3(kDotBlocks, 1), dim3(kDotThreads),
                     0, s, v->N, 1, p->rhs, v->N, p->rhs, W.partial,
                     (const int*)nullptr);
  hipLaunchKernelGGL(multi_dot_final, dim3(1), dim3(kDotThreads), 0, s,
                     (int)kDotBlocks, 1, W.partial, W.scal, 2,
                     (const int*)nullptr);
  GF_HIP(ctx, hipMemcpyAsync(h_norm + 2 * v->hidx + 1, W.scal, sizeof(double),
                             hipMemcpyDeviceToHost, s));
  return residual_issue(v);
}

// r = b - K x (compensated) and ||r|| into the solve's pinned slot
int BatchEngine::residual_issue(Solve* v) {
  gf_problem p = v->p;
  hipStream_t s = ctx->stream;
  KrylovWork& W = p->kw;
  GF_TRY(gf_bsr_resid_comp_f64(ctx, p->nnodes, p->krowp, p->kcols, p->kvals,
                               p->sol, p->rhs, W.r));
  hipLaunchKernelGGL(multi_dot_partial, dim3(kDotBlocks, 1), dim3(kDotThreads),
                     0, s, v->N, 1, W.r, v->N, W.r, W.partial,
                     (const int*)nullptr);
  hipLaunchKernelGGL(multi_dot_final, dim3(1), dim3(kDotThreads), 0, s,
                     (int)kDotBlocks, 1, W.partial, W.nrm, 2,
                     (const int*)nullptr);
  GF_HIP(ctx, hipMemcpyAsync(h_norm + 2 * v->hidx, W.nrm, sizeof(double),
                             hipMemcpyDeviceToHost, s));
  return check_launch(ctx, "batched residual");
}

// After a synchronisation: the stop tests on the residual read back, then
// V_0 = r / beta and the cycle's state
int BatchEngine::cycle_decide(Solve* v) {
  gf_problem p = v->p;
  hipStream_t s = ctx->stream;
  KrylovWork& W = p->kw;
  const gf_options& o = p->opt;
  if (v->outer == 0 && v->hist.empty()) {
    const double bn = h_norm[2 * v->hidx + 1];
    v->bnorm = bn > 0.0 ? bn : 1.0;
  }
  const double beta = h_norm[2 * v->hidx];
  v->true_rel = beta / v->bnorm;
  if (!std::isfinite(beta)) {
    v->status = set_error(ctx, GF_ERR_BREAKDOWN, "GMRES: non-finite residual");
    v->finished = true;
    return GF_OK;
  }
  if (v->true_rel <= o.rtol) {
    v->status = GF_OK;
    v->finished = true;
    return GF_OK;
  }
  if (v->it >= o.max_iters) {
    v->finished = true;
    return GF_OK;
  }
  v->hist.push_back(v->true_rel);
  const int sc = o.stall_cycles;
  if (sc > 0 && (int)v->hist.size() > sc &&
      v->true_rel > 0.9 * v->hist[v->hist.size() - 1 - sc]) {
    v->stalled = true;
    v->finished = true;
    return GF_OK;
  }
  v->outer++;
  hipLaunchKernelGGL(scale_by_inv, grid1(v->N), dim3(256), 0, s, v->N, W.r,
                     W.nrm, W.V);
  hipLaunchKernelGGL(init_g, dim3(1), dim3(64 * div_up(v->m + 1, 64)), 0, s,
                     v->m, W.nrm, W.state);
  v->hs = SlotState{};
  v->hs.j = 0;
  v->hs.limit = std::min(v->m, o.max_iters - v->it);
  v->hs.stop = 0;
  v->hs.mode = 0;
  v->hs.tol_abs = kCycleTarget * o.rtol * v->bnorm;
  if (v->pos >= 0 && !dirty)
    GF_HIP(ctx, hipMemcpyAsync(&d_state[v->pos], &v->hs, sizeof(SlotState),
                               hipMemcpyHostToDevice, s));
  return check_launch(ctx, "batched cycle start");
}

// x += M^-1 V y over the cycle's iterations (Z y when the preconditioned
// basis is kept), then the next residual
int BatchEngine::cycle_end(Solve* v) {
  gf_problem p = v->p;
  hipStream_t s = ctx->stream;
  KrylovWork& W = p->kw;
  const int j = v->hs.stop;
  v->it += j;
  hipLaunchKernelGGL(solve_upper, dim3(1), dim3(1024),
                     sizeof(double) * (size_t)(j > 0 ? j : 1), s, j, v->m,
                     W.state, W.y);
  if (W.Z) {
    hipLaunchKernelGGL(lincomb, grid1(v->N), dim3(256), 0, s, v->N, j, W.Z,
                       v->N, W.y, W.tmp);
    hipLaunchKernelGGL(axpy_from<double>, grid1(v->N), dim3(256), 0, s, v->N,
                       (const double*)W.tmp, p->sol);
    GF_TRY(check_launch(ctx, "batched cycle end"));
    return residual_issue(v);
  }
  hipLaunchKernelGGL(lincomb, grid1(v->N), dim3(256), 0, s, v->N, j, W.V,
                     v->N, W.y, W.tmp);
  const unsigned long long sent = sweep_sentinel_bits(fp64());
  if (fp64()) {
    hipLaunchKernelGGL(gather_one<double>, grid1(v->N), dim3(256), 0, s, v->N,
                       W.tmp, p->work64, p->work64b, sent);
    GF_TRY(launch_batch_sweeps<double>(s, (const SweepSlot<double>*)v->d_solo,
                                       v->d_sl, v->nsl, v->d_su, v->nsu,
                                       std::max(1, v->iv.grid), d_counters,
                                       d_err));
    hipLaunchKernelGGL(axpy_from<double>, grid1(v->N), dim3(256), 0, s, v->N,
                       (const double*)p->kw.z, p->sol);
  } else {
    hipLaunchKernelGGL(gather_one<float>, grid1(v->N), dim3(256), 0, s, v->N,
                       W.tmp, p->t32a, p->t32b, sent);
    GF_TRY(launch_batch_sweeps<float>(s, (const SweepSlot<float>*)v->d_solo,
                                      v->d_sl, v->nsl, v->d_su, v->nsu,
                                      std::max(1, v->iv.grid), d_counters,
                                      d_err));
    hipLaunchKernelGGL(axpy_from<float>, grid1(v->N), dim3(256), 0, s, v->N,
                       (const float*)p->t32c, p->sol);
  }
  GF_TRY(check_launch(ctx, "batched cycle end"));
  return residual_issue(v);
}

// x = unpermuted solution to the host, result to the waiting worker (which
// is always released, whatever fails here)
int BatchEngine::finish(Solve* v) {
  gf_problem p = v->p;
  hipStream_t s = ctx->stream;
  int st = gf_permute_vec(ctx, p->nnodes, p->iperm, p->sol, p->work64);
  if (st == GF_OK)
    st = check_hip(ctx, hipMemcpyAsync(v->job->x, p->work64,
                                       sizeof(double) * v->N,
                                       hipMemcpyDeviceToHost, s),
                   "result copy");
  if (st == GF_OK && v->ev1)
    st = check_hip(ctx, hipEventRecord(v->ev1, s), "event");
  if (st == GF_OK) st = check_hip(ctx, hipStreamSynchronize(s), "sync");
  if (st != GF_OK && v->status == GF_OK) v->status = st;
  float ms = 0.f;
  if (v->ev0 && v->ev1) hipEventElapsedTime(&ms, v->ev0, v->ev1);
  if (v->status == GF_ERR_NO_CONVERGENCE)
    set_error(ctx, v->status, "GMRES: %s after %d iterations, rel. residual "
              "%.3e > %.1e", v->stalled ? "stagnated" : "iteration limit",
              v->it, v->true_rel, p->opt.rtol);
  BatchJob* job = v->job;
  job->info.status = v->status;
  job->info.iterations = v->it;
  job->info.outer = v->outer;
  job->info.rel_resid = v->true_rel;
  job->info.t_assemble = p->t_assemble;
  job->info.t_factor = p->t_factor;
  job->info.t_solve = 1e-3 * ms;
  job->status = v->status;
  dfree(ctx, v->d_one);
  dfree(ctx, v->d_solo);
  dfree(ctx, v->d_sl);
  dfree(ctx, v->d_su);
  if (v->ev0) hipEventDestroy(v->ev0);
  if (v->ev1) hipEventDestroy(v->ev1);
  {
    std::lock_guard<std::mutex> lk(mu);
    job->done = true;
  }
  done_cv.notify_all();
  return GF_OK;
}

void BatchEngine::loop() {
  hipSetDevice(device);
  auto release = [&](Solve* v) {
    finish(v);
    if (v->hidx >= 0) hidx_used[v->hidx] = 0;
    delete v;
  };
  auto fail_all = [&](int st) {
    for (Solve* v : act) {
      v->status = st;
      release(v);
    }
    act.clear();
    dirty = true;
    // a later batch starts with a clear wait-expired flag
    hipLaunchKernelGGL(fill_one_int, dim3(1), dim3(64), 0, ctx->stream, d_err,
                       0);
    hipStreamSynchronize(ctx->stream);
  };
  for (;;) {
    std::vector<BatchJob*> fresh;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] {
        return stopping || !incoming.empty() || !act.empty();
      });
      if (stopping && incoming.empty() && act.empty()) break;
      while (!incoming.empty() && (int)(act.size() + fresh.size()) < max_slots) {
        fresh.push_back(incoming.front());
        incoming.pop_front();
      }
    }
    // new solves and finished cycles: their device work is issued first,
    // then one synchronisation serves all of their readbacks
    std::vector<Solve*> pending, started;
    for (BatchJob* j : fresh) {
      Solve* v = new Solve;
      v->job = j;
      v->p = j->p;
      for (int h = 0; h < max_slots; h++)
        if (!hidx_used[h]) {
          hidx_used[h] = 1;
          v->hidx = h;
          break;
        }
      const int st = start(v);
      if (st != GF_OK) {
        v->status = st;
        release(v);
        continue;
      }
      pending.push_back(v);
      started.push_back(v);
    }
    int st = GF_OK;
    for (Solve* v : act)
      if (v->hs.stop > 0) {
        const int cst = cycle_end(v);
        if (cst != GF_OK) {
          v->status = cst;
          v->finished = true;
        } else {
          pending.push_back(v);
        }
      }
    if (!pending.empty())
      st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (st == GF_OK)
      for (Solve* v : pending) {
        const int cst = cycle_decide(v);
        if (cst != GF_OK) {
          v->status = cst;
          v->finished = true;
        }
      }
    if (st != GF_OK) {
      for (Solve* v : started) act.push_back(v);
      fail_all(st);
      continue;
    }
    // retire finished solves, admit the fresh ones that still run
    bool changed = false;
    for (Solve*& v : act)
      if (v->finished) {
        release(v);
        v = nullptr;
        changed = true;
      }
    act.erase(std::remove(act.begin(), act.end(), nullptr), act.end());
    for (Solve* v : started) {
      if (v->finished) {
        release(v);
      } else {
        act.push_back(v);
        changed = true;
      }
    }
    if (changed) dirty = true;
    if (act.empty()) continue;
    if (dirty) st = rebuild();
    if (st == GF_OK) {
      if (use_graph)
        st = check_hip(ctx, hipGraphLaunch(gexec, ctx->stream), "graph launch");
      else
        st = issue_steps();
    }
    const int ns = (int)act.size();
    int* herr = reinterpret_cast<int*>(h_scal + 2);
    if (st == GF_OK)
      st = check_hip(ctx, hipMemcpyAsync(h_state, d_state,
                                         sizeof(SlotState) * ns,
                                         hipMemcpyDeviceToHost, ctx->stream),
                     "state copy");
    if (st == GF_OK)
      st = check_hip(ctx, hipMemcpyAsync(herr, d_err, sizeof(int),
                                         hipMemcpyDeviceToHost, ctx->stream),
                     "err copy");
    if (st == GF_OK) st = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (st == GF_OK && *herr)
      st = set_error(ctx, GF_ERR_TIMEOUT,
                     "triangular sweep: a dependency wait expired");
    if (st != GF_OK) {
      fail_all(st);
      continue;
    }
    for (int i = 0; i < ns; i++) act[i]->hs = h_state[i];
  }
}

int batch_engine_create(int device, int precision, int max_slots,
                        BatchEngine** out) {
  if (!out || (precision != 0 && precision != 2) || max_slots < 1 ||
      max_slots > 64)
    return GF_ERR_INVALID;
  auto* e = new BatchEngine;
  e->device = device;
  e->prec = precision;
  e->max_slots = max_slots;
  if (const char* g = getenv("GPUFEM_BATCH_GRAPH")) e->use_graph = atoi(g) != 0;
  if (const char* g = getenv("GPUFEM_BATCH_STEPS"))
    e->steps = std::max(1, std::min(64, atoi(g)));
  if (const char* g = getenv("GPUFEM_BATCH_WAVES_PER_CU"))
    e->waves_per_cu = std::max(1, std::min(16, atoi(g)));
  int st = context_create(device, true, &e->ctx);
  gf_context c = e->ctx;
  const size_t tslot = precision == 0 ? sizeof(SweepSlot<double>)
                                      : sizeof(SweepSlot<float>);
  if (!st) st = dmalloc(c, (void**)&e->d_kry, sizeof(KrySlot) * max_slots);
  if (!st) st = dmalloc(c, &e->d_sweep, tslot * max_slots);
  if (!st) st = dmalloc(c, (void**)&e->d_state, sizeof(SlotState) * max_slots);
  if (!st) st = dmalloc(c, (void**)&e->d_voff, sizeof(long long) * (max_slots + 1));
  if (!st) st = dmalloc(c, (void**)&e->d_roff, sizeof(long long) * (max_slots + 1));
  if (!st) st = dmalloc(c, (void**)&e->d_counters, 2 * sizeof(unsigned long long));
  if (!st) st = dmalloc(c, (void**)&e->d_err, sizeof(int));
  if (!st)
    st = check_hip(c, hipHostMalloc((void**)&e->h_state,
                                    sizeof(SlotState) * max_slots),
                   "hipHostMalloc");
  if (!st)
    st = check_hip(c, hipHostMalloc((void**)&e->h_scal, sizeof(double) * 4),
                   "hipHostMalloc");
  if (!st)
    st = check_hip(c, hipHostMalloc((void**)&e->h_norm,
                                    sizeof(double) * 2 * max_slots),
                   "hipHostMalloc");
  e->hidx_used.assign(max_slots, 0);
  if (!st) {
    hipLaunchKernelGGL(fill_one_int, dim3(1), dim3(64), 0, c->stream, e->d_err,
                       0);
    st = check_hip(c, hipStreamSynchronize(c->stream), "engine init");
  }
  if (st) {
    batch_engine_destroy(e);
    return st;
  }
  e->th = std::thread([e] { e->loop(); });
  *out = e;
  return GF_OK;
}

void batch_engine_destroy(BatchEngine* e) {
  if (!e) return;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    e->stopping = true;
  }
  e->cv.notify_all();
  if (e->th.joinable()) e->th.join();
  gf_context c = e->ctx;
  if (c) {
    if (e->gexec) hipGraphExecDestroy(e->gexec);
    if (e->graph) hipGraphDestroy(e->graph);
    dfree(c, e->d_kry);
    dfree(c, e->d_sweep);
    dfree(c, e->d_state);
    dfree(c, e->d_voff);
    dfree(c, e->d_roff);
    dfree(c, e->d_lm);
    dfree(c, e->d_um);
    dfree(c, e->d_counters);
    dfree(c, e->d_err);
    if (e->h_state) hipHostFree(e->h_state);
    if (e->h_scal) hipHostFree(e->h_scal);
    if (e->h_norm) hipHostFree(e->h_norm);
    gf_context_destroy(c);
  }
  delete e;
}

int batch_engine_solve(BatchEngine* e, BatchJob* job) {
  if (!e || !job || !job->p || !job->loads || !job->x) return GF_ERR_INVALID;
  if ((e->prec == 0) != (job->p->opt.precision == 0) ||
      job->p->opt.precision == 1)
    return GF_ERR_INVALID;
  std::unique_lock<std::mutex> lk(e->mu);
  job->done = false;
  e->incoming.push_back(job);
  e->cv.notify_all();
  e->done_cv.wait(lk, [&] { return job->done; });
  return job->status;
}

}  // namespace gf
