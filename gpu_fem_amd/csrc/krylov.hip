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
__global__ void __launch_bounds__(kDotThreads)
    multi_dot2_partial(size_t N, int nvec, const double* __restrict__ V,
                       size_t ldv, const double* __restrict__ x1,
                       const double* __restrict__ x2,
                       double* __restrict__ partial) {
  __shared__ double red[kDotThreads / 64][2 * kDotChunk + 1];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int i0 = blockIdx.y * kDotChunk;
  const int nout = 2 * nvec + 1;
  const size_t len = (N + gridDim.x - 1) / gridDim.x;
  const size_t t0 = (size_t)blockIdx.x * len;
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
    else if (c == 2 * kDotChunk && blockIdx.y == 0)
      o = 2 * nvec;
    if (o >= 0) partial[(size_t)blockIdx.x * nout + o] = v;
  }
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
__global__ void __launch_bounds__(256)
    dcgs2_coeffs(int j, int m, const double* __restrict__ d, double* Hc,
                 double* rho, double* st, double* co, double* out,
                 int finalize) {
  __shared__ double red[4];
  extern __shared__ double sh[];  // col[m + 2] | lcs[m] | lsn[m] | s[m]
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
      hc[i] = h;
      col[i] = h;
    }
    for (int i = tid; i < j - 1; i += 256) {
      lcs[i] = cs[i];
      lsn[i] = sn[i];
    }
    if (tid == 0) {
      hc[j] = r * alpha;
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
      cs[jc] = cj;
      sn[jc] = sj;
      col[jc] = cj * a + sj * c;
      col[jc + 1] = 0.0;
      g[jc + 1] = -sj * g[jc];
      g[jc] = cj * g[jc];
      out[0] = fabs(g[jc + 1]);
      out[1] = r * alpha;
    }
    __syncthreads();
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

// One pass over V[0..j-1]: q_j = (V_j - sum s_i V_i) / alpha (final) and
// V_{j+1} = (w - sum_{i<j} c_i V_i - c_j q_j) / (alpha rho)
__global__ void update2(size_t N, int j, int m, double* __restrict__ V,
                        size_t ldv, const double* __restrict__ w,
                        const double* __restrict__ co) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N) return;
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

inline dim3 grid1(size_t N) { return dim3(div_up(N, 256)); }

}  // namespace

int KrylovWork::init(gf_context c, size_t N_, int m_, bool mixed) {
  ctx = c;
  N = N_;
  m = m_;
  auto dm = [&](void** p, size_t bytes) {
    return dmalloc(ctx, p, bytes ? bytes : 8);
  };
  GF_TRY(dm((void**)&V, sizeof(double) * N * (m + 1)));
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

int gmres_solve(KrylovWork& W, const Operator& op, const double* b, double* x,
                int max_iters, int stall_cycles, double rtol, int prec,
                GmresResult* res) {
  const bool ir = prec == 1;
  gf_context ctx = W.ctx;
  hipStream_t s = ctx->stream;
  const size_t N = W.N;
  const int m = W.m;
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
      while (!stop) {
        if (issued <= limit) {
          const int js = issued;
          double* vj = W.V + (size_t)js * N;
          if (js < limit) {
            GF_TRY(op.apply_prec_a(vj, W.w, prec));  // w = A M^-1 v_j
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
        // iteration count j + 1 is reported by step j + 1
        if (issued > j + 1 && (issued - (j + 1) >= 2 || issued > limit)) {
          const int r = j + 1;
          GF_HIP(ctx, hipEventSynchronize(W.ev[r % 2]));
          const double est = W.host_scal[2 * (r % 2)];
          const double hn = W.host_scal[2 * (r % 2) + 1];
          last_est = est;
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
    hipLaunchKernelGGL(lincomb, grid1(N), dim3(256), 0, s, N, j, W.V, N, W.y,
                       W.tmp);
    GF_TRY(op.apply_prec(W.tmp, W.z, prec));
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

int Operator::apply_prec_a(const double* in, double* out, int prec) const {
  const size_t N = 6 * (size_t)n;
  if (prec == 0) {
    GF_TRY(gf_bsr_ilu_apply_f64(ilu, mlu, in, work64, work64b));
    return gf_bsr_spmv_f64(ctx, n, krowp, kcols, kvals, work64b, out);
  }
  hipLaunchKernelGGL(to_f32, grid1(N), dim3(256), 0, ctx->stream, N, in, t32a);
  GF_TRY(gf_bsr_ilu_apply_f32(ilu, mlu32, t32a, t32b, t32c));
  if (prec == 2) {
    hipLaunchKernelGGL(to_f64, grid1(N), dim3(256), 0, ctx->stream, N, t32c,
                       work64b);
    GF_TRY(check_launch(ctx, "apply_prec_a fp32 ilu"));
    return gf_bsr_spmv_f64(ctx, n, krowp, kcols, kvals, work64b, out);
  }
  GF_TRY(gf_bsr_spmv_f32(ctx, n, krowp, kcols, kvals32, t32c, t32a));
  hipLaunchKernelGGL(to_f64, grid1(N), dim3(256), 0, ctx->stream, N, t32a, out);
  return check_launch(ctx, "apply_prec_a mixed");
}

}  // namespace gf
