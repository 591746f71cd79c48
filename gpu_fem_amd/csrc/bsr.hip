// BSR (6x6 blocks) kernels: boundary conditions, value copy into the
// preconditioner pattern, vector permutation, SpMV, precision conversion.
//
//   bc_rows_kernel / bc_cols_kernel   bsr_mat.cuh:2-45 / :48-104
//   bc_vec_kernel                     vec.cuh:4-13
//   copy_pattern_kernel               bsr_mat.cuh:107-140
//   permute_kernel                    vec.cuh:16-34
//   spmv_kernel                       cusparseDbsrmv (gmres.cu:588-599)
#include <mutex>
#include <vector>

#include "common.h"

namespace gf {
namespace {

__global__ void bc_rows_kernel(int nbcs, const int* __restrict__ bcs,
                               const int* __restrict__ rowp,
                               const int* __restrict__ cols,
                               const int* __restrict__ perm, double* vals) {
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nbcs) return;
  const int dof = bcs[ib];
  const int brow = perm[dof / 6], ir = dof % 6, grow = 6 * brow + ir;
  for (int k = rowp[brow]; k < rowp[brow + 1]; k++) {
    const int gc0 = 6 * cols[k];
    double* v = vals + 36 * (size_t)k + 6 * ir;
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = (gc0 + c == grow) ? 1.0 : 0.0;
  }
}

__global__ void bc_cols_kernel(int nbcs, const int* __restrict__ bcs,
                               const int* __restrict__ trowp,
                               const int* __restrict__ tcols,
                               const int* __restrict__ tmap,
                               const int* __restrict__ perm, double* vals) {
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nbcs) return;
  const int dof = bcs[ib];
  const int bcol = perm[dof / 6], ic = dof % 6, gcol = 6 * bcol + ic;
  for (int t = trowp[bcol]; t < trowp[bcol + 1]; t++) {
    const int gr0 = 6 * tcols[t];
    double* v = vals + 36 * (size_t)tmap[t] + ic;
#pragma unroll
    for (int r = 0; r < 6; r++) v[6 * r] = (gr0 + r == gcol) ? 1.0 : 0.0;
  }
}

__global__ void bc_vec_kernel(int nbcs, const int* __restrict__ bcs,
                              const int* __restrict__ perm, double* vec) {
  const int ib = blockIdx.x * blockDim.x + threadIdx.x;
  if (ib >= nbcs) return;
  const int dof = bcs[ib];
  const size_t idx = perm ? 6 * (size_t)perm[dof / 6] + dof % 6 : (size_t)dof;
  vec[idx] = 0.0;
}

// One wave per target row, in chunks of 64 target blocks: lane l finds the
// source block of target block l of the chunk (binary search in the short
// source row, once per block), the indices go through LDS, and the chunk's
// 36 * 64 values are written with consecutive lanes on consecutive values
// (most target blocks are fill: zeros).
__global__ void __launch_bounds__(256)
    copy_pattern_kernel(int nnodes, const int* __restrict__ rowp,
                        const int* __restrict__ cols,
                        const double* __restrict__ vals,
                        const int* __restrict__ trowp,
                        const int* __restrict__ tcols, double* tvals) {
  __shared__ int s_src[4][64];
  const int w = threadIdx.x / 64;
  const int row = blockIdx.x * 4 + w;
  const int lane = threadIdx.x % 64;
  if (row >= nnodes) return;
  const int t0 = trowp[row], t1 = trowp[row + 1];
  const int s0 = rowp[row], s1 = rowp[row + 1];
  for (int c0 = t0; c0 < t1; c0 += 64) {
    const int nb = min(64, t1 - c0);
    int src = -1;
    if (lane < nb) {
      const int col = tcols[c0 + lane];
      int lo = s0, hi = s1;  // lower_bound of col in the source row
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cols[mid] < col)
          lo = mid + 1;
        else
          hi = mid;
      }
      src = (lo < s1 && cols[lo] == col) ? lo : -1;
    }
    s_src[w][lane] = src;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double* dst = tvals + 36 * (size_t)c0;
    const int nval = 36 * nb;
#pragma unroll 4
    for (int v = lane; v < nval; v += 64) {
      const int sb = s_src[w][v / 36];
      dst[v] = sb >= 0 ? vals[36 * (size_t)sb + v % 36] : 0.0;
    }
    // the next chunk's indices overwrite s_src
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

__global__ void permute_kernel(int nnodes, const int* __restrict__ perm,
                               const double* __restrict__ in, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * (size_t)nnodes) return;
  out[6 * (size_t)perm[i / 6] + i % 6] = in[i];
}

// six lanes per block row (one per scalar row r); ten rows per wave
template <typename T>
__global__ void __launch_bounds__(256)
    spmv_kernel(int nnodes, const int* __restrict__ rowp,
                const int* __restrict__ cols, const T* __restrict__ vals,
                const T* __restrict__ x, T* __restrict__ y) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane >= 60) return;
  const int row = (blockIdx.x * (blockDim.x / 64) + wave) * 10 + lane / 6;
  const int r = lane % 6;
  if (row >= nnodes) return;
  T acc = 0;
  for (int k = rowp[row]; k < rowp[row + 1]; k++) {
    const T* a = vals + 36 * (size_t)k + 6 * r;
    const T* xb = x + 6 * (size_t)cols[k];
#pragma unroll
    for (int c = 0; c < 6; c++) acc += a[c] * xb[c];
  }
  y[6 * (size_t)row + r] = acc;
}

// r = A x - b; per workgroup: the entry of largest magnitude (first index on
// ties) and the sum of squares
__global__ void __launch_bounds__(256)
    resid_partial_kernel(int nnodes, const int* __restrict__ rowp,
                         const int* __restrict__ cols,
                         const double* __restrict__ vals,
                         const double* __restrict__ x,
                         const double* __restrict__ b, double* pmax,
                         long long* pidx, double* pss) {
  __shared__ double smax[256], sss[256];
  __shared__ long long sidx[256];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int row = (blockIdx.x * (blockDim.x / 64) + wave) * 10 + lane / 6;
  const int r = lane % 6;
  double v = 0.0;
  long long idx = 0x7fffffffffffffffll;
  if (lane < 60 && row < nnodes) {
    double acc = 0.0;
    for (int k = rowp[row]; k < rowp[row + 1]; k++) {
      const double* a = vals + 36 * (size_t)k + 6 * r;
      const double* xb = x + 6 * (size_t)cols[k];
#pragma unroll
      for (int c = 0; c < 6; c++) acc += a[c] * xb[c];
    }
    idx = 6 * (long long)row + r;
    v = acc - b[idx];
  }
  smax[threadIdx.x] = v;
  sidx[threadIdx.x] = idx;
  sss[threadIdx.x] = v * v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int o = threadIdx.x + s;
      const double a0 = fabs(smax[threadIdx.x]), a1 = fabs(smax[o]);
      if (a1 > a0 || (a1 == a0 && sidx[o] < sidx[threadIdx.x])) {
        smax[threadIdx.x] = smax[o];
        sidx[threadIdx.x] = sidx[o];
      }
      sss[threadIdx.x] += sss[o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pmax[blockIdx.x] = smax[0];
    pidx[blockIdx.x] = sidx[0];
    pss[blockIdx.x] = sss[0];
  }
}

// one workgroup folds the partials in a fixed order
__global__ void __launch_bounds__(256)
    resid_final_kernel(int nparts, const double* __restrict__ pmax,
                       const long long* __restrict__ pidx,
                       const double* __restrict__ pss, double* out) {
  __shared__ double smax[256], sss[256];
  __shared__ long long sidx[256];
  double m = 0.0, ss = 0.0;
  long long id = 0x7fffffffffffffffll;
  for (int i = threadIdx.x; i < nparts; i += 256) {
    const double a = fabs(pmax[i]), am = fabs(m);
    if (a > am || (a == am && pidx[i] < id)) {
      m = pmax[i];
      id = pidx[i];
    }
    ss += pss[i];
  }
  smax[threadIdx.x] = m;
  sidx[threadIdx.x] = id;
  sss[threadIdx.x] = ss;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      const int o = threadIdx.x + s;
      const double a0 = fabs(smax[threadIdx.x]), a1 = fabs(smax[o]);
      if (a1 > a0 || (a1 == a0 && sidx[o] < sidx[threadIdx.x])) {
        smax[threadIdx.x] = smax[o];
        sidx[threadIdx.x] = sidx[o];
      }
      sss[threadIdx.x] += sss[o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smax[0];
    out[1] = sqrt(sss[0]);
  }
}

// r = b - A x with compensated accumulation (Ogita, Rump and Oishi,
// "Accurate sum and dot product", SIAM J. Sci. Comput. 26(6), 2005: error-free
// TwoProduct via FMA and TwoSum): the result is as accurate as if computed in
// twice the working precision and then rounded.  On thin shells |A||x| is
// ~1e9 |b|, so the plain FP64 residual has a rounding floor near 1e-8
// relative; the outer loop of GMRES uses this one.
__global__ void __launch_bounds__(256)
    resid_comp_kernel(int nnodes, const int* __restrict__ rowp,
                      const int* __restrict__ cols,
                      const double* __restrict__ vals,
                      const double* __restrict__ x,
                      const double* __restrict__ b, double* __restrict__ r) {
#pragma clang fp contract(off)  // the error-free transformations need
                                // separately rounded products and sums
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane >= 60) return;
  const int row = (blockIdx.x * (blockDim.x / 64) + wave) * 10 + lane / 6;
  const int rr = lane % 6;
  if (row >= nnodes) return;
  double s = b[6 * (size_t)row + rr], c = 0.0;
  for (int k = rowp[row]; k < rowp[row + 1]; k++) {
    const double* a = vals + 36 * (size_t)k + 6 * rr;
    const double* xb = x + 6 * (size_t)cols[k];
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const double p = a[m] * xb[m];
      const double e = __builtin_fma(a[m], xb[m], -p);  // a x = p + e
      const double t = s - p;                           // TwoSum(s, -p)
      const double z = t - s;
      const double err = (s - (t - z)) + (-p - z);
      s = t;
      c += err - e;
    }
  }
  r[6 * (size_t)row + rr] = s + c;
}

__global__ void convert_kernel(size_t n, const double* __restrict__ in,
                               float* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)in[i];
}

__global__ void axpby_kernel(size_t n, double a, const double* __restrict__ x,
                             double b, double* __restrict__ y) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

}  // namespace
}  // namespace gf

// Scratch of gf_bsr_resid_f64_noalloc: the per-workgroup partial results
// (largest |r| entry, its index, sum of squares) of at most
// kResidScratchGroups workgroups of 40 block rows, i.e. meshes of up to
// 40 * 2^17 = 5,242,880 nodes; 3 MiB per device, held until the process
// ends.  Each device has its own copy at its own address, looked up once per
// device with hipGetSymbolAddress (the array's name is not a device address
// in host code).
constexpr int kResidScratchGroups = 1 << 17;
static __device__ __align__(256) double scratch[3 * kResidScratchGroups];

namespace {
std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device, null until looked up

// the device address of `scratch` on `device` (looked up on first use)
int scratch_on_device(gf_context ctx, double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr.empty()) {
    int ndev = 0;
    GF_HIP(ctx, hipGetDeviceCount(&ndev));
    g_scratch_ptr.assign(ndev > 0 ? ndev : 1, nullptr);
  }
  if (ctx->device < 0 || ctx->device >= (int)g_scratch_ptr.size())
    return gf::set_error(ctx, GF_ERR_DEVICE, "no such device");
  double*& p = g_scratch_ptr[ctx->device];
  if (!p) {
    int cur = 0;
    GF_HIP(ctx, hipGetDevice(&cur));
    GF_HIP(ctx, hipSetDevice(ctx->device));
    void* addr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
    hipSetDevice(cur);
    GF_TRY(gf::check_hip(ctx, e, "hipGetSymbolAddress(scratch)"));
    p = (double*)addr;
  }
  *out = p;
  return GF_OK;
}
}  // namespace

extern "C" {

int gf_bc_apply_mat(gf_context ctx, int nbcs, const int* bcs, const int* rowp,
                    const int* cols, const int* trowp, const int* tcols,
                    const int* tmap, const int* perm, double* vals) {
  if (!ctx || nbcs < 0 || (nbcs > 0 && (!bcs || !rowp || !cols || !trowp ||
                                        !tcols || !tmap || !perm || !vals)))
    return gf::set_error(ctx, GF_ERR_INVALID, "bc_apply_mat: bad args");
  if (nbcs == 0) return GF_OK;
  const unsigned grid = gf::div_up(nbcs, 64);
  hipLaunchKernelGGL(gf::bc_rows_kernel, dim3(grid), dim3(64), 0, ctx->stream,
                     nbcs, bcs, rowp, cols, perm, vals);
  GF_TRY(gf::check_launch(ctx, "bc_rows_kernel"));
  hipLaunchKernelGGL(gf::bc_cols_kernel, dim3(grid), dim3(64), 0, ctx->stream,
                     nbcs, bcs, trowp, tcols, tmap, perm, vals);
  return gf::check_launch(ctx, "bc_cols_kernel");
}

int gf_bc_apply_vec(gf_context ctx, int nbcs, const int* bcs, const int* perm,
                    double* vec) {
  if (!ctx || nbcs < 0 || (nbcs > 0 && (!bcs || !vec)))
    return gf::set_error(ctx, GF_ERR_INVALID, "bc_apply_vec: bad args");
  if (nbcs == 0) return GF_OK;
  hipLaunchKernelGGL(gf::bc_vec_kernel, dim3(gf::div_up(nbcs, 64)), dim3(64),
                     0, ctx->stream, nbcs, bcs, perm, vec);
  return gf::check_launch(ctx, "bc_vec_kernel");
}

int gf_bsr_copy_to_pattern(gf_context ctx, int nnodes, const int* rowp,
                           const int* cols, const double* vals,
                           const int* trowp, const int* tcols, double* tvals) {
  if (!ctx || nnodes <= 0 || !rowp || !cols || !vals || !trowp || !tcols ||
      !tvals)
    return gf::set_error(ctx, GF_ERR_INVALID, "copy_to_pattern: bad args");
  hipLaunchKernelGGL(gf::copy_pattern_kernel, dim3(gf::div_up(nnodes, 4)),
                     dim3(256), 0, ctx->stream, nnodes, rowp, cols, vals,
                     trowp, tcols, tvals);
  return gf::check_launch(ctx, "copy_pattern_kernel");
}

int gf_permute_vec(gf_context ctx, int nnodes, const int* perm,
                   const double* in, double* out) {
  if (!ctx || nnodes <= 0 || !perm || !in || !out || in == out)
    return gf::set_error(ctx, GF_ERR_INVALID, "permute_vec: bad args");
  hipLaunchKernelGGL(gf::permute_kernel, dim3(gf::div_up(6 * (size_t)nnodes, 256)),
                     dim3(256), 0, ctx->stream, nnodes, perm, in, out);
  return gf::check_launch(ctx, "permute_kernel");
}

int gf_bsr_spmv_f64(gf_context ctx, int nnodes, const int* rowp,
                    const int* cols, const double* vals, const double* x,
                    double* y) {
  if (!ctx || nnodes <= 0 || !rowp || !cols || !vals || !x || !y || x == y)
    return gf::set_error(ctx, GF_ERR_INVALID, "spmv_f64: bad args");
  hipLaunchKernelGGL(gf::spmv_kernel<double>, dim3(gf::div_up(nnodes, 40)),
                     dim3(256), 0, ctx->stream, nnodes, rowp, cols, vals, x, y);
  return gf::check_launch(ctx, "spmv_kernel<double>");
}

int gf_bsr_spmv_f32(gf_context ctx, int nnodes, const int* rowp,
                    const int* cols, const float* vals, const float* x,
                    float* y) {
  if (!ctx || nnodes <= 0 || !rowp || !cols || !vals || !x || !y || x == y)
    return gf::set_error(ctx, GF_ERR_INVALID, "spmv_f32: bad args");
  hipLaunchKernelGGL(gf::spmv_kernel<float>, dim3(gf::div_up(nnodes, 40)),
                     dim3(256), 0, ctx->stream, nnodes, rowp, cols, vals, x, y);
  return gf::check_launch(ctx, "spmv_kernel<float>");
}

int gf_bsr_resid_f64(gf_context ctx, int nnodes, const int* rowp,
                     const int* cols, const double* vals, const double* x,
                     const double* b, double* max_entry, double* norm2) {
  if (!ctx || nnodes <= 0 || !rowp || !cols || !vals || !x || !b)
    return gf::set_error(ctx, GF_ERR_INVALID, "bsr_resid: bad args");
  const int nblk = (int)gf::div_up(nnodes, 40);
  char* buf = nullptr;
  const size_t part = sizeof(double) * (size_t)nblk;
  GF_TRY(gf::dmalloc(ctx, (void**)&buf, 3 * part + 2 * sizeof(double)));
  double* pmax = (double*)buf;
  long long* pidx = (long long*)(buf + part);
  double* pss = (double*)(buf + 2 * part);
  double* out = (double*)(buf + 3 * part);
  hipLaunchKernelGGL(gf::resid_partial_kernel, dim3(nblk), dim3(256), 0,
                     ctx->stream, nnodes, rowp, cols, vals, x, b, pmax, pidx,
                     pss);
  hipLaunchKernelGGL(gf::resid_final_kernel, dim3(1), dim3(256), 0,
                     ctx->stream, nblk, pmax, pidx, pss, out);
  double h[2] = {0.0, 0.0};
  int st = gf::check_launch(ctx, "resid_kernel");
  if (st == GF_OK)
    st = gf::check_hip(ctx, hipMemcpyAsync(h, out, sizeof(h),
                                           hipMemcpyDeviceToHost, ctx->stream),
                       "resid copy");
  if (st == GF_OK)
    st = gf::check_hip(ctx, hipStreamSynchronize(ctx->stream), "resid sync");
  gf::dfree(ctx, buf);
  if (st != GF_OK) return st;
  if (max_entry) *max_entry = h[0];
  if (norm2) *norm2 = h[1];
  return GF_OK;
}

// gf_bsr_resid_f64 without allocation: the partial results live in the
// library's static device array `scratch`, the two results go to device
// memory, nothing synchronises.  Calls that overlap on one device share the
// array and race.
int gf_bsr_resid_f64_noalloc(gf_context ctx, int nnodes, const int* rowp,
                             const int* cols, const double* vals,
                             const double* x, const double* b, double* d_out) {
  if (!ctx || nnodes <= 0 || !rowp || !cols || !vals || !x || !b || !d_out)
    return gf::set_error(ctx, GF_ERR_INVALID, "bsr_resid_noalloc: bad args");
  const long long nblk = (nnodes + 39LL) / 40;
  if (nblk > kResidScratchGroups)
    return gf::set_error(ctx, GF_ERR_INVALID,
                         "bsr_resid_noalloc: %d nodes exceed the scratch "
                         "limit of %d", nnodes, 40 * kResidScratchGroups);
  double* base = nullptr;
  GF_TRY(scratch_on_device(ctx, &base));
  double* pmax = base;
  long long* pidx = (long long*)(base + nblk);
  double* pss = base + 2 * nblk;
  hipLaunchKernelGGL(gf::resid_partial_kernel, dim3((unsigned)nblk), dim3(256),
                     0, ctx->stream, nnodes, rowp, cols, vals, x, b, pmax,
                     pidx, pss);
  hipLaunchKernelGGL(gf::resid_final_kernel, dim3(1), dim3(256), 0,
                     ctx->stream, (int)nblk, pmax, pidx, pss, d_out);
  return gf::check_launch(ctx, "resid_kernel (noalloc)");
}

int gf_bsr_resid_comp_f64(gf_context ctx, int nnodes, const int* rowp,
                          const int* cols, const double* vals, const double* x,
                          const double* b, double* r) {
  if (!ctx || nnodes <= 0 || !rowp || !cols || !vals || !x || !b || !r ||
      r == x)
    return gf::set_error(ctx, GF_ERR_INVALID, "resid_comp: bad args");
  hipLaunchKernelGGL(gf::resid_comp_kernel, dim3(gf::div_up(nnodes, 40)),
                     dim3(256), 0, ctx->stream, nnodes, rowp, cols, vals, x, b,
                     r);
  return gf::check_launch(ctx, "resid_comp_kernel");
}

int gf_vec_axpby_f64(gf_context ctx, size_t n, double a, const double* x,
                     double b, double* y) {
  if (!ctx || !x || !y)
    return gf::set_error(ctx, GF_ERR_INVALID, "axpby: bad args");
  if (n == 0) return GF_OK;
  hipLaunchKernelGGL(gf::axpby_kernel, dim3(gf::div_up(n, 256)), dim3(256), 0,
                     ctx->stream, n, a, x, b, y);
  return gf::check_launch(ctx, "axpby_kernel");
}

int gf_convert_f64_f32(gf_context ctx, size_t n, const double* in,
                       float* out) {
  if (!ctx || !in || !out)
    return gf::set_error(ctx, GF_ERR_INVALID, "convert: bad args");
  if (n == 0) return GF_OK;
  hipLaunchKernelGGL(gf::convert_kernel, dim3(gf::div_up(n, 256)), dim3(256),
                     0, ctx->stream, n, in, out);
  return gf::check_launch(ctx, "convert_kernel");
}

}  // extern "C"
