// Block ILU numeric factorisation and the L/U triangular sweeps on a 6x6 BSR
// pattern (replaces cusparseXbsrilu02 and cusparseXbsrsv2,
// examples/iterative/gmres.cu:374-465 and :609-625).
//
// Semantics: in-place incomplete LU on the given block pattern, equal to
// scalar ILU on the block-expanded pattern: strictly-lower blocks hold L_ik,
// the diagonal block holds unit-lower L_ii below and U_ii on/above its
// diagonal, upper blocks hold U_ij.  The L sweep has a unit diagonal, the U
// sweep a non-unit one.
//
// Schedules
//   factorisation: rows ordered by dependency level on the host (the levels
//     of the factorisation equal those of the L sweep).  Default (dataflow):
//     a few launches, each over a run of levels, in which a workgroup takes
//     the next row in level order from a counter and waits, pivot by pivot,
//     only for the pivot rows it is about to use (a per-row completion flag,
//     released after the row is written); no level waits for the whole
//     previous level.  GPUFEM_ILU_SCHEDULE=level runs one launch per level
//     instead.  The arithmetic of a row is the same either way, so the two
//     schedules give bitwise-identical factors.
//   sweeps: one persistent launch per sweep, synchronised through the data:
//     the output vector is pre-filled with a signalling-NaN sentinel, each
//     entry is written once with a relaxed agent-scope atomic store, and a
//     row polls the entries it needs with relaxed agent-scope atomic loads
//     until none holds the sentinel.  No fence and no flag array: each entry
//     validates itself, so a dependency hop costs one store and one load
//     round trip, without L2 write-back or invalidation.  Rows are handed out
//     in level order, in chunks, from a 64-bit counter, so every row waited
//     on is already held by a running wavefront and the sweep cannot
//     deadlock.  The level-by-level sweep stays available (GPUFEM_TRSV=level)
//     as an independent schedule for parity checks.
#include <algorithm>
#include <mutex>
#include <cstdlib>
#include <string>
#include <vector>

#include "common.h"
#include "krylov.h"
#include "symbolic.h"

struct gf_ilu_s {
  gf_context ctx = nullptr;
  int n = 0, nnzb = 0;
  int *rowp = nullptr, *cols = nullptr, *diag = nullptr;
  int *lrows = nullptr, *urows = nullptr;  // rows ordered by level
  std::vector<int> lptr, uptr;             // level offsets (host)
  int* pivot = nullptr;                    // first zero-pivot row (device)
  int max_row_len = 0;
  std::vector<int> lmaxlen;                // longest row of each L level
  // sync-free sweep state
  int *lchunk = nullptr, *uchunk = nullptr; // chunk offsets (see below)
  int4 *lmeta = nullptr, *umeta = nullptr;  // {row, k0, k1, crit} per
                                            // position in hand-out order
  int nlchunk = 0, nuchunk = 0;
  bool lchains = false, uchains = false;  // chunks are chains in row order
  unsigned long long* counter = nullptr;   // chunk dispenser (monotonic)
  int* err = nullptr;  // set when a bounded wait expires: [0] sweeps,
                      // [1] factorisation
  unsigned long long base = 0;             // counter value at sweep start
  int grid = 0;                            // workgroups per sweep launch
  bool level_sweeps = false;
  // dataflow factorisation: done[i] == epoch once row i is final
  int* done = nullptr;
  int epoch = 0;
  // dependency level of every chunk (host; batched sweeps merge the chunk
  // lists of several systems level by level)
  std::vector<int> lchunk_level, uchunk_level;
};

namespace gf {
namespace {

// Longest row (blocks) the factorisation takes: the column indices and two
// pivots' position lists of a row live in LDS (3 ints per block of the
// longest row, 96 KB at this limit).
constexpr int kMaxRowBlocks = 8192;

// Dataflow schedule of the factorisation (kFlow kernels).  Every workgroup
// takes exactly one ticket from the counter; ticket t factors
// level_rows[t].  Tickets follow level order, so the pivot rows a row waits
// for were handed out before it to workgroups that are already running, and
// the lowest unfinished ticket never waits: the schedule cannot deadlock.
struct Flow {
  unsigned long long* counter = nullptr;
  unsigned long long base = 0;  // counter value at the launch
  int* done = nullptr;          // done[row] == epoch: row is final
  int epoch = 0;
  int* err = nullptr;           // set when a bounded wait expires
};
constexpr unsigned kFlowSpinLimit = 1u << 23;  // polls per wait (seconds)

// The row this workgroup factors (dataflow) or blockIdx.x's row (levels).
template <bool kFlow>
__device__ __forceinline__ int flow_row(const int* level_rows, int nrows,
                                        const Flow& f, int* s_row) {
  if (!kFlow) return blockIdx.x < nrows ? level_rows[blockIdx.x] : -1;
  if (threadIdx.x == 0) {
    const long long t = (long long)(atomicAdd(f.counter, 1ull) - f.base);
    const bool ok = t >= 0 && t < nrows;
    *s_row = ok ? level_rows[t] : -1;
    // each workgroup takes one ticket of [0, nrows): anything else means
    // the counter was not where the host expected it
    if (!ok)
      __hip_atomic_store(f.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return *s_row;
}

// Wait until the pivot rows piv[0..target] are final.  s_rdy[par] holds the
// length of the prefix of piv known to be final; wavefront 0 extends it
// 64 flags per load (one ballot), sleeping between polls while the next
// pivot row is not done, then takes the agent-scope acquire that makes the
// pivot rows' values visible, and publishes the new prefix in the other
// slot, s_rdy[par ^ 1]: every thread reads s_rdy[par] to decide whether to
// wait, so it must not change until all have passed the barrier.
// Uniform: every thread of the workgroup calls it with the same target.
__device__ __forceinline__ void flow_wait(const int* piv, int npiv,
                                          int target, int* s_rdy, int& par,
                                          const Flow& f) {
  if (s_rdy[par] > target) return;
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    int rdy = s_rdy[par];
    unsigned spins = 0;
    int nap = 1;
    while (rdy <= target) {
      const int t = rdy + lane;
      const bool pend =
          t < npiv && __hip_atomic_load(f.done + piv[t], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT) != f.epoch;
      const unsigned long long m = __ballot(pend);
      if (m == 0) {
        rdy = min(rdy + 64, npiv);
        continue;
      }
      rdy += __ffsll((long long)m) - 1;
      if (rdy > target) break;
      ++spins;
      const bool failed =
          spins > kFlowSpinLimit ||
          (spins % 256 == 0 && __hip_atomic_load(f.err, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT));
      if (failed) {
        // give up: the factors are reported invalid, the grid drains
        if (lane == 0)
          __hip_atomic_store(f.err, 1, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        rdy = npiv;
        break;
      }
      for (int z = 0; z < nap; z++) __builtin_amdgcn_s_sleep(1);
      nap = nap * 2 > 8 ? 8 : nap * 2;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (lane == 0) s_rdy[par ^ 1] = rdy;
  }
  __syncthreads();
  par ^= 1;
}

// Publish row i as final: every wavefront's stores of the row are released
// at agent scope before the flag is set.
__device__ __forceinline__ void flow_publish(int i, const Flow& f) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(f.done + i, f.epoch, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// One 256-thread workgroup per row of the current level.  The row's column
// indices are staged in LDS (positions of U(k) in row i are found by binary
// search there), L_ik is kept in LDS for the update, and the next k's
// pattern bounds are loaded while the current update runs.
// 512 threads with two register pairs each measured fastest on the 400x400
// plate, AMD + ILU(16): 293 ms a factorisation (256 x 4: 311, 256 x 2: 306,
// 128 x 4: 350, 128 x 2: 366).  Forming the next pivot's L_ik while the
// current update runs (one more barrier a pivot) measured slower (312 ms).
#ifndef GPUFEM_ILU_THREADS
#define GPUFEM_ILU_THREADS 512
#endif
constexpr int kIluThreads = GPUFEM_ILU_THREADS;
template <typename T, bool kFlow>
__global__ void __launch_bounds__(kIluThreads)
    ilu_row_kernel(const int* __restrict__ level_rows, int nrows,
                   const int* __restrict__ rowp, const int* __restrict__ cols,
                   const int* __restrict__ diag, T* vals, int* pivot,
                   Flow f, int maxlen) {
  // dynamic LDS: column indices of row i, and the positions in row i of
  // the current and next pivot's U-part columns (maxlen = longest row)
  extern __shared__ int ilu_gdyn[];
  int* scol = ilu_gdyn;
  int* const pos[2] = {ilu_gdyn + maxlen, ilu_gdyn + 2 * maxlen};
  __shared__ int ub[2][2];               // U-part bounds [u0, u1) per buffer
  __shared__ T udg[2][36];               // U_kk per buffer
  __shared__ T udi[2][6];                // 1 / diagonal of U_kk per buffer
  __shared__ T lik[36];
  __shared__ int s_row, s_rdy[2];
  const int i = flow_row<kFlow>(level_rows, nrows, f, &s_row);
  if (i < 0) return;
  const int tid = threadIdx.x;
  const int r0 = rowp[i], r1 = rowp[i + 1], di = diag[i];
  for (int t = tid; t < r1 - r0; t += blockDim.x) scol[t] = cols[r0 + t];
  if (tid == 0) s_rdy[0] = 0;
  int par = 0;
  __syncthreads();
  // dataflow: pivot row scol[kl] is waited for before it is staged (kl) or
  // used (L_ik reads U_kk staged from it; the update reads U_kj)
  auto wait_for = [&](int kl) {
    if (kFlow) flow_wait(scol, di - r0, kl, s_rdy, par, f);
  };
  // everything about pivot row k (its U part, U_kk, the positions of its
  // columns in row i) is final before this row starts, so it is staged for
  // the next k while the current k's update runs
  auto stage = [&](int kk, int buf) {
    const int k = scol[kk - r0];
    const int u0 = diag[k] + 1, u1 = rowp[k + 1];
    if (tid == 0) {
      ub[buf][0] = u0;
      ub[buf][1] = u1;
    }
    if (tid < 36) {
      const T u = vals[36 * (size_t)diag[k] + tid];
      udg[buf][tid] = u;
      if (tid % 7 == 0) udi[buf][tid / 7] = T(1) / u;
    }
    for (int t = tid; t < u1 - u0; t += blockDim.x) {
      const int j = cols[u0 + t];
      int lo = kk + 1 - r0, hi = r1 - r0;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (scol[mid] < j)
          lo = mid + 1;
        else
          hi = mid;
      }
      pos[buf][t] = (lo < r1 - r0 && scol[lo] == j) ? r0 + lo : -1;
    }
  };
  if (r0 < di) {
    wait_for(0);
    stage(r0, 0);
  }
  __syncthreads();
  for (int kk = r0; kk < di; kk++) {
    const int cur = (kk - r0) & 1;
    if (kk + 1 < di) wait_for(kk + 1 - r0);
    T* Aik = vals + 36 * (size_t)kk;
    const T* Ukk = udg[cur];
    const T* Ui = udi[cur];
    // L_ik = A_ik U_kk^-1 (row r by forward substitution over columns) on
    // lanes 0-5, the next k staged by the other lanes meanwhile
    if (tid < 6) {
      const int r = tid;
      T x[6];
#pragma unroll
      for (int c = 0; c < 6; c++) {
        T v = Aik[6 * r + c];
#pragma unroll
        for (int m = 0; m < c; m++) v -= x[m] * Ukk[6 * m + c];
        x[c] = v * Ui[c];
      }
#pragma unroll
      for (int c = 0; c < 6; c++) {
        Aik[6 * r + c] = x[c];
        lik[6 * r + c] = x[c];
      }
    } else if (kk + 1 < di) {
      // lanes 6.. stage the next pivot row (tid - 6 as the local index)
      const int k = scol[kk + 1 - r0];
      const int u0 = diag[k] + 1, u1 = rowp[k + 1];
      const int tl = tid - 6, nth = blockDim.x - 6;
      if (tl == 0) {
        ub[cur ^ 1][0] = u0;
        ub[cur ^ 1][1] = u1;
      }
      if (tl < 36) {
        const T u = vals[36 * (size_t)diag[k] + tl];
        udg[cur ^ 1][tl] = u;
        if (tl % 7 == 0) udi[cur ^ 1][tl / 7] = T(1) / u;
      }
      for (int t = tl; t < u1 - u0; t += nth) {
        const int j = cols[u0 + t];
        int lo = kk + 2 - r0, hi = r1 - r0;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (scol[mid] < j)
            lo = mid + 1;
          else
            hi = mid;
        }
        pos[cur ^ 1][t] = (lo < r1 - r0 && scol[lo] == j) ? r0 + lo : -1;
      }
    }
    __syncthreads();
    // A_ij -= L_ik U_kj, kBatch entries per thread at a time: every load of
    // a batch is issued before its stores (the U_kj blocks and the targets
    // are in the same array, so the compiler would otherwise serialise each
    // entry's loads behind the previous entry's store)
    const int u0 = ub[cur][0], nu = ub[cur][1] - u0;
    constexpr int kBatch = 4;
    for (int w0 = 0; w0 < 36 * nu; w0 += kBatch * (int)blockDim.x) {
      T uk[kBatch][6], av[kBatch];
      long long dst[kBatch];
      int ee[kBatch];
#pragma unroll
      for (int bq = 0; bq < kBatch; bq++) {
        const int w = w0 + bq * blockDim.x + tid;
        const int t = w / 36, e = w % 36, c = e % 6;
        const int p = w < 36 * nu ? pos[cur][t] : -1;
        dst[bq] = p < 0 ? -1 : 36LL * p + e;
        ee[bq] = e;
        const T* Ukj = vals + 36 * (size_t)(u0 + (p < 0 ? 0 : t));
#pragma unroll
        for (int m = 0; m < 6; m++) uk[bq][m] = p < 0 ? T(0) : Ukj[6 * m + c];
        av[bq] = p < 0 ? T(0) : vals[(size_t)dst[bq]];
      }
#pragma unroll
      for (int bq = 0; bq < kBatch; bq++) {
        if (dst[bq] < 0) continue;
        const int r = ee[bq] / 6;
        T v = 0;
#pragma unroll
        for (int m = 0; m < 6; m++) v += lik[6 * r + m] * uk[bq][m];
        vals[(size_t)dst[bq]] = av[bq] - v;
      }
    }
    __syncthreads();
  }
  // factor the diagonal block (no pivoting)
  T* D = vals + 36 * (size_t)di;
  if (tid == 0) {
    for (int c = 0; c < 6; c++) {
      const T piv = D[6 * c + c];
      if (piv == T(0)) {
        atomicMin(pivot, i);
        break;
      }
      for (int r = c + 1; r < 6; r++) {
        const T l = D[6 * r + c] / piv;
        D[6 * r + c] = l;
        for (int m = c + 1; m < 6; m++) D[6 * r + m] -= l * D[6 * c + m];
      }
    }
  }
  __syncthreads();
  // U_ij = L_ii^-1 A_ij for the blocks right of the diagonal
  const int nr = r1 - (di + 1);
  for (int w = tid; w < 6 * nr; w += blockDim.x) {
    const int t = w / 6, c = w % 6;
    T* A = vals + 36 * (size_t)(di + 1 + t);
    T x[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      T s = A[6 * r + c];
#pragma unroll
      for (int m = 0; m < r; m++) s -= D[6 * r + m] * x[m];
      x[r] = s;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) A[6 * r + c] = x[r];
  }
  if (kFlow) flow_publish(i, f);
}

// Rows too long for LDS, with the pivots applied in batches.  The values of
// row i stay in global memory.  A batch takes up to nbatch consecutive
// pivots whose rows are final; wavefront 0 forms their L_ik in order, each
// applied at once to the batch's own later L-part blocks (held in LDS), and
// then one pass reads every (block, column) pair of the rest of the row that
// a pivot of the batch updates once, subtracts the batch's contributions
// L_ik U_kj in pivot order and writes it once: one read-modify-write of the
// touched part of the row per batch instead of one per pivot, and each U_kj
// column loaded once for its six rows.  The pass walks the pivots' U-part
// lists (a pair is taken by the first pivot of the batch that updates it),
// so a batch of one pivot costs what a pivot step of ilu_row_kernel does.
// Every entry sees the operations of ilu_row_kernel in the same order, so
// the factors agree bit for bit.
// Dynamic LDS: the column indices of row i; per pivot of the batch, the
// position in row i of each of its U-part columns (or -1) and, per position
// of row i, 1 + the index in the pivot's U part (0: not updated).
#ifndef GPUFEM_ILU_BATCH
#define GPUFEM_ILU_BATCH 8
#endif
constexpr int kIluBatch = GPUFEM_ILU_BATCH;
constexpr int kIluPairSlots = kIluBatch * (kIluBatch - 1) / 2;
// slot of the pivot pair (p, q), p < q < kIluBatch
__device__ __forceinline__ int batch_pair(int p, int q) {
  return p * (2 * kIluBatch - p - 1) / 2 + (q - p - 1);
}
// dynamic LDS bytes per position of the longest row, batch of nbatch
__host__ __device__ constexpr size_t batch_lds_per_pos(int nbatch) {
  return sizeof(int) + 2 * sizeof(short) * (size_t)nbatch;
}
template <typename T, bool kFlow>
__global__ void __launch_bounds__(kIluThreads)
    ilu_row_batch_kernel(const int* __restrict__ level_rows, int nrows,
                         const int* __restrict__ rowp,
                         const int* __restrict__ cols,
                         const int* __restrict__ diag, T* vals, int* pivot,
                         Flow f, int nbatch, int maxlen) {
  extern __shared__ __align__(16) int ilu_bdyn[];
  __shared__ int s_u0[kIluBatch], s_nu[kIluBatch];
  __shared__ T s_ukk[kIluBatch][36];
  __shared__ T s_udi[kIluBatch][6];  // 1 / diagonal of U_kk
  __shared__ T s_lik[kIluBatch][36];
  __shared__ T s_ab[kIluBatch][36];       // A_ik of the batch's pivots
  __shared__ T s_uq[kIluPairSlots][36];   // U_kp,kq, p < q in the batch
  __shared__ int s_row, s_rdy[2];
  const int i = flow_row<kFlow>(level_rows, nrows, f, &s_row);
  if (i < 0) return;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int r0 = rowp[i], r1 = rowp[i + 1], nb = r1 - r0, dl = diag[i] - r0;
  int* scol = ilu_bdyn;
  // hit[p * maxlen + pl] = 1 + index in U(k_p) of position pl, or 0
  unsigned short* hit = reinterpret_cast<unsigned short*>(ilu_bdyn + maxlen);
  // upos[p * maxlen + t] = position in row i of U(k_p)'s column t, or -1
  short* upos = reinterpret_cast<short*>(hit + (size_t)nbatch * maxlen);
  for (int t = tid; t < nb; t += nth) scol[t] = cols[r0 + t];
  for (int e = tid; e < nbatch * maxlen; e += nth) hit[e] = 0;
  if (tid == 0) s_rdy[0] = 0;
  int par = 0;
  __syncthreads();
  // position of column j among scol[lo..nb), or -1
  auto find = [&](int j, int lo) {
    int hi = nb;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (scol[mid] < j)
        lo = mid + 1;
      else
        hi = mid;
    }
    return (lo < nb && scol[lo] == j) ? lo : -1;
  };
  for (int kb = 0; kb < dl;) {
    int np = min(nbatch, dl - kb);
    if (kFlow) {
      // the batch: the pivots from kb on whose rows are known to be final
      flow_wait(scol, dl, kb, s_rdy, par, f);
      np = min(np, s_rdy[par] - kb);
    }
    if (tid < np) {
      const int k = scol[kb + tid];
      const int u0 = diag[k] + 1;
      s_u0[tid] = u0;
      s_nu[tid] = rowp[k + 1] - u0;
    }
    for (int e = tid; e < 36 * np; e += nth) {
      const int p = e / 36, q = e % 36;
      const T u = vals[36 * (size_t)diag[scol[kb + p]] + q];
      s_ukk[p][q] = u;
      if (q % 7 == 0) s_udi[p][q / 7] = T(1) / u;
      s_ab[p][q] = vals[36 * (size_t)(r0 + kb + p) + q];
    }
    __syncthreads();
    int tot = 0;  // U-part columns of the batch's pivots
    for (int p = 0; p < np; p++) {
      const int u0 = s_u0[p], nu = s_nu[p];
      for (int t = tid; t < nu; t += nth) {
        const int pl = find(cols[u0 + t], kb + p + 1);
        upos[p * maxlen + t] = (short)pl;
        if (pl >= 0) hit[p * maxlen + pl] = (unsigned short)(t + 1);
      }
      tot += nu;
    }
    __syncthreads();
    for (int e = tid; e < 36 * np * np; e += nth) {
      const int pq = e / 36, m = e % 36, p = pq / np, q = pq % np;
      if (q <= p) continue;
      const int v = hit[p * maxlen + kb + q];
      if (v)
        s_uq[batch_pair(p, q)][m] =
            vals[36 * (size_t)(s_u0[p] + v - 1) + m];
    }
    __syncthreads();
    if (tid < 64) {
      // wavefront 0: L_ik of each pivot, each applied to the batch's later
      // L-part blocks before the next is formed
      for (int p = 0; p < np; p++) {
        if (tid < 6) {
          const int r = tid;
          T x[6];
#pragma unroll
          for (int c = 0; c < 6; c++) {
            T v = s_ab[p][6 * r + c];
#pragma unroll
            for (int m = 0; m < c; m++) v -= x[m] * s_ukk[p][6 * m + c];
            x[c] = v * s_udi[p][c];
          }
#pragma unroll
          for (int c = 0; c < 6; c++) {
            s_ab[p][6 * r + c] = x[c];
            s_lik[p][6 * r + c] = x[c];
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (tid < 36) {
          const int r = tid / 6, c = tid % 6;
          for (int q = p + 1; q < np; q++) {
            if (!hit[p * maxlen + kb + q]) continue;
            const T* U = s_uq[batch_pair(p, q)];
            T acc = 0;
#pragma unroll
            for (int m = 0; m < 6; m++) acc += s_lik[p][6 * r + m] * U[6 * m + c];
            s_ab[q][6 * r + c] = s_ab[q][6 * r + c] - acc;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
    __syncthreads();
    for (int e = tid; e < 36 * np; e += nth)
      vals[36 * (size_t)(r0 + kb + e / 36) + e % 36] = s_ab[e / 36][e % 36];
    // the rest of the row: (block, column) pairs of the pivots' U-part
    // lists, each taken by the first pivot of the batch that updates it
    const int pl0 = kb + np;
    for (int w = tid; w < 6 * tot; w += nth) {
      // L_ik is re-read from LDS per pair (hoisted out of the loop it would
      // take 36 registers per pivot)
      __asm__ volatile("" ::: "memory");
      int p = 0, rem = w;
      while (rem >= 6 * s_nu[p]) rem -= 6 * s_nu[p++];
      const int t = rem / 6, c = rem % 6;
      const int pl = upos[p * maxlen + t];
      if (pl < pl0) continue;  // not in the row (-1), or in the batch's L part
      bool first = true;
      for (int q = 0; q < p; q++) first = first && !hit[q * maxlen + pl];
      if (!first) continue;
      int tt[kIluBatch];
#pragma unroll
      for (int q = 0; q < kIluBatch; q++)
        tt[q] = (q >= p && q < np) ? (int)hit[q * maxlen + pl] - 1 : -1;
      T* A = vals + 36 * (size_t)(r0 + pl) + c;
      T a[6];
#pragma unroll
      for (int r = 0; r < 6; r++) a[r] = A[6 * r];
#pragma unroll
      for (int q0 = 0; q0 < kIluBatch; q0 += 4) {
        T u[4][6];
#pragma unroll
        for (int qq = 0; qq < 4 && q0 + qq < kIluBatch; qq++) {
          const int q = q0 + qq;
          const T* U =
              vals + 36 * (size_t)(tt[q] >= 0 ? s_u0[q] + tt[q] : 0) + c;
#pragma unroll
          for (int m = 0; m < 6; m++) u[qq][m] = tt[q] >= 0 ? U[6 * m] : T(0);
        }
#pragma unroll
        for (int qq = 0; qq < 4 && q0 + qq < kIluBatch; qq++) {
          const int q = q0 + qq;
          if (tt[q] < 0) continue;
#pragma unroll
          for (int r = 0; r < 6; r++) {
            T v = 0;
#pragma unroll
            for (int m = 0; m < 6; m++) v += s_lik[q][6 * r + m] * u[qq][m];
            a[r] = a[r] - v;
          }
        }
      }
#pragma unroll
      for (int r = 0; r < 6; r++) A[6 * r] = a[r];
    }
    __syncthreads();
    // clear the batch's hits
    for (int p = 0; p < np; p++)
      for (int t = tid; t < s_nu[p]; t += nth) {
        const int pl = upos[p * maxlen + t];
        if (pl >= 0) hit[p * maxlen + pl] = 0;
      }
    __syncthreads();
    kb += np;
  }
  __syncthreads();
  // factor the diagonal block (no pivoting)
  T* D = vals + 36 * (size_t)(r0 + dl);
  if (tid == 0) {
    for (int c = 0; c < 6; c++) {
      const T piv = D[6 * c + c];
      if (piv == T(0)) {
        atomicMin(pivot, i);
        break;
      }
      for (int r = c + 1; r < 6; r++) {
        const T l = D[6 * r + c] / piv;
        D[6 * r + c] = l;
        for (int m = c + 1; m < 6; m++) D[6 * r + m] -= l * D[6 * c + m];
      }
    }
  }
  __syncthreads();
  // U_ij = L_ii^-1 A_ij for the blocks right of the diagonal
  const int nr = nb - (dl + 1);
  for (int w = tid; w < 6 * nr; w += nth) {
    const int t = w / 6, c = w % 6;
    T* A = D + 36 * (size_t)(1 + t);
    T x[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      T s = A[6 * r + c];
#pragma unroll
      for (int m = 0; m < r; m++) s -= D[6 * r + m] * x[m];
      x[r] = s;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) A[6 * r + c] = x[r];
  }
  if (kFlow) flow_publish(i, f);
}

// The same factorisation with row i held in LDS for the whole row: L_ik is
// formed from, and the updates A_ij -= L_ik U_kj are applied to, the LDS
// copy, so the only global accesses inside the pivot loop are reads of the
// final rows k.  Those are pipelined one pivot ahead: the U-part bounds of
// every pivot row are read into LDS once; while pivot k is processed, the
// column indices and U_kk of pivot k+2 and each thread's U_kj columns of
// pivot k+1 are loaded into registers, and the positions of pivot k+1's
// columns in row i are searched.  The row is written back once at the end.
// Arithmetic is that of ilu_row_kernel, operation for operation.
// Dynamic LDS: 36 * nb values, nb column indices, 2 * nl pivot bounds.
#ifndef GPUFEM_ILU_PAIRS
#define GPUFEM_ILU_PAIRS 2
#endif
constexpr int kIluPairs = GPUFEM_ILU_PAIRS;  // (block, column) pairs of U_kj per thread
constexpr int kIluCols = 4;   // prefetched column indices per staging thread
// Short rows (levels whose rows, and so their pivot rows' U parts, are
// short) run in 256-thread workgroups: the kernel's registers allow 12
// wavefronts a CU, one 512-thread row or three 256-thread ones, and a short
// row spends its time in per-row latency (ticket, loads of its pattern and
// pivot bounds, diagonal factorisation, write-back and release), not in its
// few pivot updates.  400x400 plate, AMD + ILU(96), FP64 factorisation:
// 135 ms with 512 threads throughout, 129 ms with rows up to 128 blocks in
// 256-thread workgroups, 123 ms up to 256 or 512 blocks.  (The staging
// threads hold 4 * (256 - 6) column indices: pivot U parts up to 1,000.)
constexpr int kIluShortThreads = 256;
constexpr int kIluShortRow = 256;  // longest row (blocks) taken that way
template <typename T, bool kFlow, int kThreads = kIluThreads>
__global__ void __launch_bounds__(kThreads)
    ilu_row_lds_kernel(const int* __restrict__ level_rows, int nrows,
                       const int* __restrict__ rowp,
                       const int* __restrict__ cols,
                       const int* __restrict__ diag, T* vals, int* pivot,
                       Flow f, int maxlen) {
  extern __shared__ __align__(16) unsigned char ilu_dyn[];
  __shared__ T udg[2][36];
  __shared__ T udi[2][6];  // 1 / diagonal of U_kk
  __shared__ T lik[36];
  __shared__ int s_row, s_rdy[2];
  const int i = flow_row<kFlow>(level_rows, nrows, f, &s_row);
  if (i < 0) return;
  const int tid = threadIdx.x, nth = blockDim.x;
  const int r0 = rowp[i], r1 = rowp[i + 1], nb = r1 - r0, dl = diag[i] - r0;
  T* row = reinterpret_cast<T*>(ilu_dyn);
  int* scol = reinterpret_cast<int*>(row + 36 * (size_t)nb);
  int* ubs = scol + nb;  // [2 * kl]: U part [u0, u1) of pivot row scol[kl]
  // local positions in row i of two pivots' U-part columns, or -1
  int* const pos[2] = {ubs + 2 * nb, ubs + 2 * nb + maxlen};
  for (int e = tid; e < 36 * nb; e += nth) row[e] = vals[36 * (size_t)r0 + e];
  for (int t = tid; t < nb; t += nth) {
    const int j = cols[r0 + t];
    scol[t] = j;
    if (t < dl) {
      ubs[2 * t] = diag[j] + 1;
      ubs[2 * t + 1] = rowp[j + 1];
    }
  }
  if (tid == 0) s_rdy[0] = 0;
  int par = 0;
  __syncthreads();
  // dataflow: the values of pivot row scol[kl] are read one (U_kj) or two
  // (U_kk) pivots ahead, so iteration kl waits for pivot kl + 2
  auto wait_for = [&](int kl) {
    if (kFlow) flow_wait(scol, dl, min(kl, dl - 1), s_rdy, par, f);
  };
  const int ts = tid - 6, ns = nth - 6;  // staging threads (tid >= 6)
  // position of column j among scol[lo..nb), or -1
  auto find = [&](int j, int lo) {
    int hi = nb;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (scol[mid] < j)
        lo = mid + 1;
      else
        hi = mid;
    }
    return (lo < nb && scol[lo] == j) ? lo : -1;
  };
  // registers: this thread's U_kj columns for the current and next pivot,
  // and (staging threads) the column indices and U_kk entry of the pivot
  // after next
  T uv[kIluPairs][6], un[kIluPairs][6];
  int pc[kIluCols];
  T pd = T(0);
  auto load_u = [&](int kl, T (&dst)[kIluPairs][6]) {
    const int u0 = kl < dl ? ubs[2 * kl] : 0;
    const int nu = kl < dl ? ubs[2 * kl + 1] - u0 : 0;
#pragma unroll
    for (int sp = 0; sp < kIluPairs; sp++) {
      const int w = tid + sp * nth;
      const bool ok = w < 6 * nu;
      const T* U =
          vals + 36 * (size_t)(u0 + (ok ? w / 6 : 0)) + (ok ? w % 6 : 0);
#pragma unroll
      for (int m = 0; m < 6; m++) dst[sp][m] = ok ? U[6 * m] : T(0);
    }
  };
  auto load_cols = [&](int kl) {
    const int u0 = kl < dl ? ubs[2 * kl] : 0;
    const int nu = kl < dl ? ubs[2 * kl + 1] - u0 : 0;
#pragma unroll
    for (int q = 0; q < kIluCols; q++) {
      const int t = ts + q * ns;
      pc[q] = (ts >= 0 && t < nu) ? cols[u0 + t] : -1;
    }
    pd = (ts >= 0 && ts < 36 && kl < dl)
             ? vals[36 * (size_t)diag[scol[kl]] + ts]
             : T(0);
  };
  // positions and U_kk of pivot kl into buffer buf from the registers
  auto stage_from_regs = [&](int kl, int buf) {
    const int u0 = ubs[2 * kl], nu = ubs[2 * kl + 1] - u0;
#pragma unroll
    for (int q = 0; q < kIluCols; q++) {
      const int t = ts + q * ns;
      if (t < nu) pos[buf][t] = find(pc[q], kl + 1);
    }
    for (int t = ts + kIluCols * ns; t < nu; t += ns)
      pos[buf][t] = find(cols[u0 + t], kl + 1);
    if (ts < 36) {
      udg[buf][ts] = pd;
      if (ts % 7 == 0) udi[buf][ts / 7] = T(1) / pd;
    }
  };
  if (dl > 0) {
    wait_for(1);
    // pivot 0 staged directly by all threads
    const int k = scol[0];
    const int u0 = ubs[0], nu = ubs[1] - u0;
    if (tid < 36) {
      const T u = vals[36 * (size_t)diag[k] + tid];
      udg[0][tid] = u;
      if (tid % 7 == 0) udi[0][tid / 7] = T(1) / u;
    }
    for (int t = tid; t < nu; t += nth) pos[0][t] = find(cols[u0 + t], 1);
    load_u(0, uv);
    load_cols(1);
  }
  __syncthreads();
  for (int kl = 0; kl < dl; kl++) {
    const int cur = kl & 1;
    const int u0 = ubs[2 * kl], nu = ubs[2 * kl + 1] - u0;
    if (kl + 1 < dl) wait_for(kl + 2);
    load_u(kl + 1, un);
    if (tid < 6) {
      const int r = tid;
      const T* Ukk = udg[cur];
      const T* Ui = udi[cur];
      T* Aik = row + 36 * kl;
      T x[6];
#pragma unroll
      for (int c = 0; c < 6; c++) {
        T v = Aik[6 * r + c];
#pragma unroll
        for (int m = 0; m < c; m++) v -= x[m] * Ukk[6 * m + c];
        x[c] = v * Ui[c];
      }
#pragma unroll
      for (int c = 0; c < 6; c++) {
        Aik[6 * r + c] = x[c];
        lik[6 * r + c] = x[c];
      }
    } else if (kl + 1 < dl) {
      stage_from_regs(kl + 1, cur ^ 1);
      load_cols(kl + 2);
    }
    __syncthreads();
    // A_ij -= L_ik U_kj on the LDS row
#pragma unroll
    for (int sp = 0; sp < kIluPairs; sp++) {
      const int w = tid + sp * nth;
      if (w < 6 * nu) {
        const int pl = pos[cur][w / 6], c = w % 6;
        if (pl >= 0) {
#pragma unroll
          for (int r = 0; r < 6; r++) {
            T v = 0;
#pragma unroll
            for (int m = 0; m < 6; m++) v += lik[6 * r + m] * uv[sp][m];
            row[36 * pl + 6 * r + c] -= v;
          }
        }
      }
    }
    for (int w = tid + kIluPairs * nth; w < 6 * nu; w += nth) {
      const int pl = pos[cur][w / 6], c = w % 6;
      if (pl < 0) continue;
      const T* U = vals + 36 * (size_t)(u0 + w / 6) + c;
      T um[6];
#pragma unroll
      for (int m = 0; m < 6; m++) um[m] = U[6 * m];
#pragma unroll
      for (int r = 0; r < 6; r++) {
        T v = 0;
#pragma unroll
        for (int m = 0; m < 6; m++) v += lik[6 * r + m] * um[m];
        row[36 * pl + 6 * r + c] -= v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int sp = 0; sp < kIluPairs; sp++)
#pragma unroll
      for (int m = 0; m < 6; m++) uv[sp][m] = un[sp][m];
  }
  // factor the diagonal block (no pivoting)
  T* D = row + 36 * dl;
  if (tid == 0) {
    for (int c = 0; c < 6; c++) {
      const T piv = D[6 * c + c];
      if (piv == T(0)) {
        atomicMin(pivot, i);
        break;
      }
      for (int r = c + 1; r < 6; r++) {
        const T l = D[6 * r + c] / piv;
        D[6 * r + c] = l;
        for (int m = c + 1; m < 6; m++) D[6 * r + m] -= l * D[6 * c + m];
      }
    }
  }
  __syncthreads();
  // U_ij = L_ii^-1 A_ij for the blocks right of the diagonal
  const int nr = nb - (dl + 1);
  for (int w = tid; w < 6 * nr; w += nth) {
    const int t = w / 6, c = w % 6;
    T* A = row + 36 * (dl + 1 + t);
    T x[6];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      T sv = A[6 * r + c];
#pragma unroll
      for (int m = 0; m < r; m++) sv -= D[6 * r + m] * x[m];
      x[r] = sv;
    }
#pragma unroll
    for (int r = 0; r < 6; r++) A[6 * r + c] = x[r];
  }
  __syncthreads();
  for (int e = tid; e < 36 * nb; e += nth) vals[36 * (size_t)r0 + e] = row[e];
  if (kFlow) flow_publish(i, f);
}

// Six lanes per row (lane = scalar row r), ten rows per wave.
template <typename T, bool kLower>
__global__ void __launch_bounds__(256)
    trsv_level_kernel(const int* __restrict__ level_rows, int nrows,
                      const int* __restrict__ rowp,
                      const int* __restrict__ cols,
                      const int* __restrict__ diag, const T* __restrict__ lu,
                      const T* __restrict__ b, T* __restrict__ x) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int li = (blockIdx.x * (blockDim.x / 64) + wave) * 10 + lane / 6;
  const int r = lane % 6;
  const bool active = lane < 60 && li < nrows;
  const int i = active ? level_rows[li] : 0;
  T s = 0;
  int di = 0;
  if (active) {
    di = diag[i];
    s = b[6 * (size_t)i + r];
    const int k0 = kLower ? rowp[i] : di + 1;
    const int k1 = kLower ? di : rowp[i + 1];
    for (int k = k0; k < k1; k++) {
      const T* a = lu + 36 * (size_t)k + 6 * r;
      const T* xk = x + 6 * (size_t)cols[k];
#pragma unroll
      for (int c = 0; c < 6; c++) s -= a[c] * xk[c];
    }
  }
  // diagonal block solve across the six lanes of the row
  const int base = lane - r;  // first lane of this row's group
  const T* D = lu + 36 * (size_t)di + 6 * r;
  if (kLower) {
#pragma unroll
    for (int c = 0; c < 5; c++) {
      const T yc = __shfl(s, base + c, 64);
      if (active && r > c) s -= D[c] * yc;
    }
  } else {
    const T dinv = active ? T(1) / D[r] : T(0);
#pragma unroll
    for (int c = 5; c >= 0; c--) {
      if (active && r == c) s *= dinv;
      const T yc = __shfl(s, base + c, 64);
      if (active && r < c) s -= D[c] * yc;
    }
  }
  if (active) x[6 * (size_t)i + r] = s;
}


constexpr int kSweepThreads = 256;        // 4 wavefronts per workgroup
// Batches of 10 blocks of a pass whose factor values are loaded into
// registers before the row waits; blocks beyond them are loaded after the
// wait, on the row's critical path.  With passes of 128 blocks
// (400x400 plate, AMD + ILU(16), L+U pair): 3 -> 4.33 ms, 4 -> 4.09,
// 6 -> 3.71, 8 -> 3.48, 10 -> 3.36, 12 -> 3.29 (256 VGPRs).  At ILU(96)
// (rows of hundreds of blocks, so most passes are full): 12 batches of a
// 128-block pass 7.32 ms, 13 (the whole pass) 6.83 ms, passes of 64 blocks
// with 7 batches (the whole pass, 178 VGPRs) 5.99 ms: no load is left
// after the wait, and the wait gathers half as many dependencies.
#ifndef GPUFEM_SWEEP_ITER
#define GPUFEM_SWEEP_ITER 7
#endif
#ifndef GPUFEM_SWEEP_MINWAVES
#define GPUFEM_SWEEP_MINWAVES 1
#endif
// Waiting on the latest-level dependency alone before gathering the others
// (GPUFEM_SWEEP_CRIT=1) cuts polling traffic but adds a round trip per row:
// measured 4.70 vs 3.54 ms a sweep pair on the 400x400 plate (AMD, ILU(12))
// and 5.53 vs 3.47 ms on the 87x87 plate, and more wavefronts per CU did not
// recover it (5.07-5.41 ms at 6-16), so it is off.
#ifndef GPUFEM_SWEEP_CRIT
#define GPUFEM_SWEEP_CRIT 0
#endif
// longest sleep between polls, x 64 cycles: L+U pair on the 400x400 plate
// (ILU(16), 4 batches) 16 -> 4.14 ms, 8 -> 4.09, 4 -> 4.06, 2 -> 4.06,
// 1 -> 4.07; with 12 batches 8 -> 3.29, 2 -> 3.26; ILU(96), entries of a
// dependency in consecutive lanes: 1 -> 5.29, 2 -> 5.26, 4 -> 5.25
#ifndef GPUFEM_SWEEP_BACKOFF
#define GPUFEM_SWEEP_BACKOFF 2
#endif
// A waiting row polls only its pending entries, one poll in flight.  Two
// interleaved streams of polls of all the row's entries, the second half a
// round trip after the first, took the L+U pair of the 400x400 plate (AMD +
// ILU(96)) from 5.69 to 8.8 ms: polling traffic delays the stores and loads
// on the critical path more than the earlier detection gains.
// wavefronts per compute unit for a sweep launch: fewer waiting wavefronts
// mean less polling and counter traffic; measured on the 400x400 plate
// (AMD + ILU(12)): 16 -> 4.37 ms per L+U pair, 4 -> 3.62, 3 -> 3.51,
// 2 -> 3.64, 1 -> 5.17; with 12 register batches (ILU(16)): 2 -> 3.55,
// 3 -> 3.26, 4 -> 3.31, 5 -> 3.61; with 64-block passes (ILU(96)): 3 -> 5.99,
// 4 -> 5.95, 6 -> 6.13, 8 -> 6.19; with the entries of a dependency in
// consecutive lanes (fewer memory requests per wavefront): 3 -> 5.26,
// 4 -> 5.12, 5 -> 5.29, 6 -> 5.30
#ifndef GPUFEM_SWEEP_WAVES_PER_CU
#define GPUFEM_SWEEP_WAVES_PER_CU 4
#endif
constexpr int kSweepIter = GPUFEM_SWEEP_ITER;  // block batches in registers
constexpr int kSlots = 10;                // blocks per batch (6 lanes each)
constexpr int kGather = 64;               // dependency blocks per pass
// blocks next to the diagonal taken as a row's last pass: ILU(96) plate,
// L+U pair 5.99 ms without, 5.90 / 5.82 / 5.84 / 5.87 / 5.89 ms with 4 / 6 /
// 8 / 16 / 24 in every row longer than that, 5.77 ms with 8 in the rows of
// more than one pass (a row of at most kGather blocks keeps its one pass);
// with the entries of a dependency in consecutive lanes 5 / 8 / 10: 5.27 /
// 5.26 / 5.26 ms
#ifndef GPUFEM_SWEEP_NEAR
#define GPUFEM_SWEEP_NEAR 8
#endif
constexpr int kNearPass = GPUFEM_SWEEP_NEAR;
static_assert(kGather % 64 == 0, "a pass spreads over whole wavefronts");
constexpr unsigned kSpinLimit = 1u << 20; // bounded polls per batch

// Sentinel values: signalling NaNs.  Arithmetic never produces a signalling
// NaN (results are quiet), so an output entry that holds the sentinel has not
// been written yet in the current sweep.
template <typename T> struct Sentinel;
template <> struct Sentinel<double> {
  using Bits = unsigned long long;
  static constexpr Bits value = 0x7FF4A5A5A5A5A5A5ull;
};
template <> struct Sentinel<float> {
  using Bits = unsigned int;
  static constexpr Bits value = 0x7FA5A5A5u;
};

template <typename T>
__device__ __forceinline__ typename Sentinel<T>::Bits load_bits(const T* p) {
  using B = typename Sentinel<T>::Bits;
  return __hip_atomic_load(reinterpret_cast<const B*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void store_value(T* p, T v) {
  using B = typename Sentinel<T>::Bits;
  __hip_atomic_store(reinterpret_cast<B*>(p), __builtin_bit_cast(B, v),
                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// lane l's value, to every lane (v_readlane: a few cycles, where a
// ds_bpermute shuffle takes an LDS round trip)
__device__ __forceinline__ double read_lane(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float read_lane(float v, int l) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_readlane(__builtin_bit_cast(unsigned, v), l));
}

template <typename T>
__global__ void fill_sentinel_kernel(size_t n, T* x) {
  using B = typename Sentinel<T>::Bits;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) reinterpret_cast<B*>(x)[i] = Sentinel<T>::value;
}

// Sync-free sweep.  The output vector is filled with the sentinel before the
// launch; every output entry is written exactly once, with a relaxed
// agent-scope atomic store, and a row reads each entry it depends on with
// relaxed agent-scope atomic loads until the value is no longer the sentinel.
// Each entry validates itself, so no fences or flags are needed.
//
// Rows are taken from a monotonic 64-bit counter in chunks, either chunks of
// rows of one level in level order (order = the level-sorted rows), or
// chains of consecutive rows in sweep order (order = null).  Either way every
// dependency of a row lies in an earlier chunk, handed out earlier to a
// running wavefront, or earlier in the same chunk, so the sweep cannot
// deadlock whatever the number of resident workgroups.
//
// Per row, in passes of up to kGather dependency blocks: the pass's factor
// blocks are loaded into registers, then its dependencies' entries (six
// consecutive lanes per dependency) are loaded together, polled until none
// holds the sentinel and parked in LDS, so a pass costs about two memory
// round trips (column indices, then dependency entries) whatever its
// length; slot t = lane / 6 (0..9) takes blocks t, t+10, ... of the pass
// with r = lane % 6 the scalar row; the 6x6 diagonal block is solved by
// lanes 0-5 (v_readlane broadcasts with one row per wavefront).  The row
// last solved by the wavefront is kept in registers (a row often depends on
// the row just before it).
//
// fill_next (lower sweep only, may be null): the entries of the row in the
// upper sweep's output are set to the sentinel once the row's rhs is read,
// which allows that output to alias the rhs.
//
// kBatched: one launch sweeps the systems of several solves.  Chunk kc of
// the launch is chunk merged[kc].y of system merged[kc].x, whose arrays are
// in slots[merged[kc].x]; the merged list takes the systems' chunks level by
// level, so every row waited on is still in an earlier chunk.
template <typename T, bool kLower, int kRows, bool kBatched = false>
__global__ void __launch_bounds__(kSweepThreads, GPUFEM_SWEEP_MINWAVES)
    trsv_syncfree_kernel(const int4* __restrict__ meta_,
                         const int* __restrict__ cols_,
                         const int* __restrict__ chunk_ptr_, int nchunks,
                         const T* __restrict__ lu_, const T* b_, T* x_,
                         T* fill_next_, unsigned long long* counter,
                         unsigned long long base, int* err,
                         const SweepSlot<T>* __restrict__ slots = nullptr,
                         const int2* __restrict__ merged = nullptr) {
  using B = typename Sentinel<T>::Bits;
  const int4* __restrict__ meta = meta_;
  const int* __restrict__ cols = cols_;
  const int* __restrict__ chunk_ptr = chunk_ptr_;
  const T* __restrict__ lu = lu_;
  const T* b = b_;
  T* x = x_;
  T* fill_next = fill_next_;
  int cur_slot = -1;
  constexpr int G = 64 / kRows;        // lanes per row
  constexpr int kS = G / 6;            // block slots per row (10 or 5)
  constexpr int kH = kGather / G;      // dependency blocks per lane per pass
  constexpr bool kUseCache = kRows == 1;
  __shared__ T xs_all[kSweepThreads / 64][kRows][kGather][6];
  const int lane = threadIdx.x % 64;
  const int grp = lane / G, gl = lane % G, g0 = grp * G;
  const int t = gl / 6, r = gl % 6;
  const bool act = gl < 6 * kS;
  T(&xs)[kGather][6] = xs_all[threadIdx.x / 64][grp];
  // the last row solved by this wavefront (one row per wavefront only): a
  // chain's row usually depends on the row just before it
  int cid = -1;
  T cx[6];
#pragma unroll
  for (int c = 0; c < 6; c++) cx[c] = T(0);
  auto cached = [&](int j) { return kUseCache && cid == j; };
  unsigned long long tk = 0;
  if (lane == 0) tk = atomicAdd(counter, 1ull);
  tk = __shfl(tk, 0, 64);
  long long kc = (long long)(tk - base);
  while (kc < nchunks) {
    // take the next chunk now: the counter round trip overlaps this chunk
    unsigned long long tn = 0;
    if (lane == 0) tn = atomicAdd(counter, 1ull);
    int lc = (int)kc;
    if constexpr (kBatched) {
      const int2 mc = merged[kc];
      lc = mc.y;
      if (*slots[mc.x].go != 1) {  // this system takes no step now
        tn = __shfl(tn, 0, 64);
        kc = (long long)(tn - base);
        continue;
      }
      if (mc.x != cur_slot) {
        cur_slot = mc.x;
        const SweepSlot<T>& S = slots[mc.x];
        meta = kLower ? S.lmeta : S.umeta;
        chunk_ptr = kLower ? S.lchunk : S.uchunk;
        cols = S.cols;
        lu = S.lu;
        b = kLower ? S.b : S.tmp;
        x = kLower ? S.tmp : S.x;
        fill_next = kLower ? S.x : nullptr;
        cid = -1;  // the cached row belongs to another system
      }
    }
    const int p0 = chunk_ptr[lc], p1 = chunk_ptr[lc + 1];
    int4 mnext = (p0 + grp < p1) ? meta[p0 + grp] : make_int4(0, 0, 0, 0);
    for (int pb = p0; pb < p1; pb += kRows) {
      // rows pb + grp of one level (independent when kRows > 1)
      const bool live = pb + grp < p1;
      const int4 m = mnext;
      if (pb + kRows + grp < p1) mnext = meta[pb + kRows + grp];
      const int i = m.x, k0 = live ? m.y : 0, k1 = live ? m.z : 0;
      const int di = kLower ? m.z : m.y - 1;
      const int crit = live ? m.w : -1;
      T s = (live && gl < 6) ? b[6 * (size_t)i + gl] : T(0);
      if (kLower && fill_next && live && gl < 6)
        reinterpret_cast<B*>(fill_next)[6 * (size_t)i + gl] =
            Sentinel<T>::value;
      T drow[6];
#pragma unroll
      for (int mm = 0; mm < 6; mm++)
        drow[mm] = (live && gl < 6) ? lu[36 * (size_t)di + 6 * gl + mm] : T(0);
      // reciprocal of this lane's diagonal entry, formed before any wait
      T dd = T(1);
#pragma unroll
      for (int mm = 0; mm < 6; mm++)
        if (gl == mm) dd = drow[mm];
      const T dinv = (!kLower && live && gl < 6) ? T(1) / dd : T(0);
      // the compiler would otherwise sink the division (a dozen dependent
      // instructions) past the waits onto the critical path: 400x400 plate,
      // AMD + ILU(96), L+U pair 5.78 -> 5.69 ms
      asm volatile("" ::"v"(dinv));
      // Per pass of up to kGather blocks: the first kSweepIter*kS blocks'
      // values are loaded before anything waits; the pass's column indices
      // are read (lane q: blocks q, q+G, ...) and the dependencies' entries
      // are loaded at once, polled until no sentinel is left, and parked in
      // LDS; then lanes (t, r) accumulate A[q][r][:] . x[col_q][:].
#if GPUFEM_SWEEP_CRIT
      // wait for the critical dependency alone (one load per row group per
      // poll) before the row's dependencies are gathered: a wavefront that
      // is levels ahead then adds little polling traffic
      {
        const int cj = (crit >= 0 && !cached(crit)) ? crit : -1;
        B v = cj >= 0 ? load_bits(x + 6 * (size_t)cj) : B(0);
        unsigned spins = 0;
        int nap = 1;
        while (__any(v == Sentinel<T>::value)) {
          if (++spins > kSpinLimit) {
            if (lane == 0)
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
            break;
          }
          for (int z = 0; z < nap; z++) __builtin_amdgcn_s_sleep(1);
          nap = nap * 2 > GPUFEM_SWEEP_BACKOFF ? GPUFEM_SWEEP_BACKOFF : nap * 2;
          if (v == Sentinel<T>::value) v = load_bits(x + 6 * (size_t)cj);
        }
      }
#endif
      T acc = T(0);
      const int nb = k1 - k0;
      // The row's latest dependency is usually its nearest row: the last
      // block of an L row, the first of a U row.  In rows of more than one
      // pass the kNearPass blocks next to the diagonal form the last pass,
      // so that little work is left once the latest dependency arrives; the
      // rest go in passes of kGather blocks, U rows taking theirs last to
      // first, so that the passes that wait longest come last and the
      // others overlap with the wait.
      const bool split = kNearPass > 0 && kRows == 1 && nb > kGather;
      const int nfar = split ? nb - kNearPass : nb;
      const int nfp = (nfar + kGather - 1) / kGather;
      int npass = nfp + (split ? 1 : 0);
      // every row group of the wavefront runs the same number of passes
#pragma unroll
      for (int o = G; o < 64; o <<= 1)
        npass = max(npass, __shfl_xor(npass, o, 64));
      for (int pi = 0; pi < npass; pi++) {
        int g0b, ng;
        if (split && pi == npass - 1) {
          g0b = kLower ? k1 - kNearPass : k0;
          ng = kNearPass;
        } else {
          const int j = kLower ? pi : nfp - 1 - pi;
          const int fb = kLower ? k0 : k0 + (split ? kNearPass : 0);
          const int fe = kLower ? (split ? k1 - kNearPass : k1) : k1;
          g0b = fb + j * kGather;
          ng = max(0, min(kGather, fe - g0b));
        }
        T a[kSweepIter][6];
#pragma unroll
        for (int it = 0; it < kSweepIter; it++) {
          const int q = kS * it + t;
          const bool ok = act && q < ng;
          const T* ap = lu + 36 * (size_t)(ok ? g0b + q : 0) + 6 * r;
#pragma unroll
          for (int c = 0; c < 6; c++) a[it][c] = ok ? ap[c] : T(0);
        }
        // The pass's dependency entries, kE per lane.  One row per
        // wavefront: lane l holds entries e = l + 64 k of the pass's 6 ng
        // (dependency e / 6, component e % 6), so the six entries of a
        // dependency sit in consecutive lanes and each load instruction
        // touches a sixth as many cache lines as with a dependency per lane;
        // several rows per wavefront: lane gl holds the six entries of
        // dependencies gl + G h.  400x400 plate, AMD + ILU(96), L+U pair:
        // 5.69 ms with a dependency per lane, 5.26 ms with its entries in
        // consecutive lanes (each poll of a waiting row is then one or two
        // requests instead of six).
        constexpr bool kT = kRows == 1;
        constexpr int kE = 6 * kH;
        int jq[kE], jc[kE], jj[kE];
        B xb[kE];
#pragma unroll
        for (int k = 0; k < kE; k++) {
          if constexpr (kT) {
            jq[k] = (lane + 64 * k) / 6;
            jc[k] = (lane + 64 * k) % 6;
          } else {
            jq[k] = gl + G * (k / 6);
            jc[k] = k % 6;
          }
          jj[k] = jq[k] < ng ? cols[g0b + jq[k]] : -1;
        }
#pragma unroll
        for (int k = 0; k < kE; k++)
          xb[k] = (jj[k] >= 0 && !cached(jj[k]))
                      ? load_bits(x + 6 * (size_t)jj[k] + jc[k])
                      : B(0);
        unsigned spins = 0;
        int nap = 1;
        for (;;) {
          bool pending = false;
#pragma unroll
          for (int k = 0; k < kE; k++)
            pending |= xb[k] == Sentinel<T>::value;
          if (!__any(pending)) break;
          if (++spins > kSpinLimit) {
            if (lane == 0)
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
            for (int k = 0; k < kE; k++) xb[k] = B(0);
            break;
          }
          // back off while the dependencies are not ready: many waiting
          // wavefronts polling at full rate load the memory system
          for (int z = 0; z < nap; z++) __builtin_amdgcn_s_sleep(1);
          nap = nap * 2 > GPUFEM_SWEEP_BACKOFF ? GPUFEM_SWEEP_BACKOFF : nap * 2;
#pragma unroll
          for (int k = 0; k < kE; k++)
            if (xb[k] == Sentinel<T>::value)
              xb[k] = load_bits(x + 6 * (size_t)jj[k] + jc[k]);
        }
        // park the dependencies in this row group's LDS slice
#pragma unroll
        for (int k = 0; k < kE; k++) {
          if (jq[k] < ng) {
            T v = __builtin_bit_cast(T, xb[k]);
            if (cached(jj[k])) {
              // cx[jc[k]] without a dynamically indexed register array
              v = cx[0];
#pragma unroll
              for (int cc = 1; cc < 6; cc++)
                if (jc[k] == cc) v = cx[cc];
            }
            xs[jq[k]][jc[k]] = v;
          }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int it = 0; it < kSweepIter; it++) {
          const int q = kS * it + t;
          if (act && q < ng) {
#pragma unroll
            for (int c = 0; c < 6; c++) acc += a[it][c] * xs[q][c];
          }
        }
        for (int q = kS * kSweepIter + t; act && q < ng; q += kS) {
          const T* ap = lu + 36 * (size_t)(g0b + q) + 6 * r;
#pragma unroll
          for (int c = 0; c < 6; c++) acc += ap[c] * xs[q][c];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
      // sum the slots of scalar row r into lane r of the group
      T rs = T(0);
#pragma unroll
      for (int u = 0; u < kS; u++) rs += __shfl(acc, g0 + 6 * u + (gl % 6), 64);
      s -= (gl < 6) ? rs : T(0);
      // the diagonal block solve is on the critical path of every level:
      // with one row per wavefront the source lane is uniform and
      // v_readlane broadcasts it (400x400 plate, AMD + ILU(96): L+U pair
      // 7.75 -> 7.34 ms against ds_bpermute shuffles)
      auto bcast = [&](T v, int l) {
        if constexpr (kRows == 1)
          return read_lane(v, l);
        else
          return __shfl(v, g0 + l, 64);
      };
      if (kLower) {
#pragma unroll
        for (int mm = 0; mm < 5; mm++) {
          const T ym = bcast(s, mm);
          if (gl < 6 && gl > mm) s -= drow[mm] * ym;
        }
      } else {
#pragma unroll
        for (int mm = 5; mm >= 0; mm--) {
          if (gl == mm) s *= dinv;
          const T ym = bcast(s, mm);
          if (gl < 6 && gl < mm) s -= drow[mm] * ym;
        }
      }
      if (live && gl < 6) store_value(x + 6 * (size_t)i + gl, s);
      if (kUseCache) {
        cid = i;
#pragma unroll
        for (int c = 0; c < 6; c++) cx[c] = read_lane(s, c);
      }
    }
    tn = __shfl(tn, 0, 64);
    kc = (long long)(tn - base);
  }
}

}  // namespace
}  // namespace gf

extern "C" {

int gf_ilu_create(gf_context ctx, int n, const int* rowp, const int* cols,
                  gf_ilu* out) {
  if (!ctx || n <= 0 || !rowp || !cols || !out)
    return gf::set_error(ctx, GF_ERR_INVALID, "ilu_create: bad args");
  std::vector<int> diag(n, -1);
  int maxlen = 0;
  for (int i = 0; i < n; i++) {
    maxlen = std::max(maxlen, rowp[i + 1] - rowp[i]);
    for (int k = rowp[i]; k < rowp[i + 1]; k++) {
      if (cols[k] < 0 || cols[k] >= n || (k > rowp[i] && cols[k] <= cols[k - 1]))
        return gf::set_error(ctx, GF_ERR_INVALID,
                             "ilu_create: row %d not sorted/in range", i);
      if (cols[k] == i) diag[i] = k;
    }
    if (diag[i] < 0)
      return gf::set_error(ctx, GF_ERR_INVALID,
                           "ilu_create: row %d has no diagonal block", i);
  }
  if (maxlen > gf::kMaxRowBlocks)
    return gf::set_error(ctx, GF_ERR_INVALID,
                         "ilu_create: row length %d exceeds %d blocks", maxlen,
                         gf::kMaxRowBlocks);
  auto* h = new gf_ilu_s;
  h->ctx = ctx;
  h->n = n;
  h->nnzb = rowp[n];
  h->max_row_len = maxlen;
  std::vector<int> lrows, urows;
  gf::level_schedule(n, rowp, cols, true, h->lptr, lrows);
  gf::level_schedule(n, rowp, cols, false, h->uptr, urows);
  h->lmaxlen.assign(h->lptr.size() - 1, 0);
  for (size_t l = 0; l + 1 < h->lptr.size(); l++)
    for (int q = h->lptr[l]; q < h->lptr[l + 1]; q++)
      h->lmaxlen[l] = std::max(h->lmaxlen[l],
                               rowp[lrows[q] + 1] - rowp[lrows[q]]);
  // chunks: runs of rows of one level, sized so that a level still spreads
  // over the resident wavefronts (one row per chunk for small levels)
  const int target_waves =
      GPUFEM_SWEEP_WAVES_PER_CU * std::max(ctx->num_cus, 1);
  const char* cenv = getenv("GPUFEM_TRSV_MINCHUNK");
  const int min_chunk = cenv ? std::max(1, atoi(cenv)) : 1;
  auto chunks = [&](const std::vector<int>& lp, std::vector<int>& out) {
    out.assign(1, 0);
    for (size_t l = 0; l + 1 < lp.size(); l++) {
      const int cnt = lp[l + 1] - lp[l];
      const int csz =
          std::max(std::min(min_chunk, (cnt + 63) / 64),
                   std::max(1, std::min(16, cnt / target_waves)));
      for (int a = lp[l]; a < lp[l + 1]; a += csz)
        out.push_back(std::min(a + csz, lp[l + 1]));
    }
  };
  // chains: maximal runs of consecutive rows (in sweep order) where each row
  // depends on the one before it
  auto chains = [&](bool lower, std::vector<int>& out) {
    out.assign(1, 0);
    for (int t = 1; t < n; t++) {
      const int i = lower ? t : n - 1 - t;
      bool dep = false;
      if (lower) {
        for (int k = rowp[i]; k < diag[i]; k++) dep |= cols[k] == i - 1;
      } else {
        for (int k = diag[i] + 1; k < rowp[i + 1]; k++) dep |= cols[k] == i + 1;
      }
      if (!dep) out.push_back(t);
    }
    out.push_back(n);
  };
  // Rows are handed out by level, in chunks of rows of one level.  Handing
  // out whole chains in row order instead (GPUFEM_TRSV=chains; each
  // wavefront forwards a chain's rows in registers) serialises the rows of a
  // chain behind each other's memory latency and measured 1.7x slower on the
  // natural-order plates.
  std::vector<int> lch, uch, lcn, ucn;
  const char* tenv = getenv("GPUFEM_TRSV");
  const bool use_chains = tenv && std::string(tenv) == "chains";
  if (use_chains) {
    chains(true, lcn);
    chains(false, ucn);
  }
  h->lchains = use_chains;
  h->uchains = use_chains;
  if (h->lchains) lch = lcn; else chunks(h->lptr, lch);
  if (h->uchains) uch = ucn; else chunks(h->uptr, uch);
  h->nlchunk = (int)lch.size() - 1;
  h->nuchunk = (int)uch.size() - 1;
  if (!use_chains) {
    auto levels_of = [](const std::vector<int>& ch, const std::vector<int>& lp,
                        std::vector<int>& out) {
      out.resize(ch.size() - 1);
      size_t l = 0;
      for (size_t c = 0; c + 1 < ch.size(); c++) {
        while (l + 2 < lp.size() && lp[l + 1] <= ch[c]) l++;
        out[c] = (int)l;
      }
    };
    levels_of(lch, h->lptr, h->lchunk_level);
    levels_of(uch, h->uptr, h->uchunk_level);
  }
  // row descriptors in hand-out order: one 16-byte load per row
  // {row, first block, end block, critical dependency}: the diagonal block
  // is the end (L) or one before the first (U); the critical dependency is
  // the row's dependency of the highest level, latest in hand-out order
  // (-1 for none), which is usually the last one to be solved
  std::vector<int4> lm(n), um(n);
  {
    std::vector<int> lev(n), ord(n);
    auto crit_of = [&](const std::vector<int>& lp, const std::vector<int>& rows,
                       bool lower, std::vector<int>& out) {
      for (size_t l = 0; l + 1 < lp.size(); l++)
        for (int q = lp[l]; q < lp[l + 1]; q++) lev[rows[q]] = (int)l;
      for (int q = 0; q < n; q++) ord[use_chains ? (lower ? q : n - 1 - q) : rows[q]] = q;
      out.assign(n, -1);
      for (int i = 0; i < n; i++) {
        const int k0 = lower ? rowp[i] : diag[i] + 1;
        const int k1 = lower ? diag[i] : rowp[i + 1];
        int best = -1;
        for (int k = k0; k < k1; k++) {
          const int j = cols[k];
          if (best < 0 || lev[j] > lev[best] ||
              (lev[j] == lev[best] && ord[j] > ord[best]))
            best = j;
        }
        out[i] = best;
      }
    };
    std::vector<int> lcrit, ucrit;
    crit_of(h->lptr, lrows, true, lcrit);
    crit_of(h->uptr, urows, false, ucrit);
    for (int p = 0; p < n; p++) {
      const int li = use_chains ? p : lrows[p];
      lm[p] = make_int4(li, rowp[li], diag[li], lcrit[li]);
      const int ui = use_chains ? n - 1 - p : urows[p];
      um[p] = make_int4(ui, diag[ui] + 1, rowp[ui + 1], ucrit[ui]);
    }
  }
  auto up = [&](int** dst, const int* src, size_t cnt) -> int {
    GF_TRY(gf::dmalloc(ctx, (void**)dst, sizeof(int) * (cnt ? cnt : 1)));
    GF_HIP(ctx, hipMemcpyAsync(*dst, src, sizeof(int) * cnt,
                               hipMemcpyHostToDevice, ctx->stream));
    return GF_OK;
  };
  int st = GF_OK;
  if ((st = up(&h->rowp, rowp, n + 1)) || (st = up(&h->cols, cols, rowp[n])) ||
      (st = up(&h->diag, diag.data(), n)) ||
      (st = up(&h->lrows, lrows.data(), n)) ||
      (st = up(&h->urows, urows.data(), n)) ||
      (st = up(&h->lchunk, lch.data(), lch.size())) ||
      (st = up((int**)&h->lmeta, (const int*)lm.data(), 4 * (size_t)n)) ||
      (st = up((int**)&h->umeta, (const int*)um.data(), 4 * (size_t)n)) ||
      (st = up(&h->uchunk, uch.data(), uch.size())) ||
      (st = gf::dmalloc(ctx, (void**)&h->pivot, sizeof(int))) ||
      (st = gf::dmalloc(ctx, (void**)&h->counter, sizeof(unsigned long long))) ||
      (st = gf::dmalloc(ctx, (void**)&h->err, 2 * sizeof(int))) ||
      (st = gf::dmalloc(ctx, (void**)&h->done, sizeof(int) * (size_t)n)) ||
      (st = gf::check_hip(ctx, hipMemsetAsync(h->done, 0, sizeof(int) * (size_t)n,
                                              ctx->stream), "memset")) ||
      (st = gf::check_hip(ctx, hipMemsetAsync(h->counter, 0, sizeof(unsigned long long),
                                              ctx->stream), "memset")) ||
      (st = gf::check_hip(ctx, hipMemsetAsync(h->err, 0, 2 * sizeof(int), ctx->stream),
                          "memset"))) {
    gf_ilu_destroy(h);
    return st;
  }
  {
    // enough wavefronts to cover the device, capped by the chunks and by
    // eight mean level widths: wavefronts further ahead than that mostly
    // wait, and on small matrices they would take the device from
    // concurrent solves on other streams
    const int nlev = (int)std::max(h->lptr.size(), h->uptr.size()) - 1;
    const int mean_w = (n + std::max(nlev, 1) - 1) / std::max(nlev, 1);
    const int waves = std::max(
        1, std::min({std::max(h->nlchunk, h->nuchunk), target_waves,
                     std::max(64, 8 * mean_w)}));
    h->grid = (waves + 3) / 4;
    const char* env = getenv("GPUFEM_TRSV");
    h->level_sweeps = env && std::string(env) == "level";
  }
  st = gf::check_hip(ctx, hipStreamSynchronize(ctx->stream), "ilu upload");
  if (st) {
    gf_ilu_destroy(h);
    return st;
  }
  *out = h;
  return GF_OK;
}

int gf_ilu_destroy(gf_ilu h) {
  if (!h) return GF_OK;
  gf::dfree(h->ctx, h->rowp);
  gf::dfree(h->ctx, h->cols);
  gf::dfree(h->ctx, h->diag);
  gf::dfree(h->ctx, h->lrows);
  gf::dfree(h->ctx, h->urows);
  gf::dfree(h->ctx, h->pivot);
  gf::dfree(h->ctx, h->lchunk);
  gf::dfree(h->ctx, h->lmeta);
  gf::dfree(h->ctx, h->umeta);
  gf::dfree(h->ctx, h->uchunk);
  gf::dfree(h->ctx, h->counter);
  gf::dfree(h->ctx, h->err);
  gf::dfree(h->ctx, h->done);
  delete h;
  return GF_OK;
}

int gf_ilu_info(gf_ilu h, int* nnzb, int* nl, int* nu) {
  if (!h) return GF_ERR_INVALID;
  if (getenv("GPUFEM_ILU_VERBOSE"))
    fprintf(stderr,
            "ilu: n=%d nnzb=%d levels L=%d U=%d %s L=%d %s U=%d\n", h->n,
            h->nnzb, (int)h->lptr.size() - 1, (int)h->uptr.size() - 1,
            h->lchains ? "chains" : "chunks", h->nlchunk,
            h->uchains ? "chains" : "chunks", h->nuchunk);
  if (nnzb) *nnzb = h->nnzb;
  if (nl) *nl = (int)h->lptr.size() - 1;
  if (nu) *nu = (int)h->uptr.size() - 1;
  return GF_OK;
}
const int* gf_ilu_device_rowp(gf_ilu h) { return h ? h->rowp : nullptr; }

int gf_ilu_check(gf_ilu h) {
  if (!h) return GF_ERR_INVALID;
  int e = 0;
  GF_HIP(h->ctx, hipMemcpyAsync(&e, h->err, sizeof(int), hipMemcpyDeviceToHost,
                                h->ctx->stream));
  GF_HIP(h->ctx, hipStreamSynchronize(h->ctx->stream));
  if (e) {
    const int zero = 0;
    hipMemcpyAsync(h->err, &zero, sizeof(int), hipMemcpyHostToDevice,
                   h->ctx->stream);
    hipStreamSynchronize(h->ctx->stream);
    return gf::set_error(h->ctx, GF_ERR_TIMEOUT,
                         "triangular sweep: a dependency wait expired");
  }
  return GF_OK;
}
const int* gf_ilu_device_cols(gf_ilu h) { return h ? h->cols : nullptr; }

}  // extern "C"

namespace gf {
unsigned long long sweep_sentinel_bits(bool fp64) {
  return fp64 ? Sentinel<double>::value : Sentinel<float>::value;
}
int* ilu_sweep_err(gf_ilu h) { return h ? h->err : nullptr; }
gf_context ilu_context(gf_ilu h) { return h ? h->ctx : nullptr; }

int ilu_view(gf_ilu h, IluView* v) {
  if (!h || !v || h->lchains || h->uchains) return GF_ERR_INVALID;
  v->n = h->n;
  v->lmeta = h->lmeta;
  v->umeta = h->umeta;
  v->lchunk = h->lchunk;
  v->uchunk = h->uchunk;
  v->cols = h->cols;
  v->lchunk_level = &h->lchunk_level;
  v->uchunk_level = &h->uchunk_level;
  v->grid = h->grid;
  return GF_OK;
}

__global__ void reset_counters_kernel(unsigned long long* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0ull;
}

// One L and one U sync-free sweep over several systems: each chunk list is
// the systems' chunks merged level by level (see IluView).  counters[0..1]
// are zeroed first by a kernel (no memset: the launches may be captured in
// a graph), so the tickets start at 0 in every replay.
template <typename T>
int launch_batch_sweeps(hipStream_t s, const SweepSlot<T>* slots,
                        const int2* lmerged, int nl, const int2* umerged,
                        int nu, int grid, unsigned long long* counters,
                        int* err) {
  hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters,
                     2);
  hipLaunchKernelGGL((trsv_syncfree_kernel<T, true, 1, true>), dim3(grid),
                     dim3(kSweepThreads), 0, s, (const int4*)nullptr,
                     (const int*)nullptr, (const int*)nullptr, nl,
                     (const T*)nullptr, (const T*)nullptr, (T*)nullptr,
                     (T*)nullptr, counters, 0ull, err, slots, lmerged);
  hipLaunchKernelGGL((trsv_syncfree_kernel<T, false, 1, true>), dim3(grid),
                     dim3(kSweepThreads), 0, s, (const int4*)nullptr,
                     (const int*)nullptr, (const int*)nullptr, nu,
                     (const T*)nullptr, (const T*)nullptr, (T*)nullptr,
                     (T*)nullptr, counters + 1, 0ull, err, slots, umerged);
  return hipGetLastError() == hipSuccess ? GF_OK : GF_ERR_DEVICE;
}
template int launch_batch_sweeps<double>(hipStream_t, const SweepSlot<double>*,
                                         const int2*, int, const int2*, int,
                                         int, unsigned long long*, int*);
template int launch_batch_sweeps<float>(hipStream_t, const SweepSlot<float>*,
                                        const int2*, int, const int2*, int,
                                        int, unsigned long long*, int*);
}  // namespace gf

namespace {
template <typename T>
int ilu_factor(gf_ilu h, T* vals, int* zero_pivot) {
  if (!h || !vals) return GF_ERR_INVALID;
  gf_context ctx = h->ctx;
  const int big = 0x7fffffff;
  GF_HIP(ctx, hipMemcpyAsync(h->pivot, &big, sizeof(int),
                             hipMemcpyHostToDevice, ctx->stream));
  // rows whose values fit in LDS (dynamic LDS per level: its longest row)
  // are factored there; the other levels run ilu_row_kernel
  static size_t lds_cap[2] = {0, 0};
  static size_t batch_cap[2] = {0, 0};
  static std::once_flag once[2];
  const int ti = sizeof(T) == 8 ? 0 : 1;
  std::call_once(once[ti], [&] {
    const void* fns[4] = {
        (const void*)gf::ilu_row_lds_kernel<T, false>,
        (const void*)gf::ilu_row_lds_kernel<T, true>,
        (const void*)gf::ilu_row_lds_kernel<T, true, gf::kIluShortThreads>,
        (const void*)gf::ilu_row_lds_kernel<T, false, gf::kIluShortThreads>};
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess) return;
    const size_t total = std::min<size_t>(prop.sharedMemPerBlock, 160 * 1024);
    size_t cap = total;
    for (const void* fn : fns) {
      hipFuncAttributes fa{};
      if (hipFuncGetAttributes(&fa, fn) != hipSuccess) return;
      if (total <= fa.sharedSizeBytes + 1024) return;
      cap = std::min(cap, total - fa.sharedSizeBytes);
    }
    for (const void* fn : fns)
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)cap) != hipSuccess) {
        hipGetLastError();
        cap = std::min<size_t>(cap, 48 * 1024);
      }
    // the global-memory kernels' column and position lists
    const void* gfns[2] = {(const void*)gf::ilu_row_kernel<T, false>,
                           (const void*)gf::ilu_row_kernel<T, true>};
    for (const void* fn : gfns)
      if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)(3 * sizeof(int) * gf::kMaxRowBlocks)) !=
          hipSuccess)
        hipGetLastError();
    const void* bfns[2] = {(const void*)gf::ilu_row_batch_kernel<T, false>,
                           (const void*)gf::ilu_row_batch_kernel<T, true>};
    size_t bcap = total;
    for (const void* fn : bfns) {
      hipFuncAttributes fa{};
      if (hipFuncGetAttributes(&fa, fn) != hipSuccess ||
          total <= fa.sharedSizeBytes) {
        bcap = 0;
        break;
      }
      bcap = std::min(bcap, total - fa.sharedSizeBytes);
    }
    for (const void* fn : bfns)
      if (bcap > 0 &&
          hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bcap) != hipSuccess) {
        hipGetLastError();
        bcap = 0;
      }
    batch_cap[ti] = bcap;
    lds_cap[ti] = cap;
  });
  const int maxlen = h->max_row_len;
  const size_t gdyn = 3 * sizeof(int) * (size_t)maxlen;
  const bool no_lds = getenv("GPUFEM_ILU_GLOBAL") != nullptr;
  // rows in global memory: pivots applied in batches (ilu_row_batch_kernel)
  // of as many pivots as the LDS holds position lists for (at most
  // kIluBatch); GPUFEM_ILU_BATCH_N lowers the batch (a test hook), and
  // GPUFEM_ILU_GLOBAL_KERNEL=pivot selects ilu_row_kernel (one pass over the
  // row per pivot)
  int nbatch = 0;
  {
    // two workgroups per CU where the LDS allows
    const size_t cap2 = std::min(batch_cap[ti], (size_t)64 * 1024);
    for (int nbt = gf::kIluBatch; nbt >= 1 && !nbatch; nbt--)
      if (gf::batch_lds_per_pos(nbt) * maxlen <= cap2) nbatch = nbt;
    for (int nbt = gf::kIluBatch; nbt >= 1 && !nbatch; nbt--)
      if (gf::batch_lds_per_pos(nbt) * maxlen <= batch_cap[ti]) nbatch = nbt;
    if (const char* benv = getenv("GPUFEM_ILU_BATCH_N"))
      nbatch = std::min(nbatch, std::max(1, atoi(benv)));
    const char* kenv = getenv("GPUFEM_ILU_GLOBAL_KERNEL");
    if (kenv && std::string(kenv) == "pivot") nbatch = 0;
  }
  const size_t bdyn = gf::batch_lds_per_pos(nbatch) * (size_t)maxlen;
  auto launch_global = [&](bool flow, const int* rows, int cnt,
                           const gf::Flow& f) {
    if (nbatch > 0) {
      if (flow)
        hipLaunchKernelGGL((gf::ilu_row_batch_kernel<T, true>), dim3(cnt),
                           dim3(gf::kIluThreads), bdyn, ctx->stream, rows,
                           cnt, h->rowp, h->cols, h->diag, vals, h->pivot, f,
                           nbatch, maxlen);
      else
        hipLaunchKernelGGL((gf::ilu_row_batch_kernel<T, false>), dim3(cnt),
                           dim3(gf::kIluThreads), bdyn, ctx->stream, rows,
                           cnt, h->rowp, h->cols, h->diag, vals, h->pivot, f,
                           nbatch, maxlen);
    } else if (flow) {
      hipLaunchKernelGGL((gf::ilu_row_kernel<T, true>), dim3(cnt),
                         dim3(gf::kIluThreads), gdyn, ctx->stream, rows, cnt,
                         h->rowp, h->cols, h->diag, vals, h->pivot, f, maxlen);
    } else {
      hipLaunchKernelGGL((gf::ilu_row_kernel<T, false>), dim3(cnt),
                         dim3(gf::kIluThreads), gdyn, ctx->stream, rows, cnt,
                         h->rowp, h->cols, h->diag, vals, h->pivot, f, maxlen);
    }
  };
  // GPUFEM_ILU_LDS_CAP (bytes) lowers the LDS budget of a row, so that one
  // factorisation mixes LDS and global-memory rows (a test hook)
  size_t cap = lds_cap[ti];
  if (const char* cenv = getenv("GPUFEM_ILU_LDS_CAP"))
    cap = std::min<size_t>(cap, (size_t)std::max(0L, atol(cenv)));
  const char* senv = getenv("GPUFEM_ILU_SCHEDULE");
  const bool short_off = getenv("GPUFEM_ILU_SHORT_OFF") != nullptr;
  const bool by_level = senv && std::string(senv) == "level";
  const int nl = (int)h->lptr.size() - 1;
  // row values, column indices and pivot bounds of the level's longest
  // row, and two position lists as long as the longest row
  auto dyn_of = [&](int l) {
    return (size_t)h->lmaxlen[l] * (36 * sizeof(T) + 3 * sizeof(int)) +
           2 * sizeof(int) * (size_t)maxlen;
  };
  if (by_level) {
    for (int l = 0; l < nl; l++) {
      const int cnt = h->lptr[l + 1] - h->lptr[l];
      const size_t dyn = dyn_of(l);
      if (!no_lds && dyn <= cap)
        hipLaunchKernelGGL((gf::ilu_row_lds_kernel<T, false>), dim3(cnt),
                           dim3(gf::kIluThreads), dyn, ctx->stream,
                           h->lrows + h->lptr[l], cnt, h->rowp, h->cols,
                           h->diag, vals, h->pivot, gf::Flow{}, maxlen);
      else
        launch_global(false, h->lrows + h->lptr[l], cnt, gf::Flow{});
    }
  } else {
    // Dataflow: one launch per phase.  Levels before the first level with a
    // row too long for LDS run the LDS kernel, in phases whose LDS need
    // stays within twice that of their first level or 40 KB (so the wide
    // early levels keep several workgroups per CU); every level from there
    // on runs the global-memory kernel in one phase (row lengths of a
    // complete factorisation oscillate around the LDS limit, and a launch
    // boundary is a barrier for every row).  Launch boundaries order the
    // phases.
    if (++h->epoch <= 0) {
      GF_HIP(ctx, hipMemsetAsync(h->done, 0, sizeof(int) * (size_t)h->n,
                                 ctx->stream));
      h->epoch = 1;
    }
    int first_global = 0;
    while (first_global < nl && !no_lds && dyn_of(first_global) <= cap)
      first_global++;
    // longest row up to each level: it bounds the U part of every pivot row
    // of the rows up to there
    std::vector<int> pmax(nl + 1, 0);
    for (int l = 0; l < nl; l++) pmax[l + 1] = std::max(pmax[l], h->lmaxlen[l]);
    int l0 = 0;
    while (l0 < nl) {
      const bool lds = l0 < first_global;
      const size_t d0 = dyn_of(l0);
      const size_t lim = std::max<size_t>(2 * d0, 40 * 1024);
      size_t dyn = d0;
      int l1 = l0 + 1;
      if (lds) {
        for (; l1 < first_global && dyn_of(l1) <= lim; l1++)
          dyn = std::max(dyn, dyn_of(l1));
      } else {
        l1 = nl;
      }
      const int r0 = h->lptr[l0], cnt = h->lptr[l1] - r0;
      gf::Flow f;
      f.counter = h->counter;
      f.base = h->base;
      f.done = h->done;
      f.epoch = h->epoch;
      f.err = h->err + 1;
      if (lds && pmax[l1] <= gf::kIluShortRow && !short_off) {
        // short rows: 256-thread workgroups, position lists as long as the
        // longest row so far
        int lmax = 0;
        for (int l = l0; l < l1; l++) lmax = std::max(lmax, h->lmaxlen[l]);
        const size_t dyn_s =
            (size_t)lmax * (36 * sizeof(T) + 3 * sizeof(int)) +
            2 * sizeof(int) * (size_t)pmax[l1];
        hipLaunchKernelGGL(
            (gf::ilu_row_lds_kernel<T, true, gf::kIluShortThreads>),
            dim3(cnt), dim3(gf::kIluShortThreads), dyn_s, ctx->stream,
            h->lrows + r0, cnt, h->rowp, h->cols, h->diag, vals, h->pivot, f,
            pmax[l1]);
      } else if (lds) {
        hipLaunchKernelGGL((gf::ilu_row_lds_kernel<T, true>), dim3(cnt),
                           dim3(gf::kIluThreads), dyn, ctx->stream,
                           h->lrows + r0, cnt, h->rowp, h->cols, h->diag, vals,
                           h->pivot, f, maxlen);
      } else {
        launch_global(true, h->lrows + r0, cnt, f);
      }
      h->base += (unsigned long long)cnt;
      l0 = l1;
    }
  }
  GF_TRY(gf::check_launch(ctx, "ilu_row_kernel"));
  int piv = big, werr = 0;
  GF_HIP(ctx, hipMemcpyAsync(&piv, h->pivot, sizeof(int),
                             hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP(ctx, hipMemcpyAsync(&werr, h->err + 1, sizeof(int),
                             hipMemcpyDeviceToHost, ctx->stream));
  GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (werr) {
    const int zero = 0;
    GF_HIP(ctx, hipMemcpyAsync(h->err + 1, &zero, sizeof(int),
                               hipMemcpyHostToDevice, ctx->stream));
    GF_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return gf::set_error(ctx, GF_ERR_TIMEOUT,
                         "ILU factorisation: a dependency wait expired");
  }
  if (zero_pivot) *zero_pivot = piv == big ? -1 : piv;
  if (piv != big)
    return gf::set_error(ctx, GF_ERR_ZERO_PIVOT,
                         "ILU: zero pivot in diagonal block of row %d", piv);
  return GF_OK;
}

template <typename T>
int ilu_apply(gf_ilu h, const T* lu, const T* b, T* tmp, T* x) {
  if (!h || !lu || !b || !tmp || !x || b == tmp || tmp == x)
    return GF_ERR_INVALID;
  gf_context ctx = h->ctx;
  if (!h->level_sweeps) {
    const unsigned long long waves = 4ull * h->grid;
    const size_t N = 6 * (size_t)h->n;
    hipLaunchKernelGGL(gf::fill_sentinel_kernel<T>, dim3(gf::div_up(N, 256)),
                       dim3(256), 0, ctx->stream, N, tmp);
    // one row per wavefront: two rows of a level per wavefront
    // (trsv_syncfree_kernel<T, L, 2>) measured slower, 4.8 vs 3.5 ms a pair
    // on the 400x400 plate
    hipLaunchKernelGGL((gf::trsv_syncfree_kernel<T, true, 1>), dim3(h->grid),
                       dim3(gf::kSweepThreads), 0, ctx->stream, h->lmeta,
                       h->cols, h->lchunk, h->nlchunk, lu, b, tmp, x,
                       h->counter, h->base, h->err);
    h->base += (unsigned long long)h->nlchunk + waves;
    hipLaunchKernelGGL((gf::trsv_syncfree_kernel<T, false, 1>), dim3(h->grid),
                       dim3(gf::kSweepThreads), 0, ctx->stream, h->umeta,
                       h->cols, h->uchunk, h->nuchunk, lu, tmp, x, (T*)nullptr,
                       h->counter, h->base, h->err);
    h->base += (unsigned long long)h->nuchunk + waves;
    return gf::check_launch(ctx, "trsv_syncfree_kernel");
  }
  const int nl = (int)h->lptr.size() - 1;
  for (int l = 0; l < nl; l++) {
    const int cnt = h->lptr[l + 1] - h->lptr[l];
    hipLaunchKernelGGL((gf::trsv_level_kernel<T, true>),
                       dim3(gf::div_up(cnt, 40)), dim3(256), 0, ctx->stream,
                       h->lrows + h->lptr[l], cnt, h->rowp, h->cols, h->diag,
                       lu, b, tmp);
  }
  const int nu = (int)h->uptr.size() - 1;
  for (int l = 0; l < nu; l++) {
    const int cnt = h->uptr[l + 1] - h->uptr[l];
    hipLaunchKernelGGL((gf::trsv_level_kernel<T, false>),
                       dim3(gf::div_up(cnt, 40)), dim3(256), 0, ctx->stream,
                       h->urows + h->uptr[l], cnt, h->rowp, h->cols, h->diag,
                       lu, tmp, x);
  }
  return gf::check_launch(ctx, "trsv_level_kernel");
}
}  // namespace

extern "C" {
int gf_bsr_ilu_factor_f64(gf_ilu h, double* vals, int* zp) {
  return ilu_factor<double>(h, vals, zp);
}
int gf_bsr_ilu_factor_f32(gf_ilu h, float* vals, int* zp) {
  return ilu_factor<float>(h, vals, zp);
}
int gf_bsr_ilu_apply_f64(gf_ilu h, const double* lu, const double* b,
                         double* tmp, double* x) {
  return ilu_apply<double>(h, lu, b, tmp, x);
}
int gf_bsr_ilu_apply_f32(gf_ilu h, const float* lu, const float* b,
                         float* tmp, float* x) {
  return ilu_apply<float>(h, lu, b, tmp, x);
}
}  // extern "C"
