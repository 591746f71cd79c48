// MITC4 Jacobian / residual / energy assembly kernels (gfx950).
//
// Replaces add_jacobian_gpu (include/base/elem_group.cuh:101-211), the
// residual path add_residual_cpu (include/base/elem_group.h:40-79) and
// add_energy (assembler.h:198-229).
//
// Layout: one 64-lane workgroup per element.
//   stage 1  lanes 0-3: node frames; lanes 4-8: tying-point geometry
//   stage 2  lanes 0-3: quadrature frames; lanes 4-8: tying normals
//   stage 3  lanes 0-23: column j of B_q (4 x 9) and D_q = detXd C B_q -> LDS
//   stage 4  all 64 lanes: a 3x3 tile of K_e = sum_q B_q^T D_q (8x8 tiles)
//   stage 5  scatter through elemIndMap (global FP64 atomics), residual
//            R_e = K_e u_e reduced through LDS.
#include <cstdlib>
#include <string>

#include <vector>

#include "common.h"
#include "mitc4.h"
#include "symbolic.h"

struct gf_asm_plan_s {
  gf_context ctx = nullptr;
  int nelems = 0, ncolors = 0;
  std::vector<int> color_ptr;
  int* elems = nullptr;  // elements in colour-major order (device)
  int* mask = nullptr;   // first-touch masks in the same order (device)
  bool zero_k = false, zero_res = false;  // untouched K blocks / rows exist
  int nnzb = 0, nnodes = 0;
};

namespace gf {
namespace {

struct ElemLDS {
  double X[12];
  double u[24];
  double data[4];
  mitc4::Geom g;
  double B[4][9][24];
  double D[4][9][24];
  double rpart[8][24];  // residual partial sums per tile column J
};

__device__ __forceinline__ void load_element(ElemLDS& s, int e, int lane,
                                             const int* conn,
                                             const double* xpts,
                                             const double* vars,
                                             const double* data) {
  if (lane < 12) {
    const int a = lane / 3, k = lane % 3;
    s.X[lane] = xpts[3 * (size_t)conn[4 * (size_t)e + a] + k];
  } else if (lane < 36) {
    const int i = lane - 12, a = i / 6, k = i % 6;
    s.u[i] = vars ? vars[6 * (size_t)conn[4 * (size_t)e + a] + k] : 0.0;
  } else if (lane < 40) {
    s.data[lane - 36] = data[4 * (size_t)e + (lane - 36)];
  }
}

__device__ __forceinline__ void element_geometry(ElemLDS& s, int lane) {
  if (lane < 4) {
    mitc4::geom_node(lane, s.X, s.g);
  } else if (lane < 9) {
    mitc4::geom_tying_x(lane - 4, s.X, s.g);
  }
  __syncthreads();
  if (lane < 4) {
    mitc4::geom_quad(lane, s.X, s.g);
  } else if (lane < 9) {
    mitc4::geom_tying_n(lane - 4, s.g);
  }
  __syncthreads();
}

__device__ __forceinline__ void element_bd(ElemLDS& s, int lane) {
  if (lane < 24) {
    mitc4::BColumnPrep prep;
    mitc4::b_column_prep(lane, s.g, prep);
#pragma unroll 1
    for (int q = 0; q < 4; q++) {
      double col[9], st[9];
      mitc4::b_column_q(lane, q, s.g, prep, col);
      mitc4::stress(s.data, col, st);
      const double sc = s.g.scale[q];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        s.B[q][k][lane] = col[k];
        s.D[q][k][lane] = sc * st[k];
      }
    }
  }
  __syncthreads();
}

// 3x3 tile (rows 3I.., cols 3J..) of K_e
__device__ __forceinline__ void element_tile(const ElemLDS& s, int I, int J,
                                             double acc[3][3]) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc[r][c] = 0.0;
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double b0 = s.B[q][k][3 * I], b1 = s.B[q][k][3 * I + 1],
                   b2 = s.B[q][k][3 * I + 2];
      const double d0 = s.D[q][k][3 * J], d1 = s.D[q][k][3 * J + 1],
                   d2 = s.D[q][k][3 * J + 2];
      acc[0][0] += b0 * d0;
      acc[0][1] += b0 * d1;
      acc[0][2] += b0 * d2;
      acc[1][0] += b1 * d0;
      acc[1][1] += b1 * d1;
      acc[1][2] += b1 * d2;
      acc[2][0] += b2 * d0;
      acc[2][1] += b2 * d1;
      acc[2][2] += b2 * d2;
    }
}

// kColored = false: every contribution is an FP64 atomic add (K and res
// zeroed by the caller).  kColored = true: the launch covers one colour of
// an element colouring (no two elements of a colour share a node), so no two
// lanes of the launch touch the same K entry or residual entry; each
// contribution is a plain store when the element is the first to touch the
// block (first-touch mask, bits 4a+b for K blocks, 16+a for residual rows)
// and a plain read-add-write otherwise.  No atomics and no zeroing pass.
template <bool kColored>
__global__ void __launch_bounds__(64)
    jacobian_kernel(int nelems, const int* __restrict__ elem_list,
                    const int* __restrict__ first_mask,
                    const int* __restrict__ conn,
                    const double* __restrict__ xpts,
                    const double* __restrict__ vars,
                    const double* __restrict__ data,
                    const int* __restrict__ elem_ind_map,
                    const int* __restrict__ perm, double* kvals, double* res) {
  __shared__ ElemLDS s;
  if ((int)blockIdx.x >= nelems) return;
  const int e = kColored ? elem_list[blockIdx.x] : (int)blockIdx.x;
  const int mask = kColored ? first_mask[blockIdx.x] : 0;
  const int lane = threadIdx.x;
  load_element(s, e, lane, conn, xpts, vars, data);
  __syncthreads();
  element_geometry(s, lane);
  element_bd(s, lane);
  const int I = lane / 8, J = lane % 8;
  double acc[3][3];
  element_tile(s, I, J, acc);
  // scatter: tile rows belong to node a = I/2, local rows 3*(I%2)..
  const int a = I / 2, b = J / 2, r0 = 3 * (I % 2), c0 = 3 * (J % 2);
  const size_t blk = (size_t)elem_ind_map[16 * (size_t)e + 4 * a + b];
  double* dst = kvals + 36 * blk;
  const bool first = (mask >> (4 * a + b)) & 1;
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double* d = &dst[6 * (r0 + r) + c0 + c];
      if (!kColored)
        unsafeAtomicAdd(d, acc[r][c]);
      else
        *d = first ? acc[r][c] : *d + acc[r][c];
    }
  if (res) {
    double p0 = 0, p1 = 0, p2 = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const double uc = s.u[3 * J + c];
      p0 += acc[0][c] * uc;
      p1 += acc[1][c] * uc;
      p2 += acc[2][c] * uc;
    }
    s.rpart[J][3 * I] = p0;
    s.rpart[J][3 * I + 1] = p1;
    s.rpart[J][3 * I + 2] = p2;
    __syncthreads();
    if (lane < 24) {
      double r = 0.0;
#pragma unroll
      for (int j = 0; j < 8; j++) r += s.rpart[j][lane];
      const int node = conn[4 * (size_t)e + lane / 6];
      const int row = perm ? perm[node] : node;
      double* d = &res[6 * (size_t)row + lane % 6];
      if (!kColored)
        unsafeAtomicAdd(d, r);
      else
        *d = ((mask >> (16 + lane / 6)) & 1) ? r : *d + r;
    }
  }
}

// ---------------------------------------------------------------------------
// Two elements per 64-lane wavefront (32 lanes each), quadrature points
// streamed through LDS: per point the element's lanes 0-23 form column j of
// B_q and D_q = detXd C B_q (the column's tying-strain part is formed once
// and kept in registers), then every lane adds two 3x3 tiles of
// B_q^T D_q.  LDS per element holds the geometry and one point's B and D
// (6.2 KB), so a compute unit keeps about three times more elements in
// flight than with all four points' B and D resident.  Scatter as in
// jacobian_kernel (atomic, or colour-ordered stores / read-add-writes).
struct ElemSlotQ {
  double X[12];
  double u[24];
  double data[4];
  mitc4::Geom g;
  double B[9][24];
  double D[9][24];
};

template <bool kColored>
__global__ void __launch_bounds__(64)
    jacobian2_kernel(int nelems, const int* __restrict__ elem_list,
                     const int* __restrict__ first_mask,
                     const int* __restrict__ conn,
                     const double* __restrict__ xpts,
                     const double* __restrict__ vars,
                     const double* __restrict__ data,
                     const int* __restrict__ elem_ind_map,
                     const int* __restrict__ perm, double* kvals,
                     double* res) {
  __shared__ ElemSlotQ S[2];
  __shared__ double rp[2][8][24];  // residual partials per tile column
  const int half = threadIdx.x / 32, l = threadIdx.x % 32;
  const int slot = blockIdx.x * 2 + half;
  const bool valid = slot < nelems;
  const int e = valid ? (kColored ? elem_list[slot] : slot) : 0;
  const int mask = (kColored && valid) ? first_mask[slot] : 0;
  ElemSlotQ& s = S[half];
  if (valid) {
    for (int v = l; v < 40; v += 32) {
      if (v < 12) {
        s.X[v] = xpts[3 * (size_t)conn[4 * (size_t)e + v / 3] + v % 3];
      } else if (v < 36) {
        const int i = v - 12;
        s.u[i] = vars ? vars[6 * (size_t)conn[4 * (size_t)e + i / 6] + i % 6]
                      : 0.0;
      } else {
        s.data[v - 36] = data[4 * (size_t)e + (v - 36)];
      }
    }
  }
  __syncthreads();
  if (valid) {
    if (l < 4)
      mitc4::geom_node(l, s.X, s.g);
    else if (l < 9)
      mitc4::geom_tying_x(l - 4, s.X, s.g);
  }
  __syncthreads();
  if (valid) {
    if (l < 4)
      mitc4::geom_quad(l, s.X, s.g);
    else if (l < 9)
      mitc4::geom_tying_n(l - 4, s.g);
  }
  __syncthreads();
  // tiles (I, J) and (I + 4, J) of the 8 x 8 tile grid
  const int I0 = l / 8, J = l % 8;
  double acc[2][3][3];
#pragma unroll
  for (int h = 0; h < 2; h++)
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[h][r][c] = 0.0;
  mitc4::BColumnPrep prep;
  if (valid && l < 24) mitc4::b_column_prep(l, s.g, prep);
#pragma unroll 1
  for (int q = 0; q < 4; q++) {
    if (valid && l < 24) {
      double col[9], st[9];
      mitc4::b_column_q(l, q, s.g, prep, col);
      mitc4::stress(s.data, col, st);
      const double sc = s.g.scale[q];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        s.B[k][l] = col[k];
        s.D[k][l] = sc * st[k];
      }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        const double d0 = s.D[k][3 * J], d1 = s.D[k][3 * J + 1],
                     d2 = s.D[k][3 * J + 2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int I = I0 + 4 * h;
          const double b0 = s.B[k][3 * I], b1 = s.B[k][3 * I + 1],
                       b2 = s.B[k][3 * I + 2];
          acc[h][0][0] += b0 * d0;
          acc[h][0][1] += b0 * d1;
          acc[h][0][2] += b0 * d2;
          acc[h][1][0] += b1 * d0;
          acc[h][1][1] += b1 * d1;
          acc[h][1][2] += b1 * d2;
          acc[h][2][0] += b2 * d0;
          acc[h][2][1] += b2 * d1;
          acc[h][2][2] += b2 * d2;
        }
      }
    }
    __syncthreads();
  }
  if (valid) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int I = I0 + 4 * h;
      const int a = I / 2, b = J / 2, r0 = 3 * (I % 2), c0 = 3 * (J % 2);
      double* dst = kvals + 36 * (size_t)elem_ind_map[16 * (size_t)e + 4 * a + b];
      const bool first = (mask >> (4 * a + b)) & 1;
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          double* d = &dst[6 * (r0 + r) + c0 + c];
          if (!kColored)
            unsafeAtomicAdd(d, acc[h][r][c]);
          else
            *d = first ? acc[h][r][c] : *d + acc[h][r][c];
        }
    }
  }
  if (!res) return;  // uniform across the workgroup
  if (valid) {
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int I = I0 + 4 * h;
#pragma unroll
      for (int r = 0; r < 3; r++) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < 3; c++) v += acc[h][r][c] * s.u[3 * J + c];
        rp[half][J][3 * I + r] = v;
      }
    }
  }
  __syncthreads();
  if (valid && l < 24) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < 8; j++) v += rp[half][j][l];
    const int node = conn[4 * (size_t)e + l / 6];
    const int row = perm ? perm[node] : node;
    double* d = &res[6 * (size_t)row + l % 6];
    if (!kColored)
      unsafeAtomicAdd(d, v);
    else
      *d = ((mask >> (16 + l / 6)) & 1) ? v : *d + v;
  }
}

// ---------------------------------------------------------------------------
// Batched Jacobian: six elements per 64-lane wavefront, ten lanes per
// element.  Lane l of an element owns one of the ten node-pair blocks
// (a <= b) of the symmetric K_e and keeps its 6x6 sum in registers; the
// quadrature points are streamed: per point the element's ten lanes form the
// 24 columns of B_q and D_q = detXd C B_q into LDS (two or three columns per
// lane) and every lane adds B_q[:,a]^T D_q[:,b] to its block.  Block (b, a)
// is the transpose of (a, b).  LDS per element: geometry, one quadrature
// point's B and D, the element residual.
constexpr int kElemsPerWave = 6;
constexpr int kLanesPerElem = 10;

struct ElemSlot {
  double X[12];
  double u[24];
  double data[4];
  mitc4::Geom g;
  double B[9][24];
  double D[9][24];
  double res[24];
};

// block index l (0..9) -> (a, b), a <= b
__device__ __forceinline__ void pair_of(int l, int& a, int& b) {
  a = l < 4 ? 0 : l < 7 ? 1 : l < 9 ? 2 : 3;
  b = a + (l - (a == 0 ? 0 : a == 1 ? 4 : a == 2 ? 7 : 9));
}

__global__ void __launch_bounds__(64)
    jacobian_batched_kernel(int nelems, const int* __restrict__ conn,
                            const double* __restrict__ xpts,
                            const double* __restrict__ vars,
                            const double* __restrict__ data,
                            const int* __restrict__ elem_ind_map,
                            const int* __restrict__ perm, double* kvals,
                            double* res) {
  __shared__ ElemSlot S[kElemsPerWave];
  const int lane = threadIdx.x;
  const int es = lane / kLanesPerElem;
  const int l = lane % kLanesPerElem;
  const int e = blockIdx.x * kElemsPerWave + es;
  const bool valid = es < kElemsPerWave && e < nelems;
  ElemSlot& s = S[es < kElemsPerWave ? es : 0];
  if (valid) {
    for (int v = l; v < 40; v += kLanesPerElem) {
      if (v < 12) {
        s.X[v] = xpts[3 * (size_t)conn[4 * (size_t)e + v / 3] + v % 3];
      } else if (v < 36) {
        const int i = v - 12;
        s.u[i] = vars ? vars[6 * (size_t)conn[4 * (size_t)e + i / 6] + i % 6]
                      : 0.0;
      } else {
        s.data[v - 36] = data[4 * (size_t)e + (v - 36)];
      }
    }
    for (int v = l; v < 24; v += kLanesPerElem) s.res[v] = 0.0;
  }
  __syncthreads();
  if (valid) {
    if (l < 4)
      mitc4::geom_node(l, s.X, s.g);
    else if (l < 9)
      mitc4::geom_tying_x(l - 4, s.X, s.g);
  }
  __syncthreads();
  if (valid) {
    if (l < 4)
      mitc4::geom_quad(l, s.X, s.g);
    else if (l < 9)
      mitc4::geom_tying_n(l - 4, s.g);
  }
  __syncthreads();
  int a, b;
  pair_of(l, a, b);
  double acc[6][6];
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = 0; c < 6; c++) acc[r][c] = 0.0;
#pragma unroll 1
  for (int q = 0; q < 4; q++) {
    if (valid) {
#pragma unroll 1
      for (int j = l; j < 24; j += kLanesPerElem) {
        mitc4::BColumnPrep prep;
        mitc4::b_column_prep(j, s.g, prep);
        double col[9], st[9];
        mitc4::b_column_q(j, q, s.g, prep, col);
        mitc4::stress(s.data, col, st);
        const double sc = s.g.scale[q];
#pragma unroll
        for (int k = 0; k < 9; k++) {
          s.B[k][j] = col[k];
          s.D[k][j] = sc * st[k];
        }
      }
    }
    __syncthreads();
    if (valid) {
#pragma unroll
      for (int k = 0; k < 9; k++) {
        double ba[6], db[6];
#pragma unroll
        for (int r = 0; r < 6; r++) ba[r] = s.B[k][6 * a + r];
#pragma unroll
        for (int c = 0; c < 6; c++) db[c] = s.D[k][6 * b + c];
#pragma unroll
        for (int r = 0; r < 6; r++)
#pragma unroll
          for (int c = 0; c < 6; c++) acc[r][c] += ba[r] * db[c];
      }
    }
    __syncthreads();
  }
  if (valid) {
    const size_t blk = (size_t)elem_ind_map[16 * (size_t)e + 4 * a + b];
    double* dst = kvals + 36 * blk;
#pragma unroll
    for (int r = 0; r < 6; r++)
#pragma unroll
      for (int c = 0; c < 6; c++) unsafeAtomicAdd(&dst[6 * r + c], acc[r][c]);
    if (a != b) {
      const size_t blt = (size_t)elem_ind_map[16 * (size_t)e + 4 * b + a];
      double* dt = kvals + 36 * blt;
#pragma unroll
      for (int r = 0; r < 6; r++)
#pragma unroll
        for (int c = 0; c < 6; c++)
          unsafeAtomicAdd(&dt[6 * c + r], acc[r][c]);
    }
  }
  if (!res) return;  // uniform across the workgroup
  if (valid) {
    // R_e = K_e u_e: rows of node a from block (a, b), rows of node b from
    // its transpose
#pragma unroll
    for (int r = 0; r < 6; r++) {
      double v = 0.0;
#pragma unroll
      for (int c = 0; c < 6; c++) v += acc[r][c] * s.u[6 * b + c];
      atomicAdd(&s.res[6 * a + r], v);
    }
    if (a != b) {
#pragma unroll
      for (int c = 0; c < 6; c++) {
        double v = 0.0;
#pragma unroll
        for (int r = 0; r < 6; r++) v += acc[r][c] * s.u[6 * a + r];
        atomicAdd(&s.res[6 * b + c], v);
      }
    }
  }
  __syncthreads();
  if (valid) {
    for (int v = l; v < 24; v += kLanesPerElem) {
      const int node = conn[4 * (size_t)e + v / 6];
      const int row = perm ? perm[node] : node;
      unsafeAtomicAdd(&res[6 * (size_t)row + v % 6], s.res[v]);
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 MFMA Jacobian: four elements per 64-lane wavefront, one per 16-lane
// block of v_mfma_f64_4x4x4_4b_f64 (four independent 4x4x4 products per
// instruction).  K_e = sum_q detXd_q B_q^T C B_q is one 24 x 24 x 36 product
// with the 36 = 9 strains x 4 quadrature points stacked strain-major: k-block
// s (4 wide) holds strain s at the four points.  Lane (i, q) of an element
// (its A row i and k index q, see mfma_d_row) forms the columns j = i + 4J
// (J = 0..5) of B_q in registers -- exactly the MFMA A operand
// A[i][q] = B_q[s][4I + i] of tile row I and the B operand
// D[q][j] = detXd_q (C B_q)[s][4J + j] of tile column J -- so B and D never
// pass through LDS.  Only the 21 tiles on and above the diagonal of the
// 6 x 6 tile grid are formed (K_e is symmetric): 9 x 21 MFMAs per four
// elements, 24,192 flops per element.  K_e is then staged in LDS and
// written as whole 6 x 6 blocks, consecutive lanes on consecutive entries
// of a block: plain stores (first element of the colour order to touch the
// block) and FP64 adds without return.  In the colour-ordered launches no
// two elements of a launch share a node, so each add is ordered by the
// launch sequence and the sum is deterministic; the uncoloured entry point
// adds in any order.
constexpr int kMfmaElems = 4;  // elements per wavefront (MFMA blocks)

struct ElemMfma {
  double X[12];
  double u[24];
  double data[4];
  union {
    struct {
      mitc4::Geom g;
      mitc4::BColumnPrep prep[24];  // quadrature-independent column parts
    };
    double Ke[24][25];  // padded rows
  };
};

// Lane layout of v_mfma_f64_4x4x4_4b_f64 on gfx950 (measured): the block of
// lane L is (L / 4) % 4; its A operand is A[L % 4][L / 16], its B operand
// B[L / 16][L % 4] and its result D[L / 16][L % 4].  With the index
// l16 = L % 4 + 4 (L / 16) inside the block, a lane's A row, B column and
// D column are l16 % 4 and its k index and D row l16 / 4.
__device__ __forceinline__ int mfma_d_row(int l16) { return l16 >> 2; }
__device__ __forceinline__ int mfma_d_col(int l16) { return l16 & 3; }

// Row s of detXd C (stress of a2d/a2disotropicshellstress.h:13-51) applied
// to the strain column e[0..8].
struct StressRows {
  double Q0, Q1, G, am, ab, ad, As;
  __device__ __forceinline__ void init(const double d[4]) {
    const double E = d[0], nu = d[1], t = d[2], to = d[3];
    const double D = E / (1.0 - nu * nu);
    Q0 = D;
    Q1 = nu * D;
    G = E / 2.0 / (1.0 + nu);
    am = t;
    ab = -t * t * to;
    ad = t * t * t / 12.0 + t * t * t * to * to;
    As = (5.0 / 6.0) * t * G;
  }
  __device__ __forceinline__ double row(int s, const double e[9]) const {
    switch (s) {
      case 0: return am * (Q0 * e[0] + Q1 * e[1]) + ab * (Q0 * e[3] + Q1 * e[4]);
      case 1: return am * (Q1 * e[0] + Q0 * e[1]) + ab * (Q1 * e[3] + Q0 * e[4]);
      case 2: return am * (G * e[2]) + ab * (G * e[5]);
      case 3: return ab * (Q0 * e[0] + Q1 * e[1]) + ad * (Q0 * e[3] + Q1 * e[4]);
      case 4: return ab * (Q1 * e[0] + Q0 * e[1]) + ad * (Q1 * e[3] + Q0 * e[4]);
      case 5: return ab * (G * e[2]) + ad * (G * e[5]);
      case 6: return As * e[6];
      case 7: return As * e[7];
      default: return 10.0 * As * e[8];
    }
  }
};

#ifndef GPUFEM_MFMA_MINWAVES
#define GPUFEM_MFMA_MINWAVES 2
#endif
template <bool kColored>
__global__ void __launch_bounds__(64, GPUFEM_MFMA_MINWAVES)
    jacobian_mfma_kernel(int nelems, const int* __restrict__ elem_list,
                         const int* __restrict__ first_mask,
                         const int* __restrict__ conn,
                         const double* __restrict__ xpts,
                         const double* __restrict__ vars,
                         const double* __restrict__ data,
                         const int* __restrict__ elem_ind_map,
                         const int* __restrict__ perm, double* kvals,
                         double* res) {
  __shared__ ElemMfma S[kMfmaElems];
  __shared__ int s_blk[kMfmaElems][16];
  __shared__ int s_mask[kMfmaElems];
  const int lane = threadIdx.x;
  // MFMA block (element) of the lane and its index 0..15 inside the block
  const int grp = (lane >> 2) & 3, l16 = (lane & 3) | ((lane >> 4) << 2);
  const int slot = blockIdx.x * kMfmaElems + grp;
  const bool valid = slot < nelems;
  const int e = valid ? (kColored ? elem_list[slot] : slot) : 0;
  ElemMfma& s = S[grp];
  if (valid) {
    for (int v = l16; v < 40; v += 16) {
      if (v < 12) {
        s.X[v] = xpts[3 * (size_t)conn[4 * (size_t)e + v / 3] + v % 3];
      } else if (v < 36) {
        const int w = v - 12;
        s.u[w] = vars ? vars[6 * (size_t)conn[4 * (size_t)e + w / 6] + w % 6]
                      : 0.0;
      } else {
        s.data[v - 36] = data[4 * (size_t)e + (v - 36)];
      }
    }
    s_blk[grp][l16] = elem_ind_map[16 * (size_t)e + l16];
    if (l16 == 0) s_mask[grp] = kColored ? first_mask[slot] : 0;
  }
  __syncthreads();
  if (valid) {
    if (l16 < 4)
      mitc4::geom_node(l16, s.X, s.g);
    else if (l16 < 9)
      mitc4::geom_tying_x(l16 - 4, s.X, s.g);
  }
  __syncthreads();
  if (valid) {
    if (l16 < 4)
      mitc4::geom_quad(l16, s.X, s.g);
    else if (l16 < 9)
      mitc4::geom_tying_n(l16 - 4, s.g);
  }
  __syncthreads();
  // the quadrature-independent part of each column, once per column
  if (valid)
    for (int j = l16; j < 24; j += 16) mitc4::b_column_prep(j, s.g, s.prep[j]);
  __syncthreads();
  // B_q columns j = i + 4J of this lane's quadrature point q, all 9 strains
  const int i = l16 & 3, q = (l16 >> 2) & 3;
  double Bv[6][9];
  StressRows C{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double sc = 0.0;
  if (valid) {
    C.init(s.data);
    sc = s.g.scale[q];
#pragma unroll
    for (int J = 0; J < 6; J++)
      mitc4::b_column_q(i + 4 * J, q, s.g, s.prep[i + 4 * J], Bv[J]);
  } else {
    // a block past the last element: zeros in, results discarded
#pragma unroll
    for (int J = 0; J < 6; J++)
#pragma unroll
      for (int k = 0; k < 9; k++) Bv[J][k] = 0.0;
  }
  // 21 tiles (I <= J) of the 6 x 6 grid of 4 x 4 tiles
  double acc[21];
#pragma unroll
  for (int t = 0; t < 21; t++) acc[t] = 0.0;
#pragma unroll
  for (int st = 0; st < 9; st++) {
    double dv[6];
#pragma unroll
    for (int J = 0; J < 6; J++) dv[J] = sc * C.row(st, Bv[J]);
    int t = 0;
#pragma unroll
    for (int I = 0; I < 6; I++)
#pragma unroll
      for (int J = I; J < 6; J++, t++)
        acc[t] = __builtin_amdgcn_mfma_f64_4x4x4f64(Bv[I][st], dv[J], acc[t],
                                                    0, 0, 0);
  }
  __syncthreads();  // geometry no longer read: Ke reuses its LDS
  if (valid) {
    const int dr = mfma_d_row(l16), dc = mfma_d_col(l16);
    int t = 0;
#pragma unroll
    for (int I = 0; I < 6; I++)
#pragma unroll
      for (int J = I; J < 6; J++, t++) {
        s.Ke[4 * I + dr][4 * J + dc] = acc[t];
        if (J != I) s.Ke[4 * J + dc][4 * I + dr] = acc[t];
      }
  }
  __syncthreads();
  // whole 6 x 6 blocks: entry f of the wavefront's 4 x 16 x 36 block values
#pragma unroll 4
  for (int f = lane; f < kMfmaElems * 576; f += 64) {
    const int g = f / 576, rem = f % 576, ab = rem / 36, rc = rem % 36;
    if (blockIdx.x * kMfmaElems + g >= nelems) continue;
    const int a = ab >> 2, b = ab & 3, r = rc / 6, c = rc % 6;
    const double v = S[g].Ke[6 * a + r][6 * b + c];
    double* d = kvals + 36 * (size_t)s_blk[g][ab] + rc;
    // colour-ordered: no other element of this launch touches the block,
    // so the add without return is ordered by the launches alone (the same
    // sum, bit for bit, as a read-add-write, without the read's latency)
    if (kColored && ((s_mask[g] >> ab) & 1))
      *d = v;
    else
      unsafeAtomicAdd(d, v);
  }
  if (!res) return;  // uniform across the workgroup
  if (valid) {
    for (int row = l16; row < 24; row += 16) {
      double v = 0.0;
#pragma unroll
      for (int c = 0; c < 24; c++) v += s.Ke[row][c] * s.u[c];
      const int node = conn[4 * (size_t)e + row / 6];
      const int prow = perm ? perm[node] : node;
      double* d = &res[6 * (size_t)prow + row % 6];
      if (!kColored)
        unsafeAtomicAdd(d, v);
      else
        *d = ((s_mask[grp] >> (16 + row / 6)) & 1) ? v : *d + v;
    }
  }
}

// ---------------------------------------------------------------------------
// Nonlinear MITC4 (IsotropicShell<..., isNonlinear = true>, physics.h:6-17):
// NonlinearShellStrainCore (a2d/a2dshellstrain.h:72-93) adds to the bending
// strains the quadratic terms of the local displacement gradients u0x, u1x
// (basis.h:470-492),
//   e3 += u0x(:,0).u1x(:,0),  e4 += u0x(:,1).u1x(:,1),
//   e5 += u0x(:,0).u1x(:,1) + u1x(:,0).u0x(:,1);
// the tying strains and the director stay linear (physics.h:214-219,
// 286-300).  With G0, G1 the maps of the element DOFs to the entries
// (r, c < 2) of u0x, u1x (mitc4::grad_columns_q), B_q(u) = dE_q/du and
// S_q = detXd_q C E_q(u):
//   R_e(u) = sum_q B_q(u)^T S_q                         (the gradient of
//   U = sum_q 1/2 detXd_q E^T C E, physics.h:18-36),
//   K_e(u) = sum_q detXd_q B_q(u)^T C B_q(u) + G0^T W_q G1 + G1^T W_q G0,
// W_q holding S_q's bending stresses on the gradient-entry pairs each
// quadratic term multiplies: K_e is the exact Hessian of U.  (The
// reference's own nonlinear Jacobian routes its displacement gradients'
// second-order terms into the residual, elem_group.h:286-289, and
// overwrites the u0x sensitivity with the u1x one, basis.h:565-567.)
// One element per 64-lane workgroup; per quadrature point lanes 0-23 form
// column j of the linear B, G0 and G1 into LDS, 21 lanes form B u, G0 u and
// G1 u, lane 0 the strain and stress, lanes 0-23 the column of B(u), C B(u)
// and of W G1, W G0, and every lane adds a 3x3 tile of K_e.  Scatter as in
// jacobian_kernel; R_e into res (6 perm[node] + k), the energy into
// *energy; kvals, res or energy may be null.
struct ElemNL {
  double X[12];
  double u[24];
  double data[4];
  mitc4::Geom g;
  double B[9][24];   // B_q(u)
  double D[9][24];   // detXd C B_q(u)
  double G0[6][24];
  double G1[6][24];
  double P0[6][24];  // W G1
  double P1[6][24];  // W G0
  double eu[21];     // B_lin u, u0x entries, u1x entries
  double S[9];       // detXd C E(u)
};

template <bool kColored>
__global__ void __launch_bounds__(64)
    tangent_kernel(int nelems, const int* __restrict__ elem_list,
                   const int* __restrict__ first_mask,
                   const int* __restrict__ conn,
                   const double* __restrict__ xpts,
                   const double* __restrict__ vars,
                   const double* __restrict__ data,
                   const int* __restrict__ elem_ind_map,
                   const int* __restrict__ perm, double* kvals, double* res,
                   double* energy) {
  __shared__ ElemNL s;
  if ((int)blockIdx.x >= nelems) return;
  const int e = kColored ? elem_list[blockIdx.x] : (int)blockIdx.x;
  const int mask = kColored ? first_mask[blockIdx.x] : 0;
  const int lane = threadIdx.x;
  if (lane < 12) {
    s.X[lane] = xpts[3 * (size_t)conn[4 * (size_t)e + lane / 3] + lane % 3];
  } else if (lane < 36) {
    const int i = lane - 12;
    s.u[i] = vars[6 * (size_t)conn[4 * (size_t)e + i / 6] + i % 6];
  } else if (lane < 40) {
    s.data[lane - 36] = data[4 * (size_t)e + (lane - 36)];
  }
  __syncthreads();
  if (lane < 4)
    mitc4::geom_node(lane, s.X, s.g);
  else if (lane < 9)
    mitc4::geom_tying_x(lane - 4, s.X, s.g);
  __syncthreads();
  if (lane < 4)
    mitc4::geom_quad(lane, s.X, s.g);
  else if (lane < 9)
    mitc4::geom_tying_n(lane - 4, s.g);
  __syncthreads();
  mitc4::BColumnPrep prep;
  if (lane < 24) mitc4::b_column_prep(lane, s.g, prep);
  const int I = lane / 8, J = lane % 8;
  double acc[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) acc[r][c] = 0.0;
  double racc = 0.0, uacc = 0.0;
#pragma unroll 1
  for (int q = 0; q < 4; q++) {
    const double sc = s.g.scale[q];
    // linear B column and the gradient maps of DOF lane
    double col[9], g0[6], g1[6];
    if (lane < 24) {
      mitc4::b_column_q(lane, q, s.g, prep, col);
      mitc4::grad_columns_q(lane, q, s.g, g0, g1);
#pragma unroll
      for (int k = 0; k < 9; k++) s.B[k][lane] = col[k];
#pragma unroll
      for (int m = 0; m < 6; m++) {
        s.G0[m][lane] = g0[m];
        s.G1[m][lane] = g1[m];
      }
    }
    __syncthreads();
    if (lane < 21) {
      const double* row = lane < 9 ? s.B[lane]
                          : lane < 15 ? s.G0[lane - 9] : s.G1[lane - 15];
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < 24; j++) v += row[j] * s.u[j];
      s.eu[lane] = v;
    }
    __syncthreads();
    if (lane == 0) {
      const double* U0 = &s.eu[9];
      const double* U1 = &s.eu[15];
      double E[9], st[9];
#pragma unroll
      for (int k = 0; k < 9; k++) E[k] = s.eu[k];
      E[3] += U0[0] * U1[0] + U0[2] * U1[2] + U0[4] * U1[4];
      E[4] += U0[1] * U1[1] + U0[3] * U1[3] + U0[5] * U1[5];
      E[5] += U0[0] * U1[1] + U0[2] * U1[3] + U0[4] * U1[5] +
              U1[0] * U0[1] + U1[2] * U0[3] + U1[4] * U0[5];
      mitc4::stress(s.data, E, st);
      double es = 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        s.S[k] = sc * st[k];
        es += E[k] * st[k];
      }
      uacc += 0.5 * sc * es;
    }
    __syncthreads();
    if (lane < 24) {
      const double* U0 = &s.eu[9];
      const double* U1 = &s.eu[15];
      // dE/du_lane: the linear column plus the quadratic terms' derivative
#pragma unroll
      for (int r = 0; r < 3; r++) {
        col[3] += g0[2 * r] * U1[2 * r] + U0[2 * r] * g1[2 * r];
        col[4] += g0[2 * r + 1] * U1[2 * r + 1] + U0[2 * r + 1] * g1[2 * r + 1];
        col[5] += g0[2 * r] * U1[2 * r + 1] + U0[2 * r] * g1[2 * r + 1] +
                  g1[2 * r] * U0[2 * r + 1] + U1[2 * r] * g0[2 * r + 1];
      }
      double st[9];
      mitc4::stress(s.data, col, st);
      double rv = 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) {
        s.B[k][lane] = col[k];
        s.D[k][lane] = sc * st[k];
        rv += col[k] * s.S[k];
      }
      racc += rv;
      const double S3 = s.S[3], S4 = s.S[4], S5 = s.S[5];
#pragma unroll
      for (int r = 0; r < 3; r++) {
        s.P0[2 * r][lane] = S3 * g1[2 * r] + S5 * g1[2 * r + 1];
        s.P0[2 * r + 1][lane] = S5 * g1[2 * r] + S4 * g1[2 * r + 1];
        s.P1[2 * r][lane] = S3 * g0[2 * r] + S5 * g0[2 * r + 1];
        s.P1[2 * r + 1][lane] = S5 * g0[2 * r] + S4 * g0[2 * r + 1];
      }
    }
    __syncthreads();
    if (kvals) {
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const int i = 3 * I + r, j = 3 * J + c;
          double v = 0.0;
#pragma unroll
          for (int k = 0; k < 9; k++) v += s.B[k][i] * s.D[k][j];
#pragma unroll
          for (int m = 0; m < 6; m++)
            v += s.G0[m][i] * s.P0[m][j] + s.G1[m][i] * s.P1[m][j];
          acc[r][c] += v;
        }
    }
    __syncthreads();
  }
  if (kvals) {
    const int a = I / 2, b = J / 2, r0 = 3 * (I % 2), c0 = 3 * (J % 2);
    double* dst =
        kvals + 36 * (size_t)elem_ind_map[16 * (size_t)e + 4 * a + b];
    const bool first = (mask >> (4 * a + b)) & 1;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        double* d = &dst[6 * (r0 + r) + c0 + c];
        if (!kColored)
          unsafeAtomicAdd(d, acc[r][c]);
        else
          *d = first ? acc[r][c] : *d + acc[r][c];
      }
  }
  if (res && lane < 24) {
    const int node = conn[4 * (size_t)e + lane / 6];
    const int row = perm ? perm[node] : node;
    double* d = &res[6 * (size_t)row + lane % 6];
    if (!kColored)
      unsafeAtomicAdd(d, racc);
    else
      *d = ((mask >> (16 + lane / 6)) & 1) ? racc : *d + racc;
  }
  if (energy && lane == 0) unsafeAtomicAdd(energy, uacc);
}

// Residual only: R_e = sum_q D_q^T (B_q u) (C symmetric).
__global__ void __launch_bounds__(64)
    residual_kernel(int nelems, const int* __restrict__ conn,
                    const double* __restrict__ xpts,
                    const double* __restrict__ vars,
                    const double* __restrict__ data,
                    const int* __restrict__ perm, double* res) {
  __shared__ ElemLDS s;
  __shared__ double Eu[4][9];
  const int e = blockIdx.x;
  const int lane = threadIdx.x;
  if (e >= nelems) return;
  load_element(s, e, lane, conn, xpts, vars, data);
  __syncthreads();
  element_geometry(s, lane);
  element_bd(s, lane);
  if (lane < 36) {
    const int q = lane / 9, k = lane % 9;
    double acc = 0.0;
    for (int j = 0; j < 24; j++) acc += s.B[q][k][j] * s.u[j];
    Eu[q][k] = acc;
  }
  __syncthreads();
  if (lane < 24) {
    double r = 0.0;
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int k = 0; k < 9; k++) r += s.D[q][k][lane] * Eu[q][k];
    const int node = conn[4 * (size_t)e + lane / 6];
    const int row = perm ? perm[node] : node;
    unsafeAtomicAdd(&res[6 * (size_t)row + lane % 6], r);
  }
}

// Energy U = sum_e sum_q 1/2 detXd (B_q u) . C (B_q u)
__global__ void __launch_bounds__(64)
    energy_kernel(int nelems, const int* __restrict__ conn,
                  const double* __restrict__ xpts,
                  const double* __restrict__ vars,
                  const double* __restrict__ data, double* energy) {
  __shared__ ElemLDS s;
  __shared__ double Eu[4][9];
  const int e = blockIdx.x;
  const int lane = threadIdx.x;
  if (e >= nelems) return;
  load_element(s, e, lane, conn, xpts, vars, data);
  __syncthreads();
  element_geometry(s, lane);
  element_bd(s, lane);
  if (lane < 36) {
    const int q = lane / 9, k = lane % 9;
    double acc = 0.0;
    for (int j = 0; j < 24; j++) acc += s.B[q][k][j] * s.u[j];
    Eu[q][k] = acc;
  }
  __syncthreads();
  if (lane < 4) {
    double st[9];
    mitc4::stress(s.data, Eu[lane], st);
    double u = 0.0;
    for (int k = 0; k < 9; k++) u += Eu[lane][k] * st[k];
    unsafeAtomicAdd(energy, 0.5 * s.g.scale[lane] * u);
  }
}

}  // namespace
}  // namespace gf

namespace {
// Jacobian kernel: the FP64 MFMA kernel by default; GPUFEM_ASSEMBLY=scalar
// selects the two-elements-per-wavefront scalar kernel (jacobian2_kernel),
// =element the one-element-per-wavefront kernel, =batched the
// six-elements-per-wavefront one (atomic entry point only)
enum { kAsmMfma = 0, kAsmElement = 1, kAsmBatched = 2, kAsmScalar = 3 };
int assembly_variant() {
  static const int v = [] {
    const char* e = getenv("GPUFEM_ASSEMBLY");
    if (e && std::string(e) == "element") return (int)kAsmElement;
    if (e && std::string(e) == "batched") return (int)kAsmBatched;
    if (e && std::string(e) == "scalar") return (int)kAsmScalar;
    return (int)kAsmMfma;
  }();
  return v;
}
}  // namespace

extern "C" {

int gf_assemble_jacobian_mitc4_f64(gf_context ctx, int nelems, int nnodes,
                                   const int* conn, const double* xpts,
                                   const double* vars, const double* data,
                                   const int* elem_ind_map, const int* perm,
                                   double* kvals, double* res) {
  if (!ctx || nelems < 0 || nnodes <= 0 || !conn || !xpts || !data ||
      !elem_ind_map || !kvals)
    return gf::set_error(ctx, GF_ERR_INVALID, "assemble_jacobian: bad args");
  if (nelems == 0) return GF_OK;
  if (assembly_variant() == kAsmMfma) {
    hipLaunchKernelGGL(gf::jacobian_mfma_kernel<false>,
                       dim3(gf::div_up(nelems, gf::kMfmaElems)), dim3(64), 0,
                       ctx->stream, nelems, (const int*)nullptr,
                       (const int*)nullptr, conn, xpts, vars, data,
                       elem_ind_map, perm, kvals, res);
    return gf::check_launch(ctx, "jacobian_mfma_kernel");
  }
  if (assembly_variant() == kAsmElement) {
    hipLaunchKernelGGL(gf::jacobian_kernel<false>, dim3(nelems), dim3(64), 0,
                       ctx->stream, nelems, (const int*)nullptr,
                       (const int*)nullptr, conn, xpts, vars, data,
                       elem_ind_map, perm, kvals, res);
    return gf::check_launch(ctx, "jacobian_kernel");
  }
  if (assembly_variant() == kAsmScalar) {
    hipLaunchKernelGGL(gf::jacobian2_kernel<false>,
                       dim3(gf::div_up(nelems, 2)), dim3(64), 0, ctx->stream,
                       nelems, (const int*)nullptr, (const int*)nullptr, conn,
                       xpts, vars, data, elem_ind_map, perm, kvals, res);
    return gf::check_launch(ctx, "jacobian2_kernel");
  }
  hipLaunchKernelGGL(gf::jacobian_batched_kernel,
                     dim3(gf::div_up(nelems, gf::kElemsPerWave)), dim3(64), 0,
                     ctx->stream, nelems, conn, xpts, vars, data, elem_ind_map,
                     perm, kvals, res);
  return gf::check_launch(ctx, "jacobian_batched_kernel");
}


int gf_asm_plan_create(gf_context ctx, int nelems, int nnodes,
                       const int* conn, const int* elem_ind_map, int nnzb,
                       gf_asm_plan* out) {
  if (!ctx || nelems <= 0 || nnodes <= 0 || !conn || !elem_ind_map ||
      nnzb <= 0 || !out)
    return gf::set_error(ctx, GF_ERR_INVALID, "asm_plan_create: bad args");
  auto* pl = new gf_asm_plan_s;
  pl->ctx = ctx;
  pl->nelems = nelems;
  pl->nnzb = nnzb;
  pl->nnodes = nnodes;
  std::vector<int> elems;
  gf::element_coloring(nelems, nnodes, 4, conn, pl->color_ptr, elems);
  pl->ncolors = (int)pl->color_ptr.size() - 1;
  // first-touch masks in processing order (colour-major)
  std::vector<int> mask(nelems, 0);
  std::vector<char> kt(nnzb, 0), rt(nnodes, 0);
  for (int p = 0; p < nelems; p++) {
    const int e = elems[p];
    int m = 0;
    for (int ab = 0; ab < 16; ab++) {
      const int blk = elem_ind_map[16 * (size_t)e + ab];
      if (blk < 0 || blk >= nnzb) {
        delete pl;
        return gf::set_error(ctx, GF_ERR_INVALID,
                             "asm_plan_create: element %d maps outside K", e);
      }
      if (!kt[blk]) {
        kt[blk] = 1;
        m |= 1 << ab;
      }
    }
    for (int a = 0; a < 4; a++) {
      const int nd = conn[4 * (size_t)e + a];
      if (!rt[nd]) {
        rt[nd] = 1;
        m |= 1 << (16 + a);
      }
    }
    mask[p] = m;
  }
  for (int k = 0; k < nnzb && !pl->zero_k; k++) pl->zero_k = !kt[k];
  for (int i = 0; i < nnodes && !pl->zero_res; i++) pl->zero_res = !rt[i];
  int st = gf::dmalloc(ctx, (void**)&pl->elems, sizeof(int) * nelems);
  if (!st) st = gf::dmalloc(ctx, (void**)&pl->mask, sizeof(int) * nelems);
  if (st) {
    gf_asm_plan_destroy(pl);
    return st;
  }
  // copies on the context's stream (a plain hipMemcpy would also wait for
  // the other streams of the device); the host vectors live until the sync
  if (hipMemcpyAsync(pl->elems, elems.data(), sizeof(int) * nelems,
                     hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipMemcpyAsync(pl->mask, mask.data(), sizeof(int) * nelems,
                     hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess) {
    gf_asm_plan_destroy(pl);
    return gf::set_error(ctx, GF_ERR_DEVICE, "asm_plan_create: upload");
  }
  *out = pl;
  return GF_OK;
}

int gf_asm_plan_destroy(gf_asm_plan pl) {
  if (!pl) return GF_OK;
  gf::dfree(pl->ctx, pl->elems);
  gf::dfree(pl->ctx, pl->mask);
  delete pl;
  return GF_OK;
}

int gf_asm_plan_colors(gf_asm_plan pl) { return pl ? pl->ncolors : -1; }

int gf_assemble_jacobian_colored_f64(gf_context ctx, gf_asm_plan pl,
                                     const int* conn, const double* xpts,
                                     const double* vars, const double* data,
                                     const int* elem_ind_map, const int* perm,
                                     double* kvals, double* res) {
  if (!ctx || !pl || !conn || !xpts || !data || !elem_ind_map || !kvals)
    return gf::set_error(ctx, GF_ERR_INVALID, "assemble_colored: bad args");
  if (pl->zero_k)
    GF_HIP(ctx, hipMemsetAsync(kvals, 0, sizeof(double) * 36 * (size_t)pl->nnzb,
                               ctx->stream));
  if (res && pl->zero_res)
    GF_HIP(ctx, hipMemsetAsync(res, 0, sizeof(double) * 6 * (size_t)pl->nnodes,
                               ctx->stream));
  for (int c = 0; c < pl->ncolors; c++) {
    const int c0 = pl->color_ptr[c], cnt = pl->color_ptr[c + 1] - c0;
    if (assembly_variant() == kAsmMfma)
      hipLaunchKernelGGL(gf::jacobian_mfma_kernel<true>,
                         dim3(gf::div_up(cnt, gf::kMfmaElems)), dim3(64), 0,
                         ctx->stream, cnt, pl->elems + c0, pl->mask + c0,
                         conn, xpts, vars, data, elem_ind_map, perm, kvals,
                         res);
    else if (assembly_variant() == kAsmElement)
      hipLaunchKernelGGL(gf::jacobian_kernel<true>, dim3(cnt), dim3(64), 0,
                         ctx->stream, cnt, pl->elems + c0, pl->mask + c0,
                         conn, xpts, vars, data, elem_ind_map, perm, kvals,
                         res);
    else
      hipLaunchKernelGGL(gf::jacobian2_kernel<true>, dim3(gf::div_up(cnt, 2)),
                         dim3(64), 0, ctx->stream, cnt, pl->elems + c0,
                         pl->mask + c0, conn, xpts, vars, data, elem_ind_map,
                         perm, kvals, res);
  }
  return gf::check_launch(ctx, "jacobian_kernel<colored>");
}

int gf_assemble_tangent_mitc4_f64(gf_context ctx, int nelems, int nnodes,
                                  const int* conn, const double* xpts,
                                  const double* vars, const double* data,
                                  const int* elem_ind_map, const int* perm,
                                  double* kvals, double* res,
                                  double* energy) {
  if (!ctx || nelems < 0 || nnodes <= 0 || !conn || !xpts || !vars ||
      !data || (kvals && !elem_ind_map))
    return gf::set_error(ctx, GF_ERR_INVALID, "assemble_tangent: bad args");
  if (energy)
    GF_HIP(ctx, hipMemsetAsync(energy, 0, sizeof(double), ctx->stream));
  if (nelems == 0) return GF_OK;
  hipLaunchKernelGGL(gf::tangent_kernel<false>, dim3(nelems), dim3(64), 0,
                     ctx->stream, nelems, (const int*)nullptr,
                     (const int*)nullptr, conn, xpts, vars, data, elem_ind_map,
                     perm, kvals, res, energy);
  return gf::check_launch(ctx, "tangent_kernel");
}

int gf_assemble_tangent_colored_f64(gf_context ctx, gf_asm_plan pl,
                                    const int* conn, const double* xpts,
                                    const double* vars, const double* data,
                                    const int* elem_ind_map, const int* perm,
                                    double* kvals, double* res) {
  if (!ctx || !pl || !conn || !xpts || !vars || !data || !elem_ind_map ||
      !kvals)
    return gf::set_error(ctx, GF_ERR_INVALID, "assemble_tangent_colored: bad args");
  if (pl->zero_k)
    GF_HIP(ctx, hipMemsetAsync(kvals, 0, sizeof(double) * 36 * (size_t)pl->nnzb,
                               ctx->stream));
  if (res && pl->zero_res)
    GF_HIP(ctx, hipMemsetAsync(res, 0, sizeof(double) * 6 * (size_t)pl->nnodes,
                               ctx->stream));
  for (int c = 0; c < pl->ncolors; c++) {
    const int c0 = pl->color_ptr[c], cnt = pl->color_ptr[c + 1] - c0;
    hipLaunchKernelGGL(gf::tangent_kernel<true>, dim3(cnt), dim3(64), 0,
                       ctx->stream, cnt, pl->elems + c0, pl->mask + c0, conn,
                       xpts, vars, data, elem_ind_map, perm, kvals, res,
                       (double*)nullptr);
  }
  return gf::check_launch(ctx, "tangent_kernel<colored>");
}

int gf_assemble_residual_mitc4_f64(gf_context ctx, int nelems, int nnodes,
                                   const int* conn, const double* xpts,
                                   const double* vars, const double* data,
                                   const int* perm, double* res) {
  if (!ctx || nelems < 0 || nnodes <= 0 || !conn || !xpts || !vars ||
      !data || !res)
    return gf::set_error(ctx, GF_ERR_INVALID, "assemble_residual: bad args");
  if (nelems == 0) return GF_OK;
  hipLaunchKernelGGL(gf::residual_kernel, dim3(nelems), dim3(64), 0,
                     ctx->stream, nelems, conn, xpts, vars, data, perm, res);
  return gf::check_launch(ctx, "residual_kernel");
}

int gf_energy_mitc4_f64(gf_context ctx, int nelems, const int* conn,
                        const double* xpts, const double* vars,
                        const double* data, double* energy) {
  if (!ctx || nelems < 0 || !conn || !xpts || !vars || !data || !energy)
    return gf::set_error(ctx, GF_ERR_INVALID, "energy: bad args");
  GF_HIP(ctx, hipMemsetAsync(energy, 0, sizeof(double), ctx->stream));
  if (nelems == 0) return GF_OK;
  hipLaunchKernelGGL(gf::energy_kernel, dim3(nelems), dim3(64), 0,
                     ctx->stream, nelems, conn, xpts, vars, data, energy);
  return gf::check_launch(ctx, "energy_kernel");
}

}  // extern "C"
