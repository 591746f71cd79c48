// Device-side MITC4 shell element algebra (linear strains, linearized
// rotation director, no reference axis, 2x2 Gauss quadrature).
//
// The reference evaluates the element with forward/reverse automatic
// differentiation per Jacobian column (include/shell/elem_group.h:183-303).
// For the linear strain the strain map E_q(u) = B_q u is linear, so here B_q
// is formed in closed form, column by column, from per-element geometry:
//
//  tying strains (include/shell/physics.h:95-209), points
//    g11 @ (0,-1),(0,1): Xxi . Uxi          g22 @ (-1,0),(1,0): Xeta . Ueta
//    g12 @ (0,0): (Uxi.Xeta + Ueta.Xxi)/2
//    g23 @ (-1,0),(1,0): (Xeta.d0 + n0.Ueta)/2
//    g13 @ (0,-1),(0,1): (Xxi.d0 + n0.Uxi)/2,    d_a = q_a x fn_a
//    so d/du_a g11 = N_a,xi Xxi ; d/dq_a g23 = N_a (fn_a x Xeta) / 2 ; ...
//  e0ty = A^T G A with A = Xd^-1 T (basis.h:495-499,
//    a2d/a2dsymmatrotateframe.h:241-249)
//  u1x(k,j) = t_k.(d0,xi A0j + d0,eta A1j + Uxi Z0j + Ueta Z1j + d0 Z2j),
//    Z = -Xd^-1 Xdz Xd^-1 T (basis.h:470-492)
//  drill (shell_utils.h:42-104, director.h:56-82):
//    et = sum_a N_a [ -n_a . q_a + (Uxi(a) . v_a + Ueta(a) . w_a) / 2 ],
//    n_a = t_a0 x t_a1, v_a = W00 t_a1 - W01 t_a0, w_a = W10 t_a1 - W11 t_a0,
//    W = Xd_a^-1 T_a at node a
//  strain E (a2d/a2dshellstrain.h:13-30) and stress C E
//    (a2d/a2disotropicshellstress.h:13-51), K_e = sum_q detXd_q B^T C B.
#pragma once
#include <hip/hip_runtime.h>

namespace gf {
namespace mitc4 {

constexpr double kRt3 = 0.577350269189626;  // quadrature.h:9 literal

__device__ __forceinline__ void shape(double xi, double eta, double N[4],
                                      double Nxi[4], double Neta[4]) {
  const double na0 = 0.5 * (1.0 - xi), na1 = 0.5 * (1.0 + xi);
  const double nb0 = 0.5 * (1.0 - eta), nb1 = 0.5 * (1.0 + eta);
  N[0] = na0 * nb0;
  N[1] = na1 * nb0;
  N[2] = na0 * nb1;
  N[3] = na1 * nb1;
  Nxi[0] = -0.5 * nb0;
  Nxi[1] = 0.5 * nb0;
  Nxi[2] = -0.5 * nb1;
  Nxi[3] = 0.5 * nb1;
  Neta[0] = -0.5 * na0;
  Neta[1] = -0.5 * na1;
  Neta[2] = 0.5 * na0;
  Neta[3] = 0.5 * na1;
}

__device__ __forceinline__ void cross(const double a[3], const double b[3],
                                      double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
// row-major 3x3 inverse by adjugate; returns the determinant
__device__ __forceinline__ double inv3(const double A[9], double Ai[9]) {
  const double d = A[8] * (A[0] * A[4] - A[3] * A[1]) -
                   A[7] * (A[0] * A[5] - A[3] * A[2]) +
                   A[6] * (A[1] * A[5] - A[2] * A[4]);
  const double id = 1.0 / d;
  Ai[0] = (A[4] * A[8] - A[5] * A[7]) * id;
  Ai[1] = -(A[1] * A[8] - A[2] * A[7]) * id;
  Ai[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  Ai[3] = -(A[3] * A[8] - A[5] * A[6]) * id;
  Ai[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  Ai[5] = -(A[0] * A[5] - A[2] * A[3]) * id;
  Ai[6] = (A[3] * A[7] - A[4] * A[6]) * id;
  Ai[7] = -(A[0] * A[7] - A[1] * A[6]) * id;
  Ai[8] = (A[0] * A[4] - A[1] * A[3]) * id;
  return d;
}
__device__ __forceinline__ void mm3(const double* A, const double* B,
                                    double* C) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      C[3 * i + j] =
          A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}
// shell transform without reference axis: t1 from Xxi, normal n0
// (shell_utils.h:4-40); T columns (t1, t2, nhat), row-major
__device__ __forceinline__ void transform(const double Xxi[3],
                                          const double n0[3], double T[9]) {
  const double ni = 1.0 / sqrt(dot3(n0, n0));
  const double nh[3] = {n0[0] * ni, n0[1] * ni, n0[2] * ni};
  const double d = dot3(nh, Xxi);
  double t1[3] = {Xxi[0] - d * nh[0], Xxi[1] - d * nh[1], Xxi[2] - d * nh[2]};
  const double ti = 1.0 / sqrt(dot3(t1, t1));
  t1[0] *= ti;
  t1[1] *= ti;
  t1[2] *= ti;
  double t2[3];
  cross(nh, t1, t2);
  for (int i = 0; i < 3; i++) {
    T[3 * i] = t1[i];
    T[3 * i + 1] = t2[i];
    T[3 * i + 2] = nh[i];
  }
}

// Tying-point locations: 0:(0,-1) 1:(0,1) 2:(-1,0) 3:(1,0) 4:(0,0).
// ety index -> location: g11 {0,1}, g22 {2,3}, g12 {4}, g23 {2,3}, g13 {0,1}.
__device__ __forceinline__ void tying_loc(int l, double& xi, double& eta) {
  xi = (l == 2) ? -1.0 : (l == 3 ? 1.0 : 0.0);
  eta = (l == 0) ? -1.0 : (l == 1 ? 1.0 : 0.0);
}

// Per-element geometry shared by the lanes that form B.
struct Geom {
  double fn[4][3];     // node normals
  double dn[4][3];     // drill normal n_a = t_a0 x t_a1
  double dv[4][3];     // drill v_a
  double dw[4][3];     // drill w_a
  double tXxi[5][3];   // Xxi at tying locations
  double tXeta[5][3];  // Xeta at tying locations
  double tn0[5][3];    // interpolated normal at tying locations
  double A[4][9];      // Xd^-1 T at quadrature points
  double Z[4][9];      // -Xd^-1 Xdz Xd^-1 T
  double T[4][6];      // t0, t1 (first two columns of T) per quadrature point
  double coef[4][25];  // e0ty strains (E0,E1,E2,E6,E7) vs (g11,g22,g12,g23,g13)
  double scale[4];     // detXd * weight
};

// Node-point stage (one lane per node a = 0..3).
__device__ __forceinline__ void geom_node(int a, const double X[12],
                                          Geom& g) {
  const double xi = -1.0 + 2.0 * (a % 2), eta = -1.0 + 2.0 * (a / 2);
  double N[4], Nx[4], Ne[4];
  shape(xi, eta, N, Nx, Ne);
  double Xxi[3] = {0, 0, 0}, Xeta[3] = {0, 0, 0};
  for (int b = 0; b < 4; b++)
    for (int k = 0; k < 3; k++) {
      Xxi[k] += Nx[b] * X[3 * b + k];
      Xeta[k] += Ne[b] * X[3 * b + k];
    }
  double fn[3];
  cross(Xxi, Xeta, fn);
  const double inv = 1.0 / sqrt(dot3(fn, fn));
  for (int k = 0; k < 3; k++) fn[k] *= inv;
  double T[9], Xd[9], Xdi[9], W[9];
  transform(Xxi, fn, T);
  for (int i = 0; i < 3; i++) {
    Xd[3 * i] = Xxi[i];
    Xd[3 * i + 1] = Xeta[i];
    Xd[3 * i + 2] = fn[i];
  }
  inv3(Xd, Xdi);
  mm3(Xdi, T, W);
  const double t0[3] = {T[0], T[3], T[6]}, t1[3] = {T[1], T[4], T[7]};
  double n[3];
  cross(t0, t1, n);
  for (int k = 0; k < 3; k++) {
    g.fn[a][k] = fn[k];
    g.dn[a][k] = n[k];
    g.dv[a][k] = W[0] * t1[k] - W[1] * t0[k];
    g.dw[a][k] = W[3] * t1[k] - W[4] * t0[k];
  }
}

// Tying-location stage, geometry part (lane per location l = 0..4).
__device__ __forceinline__ void geom_tying_x(int l, const double X[12],
                                             Geom& g) {
  double xi, eta, N[4], Nx[4], Ne[4];
  tying_loc(l, xi, eta);
  shape(xi, eta, N, Nx, Ne);
  for (int k = 0; k < 3; k++) {
    double sx = 0, se = 0;
    for (int b = 0; b < 4; b++) {
      sx += Nx[b] * X[3 * b + k];
      se += Ne[b] * X[3 * b + k];
    }
    g.tXxi[l][k] = sx;
    g.tXeta[l][k] = se;
  }
}

// Tying-location stage, normal part (needs node normals).
__device__ __forceinline__ void geom_tying_n(int l, Geom& g) {
  double xi, eta, N[4], Nx[4], Ne[4];
  tying_loc(l, xi, eta);
  shape(xi, eta, N, Nx, Ne);
  for (int k = 0; k < 3; k++) {
    double s = 0;
    for (int b = 0; b < 4; b++) s += N[b] * g.fn[b][k];
    g.tn0[l][k] = s;
  }
}

// Quadrature-point stage (lane per point q = 0..3; needs node normals).
__device__ __forceinline__ void geom_quad(int q, const double X[12],
                                          Geom& g) {
  const double xi = (q % 2 == 0) ? -kRt3 : kRt3;
  const double eta = (q / 2 == 0) ? -kRt3 : kRt3;
  double N[4], Nx[4], Ne[4];
  shape(xi, eta, N, Nx, Ne);
  double Xxi[3], Xeta[3], n0[3], nxi[3], neta[3];
  for (int k = 0; k < 3; k++) {
    double a = 0, b = 0, c = 0, d = 0, e = 0;
    for (int m = 0; m < 4; m++) {
      a += Nx[m] * X[3 * m + k];
      b += Ne[m] * X[3 * m + k];
      c += N[m] * g.fn[m][k];
      d += Nx[m] * g.fn[m][k];
      e += Ne[m] * g.fn[m][k];
    }
    Xxi[k] = a;
    Xeta[k] = b;
    n0[k] = c;
    nxi[k] = d;
    neta[k] = e;
  }
  double Xd[9], Xdz[9], T[9], Xdi[9], A[9], tmp[9], Z[9];
  for (int i = 0; i < 3; i++) {
    Xd[3 * i] = Xxi[i];
    Xd[3 * i + 1] = Xeta[i];
    Xd[3 * i + 2] = n0[i];
    Xdz[3 * i] = nxi[i];
    Xdz[3 * i + 1] = neta[i];
    Xdz[3 * i + 2] = 0.0;
  }
  transform(Xxi, n0, T);
  const double det = inv3(Xd, Xdi);
  mm3(Xdi, T, A);
  mm3(Xdi, Xdz, tmp);
  mm3(tmp, A, Z);
  for (int i = 0; i < 9; i++) {
    g.A[q][i] = A[i];
    g.Z[q][i] = -Z[i];
  }
  g.T[q][0] = T[0];
  g.T[q][1] = T[3];
  g.T[q][2] = T[6];
  g.T[q][3] = T[1];
  g.T[q][4] = T[4];
  g.T[q][5] = T[7];
  g.scale[q] = det * 1.0;
  // e0ty(i,j) = A0i A0j g11 + A1i A1j g22 + (A0i A1j + A1i A0j) g12
  //           + (A1i A2j + A2i A1j) g23 + (A0i A2j + A2i A0j) g13
  const int ii[5] = {0, 1, 0, 1, 0}, jj[5] = {0, 1, 1, 2, 2};
  const double fac[5] = {1.0, 1.0, 2.0, 2.0, 2.0};
  for (int m = 0; m < 5; m++) {
    const int i = ii[m], j = jj[m];
    const double A0i = A[i], A1i = A[3 + i], A2i = A[6 + i];
    const double A0j = A[j], A1j = A[3 + j], A2j = A[6 + j];
    g.coef[q][5 * m + 0] = fac[m] * A0i * A0j;
    g.coef[q][5 * m + 1] = fac[m] * A1i * A1j;
    g.coef[q][5 * m + 2] = fac[m] * (A0i * A1j + A1i * A0j);
    g.coef[q][5 * m + 3] = fac[m] * (A1i * A2j + A2i * A1j);
    g.coef[q][5 * m + 4] = fac[m] * (A0i * A2j + A2i * A0j);
  }
}

// Shape function of node a and its derivatives at (xi, eta):
// N_a = (1 + xa xi)(1 + ea eta) / 4 with (xa, ea) the node's corner.
__device__ __forceinline__ void shape_a(int a, double xi, double eta,
                                        double& N, double& Nx, double& Ne) {
  const double xa = (a & 1) ? 1.0 : -1.0, ea = (a & 2) ? 1.0 : -1.0;
  const double fx = 0.5 * (1.0 + xa * xi), fe = 0.5 * (1.0 + ea * eta);
  N = fx * fe;
  Nx = 0.5 * xa * fe;
  Ne = 0.5 * ea * fx;
}

// Quadrature-point independent part of column j (DOF 6a + k): the nine
// tying-strain derivatives and the drill contributions at the four nodes.
struct BColumnPrep {
  double ety[9];
  double drill_node[4];
};

__device__ __forceinline__ void b_column_prep(int j, const Geom& g,
                                              BColumnPrep& p) {
  const int a = j / 6, k = j % 6;
  const bool rot = k >= 3;
  const int kk = rot ? k - 3 : k;
  double N[5], Nx[5], Ne[5];  // node a's shape functions at the 5 locations
#pragma unroll
  for (int l = 0; l < 5; l++) {
    double xi, eta;
    tying_loc(l, xi, eta);
    shape_a(a, xi, eta, N[l], Nx[l], Ne[l]);
  }
  double* ety = p.ety;
  if (!rot) {
    ety[0] = Nx[0] * g.tXxi[0][kk];
    ety[1] = Nx[1] * g.tXxi[1][kk];
    ety[2] = Ne[2] * g.tXeta[2][kk];
    ety[3] = Ne[3] * g.tXeta[3][kk];
    ety[4] = 0.5 * (Nx[4] * g.tXeta[4][kk] + Ne[4] * g.tXxi[4][kk]);
    ety[5] = 0.5 * Ne[2] * g.tn0[2][kk];
    ety[6] = 0.5 * Ne[3] * g.tn0[3][kk];
    ety[7] = 0.5 * Nx[0] * g.tn0[0][kk];
    ety[8] = 0.5 * Nx[1] * g.tn0[1][kk];
  } else {
    double c[3];
    ety[0] = ety[1] = ety[2] = ety[3] = ety[4] = 0.0;
    cross(g.fn[a], g.tXeta[2], c);
    ety[5] = 0.5 * N[2] * c[kk];
    cross(g.fn[a], g.tXeta[3], c);
    ety[6] = 0.5 * N[3] * c[kk];
    cross(g.fn[a], g.tXxi[0], c);
    ety[7] = 0.5 * N[0] * c[kk];
    cross(g.fn[a], g.tXxi[1], c);
    ety[8] = 0.5 * N[1] * c[kk];
  }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    if (rot) {
      p.drill_node[b] = 0.0;
      continue;
    }
    // N_a,xi at node b is +-1/2 when a and b share the eta row, else 0
    const double xa = (a & 1) ? 1.0 : -1.0, ea = (a & 2) ? 1.0 : -1.0;
    const double nxi = ((a >> 1) == (b >> 1)) ? 0.5 * xa : 0.0;
    const double neta = ((a & 1) == (b & 1)) ? 0.5 * ea : 0.0;
    p.drill_node[b] = 0.5 * (nxi * g.dv[b][kk] + neta * g.dw[b][kk]);
  }
}

// Column j of B_q at quadrature point q: out[s], s = 0..8.
__device__ __forceinline__ void b_column_q(int j, int q, const Geom& g,
                                           const BColumnPrep& p,
                                           double out[9]) {
  const int a = j / 6, k = j % 6;
  const bool rot = k >= 3;
  const int kk = rot ? k - 3 : k;
  const double xi = (q & 1) ? kRt3 : -kRt3;
  const double eta = (q & 2) ? kRt3 : -kRt3;
  double Na, Nxa, Nea;
  shape_a(a, xi, eta, Na, Nxa, Nea);
  const double* ety = p.ety;
  // interpolated tying strains (basis.h:314-368)
  double gt[5];
  gt[0] = 0.5 * (1.0 - eta) * ety[0] + 0.5 * (1.0 + eta) * ety[1];
  gt[1] = 0.5 * (1.0 - xi) * ety[2] + 0.5 * (1.0 + xi) * ety[3];
  gt[2] = ety[4];
  gt[3] = 0.5 * (1.0 - xi) * ety[5] + 0.5 * (1.0 + xi) * ety[6];
  gt[4] = 0.5 * (1.0 - eta) * ety[7] + 0.5 * (1.0 + eta) * ety[8];
  const double* cf = g.coef[q];
  double e[5];
#pragma unroll
  for (int m = 0; m < 5; m++)
    e[m] = cf[5 * m] * gt[0] + cf[5 * m + 1] * gt[1] + cf[5 * m + 2] * gt[2] +
           cf[5 * m + 3] * gt[3] + cf[5 * m + 4] * gt[4];
  // bending: derivative of u1x(k2, j2)
  const double* A = g.A[q];
  const double* Z = g.Z[q];
  double u00, u11, u01, u10;
  if (!rot) {
    const double c0 = Nxa * Z[0] + Nea * Z[3];
    const double c1 = Nxa * Z[1] + Nea * Z[4];
    const double t0k = g.T[q][kk], t1k = g.T[q][3 + kk];
    u00 = c0 * t0k;
    u11 = c1 * t1k;
    u01 = c1 * t0k;
    u10 = c0 * t1k;
  } else {
    const double c0 = Nxa * A[0] + Nea * A[3] + Na * Z[6];
    const double c1 = Nxa * A[1] + Nea * A[4] + Na * Z[7];
    const double t0[3] = {g.T[q][0], g.T[q][1], g.T[q][2]};
    const double t1[3] = {g.T[q][3], g.T[q][4], g.T[q][5]};
    double f0[3], f1[3];
    cross(g.fn[a], t0, f0);
    cross(g.fn[a], t1, f1);
    u00 = c0 * f0[kk];
    u11 = c1 * f1[kk];
    u01 = c1 * f0[kk];
    u10 = c0 * f1[kk];
  }
  double et;
  if (!rot) {
    double Nb[4], d1, d2;
#pragma unroll
    for (int b = 0; b < 4; b++) shape_a(b, xi, eta, Nb[b], d1, d2);
    et = Nb[0] * p.drill_node[0] + Nb[1] * p.drill_node[1] +
         Nb[2] * p.drill_node[2] + Nb[3] * p.drill_node[3];
  } else {
    et = -Na * g.dn[a][kk];
  }
  out[0] = e[0];
  out[1] = e[1];
  out[2] = e[2];
  out[3] = u00;
  out[4] = u11;
  out[5] = u01 + u10;
  out[6] = e[3];
  out[7] = e[4];
  out[8] = et;
}

// Column j of the local displacement gradients at quadrature point q, the
// entries (r, c), c = 0, 1, of u0x = T^T [U_xi | U_eta | d0] Xd^-1 T and
// u1x = T^T ([d0_xi | d0_eta | 0] Xd^-1 T + [U_xi | U_eta | d0] Z)
// (basis.h:470-492), in the order (0,0) (0,1) (1,0) (1,1) (2,0) (2,1):
//   displacement DOF (a, k):  u0x_rc = t_r[k] (N_a,xi A_0c + N_a,eta A_1c)
//                             u1x_rc = t_r[k] (N_a,xi Z_0c + N_a,eta Z_1c)
//   rotation DOF (a, k):      u0x_rc = (fn_a x t_r)[k] N_a A_2c
//                             u1x_rc = (fn_a x t_r)[k] (N_a,xi A_0c +
//                                       N_a,eta A_1c + N_a Z_2c)
// with A = Xd^-1 T, Z = -Xd^-1 Xdz Xd^-1 T and t_r the columns of T
// (d_a = q_a x fn_a, director.h:98-107).  Entries (0,0), (1,1), (0,1) + (1,0)
// of u1x are the linear bending strains of b_column_q.
__device__ __forceinline__ void grad_columns_q(int j, int q, const Geom& g,
                                               double g0[6], double g1[6]) {
  const int a = j / 6, k = j % 6;
  const bool rot = k >= 3;
  const int kk = rot ? k - 3 : k;
  const double xi = (q & 1) ? kRt3 : -kRt3;
  const double eta = (q & 2) ? kRt3 : -kRt3;
  double Na, Nxa, Nea;
  shape_a(a, xi, eta, Na, Nxa, Nea);
  const double* A = g.A[q];
  const double* Z = g.Z[q];
  const double t0[3] = {g.T[q][0], g.T[q][1], g.T[q][2]};
  const double t1[3] = {g.T[q][3], g.T[q][4], g.T[q][5]};
  double t2[3];
  cross(t0, t1, t2);  // the normal column of the orthonormal T
  double f[3];        // t_r[k] or (fn_a x t_r)[k]
  if (!rot) {
    f[0] = t0[kk];
    f[1] = t1[kk];
    f[2] = t2[kk];
  } else {
    double c[3];
    cross(g.fn[a], t0, c);
    f[0] = c[kk];
    cross(g.fn[a], t1, c);
    f[1] = c[kk];
    cross(g.fn[a], t2, c);
    f[2] = c[kk];
  }
#pragma unroll
  for (int c = 0; c < 2; c++) {
    const double pa = rot ? Na * A[6 + c] : Nxa * A[c] + Nea * A[3 + c];
    const double pz = rot ? Nxa * A[c] + Nea * A[3 + c] + Na * Z[6 + c]
                          : Nxa * Z[c] + Nea * Z[3 + c];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      g0[2 * r + c] = f[r] * pa;
      g1[2 * r + c] = f[r] * pz;
    }
  }
}

// S = C E for one material point (a2d/a2disotropicshellstress.h:13-51).
__device__ __forceinline__ void stress(const double d[4], const double e[9],
                                       double s[9]) {
  const double E = d[0], nu = d[1], t = d[2], to = d[3];
  const double D = E / (1.0 - nu * nu);
  const double Q0 = D, Q1 = nu * D, G = E / 2.0 / (1.0 + nu);
  const double am = t, ab = -t * t * to, ad = t * t * t / 12.0 + t * t * t * to * to;
  const double m0 = Q0 * e[0] + Q1 * e[1], m1 = Q1 * e[0] + Q0 * e[1],
               m2 = G * e[2];
  const double b0 = Q0 * e[3] + Q1 * e[4], b1 = Q1 * e[3] + Q0 * e[4],
               b2 = G * e[5];
  s[0] = am * m0 + ab * b0;
  s[1] = am * m1 + ab * b1;
  s[2] = am * m2 + ab * b2;
  s[3] = ab * m0 + ad * b0;
  s[4] = ab * m1 + ad * b1;
  s[5] = ab * m2 + ad * b2;
  const double As = (5.0 / 6.0) * t * G;
  s[6] = As * e[6];
  s[7] = As * e[7];
  s[8] = 10.0 * As * e[8];
}

}  // namespace mitc4
}  // namespace gf
