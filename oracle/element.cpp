// Oracle (test infrastructure only): linear MITC4 shell element restated from
// the reference's forward maps.  The Jacobian is formed the way the reference's
// add_element_quadpt_jacobian_col does it for a linear strain: a forward
// ("hforward") pass with a unit DOF vector gives B e_j, and the stress/energy
// stage contracts it, K_ij = sum_q detXd_q w_q (B e_i)^T C (B e_j).
//
// Reference files restated here:
//   include/shell/quadrature.h:5-18      QuadLinearQuadrature (rt3 literal)
//   include/shell/basis.h:55-136         Lagrange shape functions, frames
//   include/shell/basis.h:138-142        node points (tensor order)
//   include/shell/basis.h:235-287        tying points / tying interpolation
//   include/shell/basis.h:292-311        ShellComputeNodeNormals
//   include/shell/basis.h:314-368        interpTyingStrain
//   include/shell/basis.h:428-500        ShellComputeDispGrad
//   include/shell/shell_utils.h:4-40     ShellComputeTransform (no ref axis)
//   include/shell/shell_utils.h:42-104   ShellComputeDrillStrain
//   include/shell/director.h:56-107      LinearizedRotation (C = I - q^x,
//                                         drill strain, director q x n)
//   include/shell/physics.h:95-209       computeTyingStrain
//   include/shell/a2d/a2dshellstrain.h:13-30        LinearShellStrainCore
//   include/shell/a2d/a2disotropicshellstress.h:13-51  stress
//   include/shell/a2d/a2dsymmatrotateframe.h:241-249  e0ty = A^T G A
//   include/shell/data.h:11-12,33-43     shear factor 5/6, drill reg. 10
#include <cmath>
#include <cstring>

#include "oracle.h"

namespace {

const double kRt3 = 0.577350269189626;  // quadrature.h:9 (truncated literal)

void shape(double xi, double eta, double N[4], double Nxi[4], double Neta[4]) {
  // basis.h:105-129 (lagrangeLobatto2DGrad, order 2)
  double na[2] = {0.5 * (1.0 - xi), 0.5 * (1.0 + xi)};
  double nb[2] = {0.5 * (1.0 - eta), 0.5 * (1.0 + eta)};
  double dna[2] = {-0.5, 0.5}, dnb[2] = {-0.5, 0.5};
  for (int ie = 0; ie < 2; ie++)
    for (int ix = 0; ix < 2; ix++) {
      N[2 * ie + ix] = na[ix] * nb[ie];
      Nxi[2 * ie + ix] = dna[ix] * nb[ie];
      Neta[2 * ie + ix] = na[ix] * dnb[ie];
    }
}

// interpFields / interpFieldsGrad (basis.h:136-176): stride = values per node
void interp(const double pt[2], const double* v, int stride, double f[3]) {
  double N[4], Nx[4], Ne[4];
  shape(pt[0], pt[1], N, Nx, Ne);
  for (int k = 0; k < 3; k++) {
    f[k] = 0.0;
    for (int a = 0; a < 4; a++) f[k] += N[a] * v[a * stride + k];
  }
}
void interp_grad(const double pt[2], const double* v, int stride, double fx[3],
                 double fe[3]) {
  double N[4], Nx[4], Ne[4];
  shape(pt[0], pt[1], N, Nx, Ne);
  for (int k = 0; k < 3; k++) {
    fx[k] = 0.0;
    fe[k] = 0.0;
    for (int a = 0; a < 4; a++) {
      fx[k] += Nx[a] * v[a * stride + k];
      fe[k] += Ne[a] * v[a * stride + k];
    }
  }
}

void cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
double dot(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
// columns a, b, c -> row-major 3x3 (basis.h:131-137 assembleFrame)
void frame(const double a[3], const double b[3], const double c[3],
           double F[9]) {
  for (int i = 0; i < 3; i++) {
    F[3 * i] = a[i];
    F[3 * i + 1] = b[i];
    F[3 * i + 2] = c[i];
  }
}
// C = op(A) op(B), 3x3 row-major
void mm(const double* A, const double* B, double* C, bool tA = false,
        bool tB = false) {
  double R[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int k = 0; k < 3; k++) {
        double a = tA ? A[3 * k + i] : A[3 * i + k];
        double b = tB ? B[3 * j + k] : B[3 * k + j];
        s += a * b;
      }
      R[3 * i + j] = s;
    }
  std::memcpy(C, R, sizeof(R));
}
double det3(const double A[9]) {
  return A[8] * (A[0] * A[4] - A[3] * A[1]) -
         A[7] * (A[0] * A[5] - A[3] * A[2]) +
         A[6] * (A[1] * A[5] - A[2] * A[4]);
}
void inv3(const double A[9], double Ai[9]) {
  double d = det3(A);
  double id = 1.0 / d;
  Ai[0] = (A[4] * A[8] - A[5] * A[7]) * id;
  Ai[1] = -(A[1] * A[8] - A[2] * A[7]) * id;
  Ai[2] = (A[1] * A[5] - A[2] * A[4]) * id;
  Ai[3] = -(A[3] * A[8] - A[5] * A[6]) * id;
  Ai[4] = (A[0] * A[8] - A[2] * A[6]) * id;
  Ai[5] = -(A[0] * A[5] - A[2] * A[3]) * id;
  Ai[6] = (A[3] * A[7] - A[4] * A[6]) * id;
  Ai[7] = -(A[0] * A[7] - A[1] * A[6]) * id;
  Ai[8] = (A[0] * A[4] - A[1] * A[3]) * id;
}

// ShellComputeTransform without reference axis (shell_utils.h:4-40)
void transform(const double dXdxi[3], const double n0[3], double T[9]) {
  double nn = std::sqrt(dot(n0, n0));
  double nhat[3] = {n0[0] / nn, n0[1] / nn, n0[2] / nn};
  double d = dot(nhat, dXdxi);
  double tmp[3];
  for (int i = 0; i < 3; i++) tmp[i] = 1.0 * dXdxi[i] + (-d) * nhat[i];
  double tn = std::sqrt(dot(tmp, tmp));
  double t1[3] = {tmp[0] / tn, tmp[1] / tn, tmp[2] / tn};
  double t2[3];
  cross(nhat, t1, t2);
  for (int i = 0; i < 3; i++) {
    T[3 * i] = t1[i];
    T[3 * i + 1] = t2[i];
    T[3 * i + 2] = nhat[i];
  }
}

void node_point(int a, double pt[2]) {  // basis.h:138-142
  pt[0] = -1.0 + 2.0 * (a % 2);
  pt[1] = -1.0 + 2.0 * (a / 2);
}

// ShellComputeNodeNormals (basis.h:292-311)
void node_normals(const double X[12], double fn[12]) {
  for (int a = 0; a < 4; a++) {
    double pt[2], Xxi[3], Xeta[3], t[3];
    node_point(a, pt);
    interp_grad(pt, X, 3, Xxi, Xeta);
    cross(Xxi, Xeta, t);
    double nrm = std::sqrt(dot(t, t));
    for (int k = 0; k < 3; k++) fn[3 * a + k] = (1.0 / nrm) * t[k];
  }
}

// ShellComputeDrillStrain (shell_utils.h:42-104) with LinearizedRotation
double drill_strain(const double pt[2], const double X[12], const double v[24],
                    const double fn[12]) {
  double etn[4];
  for (int a = 0; a < 4; a++) {
    double np[2], Xxi[3], Xeta[3], Xd[9], T[9];
    node_point(a, np);
    interp_grad(np, X, 3, Xxi, Xeta);
    frame(Xxi, Xeta, &fn[3 * a], Xd);
    transform(Xxi, &fn[3 * a], T);
    double uxi[3], ueta[3], zero[3] = {0, 0, 0}, u0xn[9];
    interp_grad(np, v, 6, uxi, ueta);
    frame(uxi, ueta, zero, u0xn);
    // C = I - q^x  (director.h:56-65 setMatSkew(-1, q) + unit diagonal)
    const double* q = &v[6 * a + 3];
    double C[9] = {1.0, q[2], -q[1], -q[2], 1.0, q[0], q[1], -q[0], 1.0};
    double tmp[9], Ct[9];
    mm(T, C, tmp, true, false);
    mm(tmp, T, Ct);
    double Xdinv[9], XdinvTn[9];
    inv3(Xd, Xdinv);
    mm(Xdinv, T, XdinvTn);
    mm(u0xn, XdinvTn, tmp);
    mm(T, tmp, u0xn, true, false);
    etn[a] = 0.5 * (Ct[3] + u0xn[3] - Ct[1] - u0xn[1]);  // director.h:79-82
  }
  double N[4], Nx[4], Ne[4];
  shape(pt[0], pt[1], N, Nx, Ne);
  double et = 0.0;
  for (int a = 0; a < 4; a++) et += N[a] * etn[a];
  return et;
}

// computeTyingStrain (physics.h:95-209); order g11(2) g22(2) g12(1) g23(2) g13(2)
void tying_strain(const double X[12], const double fn[12], const double v[24],
                  const double d[12], double ety[9]) {
  double Xxi[3], Xeta[3], Uxi[3], Ueta[3], d0[3], n0[3];
  for (int t = 0; t < 2; t++) {  // g11 at (0, -1), (0, 1)
    double pt[2] = {0.0, t == 0 ? -1.0 : 1.0};
    interp_grad(pt, X, 3, Xxi, Xeta);
    interp_grad(pt, v, 6, Uxi, Ueta);
    ety[0 + t] = dot(Uxi, Xxi);
  }
  for (int t = 0; t < 2; t++) {  // g22 at (-1, 0), (1, 0)
    double pt[2] = {t == 0 ? -1.0 : 1.0, 0.0};
    interp_grad(pt, X, 3, Xxi, Xeta);
    interp_grad(pt, v, 6, Uxi, Ueta);
    ety[2 + t] = dot(Ueta, Xeta);
  }
  {  // g12 at (0, 0)
    double pt[2] = {0.0, 0.0};
    interp_grad(pt, X, 3, Xxi, Xeta);
    interp_grad(pt, v, 6, Uxi, Ueta);
    ety[4] = 0.5 * (dot(Uxi, Xeta) + dot(Ueta, Xxi));
  }
  for (int t = 0; t < 2; t++) {  // g23 at (-1, 0), (1, 0)
    double pt[2] = {t == 0 ? -1.0 : 1.0, 0.0};
    interp_grad(pt, X, 3, Xxi, Xeta);
    interp_grad(pt, v, 6, Uxi, Ueta);
    interp(pt, d, 3, d0);
    interp(pt, fn, 3, n0);
    ety[5 + t] = 0.5 * (dot(Xeta, d0) + dot(n0, Ueta));
  }
  for (int t = 0; t < 2; t++) {  // g13 at (0, -1), (0, 1)
    double pt[2] = {0.0, t == 0 ? -1.0 : 1.0};
    interp_grad(pt, X, 3, Xxi, Xeta);
    interp_grad(pt, v, 6, Uxi, Ueta);
    interp(pt, d, 3, d0);
    interp(pt, fn, 3, n0);
    ety[7 + t] = 0.5 * (dot(Xxi, d0) + dot(n0, Uxi));
  }
}

// interpTyingStrain (basis.h:314-368) -> packed SymMat [11,12,13,22,23,33]
void interp_tying(const double pt[2], const double ety[9], double g[6]) {
  double nxi[2] = {0.5 * (1.0 - pt[0]), 0.5 * (1.0 + pt[0])};
  double net[2] = {0.5 * (1.0 - pt[1]), 0.5 * (1.0 + pt[1])};
  g[0] = net[0] * ety[0] + net[1] * ety[1];  // g11
  g[3] = nxi[0] * ety[2] + nxi[1] * ety[3];  // g22
  g[1] = ety[4];                             // g12
  g[4] = nxi[0] * ety[5] + nxi[1] * ety[6];  // g23
  g[2] = net[0] * ety[7] + net[1] * ety[8];  // g13
  g[5] = 0.0;
}

// SymMatRotateFrame: C = A^T B A with packed symmetric B, C
void sym_rotate(const double A[9], const double g[6], double e[6]) {
  double G[9] = {g[0], g[1], g[2], g[1], g[3], g[4], g[2], g[4], g[5]};
  double tmp[9], E[9];
  mm(A, G, tmp, true, false);
  mm(tmp, A, E);
  e[0] = E[0];
  e[1] = E[1];
  e[2] = E[2];
  e[3] = E[4];
  e[4] = E[5];
  e[5] = E[8];
}

}  // namespace

namespace orc_internal {
// Strain vector E (9) at quadrature point iq and the local displacement
// gradients u0x, u1x (3 x 3 row-major, basis.h:470-492); returns
// detXd * weight.  Follows add_element_quadpt_residual's forward block
// (include/shell/elem_group.h:123-160).
double strain_full(int iq, const double X[12], const double v[24],
                   double E[9], double u0xl[9], double u1xl[9]) {
  double pt[2] = {iq % 2 == 0 ? -kRt3 : kRt3, iq / 2 == 0 ? -kRt3 : kRt3};
  double weight = 1.0 * 1.0;
  double fn[12];
  node_normals(X, fn);
  double et = drill_strain(pt, X, v, fn);
  double d[12];
  for (int a = 0; a < 4; a++) cross(&v[6 * a + 3], &fn[3 * a], &d[3 * a]);
  double ety[9];
  tying_strain(X, fn, v, d, ety);

  // ShellComputeDispGrad (basis.h:428-500)
  double Xxi[3], Xeta[3], nxi[3], neta[3], n0[3], zero[3] = {0, 0, 0};
  interp(pt, fn, 3, n0);
  interp_grad(pt, X, 3, Xxi, Xeta);
  interp_grad(pt, fn, 3, nxi, neta);
  double Xd[9], Xdz[9], T[9];
  frame(Xxi, Xeta, n0, Xd);
  frame(nxi, neta, zero, Xdz);
  transform(Xxi, n0, T);
  double d0[3], d0xi[3], d0eta[3], u0xi[3], u0eta[3];
  interp(pt, d, 3, d0);
  interp_grad(pt, d, 3, d0xi, d0eta);
  interp_grad(pt, v, 6, u0xi, u0eta);
  double u0x[9], u1x[9];
  frame(u0xi, u0eta, d0, u0x);
  frame(d0xi, d0eta, zero, u1x);
  double Xdinv[9], XdinvT[9], XdinvzT[9], tmp[9], tmp2[9];
  inv3(Xd, Xdinv);
  double detXd = det3(Xd);
  mm(Xdinv, T, XdinvT);
  mm(Xdinv, Xdz, tmp);
  for (int i = 0; i < 9; i++) tmp[i] *= -1.0;
  mm(tmp, XdinvT, XdinvzT);
  mm(u1x, XdinvT, tmp);
  mm(u0x, XdinvzT, tmp2);
  for (int i = 0; i < 9; i++) tmp[i] += tmp2[i];
  mm(T, tmp, u1x, true, false);
  // u0x = T^T u0d Xd^-1 T (basis.h:490-492)
  mm(u0x, XdinvT, tmp);
  mm(T, tmp, u0xl, true, false);
  for (int i = 0; i < 9; i++) u1xl[i] = u1x[i];
  double g[6], e0ty[6];
  interp_tying(pt, ety, g);
  sym_rotate(XdinvT, g, e0ty);

  // LinearShellStrainCore (a2dshellstrain.h:13-30)
  E[0] = e0ty[0];
  E[1] = e0ty[3];
  E[2] = 2.0 * e0ty[1];
  E[3] = u1x[0];
  E[4] = u1x[4];
  E[5] = u1x[1] + u1x[3];
  E[6] = 2.0 * e0ty[4];
  E[7] = 2.0 * e0ty[2];
  E[8] = et;
  return detXd * weight;
}

// NonlinearShellStrainCore (a2d/a2dshellstrain.h:72-93): the bending
// strains gain the quadratic terms of the displacement gradients (row-major
// u0x, u1x): k11 += u0x(:,0).u1x(:,0), k22 += u0x(:,1).u1x(:,1),
// k12 += u0x(:,0).u1x(:,1) + u1x(:,0).u0x(:,1).  The tying strains and the
// director stay linear (physics.h:214-219, 286-300).
void add_nonlinear(const double u0x[9], const double u1x[9], double E[9]) {
  E[3] += u0x[0] * u1x[0] + u0x[3] * u1x[3] + u0x[6] * u1x[6];
  E[4] += u0x[1] * u1x[1] + u0x[4] * u1x[4] + u0x[7] * u1x[7];
  E[5] += u0x[0] * u1x[1] + u0x[3] * u1x[4] + u0x[6] * u1x[7] +
          u1x[0] * u0x[1] + u1x[3] * u0x[4] + u1x[6] * u0x[7];
}
}  // namespace orc_internal

extern "C" double orc_mitc4_strain(int iq, const double X[12],
                                   const double v[24], double E[9]) {
  double u0x[9], u1x[9];
  return orc_internal::strain_full(iq, X, v, E, u0x, u1x);
}

extern "C" double orc_mitc4_strain_nl(int iq, const double X[12],
                                      const double v[24], double E[9],
                                      double u0x[9], double u1x[9]) {
  const double sc = orc_internal::strain_full(iq, X, v, E, u0x, u1x);
  orc_internal::add_nonlinear(u0x, u1x, E);
  return sc;
}

namespace {
// IsotropicShellStressCore (a2disotropicshellstress.h:13-51)
void stress(const double data[4], const double e[9], double s[9]) {
  double E = data[0], nu = data[1], t = data[2], to = data[3];
  double D = E / (1.0 - nu * nu);
  double C[6] = {D, nu * D, 0.0, D, 0.0, E / 2.0 / (1.0 + nu)};
  auto symv = [&](double sc, const double* x, double* y, bool add) {
    double y0 = sc * (C[0] * x[0] + C[1] * x[1] + C[2] * x[2]);
    double y1 = sc * (C[1] * x[0] + C[3] * x[1] + C[4] * x[2]);
    double y2 = sc * (C[2] * x[0] + C[4] * x[1] + C[5] * x[2]);
    if (add) {
      y[0] += y0;
      y[1] += y1;
      y[2] += y2;
    } else {
      y[0] = y0;
      y[1] = y1;
      y[2] = y2;
    }
  };
  symv(t, e, s, false);
  symv(-t * t * to, &e[3], s, true);
  symv(-t * t * to, e, &s[3], false);
  symv(t * t * t / 12.0, &e[3], &s[3], true);
  symv(t * t * t * to * to, &e[3], &s[3], true);
  double As = (5.0 / 6.0) * t * C[5];
  s[6] = As * e[6];
  s[7] = As * e[7];
  s[8] = 10.0 * As * e[8];
}
}  // namespace

extern "C" void orc_mitc4_element(const double X[12], const double vars[24],
                                  const double data[4], double* K, double* R,
                                  double* energy) {
  double Bq[4][24][9], Eu[4][9], scale[4];
  double unit[24];
  for (int q = 0; q < 4; q++) {
    scale[q] = orc_mitc4_strain(q, X, vars, Eu[q]);
    for (int j = 0; j < 24; j++) {
      std::memset(unit, 0, sizeof(unit));
      unit[j] = 1.0;
      orc_mitc4_strain(q, X, unit, Bq[q][j]);
    }
  }
  double U = 0.0;
  if (K) std::memset(K, 0, sizeof(double) * 576);
  if (R) std::memset(R, 0, sizeof(double) * 24);
  for (int q = 0; q < 4; q++) {
    double Su[9];
    stress(data, Eu[q], Su);
    double eS = 0.0;
    for (int s = 0; s < 9; s++) eS += Eu[q][s] * Su[s];
    U += 0.5 * scale[q] * eS;
    for (int j = 0; j < 24; j++) {
      double Sj[9];
      stress(data, Bq[q][j], Sj);
      if (K) {
        for (int i = 0; i < 24; i++) {
          double acc = 0.0;
          for (int s = 0; s < 9; s++) acc += Bq[q][i][s] * Sj[s];
          K[24 * i + j] += scale[q] * acc;
        }
      }
      if (R) {
        double acc = 0.0;
        for (int s = 0; s < 9; s++) acc += Bq[q][j][s] * Su[s];
        R[j] += scale[q] * acc;
      }
    }
  }
  if (energy) *energy = U;
}

// Nonlinear element (IsotropicShell<..., isNonlinear = true>, physics.h:6-17):
// the energy U = sum_q 1/2 detXd_q E_q(u)^T C E_q(u) with E_q from
// NonlinearShellStrainCore, its gradient R and its Hessian K, formed from the
// linear maps (unit DOF vectors through the forward maps) of the linear
// strain, u0x and u1x:
//   dE_s/du_j = L_sj + [s = 3..5] d(quadratic)/du_j,
//   d2E_s/du_i du_j = sum over the pairs (m, m') the quadratic term of s
//                     multiplies of u0x_m(e_i) u1x_m'(e_j) + u0x_m(e_j) u1x_m'(e_i),
//   R_j = sum_q detXd_q (dE/du_j) . C E,
//   K_ij = sum_q detXd_q [(dE/du_i) . C (dE/du_j) + sum_s (C E)_s d2E_s/du_i du_j].
extern "C" void orc_mitc4_element_nl(const double X[12], const double vars[24],
                                     const double data[4], double* K,
                                     double* R, double* energy) {
  if (K) std::memset(K, 0, sizeof(double) * 576);
  if (R) std::memset(R, 0, sizeof(double) * 24);
  double U = 0.0;
  // (m, m') pairs of the quadratic terms of E3, E4, E5 (row-major indices)
  const int p3[3][2] = {{0, 0}, {3, 3}, {6, 6}};
  const int p4[3][2] = {{1, 1}, {4, 4}, {7, 7}};
  const int p5[6][2] = {{0, 1}, {3, 4}, {6, 7}, {1, 0}, {4, 3}, {7, 6}};
  for (int q = 0; q < 4; q++) {
    double Eu[9], U0[9], U1[9];
    const double sc = orc_mitc4_strain_nl(q, X, vars, Eu, U0, U1);
    double L[24][9], G0[24][9], G1[24][9], unit[24];
    for (int j = 0; j < 24; j++) {
      std::memset(unit, 0, sizeof(unit));
      unit[j] = 1.0;
      orc_internal::strain_full(q, X, unit, L[j], G0[j], G1[j]);
    }
    // dE/du_j
    double Bt[24][9];
    for (int j = 0; j < 24; j++) {
      for (int s = 0; s < 9; s++) Bt[j][s] = L[j][s];
      for (int t = 0; t < 3; t++) {
        Bt[j][3] += G0[j][p3[t][0]] * U1[p3[t][1]] + U0[p3[t][0]] * G1[j][p3[t][1]];
        Bt[j][4] += G0[j][p4[t][0]] * U1[p4[t][1]] + U0[p4[t][0]] * G1[j][p4[t][1]];
      }
      for (int t = 0; t < 6; t++)
        Bt[j][5] += G0[j][p5[t][0]] * U1[p5[t][1]] + U0[p5[t][0]] * G1[j][p5[t][1]];
    }
    double Su[9];
    stress(data, Eu, Su);
    double eS = 0.0;
    for (int s = 0; s < 9; s++) eS += Eu[s] * Su[s];
    U += 0.5 * sc * eS;
    if (R)
      for (int j = 0; j < 24; j++) {
        double acc = 0.0;
        for (int s = 0; s < 9; s++) acc += Bt[j][s] * Su[s];
        R[j] += sc * acc;
      }
    if (K)
      for (int j = 0; j < 24; j++) {
        double Sj[9];
        stress(data, Bt[j], Sj);
        for (int i = 0; i < 24; i++) {
          double acc = 0.0;
          for (int s = 0; s < 9; s++) acc += Bt[i][s] * Sj[s];
          double geo = 0.0;
          for (int t = 0; t < 3; t++) {
            geo += Su[3] * (G0[i][p3[t][0]] * G1[j][p3[t][1]] +
                            G0[j][p3[t][0]] * G1[i][p3[t][1]]);
            geo += Su[4] * (G0[i][p4[t][0]] * G1[j][p4[t][1]] +
                            G0[j][p4[t][0]] * G1[i][p4[t][1]]);
          }
          for (int t = 0; t < 6; t++)
            geo += Su[5] * (G0[i][p5[t][0]] * G1[j][p5[t][1]] +
                            G0[j][p5[t][0]] * G1[i][p5[t][1]]);
          K[24 * i + j] += sc * (acc + geo);
        }
      }
  }
  if (energy) *energy = U;
}
