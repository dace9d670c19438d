/*
 * geofem_oracle.c -- FP64 CPU restatement of geo_simulator's Tet10
 * assemble-and-solve path.  TEST INFRASTRUCTURE ONLY (see the header).
 *
 * Parallel structure mirrors the reference: element kernels run in parallel
 * (OpenMP here, Rayon par_iter there: src/fem/assembly.rs:77-104,399-410);
 * the triplet merge, CSR conversion, Dirichlet elimination and CG are serial
 * (src/fem/assembly.rs:107-116,168-230, src/linalg/iterative.rs:81-110).
 */
#include "geofem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Quadrature tables: src/fem/quadrature.rs:13-104 (literal constants).      */
/* ------------------------------------------------------------------------ */
#define QA 0.5854101966249685
#define QB 0.1381966011250105
static const double Q1_PTS[1][4] = {{0.25, 0.25, 0.25, 0.25}};
static const double Q1_W[1] = {1.0 / 6.0};
static const double Q4_PTS[4][4] = {
    {QA, QB, QB, QB}, {QB, QA, QB, QB}, {QB, QB, QA, QB}, {QB, QB, QB, QA}};
static const double Q4_W[4] = {1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0, 1.0 / 24.0};
static const double Q5_PTS[5][4] = {{0.25, 0.25, 0.25, 0.25},
                                    {1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 0.5},
                                    {1.0 / 6.0, 1.0 / 6.0, 0.5, 1.0 / 6.0},
                                    {1.0 / 6.0, 0.5, 1.0 / 6.0, 1.0 / 6.0},
                                    {0.5, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0}};
static const double Q5_W[5] = {-2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0,
                               3.0 / 40.0};
#define A1 0.25
#define A2 0.0714285714285714
#define B2 0.7857142857142857
#define A3 0.3994035761667992
#define B3 0.1005964238332008
static const double Q11_PTS[11][4] = {
    {A1, A1, A1, A1}, {A2, A2, A2, B2}, {A2, A2, B2, A2}, {A2, B2, A2, A2},
    {B2, A2, A2, A2}, {A3, A3, B3, B3}, {A3, B3, A3, B3}, {A3, B3, B3, A3},
    {B3, A3, A3, B3}, {B3, A3, B3, A3}, {B3, B3, A3, A3}};
static const double Q11_W[11] = {
    -0.01315555555555556, 0.007622222222222222, 0.007622222222222222,
    0.007622222222222222, 0.007622222222222222, 0.024888888888888888,
    0.024888888888888888, 0.024888888888888888, 0.024888888888888888,
    0.024888888888888888, 0.024888888888888888};

int or_quad_rule(int npts, const double **pts, const double **weights) {
  switch (npts) {
  case 1: *pts = &Q1_PTS[0][0]; *weights = Q1_W; return OR_OK;
  case 4: *pts = &Q4_PTS[0][0]; *weights = Q4_W; return OR_OK;
  case 5: *pts = &Q5_PTS[0][0]; *weights = Q5_W; return OR_OK;
  case 11: *pts = &Q11_PTS[0][0]; *weights = Q11_W; return OR_OK;
  default: return OR_ERR_INVALID_ARG;
  }
}

/* ------------------------------------------------------------------------ */
/* Tet10 basis: src/fem/basis.rs                                             */
/* ------------------------------------------------------------------------ */

/* src/fem/basis.rs:56-73 */
void or_shape_functions(const double L[4], double N[10]) {
  const double L0 = L[0], L1 = L[1], L2 = L[2], L3 = L[3];
  N[0] = L0 * (2.0 * L0 - 1.0);
  N[1] = L1 * (2.0 * L1 - 1.0);
  N[2] = L2 * (2.0 * L2 - 1.0);
  N[3] = L3 * (2.0 * L3 - 1.0);
  N[4] = 4.0 * L0 * L1;
  N[5] = 4.0 * L1 * L2;
  N[6] = 4.0 * L2 * L0;
  N[7] = 4.0 * L0 * L3;
  N[8] = 4.0 * L1 * L3;
  N[9] = 4.0 * L2 * L3;
}

/* src/fem/basis.rs:85-110 */
void or_shape_derivs_bary(const double L[4], double dN[10][4]) {
  const double L0 = L[0], L1 = L[1], L2 = L[2], L3 = L[3];
  memset(dN, 0, sizeof(double) * 40);
  dN[0][0] = 4.0 * L0 - 1.0;
  dN[1][1] = 4.0 * L1 - 1.0;
  dN[2][2] = 4.0 * L2 - 1.0;
  dN[3][3] = 4.0 * L3 - 1.0;
  dN[4][0] = 4.0 * L1; dN[4][1] = 4.0 * L0;
  dN[5][1] = 4.0 * L2; dN[5][2] = 4.0 * L1;
  dN[6][0] = 4.0 * L2; dN[6][2] = 4.0 * L0;
  dN[7][0] = 4.0 * L3; dN[7][3] = 4.0 * L0;
  dN[8][1] = 4.0 * L3; dN[8][3] = 4.0 * L1;
  dN[9][2] = 4.0 * L3; dN[9][3] = 4.0 * L2;
}

/* src/fem/basis.rs:196-223: J[r][c] = sum_i x_i[r] * (dN_i/dL_{c+1} - dN_i/dL_0),
 * accumulated over the 10 nodes in order. */
void or_jacobian(const double L[4], const double nodes[10][3], double J[3][3]) {
  double dN[10][4];
  or_shape_derivs_bary(L, dN);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) J[r][c] = 0.0;
  for (int i = 0; i < 10; ++i) {
    const double d1 = dN[i][1] - dN[i][0];
    const double d2 = dN[i][2] - dN[i][0];
    const double d3 = dN[i][3] - dN[i][0];
    for (int r = 0; r < 3; ++r) {
      const double xr = nodes[i][r];
      J[r][0] += xr * d1;
      J[r][1] += xr * d2;
      J[r][2] += xr * d3;
    }
  }
}

/* 3x3 determinant by cofactor expansion along the first row (the textbook
 * form; nalgebra's exact rounding order is parity unpinned). */
double or_det3(const double m[3][3]) {
  const double minor_12_23 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  const double minor_11_23 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  const double minor_11_22 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  return m[0][0] * minor_12_23 - m[0][1] * minor_11_23 + m[0][2] * minor_11_22;
}

/* Adjugate / determinant inverse; fails (returns nonzero) when det == 0,
 * matching `J.try_inverse().expect("Singular Jacobian")` (src/fem/basis.rs:247). */
int or_inv3(const double m[3][3], double o[3][3]) {
  const double minor_12_23 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  const double minor_11_23 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  const double minor_11_22 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  const double det =
      m[0][0] * minor_12_23 - m[0][1] * minor_11_23 + m[0][2] * minor_11_22;
  if (det == 0.0) return OR_ERR_SINGULAR_JACOBIAN;
  o[0][0] = minor_12_23 / det;
  o[0][1] = (m[0][2] * m[2][1] - m[2][2] * m[0][1]) / det;
  o[0][2] = (m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det;
  o[1][0] = -minor_11_23 / det;
  o[1][1] = (m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det;
  o[1][2] = (m[0][2] * m[1][0] - m[1][2] * m[0][0]) / det;
  o[2][0] = minor_11_22 / det;
  o[2][1] = (m[0][1] * m[2][0] - m[2][1] * m[0][0]) / det;
  o[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det;
  return OR_OK;
}

/* src/fem/basis.rs:236-276: grad N_i = J^{-T} (dN_i/dL1-dN_i/dL0, ...). */
int or_shape_derivs_cart(const double L[4], const double nodes[10][3],
                         double dNdx[10][3]) {
  double dN[10][4], J[3][3], Ji[3][3];
  or_shape_derivs_bary(L, dN);
  or_jacobian(L, nodes, J);
  if (or_inv3(J, Ji) != OR_OK) return OR_ERR_SINGULAR_JACOBIAN;
  for (int i = 0; i < 10; ++i) {
    const double v0 = dN[i][1] - dN[i][0];
    const double v1 = dN[i][2] - dN[i][0];
    const double v2 = dN[i][3] - dN[i][0];
    for (int r = 0; r < 3; ++r) /* (J^-T v)[r] = sum_m Jinv[m][r] v[m] */
      dNdx[i][r] = Ji[0][r] * v0 + Ji[1][r] * v1 + Ji[2][r] * v2;
  }
  return OR_OK;
}

/* src/fem/basis.rs:352-361 */
double or_element_volume(const double nodes[10][3]) {
  double vol = 0.0, J[3][3];
  for (int q = 0; q < 4; ++q) {
    or_jacobian(Q4_PTS[q], nodes, J);
    vol += fabs(or_det3(J)) * Q4_W[q];
  }
  return vol;
}

/* ------------------------------------------------------------------------ */
/* Element kernels                                                           */
/* ------------------------------------------------------------------------ */

/* src/fem/element.rs:23-61 (4-point rule; any rule for the rule-invariance
 * check). K[i][j] += (k * grad_dot) * w, w = weight * |det J|. */
int or_thermal_stiffness_rule(const double nodes[10][3], double k, int npts,
                              double K[10][10]) {
  const double *pts, *wts;
  if (or_quad_rule(npts, &pts, &wts) != OR_OK) return OR_ERR_INVALID_ARG;
  memset(K, 0, sizeof(double) * 100);
  for (int q = 0; q < npts; ++q) {
    const double *L = pts + 4 * q;
    double dNdx[10][3], J[3][3];
    if (or_shape_derivs_cart(L, nodes, dNdx) != OR_OK)
      return OR_ERR_SINGULAR_JACOBIAN;
    or_jacobian(L, nodes, J);
    const double w = wts[q] * fabs(or_det3(J));
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) {
        const double gd = dNdx[i][0] * dNdx[j][0] + dNdx[i][1] * dNdx[j][1] +
                          dNdx[i][2] * dNdx[j][2];
        K[i][j] += k * gd * w;
      }
  }
  return OR_OK;
}

int or_thermal_stiffness(const double nodes[10][3], double k, double K[10][10]) {
  return or_thermal_stiffness_rule(nodes, k, 4, K);
}

/* src/fem/element.rs:74-105: M_ij += rho_cp * N_i * N_j * w over the 5-point
 * rule, w = w_q |det J|, rho_cp = density * specific_heat (no inverse of J,
 * so no singular-Jacobian panic). */
void or_thermal_mass(const double nodes[10][3], double density, double specific_heat,
                     double M[10][10]) {
  const double *pts, *wts;
  or_quad_rule(5, &pts, &wts);
  const double rho_cp = density * specific_heat;
  memset(M, 0, sizeof(double) * 100);
  for (int q = 0; q < 5; ++q) {
    const double *L = pts + 4 * q;
    double N[10], J[3][3];
    or_shape_functions(L, N);
    or_jacobian(L, nodes, J);
    const double w = wts[q] * fabs(or_det3(J));
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) M[i][j] += rho_cp * N[i] * N[j] * w;
  }
}

/* src/fem/element.rs:117-140: f_i += (N_i * Q) * w. */
int or_thermal_load(const double nodes[10][3], double Q, double fe[10]) {
  for (int i = 0; i < 10; ++i) fe[i] = 0.0;
  for (int q = 0; q < 4; ++q) {
    double N[10], J[3][3];
    or_shape_functions(Q4_PTS[q], N);
    or_jacobian(Q4_PTS[q], nodes, J);
    const double w = Q4_W[q] * fabs(or_det3(J));
    for (int i = 0; i < 10; ++i) fe[i] += N[i] * Q * w;
  }
  return OR_OK;
}

/* src/mechanics/constitutive.rs:30-41 (validation) and :63-102 (D). */
int or_constitutive(double E, double nu, double D[6][6]) {
  if (!(E > 0.0)) return OR_ERR_INVALID_MATERIAL;
  if (!(nu > -1.0 && nu < 0.5)) return OR_ERR_INVALID_MATERIAL;
  const double factor = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double diag = 1.0 - nu, off = nu, shear = (1.0 - 2.0 * nu) / 2.0;
  memset(D, 0, sizeof(double) * 36);
  D[0][0] = diag; D[0][1] = off;  D[0][2] = off;
  D[1][0] = off;  D[1][1] = diag; D[1][2] = off;
  D[2][0] = off;  D[2][1] = off;  D[2][2] = diag;
  D[3][3] = shear; D[4][4] = shear; D[5][5] = shear;
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) D[a][b] = D[a][b] * factor;
  return OR_OK;
}

/* src/mechanics/strain.rs:38-71: Voigt rows [xx,yy,zz,xy,yz,zx]. */
void or_b_matrix(const double dNdx[10][3], double B[6][30]) {
  memset(B, 0, sizeof(double) * 180);
  for (int i = 0; i < 10; ++i) {
    const int c = 3 * i;
    const double dx = dNdx[i][0], dy = dNdx[i][1], dz = dNdx[i][2];
    B[0][c + 0] = dx;
    B[1][c + 1] = dy;
    B[2][c + 2] = dz;
    B[3][c + 0] = dy; B[3][c + 1] = dx;
    B[4][c + 1] = dz; B[4][c + 2] = dy;
    B[5][c + 2] = dx; B[5][c + 0] = dz;
  }
}

/* src/mechanics/element.rs:33-65: dense DB = D*B, BtDB = B^T*(DB),
 * K += w * BtDB, products accumulated in ascending inner index. */
static int elastic_stiffness_D(const double nodes[10][3], const double D[6][6],
                               double K[30][30]) {
  memset(K, 0, sizeof(double) * 900);
  for (int q = 0; q < 4; ++q) {
    const double *L = Q4_PTS[q];
    double J[3][3], dNdx[10][3], B[6][30], DB[6][30];
    or_jacobian(L, nodes, J);
    const double det = or_det3(J);
    if (or_shape_derivs_cart(L, nodes, dNdx) != OR_OK)
      return OR_ERR_SINGULAR_JACOBIAN;
    or_b_matrix(dNdx, B);
    const double w = Q4_W[q] * fabs(det);
    for (int kk = 0; kk < 6; ++kk)
      for (int c = 0; c < 30; ++c) {
        double s = D[kk][0] * B[0][c];
        for (int m = 1; m < 6; ++m) s += D[kk][m] * B[m][c];
        DB[kk][c] = s;
      }
    for (int a = 0; a < 30; ++a)
      for (int c = 0; c < 30; ++c) {
        double s = B[0][a] * DB[0][c];
        for (int kk = 1; kk < 6; ++kk) s += B[kk][a] * DB[kk][c];
        K[a][c] += w * s;
      }
  }
  return OR_OK;
}

int or_elastic_stiffness(const double nodes[10][3], double E, double nu,
                         double K[30][30]) {
  double D[6][6];
  int st = or_constitutive(E, nu, D);
  if (st != OR_OK) return st;
  return elastic_stiffness_D(nodes, D, K);
}

/* NewtonianViscosity::new / constitutive_matrix
 * (src/mechanics/constitutive.rs:179-186, 227-255): lambda_v = zeta - 2mu/3
 * on the normal block, 2mu added on its diagonal and on the shear diagonal. */
int or_viscous_constitutive(double mu, double zeta, double D[6][6]) {
  if (!(mu > 0.0)) return OR_ERR_INVALID_MATERIAL;
  const double f2mu = 2.0 * mu;
  const double lv = zeta - (2.0 * mu / 3.0);
  memset(D, 0, sizeof(double) * 36);
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) D[i][j] = lv;
    D[i][i] += f2mu;
  }
  D[3][3] = f2mu; D[4][4] = f2mu; D[5][5] = f2mu;
  return OR_OK;
}

/* ElasticityElement::viscosity_matrix (src/mechanics/element.rs:87-118):
 * the elastic element's B^T D B with the viscous D. */
int or_viscosity_stiffness(const double nodes[10][3], double mu, double zeta,
                           double K[30][30]) {
  double D[6][6];
  int st = or_viscous_constitutive(mu, zeta, D);
  if (st != OR_OK) return st;
  return elastic_stiffness_D(nodes, D, K);
}

/* ElasticityElement::maxwell_viscoelastic (src/mechanics/element.rs:257-295):
 * K_e with the elastic D, and f_history -= w * B^T sigma_n (B^T sigma as a
 * dense product, ascending Voigt row). */
int or_maxwell_element(const double nodes[10][3], double E, double nu,
                       const double sigma[6], double K[30][30], double fe[30]) {
  double D[6][6];
  int st = or_constitutive(E, nu, D);
  if (st != OR_OK) return st;
  st = elastic_stiffness_D(nodes, D, K);
  if (st != OR_OK) return st;
  for (int a = 0; a < 30; ++a) fe[a] = 0.0;
  for (int q = 0; q < 4; ++q) {
    const double *L = Q4_PTS[q];
    double J[3][3], dNdx[10][3], B[6][30];
    or_jacobian(L, nodes, J);
    const double w = Q4_W[q] * fabs(or_det3(J));
    if (or_shape_derivs_cart(L, nodes, dNdx) != OR_OK) return OR_ERR_SINGULAR_JACOBIAN;
    or_b_matrix(dNdx, B);
    for (int a = 0; a < 30; ++a) {
      double bs = B[0][a] * sigma[0];
      for (int k = 1; k < 6; ++k) bs += B[k][a] * sigma[k];
      fe[a] -= w * bs;
    }
  }
  return OR_OK;
}

/* update_stresses_maxwell for one element (src/mechanics/stress_update.rs:
 * 53-145): the strain increment B du averaged over the 4-point rule's points,
 * the elastic predictor, the deviatoric part relaxed by 1/(1 + dt/tau_M). */
int or_maxwell_stress(const double nodes[10][3], double E, double nu, double visc,
                      double dt, const double du[30], const double sn[6], double out[6]) {
  const double G = E / (2.0 * (1.0 + nu));
  const double tau = visc / (E / (2.0 * (1.0 + nu)));
  const double factor = 1.0 / (1.0 + dt / tau);
  double d[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < 4; ++q) {
    double dNdx[10][3], B[6][30];
    if (or_shape_derivs_cart(Q4_PTS[q], nodes, dNdx) != OR_OK) return OR_ERR_SINGULAR_JACOBIAN;
    or_b_matrix(dNdx, B);
    for (int v = 0; v < 6; ++v) {
      double s = B[v][0] * du[0];
      for (int m = 1; m < 30; ++m) s += B[v][m] * du[m];
      d[v] += s;
    }
  }
  for (int v = 0; v < 6; ++v) d[v] /= 4.0;
  const double Kb = E / (3.0 * (1.0 - 2.0 * nu));
  const double lambda = Kb - 2.0 * G / 3.0;
  double t[6];
  t[0] = sn[0] + lambda * (d[0] + d[1] + d[2]) + 2.0 * G * d[0];
  t[1] = sn[1] + lambda * (d[0] + d[1] + d[2]) + 2.0 * G * d[1];
  t[2] = sn[2] + lambda * (d[0] + d[1] + d[2]) + 2.0 * G * d[2];
  t[3] = sn[3] + G * d[3];
  t[4] = sn[4] + G * d[4];
  t[5] = sn[5] + G * d[5];
  const double p = (t[0] + t[1] + t[2]) / 3.0;
  double sdev[6] = {t[0] - p, t[1] - p, t[2] - p, t[3], t[4], t[5]};
  for (int v = 0; v < 6; ++v) sdev[v] = sdev[v] * factor;
  out[0] = sdev[0] + p;
  out[1] = sdev[1] + p;
  out[2] = sdev[2] + p;
  out[3] = sdev[3];
  out[4] = sdev[4];
  out[5] = sdev[5];
  return OR_OK;
}

/* src/mechanics/body_force.rs:49-84: f[3i+c] += (w * N_i) * (rho g)_c. */
int or_gravity_load(const double nodes[10][3], double rho, const double g[3],
                    double fe[30]) {
  const double bf[3] = {rho * g[0], rho * g[1], rho * g[2]};
  for (int i = 0; i < 30; ++i) fe[i] = 0.0;
  for (int q = 0; q < 4; ++q) {
    double N[10], J[3][3];
    or_shape_functions(Q4_PTS[q], N);
    or_jacobian(Q4_PTS[q], nodes, J);
    const double w = Q4_W[q] * fabs(or_det3(J));
    for (int i = 0; i < 10; ++i)
      for (int c = 0; c < 3; ++c) fe[3 * i + c] += w * N[i] * bf[c];
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Mesh generator: src/mesh_generator_improved.rs:19-169                     */
/* ------------------------------------------------------------------------ */
void or_cube_sizes(int64_t nx, int64_t ny, int64_t nz, int64_t *n_nodes,
                   int64_t *n_elems) {
  *n_nodes = (2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1);
  *n_elems = 6 * nx * ny * nz;
}

typedef struct {
  uint64_t *keys; /* (lo << 32 | hi) + 1, 0 = empty */
  int64_t *vals;
  uint64_t mask;
} edge_map;

static uint64_t mix64(uint64_t x) {
  x ^= x >> 33; x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33; return x;
}

/* get-or-create the midpoint of edge (v1,v2): edge cache of
 * src/mesh_generator_improved.rs:42-57 (new nodes appended in first-touch
 * order, coordinates (p1+p2)/2). */
static int64_t edge_mid(edge_map *em, double *xyz, int64_t *n, int64_t v1,
                        int64_t v2) {
  const int64_t lo = v1 < v2 ? v1 : v2, hi = v1 < v2 ? v2 : v1;
  const uint64_t key = (((uint64_t)lo << 32) | (uint64_t)hi) + 1;
  uint64_t h = mix64(key) & em->mask;
  while (em->keys[h] != 0) {
    if (em->keys[h] == key) return em->vals[h];
    h = (h + 1) & em->mask;
  }
  const int64_t id = (*n)++;
  xyz[3 * id + 0] = (xyz[3 * v1 + 0] + xyz[3 * v2 + 0]) / 2.0;
  xyz[3 * id + 1] = (xyz[3 * v1 + 1] + xyz[3 * v2 + 1]) / 2.0;
  xyz[3 * id + 2] = (xyz[3 * v1 + 2] + xyz[3 * v2 + 2]) / 2.0;
  em->keys[h] = key;
  em->vals[h] = id;
  return id;
}

int or_generate_cube(int64_t nx, int64_t ny, int64_t nz, double lx, double ly,
                     double lz, double *xyz, int64_t *conn) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return OR_ERR_INVALID_ARG;
  int64_t n_nodes, n_elems;
  or_cube_sizes(nx, ny, nz, &n_nodes, &n_elems);
  const double dx = lx / (double)nx, dy = ly / (double)ny, dz = lz / (double)nz;
  int64_t n = 0;
  for (int64_t iz = 0; iz <= nz; ++iz)
    for (int64_t iy = 0; iy <= ny; ++iy)
      for (int64_t ix = 0; ix <= nx; ++ix) {
        xyz[3 * n + 0] = (double)ix * dx;
        xyz[3 * n + 1] = (double)iy * dy;
        xyz[3 * n + 2] = (double)iz * dz;
        ++n;
      }
  uint64_t cap = 1;
  while (cap < (uint64_t)(2 * n_nodes)) cap <<= 1;
  edge_map em;
  em.keys = (uint64_t *)calloc(cap, sizeof(uint64_t));
  em.vals = (int64_t *)malloc(cap * sizeof(int64_t));
  em.mask = cap - 1;
  if (!em.keys || !em.vals) { free(em.keys); free(em.vals); return OR_ERR_OOM; }
  const int64_t sx = 1, sy = nx + 1, sz = (nx + 1) * (ny + 1);
  int64_t e = 0;
  for (int64_t iz = 0; iz < nz; ++iz)
    for (int64_t iy = 0; iy < ny; ++iy)
      for (int64_t ix = 0; ix < nx; ++ix) {
        const int64_t b = ix * sx + iy * sy + iz * sz;
        const int64_t v000 = b, v100 = b + sx, v010 = b + sy, v110 = b + sx + sy;
        const int64_t v001 = b + sz, v101 = b + sx + sz, v011 = b + sy + sz,
                      v111 = b + sx + sy + sz;
        /* src/mesh_generator_improved.rs:105-118 */
        const int64_t tets[6][4] = {{v000, v100, v110, v111},
                                    {v000, v100, v111, v101},
                                    {v000, v010, v110, v111},
                                    {v000, v010, v111, v011},
                                    {v000, v001, v111, v101},
                                    {v000, v001, v111, v011}};
        for (int t = 0; t < 6; ++t) {
          const int64_t v0 = tets[t][0], v1 = tets[t][1];
          int64_t v2 = tets[t][2], v3 = tets[t][3];
          /* orientation fix: src/mesh_generator_improved.rs:139-155 */
          const double *p0 = xyz + 3 * v0, *p1 = xyz + 3 * v1,
                       *p2 = xyz + 3 * v2, *p3 = xyz + 3 * v3;
          const double e1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
          const double e2[3] = {p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]};
          const double e3[3] = {p3[0] - p0[0], p3[1] - p0[1], p3[2] - p0[2]};
          const double cx = e2[1] * e3[2] - e2[2] * e3[1];
          const double cy = e2[2] * e3[0] - e2[0] * e3[2];
          const double cz = e2[0] * e3[1] - e2[1] * e3[0];
          const double det = e1[0] * cx + e1[1] * cy + e1[2] * cz;
          if (det < 0.0) { const int64_t tmp = v2; v2 = v3; v3 = tmp; }
          /* edges in the order e01,e12,e20,e03,e13,e23 (:158-163) */
          const int64_t e01 = edge_mid(&em, xyz, &n, v0, v1);
          const int64_t e12 = edge_mid(&em, xyz, &n, v1, v2);
          const int64_t e20 = edge_mid(&em, xyz, &n, v2, v0);
          const int64_t e03 = edge_mid(&em, xyz, &n, v0, v3);
          const int64_t e13 = edge_mid(&em, xyz, &n, v1, v3);
          const int64_t e23 = edge_mid(&em, xyz, &n, v2, v3);
          int64_t *c = conn + 10 * e;
          c[0] = v0; c[1] = v1; c[2] = v2; c[3] = v3;
          c[4] = e01; c[5] = e12; c[6] = e20; c[7] = e03; c[8] = e13; c[9] = e23;
          ++e;
        }
      }
  free(em.keys);
  free(em.vals);
  return (n == n_nodes && e == n_elems) ? OR_OK : OR_ERR_INVALID_ARG;
}

/* ------------------------------------------------------------------------ */
/* Triplets -> CSR with duplicates summed (sprs TriMat::to_csr semantics,    */
/* src/fem/assembly.rs:57,116,229).  Summation order per slot: triplet       */
/* insertion order (ascending element index for the assemblers).             */
/* ------------------------------------------------------------------------ */
void or_csr_free(or_csr *m) {
  if (!m) return;
  free(m->indptr); free(m->indices); free(m->data);
  m->indptr = NULL; m->indices = NULL; m->data = NULL;
  m->rows = m->cols = m->nnz = 0;
}

typedef struct { int64_t col; double val; } colval;

static int cmp_colval(const void *a, const void *b) {
  const int64_t x = ((const colval *)a)->col, y = ((const colval *)b)->col;
  return (x > y) - (x < y);
}

/* trip_row/trip_col as int32 (when small) or int64; accessor by macro */
static int triplets_to_csr(int64_t rows, int64_t cols, int64_t nt,
                           const int32_t *ti32, const int32_t *tj32,
                           const int64_t *ti64, const int64_t *tj64,
                           const double *tv, or_csr *out) {
#define TI(t) (ti32 ? (int64_t)ti32[t] : ti64[t])
#define TJ(t) (tj32 ? (int64_t)tj32[t] : tj64[t])
  memset(out, 0, sizeof(*out));
  int64_t *rstart = (int64_t *)calloc((size_t)rows + 1, sizeof(int64_t));
  int64_t *perm = (int64_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int64_t));
  int64_t *pos = (int64_t *)malloc((size_t)(cols > 0 ? cols : 1) * sizeof(int64_t));
  colval *acc = NULL;
  int64_t acc_cap = 0;
  if (!rstart || !perm || !pos) goto oom;
  for (int64_t t = 0; t < nt; ++t) {
    const int64_t i = TI(t), j = TJ(t);
    if (i < 0 || i >= rows || j < 0 || j >= cols) {
      free(rstart); free(perm); free(pos);
      return OR_ERR_INVALID_ARG;
    }
    rstart[i + 1]++;
  }
  for (int64_t r = 0; r < rows; ++r) rstart[r + 1] += rstart[r];
  {
    int64_t *cur = (int64_t *)malloc((size_t)(rows > 0 ? rows : 1) * sizeof(int64_t));
    if (!cur) goto oom;
    memcpy(cur, rstart, (size_t)rows * sizeof(int64_t));
    for (int64_t t = 0; t < nt; ++t) perm[cur[TI(t)]++] = t; /* stable */
    free(cur);
  }
  for (int64_t c = 0; c < cols; ++c) pos[c] = -1;
  out->rows = rows;
  out->cols = cols;
  out->indptr = (int64_t *)malloc((size_t)(rows + 1) * sizeof(int64_t));
  out->indices = (int64_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int64_t));
  out->data = (double *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(double));
  if (!out->indptr || !out->indices || !out->data) goto oom;
  int64_t nnz = 0;
  out->indptr[0] = 0;
  for (int64_t r = 0; r < rows; ++r) {
    const int64_t b = rstart[r], e = rstart[r + 1];
    if (e - b > acc_cap) {
      acc_cap = (e - b) * 2;
      free(acc);
      acc = (colval *)malloc((size_t)acc_cap * sizeof(colval));
      if (!acc) goto oom;
    }
    int64_t nu = 0;
    for (int64_t k = b; k < e; ++k) {
      const int64_t t = perm[k];
      const int64_t j = TJ(t);
      if (pos[j] < 0) {
        pos[j] = nu;
        acc[nu].col = j;
        acc[nu].val = tv[t];
        ++nu;
      } else {
        acc[pos[j]].val += tv[t];
      }
    }
    for (int64_t u = 0; u < nu; ++u) pos[acc[u].col] = -1;
    qsort(acc, (size_t)nu, sizeof(colval), cmp_colval);
    for (int64_t u = 0; u < nu; ++u) {
      out->indices[nnz] = acc[u].col;
      out->data[nnz] = acc[u].val;
      ++nnz;
    }
    out->indptr[r + 1] = nnz;
  }
  out->nnz = nnz;
  if (nnz > 0) {
    int64_t *ni = (int64_t *)realloc(out->indices, (size_t)nnz * sizeof(int64_t));
    double *nd = (double *)realloc(out->data, (size_t)nnz * sizeof(double));
    if (ni) out->indices = ni;
    if (nd) out->data = nd;
  }
  free(rstart); free(perm); free(pos); free(acc);
  return OR_OK;
oom:
  free(rstart); free(perm); free(pos); free(acc);
  or_csr_free(out);
  return OR_ERR_OOM;
#undef TI
#undef TJ
}

int or_csr_from_triplets(int64_t rows, int64_t cols, int64_t ntrip,
                         const int64_t *ti, const int64_t *tj,
                         const double *tv, or_csr *out) {
  return triplets_to_csr(rows, cols, ntrip, NULL, NULL, ti, tj, tv, out);
}

static void gather_nodes(const double *xyz, const int64_t *c, double nodes[10][3]) {
  for (int a = 0; a < 10; ++a)
    for (int r = 0; r < 3; ++r) nodes[a][r] = xyz[3 * c[a] + r];
}

static int check_conn(int64_t n_nodes, int64_t n_elems, const int64_t *conn) {
  for (int64_t i = 0; i < 10 * n_elems; ++i)
    if (conn[i] < 0 || conn[i] >= n_nodes) return OR_ERR_INVALID_ARG;
  return OR_OK;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* src/fem/assembly.rs:69-117: parallel element kernels (Rayon map/collect ->
 * OpenMP loop into an element-ordered array), then a serial element-major
 * triplet merge and to_csr.  DOF = node id (:95-97). */
static int scatter_scalar(int64_t n_nodes, int64_t n_elems, const int64_t *conn,
                          const double *Ke, or_csr *out);

int or_assemble_thermal(int64_t n_nodes, const double *xyz, int64_t n_elems,
                        const int64_t *conn, double k, const double *k_elem,
                        int nthreads, or_csr *out) {
  memset(out, 0, sizeof(*out));
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  double *Ke = (double *)malloc((size_t)(n_elems > 0 ? n_elems : 1) * 100 * sizeof(double));
  if (!Ke) return OR_ERR_OOM;
  int err = OR_OK;
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3];
    gather_nodes(xyz, conn + 10 * e, nodes);
    const double ke = k_elem ? k_elem[e] : k;
    if (or_thermal_stiffness(nodes, ke, (double(*)[10])(Ke + 100 * e)) != OR_OK) {
#pragma omp atomic write
      err = OR_ERR_SINGULAR_JACOBIAN;
    }
  }
  if (err != OR_OK) { free(Ke); return err; }
  err = scatter_scalar(n_nodes, n_elems, conn, Ke, out);
  free(Ke);
  return err;
}

/* The serial triplet scatter of src/fem/assembly.rs:107-116 (DOF = node id)
 * for 10x10 element matrices Ke[e][10][10]. */
static int scatter_scalar(int64_t n_nodes, int64_t n_elems, const int64_t *conn,
                          const double *Ke, or_csr *out) {
  int err;
  const int64_t nt = 100 * n_elems;
  const int small = n_nodes < 2147483647LL;
  int32_t *ti32 = NULL, *tj32 = NULL;
  int64_t *ti64 = NULL, *tj64 = NULL;
  if (small) {
    ti32 = (int32_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int32_t));
    tj32 = (int32_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int32_t));
  } else {
    ti64 = (int64_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int64_t));
    tj64 = (int64_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int64_t));
  }
  if ((small && (!ti32 || !tj32)) || (!small && (!ti64 || !tj64))) {
    free(ti32); free(tj32); free(ti64); free(tj64);
    return OR_ERR_OOM;
  }
  int64_t t = 0;
  for (int64_t e = 0; e < n_elems; ++e) {
    const int64_t *c = conn + 10 * e;
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) {
        if (small) { ti32[t] = (int32_t)c[i]; tj32[t] = (int32_t)c[j]; }
        else { ti64[t] = c[i]; tj64[t] = c[j]; }
        ++t;
      }
  }
  err = triplets_to_csr(n_nodes, n_nodes, nt, ti32, tj32, ti64, tj64, Ke, out);
  free(ti32); free(tj32); free(ti64); free(tj64);
  return err;
}

/* Global thermal mass: element.rs:74-105 per element (parallel), scattered
 * as the stiffness is (no reference counterpart at the global level). */
int or_assemble_thermal_mass(int64_t n_nodes, const double *xyz, int64_t n_elems,
                             const int64_t *conn, double density, double specific_heat,
                             const double *rho_cp_elem, int nthreads, or_csr *out) {
  memset(out, 0, sizeof(*out));
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  double *Me = (double *)malloc((size_t)(n_elems > 0 ? n_elems : 1) * 100 * sizeof(double));
  if (!Me) return OR_ERR_OOM;
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3];
    gather_nodes(xyz, conn + 10 * e, nodes);
    if (rho_cp_elem)
      or_thermal_mass(nodes, rho_cp_elem[e], 1.0, (double(*)[10])(Me + 100 * e));
    else
      or_thermal_mass(nodes, density, specific_heat, (double(*)[10])(Me + 100 * e));
  }
  const int err = scatter_scalar(n_nodes, n_elems, conn, Me, out);
  free(Me);
  return err;
}

/* src/fem/assembly.rs:389-437: parallel 30x30 element matrices, then serial
 * scatter with global_dof(node,c) = 3*node + c (src/fem/dof.rs:87-106). */
/* Serial 900-triplet scatter of 30x30 element matrices with
 * global_dof(node,c) = 3*node + c (src/fem/dof.rs:87-106,
 * src/fem/assembly.rs:415-434), then the CSR conversion. Frees Ke. */
static int scatter_dof3(int64_t n_nodes, int64_t n_elems, const int64_t *conn,
                        double *Ke, or_csr *out) {
  const int64_t ndof = 3 * n_nodes;
  const int64_t nt = 900 * n_elems;
  const int small = ndof < 2147483647LL;
  int32_t *ti32 = NULL, *tj32 = NULL;
  int64_t *ti64 = NULL, *tj64 = NULL;
  double *tv = (double *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(double));
  if (small) {
    ti32 = (int32_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int32_t));
    tj32 = (int32_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int32_t));
  } else {
    ti64 = (int64_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int64_t));
    tj64 = (int64_t *)malloc((size_t)(nt > 0 ? nt : 1) * sizeof(int64_t));
  }
  if (!tv || (small && (!ti32 || !tj32)) || (!small && (!ti64 || !tj64))) {
    free(Ke); free(tv); free(ti32); free(tj32); free(ti64); free(tj64);
    return OR_ERR_OOM;
  }
  int64_t t = 0;
  for (int64_t e = 0; e < n_elems; ++e) {
    const int64_t *c = conn + 10 * e;
    const double *K = Ke + 900 * e;
    for (int i = 0; i < 10; ++i)
      for (int di = 0; di < 3; ++di) {
        const int64_t gi = 3 * c[i] + di;
        for (int j = 0; j < 10; ++j)
          for (int dj = 0; dj < 3; ++dj) {
            const int64_t gj = 3 * c[j] + dj;
            if (small) { ti32[t] = (int32_t)gi; tj32[t] = (int32_t)gj; }
            else { ti64[t] = gi; tj64[t] = gj; }
            tv[t] = K[(3 * i + di) * 30 + 3 * j + dj];
            ++t;
          }
      }
  }
  free(Ke);
  const int err = triplets_to_csr(ndof, ndof, nt, ti32, tj32, ti64, tj64, tv, out);
  free(tv); free(ti32); free(tj32); free(ti64); free(tj64);
  return err;
}

/* The element matrices of one isotropic Voigt law (kind 0: elastic D(E, nu),
 * E per element when e0; kind 1: viscous D(mu, zeta), mu / zeta per element
 * when e0 / e1), in parallel as Rayon's par_iter; optional stress-history
 * loads (kind 0 only) into Fe. */
static int element_matrices_dof3(const double *xyz, int64_t n_elems, const int64_t *conn,
                                 int kind, double p0, double p1, const double *e0,
                                 const double *e1, const double *sigma, double *Ke,
                                 double *Fe, int nthreads) {
  double D[6][6];
  for (int64_t e = 0; e < (e0 || e1 ? n_elems : 1); ++e) {
    const double a = e0 ? e0[e] : p0, b = e1 ? e1[e] : p1;
    const int st = kind == 0 ? or_constitutive(a, b, D) : or_viscous_constitutive(a, b, D);
    if (st != OR_OK) return st;
  }
  int err = OR_OK;
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3], De[6][6];
    gather_nodes(xyz, conn + 10 * e, nodes);
    const double a = e0 ? e0[e] : p0, b = e1 ? e1[e] : p1;
    int st;
    if (sigma) {
      st = or_maxwell_element((const double(*)[3])nodes, a, b, sigma + 6 * e,
                              (double(*)[30])(Ke + 900 * e), Fe + 30 * e);
    } else {
      if (kind == 0) or_constitutive(a, b, De);
      else or_viscous_constitutive(a, b, De);
      st = elastic_stiffness_D(nodes, (const double(*)[6])De, (double(*)[30])(Ke + 900 * e));
    }
    if (st != OR_OK) {
#pragma omp atomic write
      err = OR_ERR_SINGULAR_JACOBIAN;
    }
  }
  return err;
}

/* src/fem/assembly.rs:389-437: parallel 30x30 element matrices, then serial
 * scatter with global_dof(node,c) = 3*node + c (src/fem/dof.rs:87-106). */
int or_assemble_elastic(int64_t n_nodes, const double *xyz, int64_t n_elems,
                        const int64_t *conn, double E, double nu,
                        const double *E_elem, int nthreads, or_csr *out) {
  memset(out, 0, sizeof(*out));
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  double *Ke = (double *)malloc((size_t)(n_elems > 0 ? n_elems : 1) * 900 * sizeof(double));
  if (!Ke) return OR_ERR_OOM;
  const int err = element_matrices_dof3(xyz, n_elems, conn, 0, E, nu, E_elem, NULL, NULL, Ke,
                                        NULL, nthreads);
  if (err != OR_OK) { free(Ke); return err; }
  return scatter_dof3(n_nodes, n_elems, conn, Ke, out);
}

/* Assembler::assemble_viscosity_stiffness_parallel (src/fem/assembly.rs:646-694):
 * the same parallel element loop and serial scatter with the viscous D;
 * mu_elem / zeta_elem (nullable) replace mu / zeta per element. */
int or_assemble_viscosity(int64_t n_nodes, const double *xyz, int64_t n_elems,
                          const int64_t *conn, double mu, double zeta,
                          const double *mu_elem, const double *zeta_elem, int nthreads,
                          or_csr *out) {
  memset(out, 0, sizeof(*out));
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  double *Ke = (double *)malloc((size_t)(n_elems > 0 ? n_elems : 1) * 900 * sizeof(double));
  if (!Ke) return OR_ERR_OOM;
  const int err = element_matrices_dof3(xyz, n_elems, conn, 1, mu, zeta, mu_elem, zeta_elem,
                                        NULL, Ke, NULL, nthreads);
  if (err != OR_OK) { free(Ke); return err; }
  return scatter_dof3(n_nodes, n_elems, conn, Ke, out);
}

/* Assembler::assemble_maxwell_viscoelastic_parallel (src/fem/assembly.rs:959-1020):
 * element (K_e, f_e) in parallel (MaxwellViscoelasticity::new's checks,
 * src/mechanics/constitutive.rs:312-325), then K scattered and f_history
 * summed serially in element order.  dt is unused, as in the reference. */
int or_assemble_maxwell(int64_t n_nodes, const double *xyz, int64_t n_elems,
                        const int64_t *conn, double E, double nu, double visc, double dt,
                        const double *stress, int nthreads, or_csr *out, double *f_hist) {
  (void)dt;
  memset(out, 0, sizeof(*out));
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  if (!(visc > 0.0)) return OR_ERR_INVALID_MATERIAL;
  double *Ke = (double *)malloc((size_t)(n_elems > 0 ? n_elems : 1) * 900 * sizeof(double));
  double *Fe = (double *)malloc((size_t)(n_elems > 0 ? n_elems : 1) * 30 * sizeof(double));
  if (!Ke || !Fe) { free(Ke); free(Fe); return OR_ERR_OOM; }
  const int err = element_matrices_dof3(xyz, n_elems, conn, 0, E, nu, NULL, NULL, stress, Ke,
                                        Fe, nthreads);
  if (err != OR_OK) { free(Ke); free(Fe); return err; }
  for (int64_t i = 0; i < 3 * n_nodes; ++i) f_hist[i] = 0.0;
  for (int64_t e = 0; e < n_elems; ++e)
    for (int i = 0; i < 10; ++i)
      for (int c = 0; c < 3; ++c) f_hist[3 * conn[10 * e + i] + c] += Fe[30 * e + 3 * i + c];
  free(Fe);
  return scatter_dof3(n_nodes, n_elems, conn, Ke, out);
}

/* update_stresses_maxwell (src/mechanics/stress_update.rs:39-147) over all
 * elements; u_n, u_next indexed by DOF 3 node + c. */
int or_update_stresses_maxwell(int64_t n_nodes, const double *xyz, int64_t n_elems,
                               const int64_t *conn, double E, double nu, double visc,
                               double dt, const double *u_n, const double *u_next,
                               const double *sig_in, double *sig_out) {
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  int err = OR_OK;
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3], du[30];
    const int64_t *c = conn + 10 * e;
    gather_nodes(xyz, c, nodes);
    for (int i = 0; i < 10; ++i)
      for (int k = 0; k < 3; ++k) du[3 * i + k] = u_next[3 * c[i] + k] - u_n[3 * c[i] + k];
    if (or_maxwell_stress((const double(*)[3])nodes, E, nu, visc, dt, du, sig_in + 6 * e,
                          sig_out + 6 * e) != OR_OK)
      err = OR_ERR_SINGULAR_JACOBIAN;
  }
  return err;
}

/* src/fem/assembly.rs:130-154 (serial, element order). */
int or_assemble_thermal_load(int64_t n_nodes, const double *xyz,
                             int64_t n_elems, const int64_t *conn, double Q,
                             double *f) {
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  for (int64_t i = 0; i < n_nodes; ++i) f[i] = 0.0;
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3], fe[10];
    gather_nodes(xyz, conn + 10 * e, nodes);
    or_thermal_load(nodes, Q, fe);
    for (int i = 0; i < 10; ++i) f[conn[10 * e + i]] += fe[i];
  }
  return OR_OK;
}

/* src/fem/assembly.rs:906-937 (serial, element order). */
int or_assemble_gravity_load(int64_t n_nodes, const double *xyz,
                             int64_t n_elems, const int64_t *conn, double rho,
                             const double g[3], double *f) {
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  for (int64_t i = 0; i < 3 * n_nodes; ++i) f[i] = 0.0;
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3], fe[30];
    gather_nodes(xyz, conn + 10 * e, nodes);
    or_gravity_load(nodes, rho, g, fe);
    for (int i = 0; i < 10; ++i)
      for (int c = 0; c < 3; ++c) f[3 * conn[10 * e + i] + c] += fe[3 * i + c];
  }
  return OR_OK;
}

/* src/fem/assembly.rs:168-230: scaled Dirichlet elimination.
 * c = max(1, max_i |K_ii|); free rows: f_j -= K_ji g_i over constrained
 * columns in column order; constrained row -> c e_i, f_i = g_i c; free rows
 * keep only free columns. */
int or_apply_dirichlet(const or_csr *K, const double *f, const uint8_t *is_dir,
                       const double *values, or_csr *Kbc, double *fbc) {
  memset(Kbc, 0, sizeof(*Kbc));
  const int64_t n = K->rows;
  double max_diag = 1.0;
  for (int64_t r = 0; r < n; ++r)
    for (int64_t k = K->indptr[r]; k < K->indptr[r + 1]; ++k)
      if (K->indices[k] == r) {
        const double a = fabs(K->data[k]);
        if (a > max_diag) max_diag = a;
      }
  const double c = max_diag;
  for (int64_t r = 0; r < n; ++r) fbc[r] = f[r];
  for (int64_t r = 0; r < n; ++r) {
    if (is_dir[r]) continue;
    for (int64_t k = K->indptr[r]; k < K->indptr[r + 1]; ++k) {
      const int64_t j = K->indices[k];
      if (is_dir[j]) fbc[r] -= K->data[k] * values[j];
    }
  }
  /* second pass: rows are already sorted/unique, so the rebuilt TriMat's
   * to_csr is a filtered copy. */
  int64_t nnz = 0;
  for (int64_t r = 0; r < n; ++r) {
    if (is_dir[r]) { nnz += 1; continue; }
    for (int64_t k = K->indptr[r]; k < K->indptr[r + 1]; ++k)
      if (!is_dir[K->indices[k]]) ++nnz;
  }
  Kbc->rows = n;
  Kbc->cols = K->cols;
  Kbc->nnz = nnz;
  Kbc->indptr = (int64_t *)malloc((size_t)(n + 1) * sizeof(int64_t));
  Kbc->indices = (int64_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
  Kbc->data = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  if (!Kbc->indptr || !Kbc->indices || !Kbc->data) { or_csr_free(Kbc); return OR_ERR_OOM; }
  int64_t p = 0;
  Kbc->indptr[0] = 0;
  for (int64_t r = 0; r < n; ++r) {
    if (is_dir[r]) {
      Kbc->indices[p] = r;
      Kbc->data[p] = c;
      ++p;
      fbc[r] = values[r] * c;
    } else {
      for (int64_t k = K->indptr[r]; k < K->indptr[r + 1]; ++k) {
        const int64_t j = K->indices[k];
        if (!is_dir[j]) {
          Kbc->indices[p] = j;
          Kbc->data[p] = K->data[k];
          ++p;
        }
      }
    }
    Kbc->indptr[r + 1] = p;
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Neumann faces: src/fem/boundary.rs, src/fem/assembly.rs:248-273          */
/* ------------------------------------------------------------------------ */

/* Tet10Face::from_element_face (src/fem/boundary.rs:72-115): element-local
 * node of face node k (vertices, then edges 01, 12, 20 of the triangle). */
static const int FACE_NODES[4][6] = {{1, 2, 3, 5, 9, 8},
                                     {0, 2, 3, 6, 9, 7},
                                     {0, 1, 3, 4, 8, 7},
                                     {0, 1, 2, 4, 5, 6}};

int or_face_nodes(int local_face, int out[6]) {
  if (local_face < 0 || local_face > 3) return OR_ERR_INVALID_ARG;
  for (int k = 0; k < 6; ++k) out[k] = FACE_NODES[local_face][k];
  return OR_OK;
}

/* dx/dL1 and dx/dL2 of the quadratic triangle at (l1, l2)
 * (src/fem/boundary.rs:214-225 derivative table) */
static void face_tangents(const double c[6][3], double l1, double l2, double t1[3],
                          double t2[3]) {
  const double l3 = 1.0 - l1 - l2;
  const double d[6][2] = {{4.0 * l1 - 1.0, 0.0},
                          {0.0, 4.0 * l2 - 1.0},
                          {-4.0 * l3 + 1.0, -4.0 * l3 + 1.0},
                          {4.0 * l2, 4.0 * l1},
                          {-4.0 * l2, 4.0 * (l3 - l2)},
                          {4.0 * (l3 - l1), -4.0 * l1}};
  for (int r = 0; r < 3; ++r) t1[r] = t2[r] = 0.0;
  for (int i = 0; i < 6; ++i)
    for (int r = 0; r < 3; ++r) {
      t1[r] += c[i][r] * d[i][0];
      t2[r] += c[i][r] * d[i][1];
    }
}

static double cross_norm(const double a[3], const double b[3]) {
  const double x = a[1] * b[2] - a[2] * b[1];
  const double y = a[2] * b[0] - a[0] * b[2];
  const double z = a[0] * b[1] - a[1] * b[0];
  return sqrt(x * x + y * y + z * z);
}

static const double FACE_QP[3][3] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                                     {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                                     {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};

/* Tet10Face::area (src/fem/boundary.rs:118-182): 3-point rule of |dx/dL1 x dx/dL2|. */
double or_face_area(const double c[6][3]) {
  double area = 0.0;
  for (int q = 0; q < 3; ++q) {
    double t1[3], t2[3];
    face_tangents(c, FACE_QP[q][0], FACE_QP[q][1], t1, t2);
    area += FACE_QP[q][2] * cross_norm(t1, t2);
  }
  return area;
}

/* Tet10Face::integrate_flux (src/fem/boundary.rs:187-253):
 * f_i += w N_i flux |dx/dL1 x dx/dL2| over the 3-point rule. */
void or_face_flux(const double c[6][3], double flux, double out[6]) {
  for (int i = 0; i < 6; ++i) out[i] = 0.0;
  for (int q = 0; q < 3; ++q) {
    const double l1 = FACE_QP[q][0], l2 = FACE_QP[q][1], w = FACE_QP[q][2];
    const double l3 = 1.0 - l1 - l2;
    const double n[6] = {l1 * (2.0 * l1 - 1.0), l2 * (2.0 * l2 - 1.0), l3 * (2.0 * l3 - 1.0),
                         4.0 * l1 * l2,         4.0 * l2 * l3,         4.0 * l3 * l1};
    double t1[3], t2[3];
    face_tangents(c, l1, l2, t1, t2);
    const double det = cross_norm(t1, t2);
    for (int i = 0; i < 6; ++i) out[i] += w * n[i] * flux * det;
  }
}

/* Assembler::apply_neumann_bcs (src/fem/assembly.rs:248-273): f[node] +=
 * flux contribution, face by face in the given order (the reference walks a
 * hash map, whose order is not fixed). */
int or_apply_neumann(int64_t n_nodes, const double *xyz, int64_t n_elems,
                     const int64_t *conn, int64_t n_faces, const int64_t *elem_ids,
                     const int32_t *local_faces, const double *fluxes, double *f) {
  for (int64_t k = 0; k < n_faces; ++k) {
    if (elem_ids[k] < 0 || elem_ids[k] >= n_elems) return OR_ERR_INVALID_ARG;
    if (local_faces[k] < 0 || local_faces[k] > 3) return OR_ERR_INVALID_ARG;
    const int64_t *c = conn + 10 * elem_ids[k];
    double coords[6][3], contrib[6];
    int64_t nodes[6];
    for (int i = 0; i < 6; ++i) {
      nodes[i] = c[FACE_NODES[local_faces[k]][i]];
      if (nodes[i] < 0 || nodes[i] >= n_nodes) return OR_ERR_INVALID_ARG;
      for (int r = 0; r < 3; ++r) coords[i][r] = xyz[3 * nodes[i] + r];
    }
    or_face_flux(coords, fluxes[k], contrib);
    for (int i = 0; i < 6; ++i) f[nodes[i]] += contrib[i];
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Linear algebra                                                            */
/* ------------------------------------------------------------------------ */

/* src/linalg/solver.rs:62-73: row-serial sequential sum. */
void or_spmv(const or_csr *A, const double *x, double *y) {
  for (int64_t r = 0; r < A->rows; ++r) {
    double s = 0.0;
    for (int64_t k = A->indptr[r]; k < A->indptr[r + 1]; ++k)
      s += A->data[k] * x[A->indices[k]];
    y[r] = s;
  }
}

/* src/linalg/preconditioner.rs:37-51: 1/A_ii if stored and |A_ii| > 1e-14,
 * else 1. */
void or_jacobi_diag_inv(const or_csr *A, double *dinv) {
  for (int64_t r = 0; r < A->rows; ++r) {
    dinv[r] = 1.0;
    for (int64_t k = A->indptr[r]; k < A->indptr[r + 1]; ++k)
      if (A->indices[k] == r) {
        if (fabs(A->data[k]) > 1e-14) dinv[r] = 1.0 / A->data[k];
        break;
      }
  }
}

static double norm2(const double *v, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += v[i] * v[i];
  return sqrt(s);
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* src/linalg/iterative.rs:46-119 (+ :123-131 for the preconditioner choice),
 * statement by statement.  The iteration counter excludes the converging
 * pass; the recursive residual is never recomputed. */
/* Timing column only, not part of the restatement: the loop of or_cg_solve
 * with its SpMV rows and vector updates split over nthreads OpenMP threads
 * and its dot products as OpenMP reductions (their summation order depends
 * on the thread count, so iterates agree with or_cg_solve to rounding, not
 * bit for bit).  It gives the CPU side a parallel CG beside the reference's
 * serial one. */
int or_cg_solve_omp(const or_csr *A, const double *b, int64_t max_iter, double tol,
                    double abs_tol, int nthreads, double *x, or_stats *st) {
  const int64_t n = A->rows;
  const double t0 = now_s();
  set_threads(nthreads);
  double bb = 0.0;
#pragma omp parallel for reduction(+ : bb) schedule(static)
  for (int64_t i = 0; i < n; ++i) bb += b[i] * b[i];
  const double b_norm = sqrt(bb);
  memset(st, 0, sizeof(*st));
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  if (b_norm < 1e-25) {
    st->converged = 1;
    st->solve_time = now_s() - t0;
    return OR_OK;
  }
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *z = (double *)malloc((size_t)n * sizeof(double));
  double *p = (double *)malloc((size_t)n * sizeof(double));
  double *ap = (double *)malloc((size_t)n * sizeof(double));
  double *dinv = (double *)malloc((size_t)n * sizeof(double));
  if (!r || !z || !p || !ap || !dinv) {
    free(r); free(z); free(p); free(ap); free(dinv);
    return OR_ERR_OOM;
  }
  or_jacobi_diag_inv(A, dinv);
  double rz = 0.0;
#pragma omp parallel for reduction(+ : rz) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i];
    z[i] = r[i] * dinv[i];
    p[i] = z[i];
    rz += r[i] * z[i];
  }
  int64_t it = 0;
  int converged = 0;
  double final_res = b_norm;
  while (it < max_iter) {
    double pap = 0.0;
#pragma omp parallel for reduction(+ : pap) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t k = A->indptr[i]; k < A->indptr[i + 1]; ++k) s += A->data[k] * p[A->indices[k]];
      ap[i] = s;
      pap += p[i] * s;
    }
    if (fabs(pap) < 1e-30) break;
    const double alpha = rz / pap;
    double rr = 0.0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
      rr += r[i] * r[i];
    }
    const double rn = sqrt(rr);
    final_res = rn;
    if (rn < tol * b_norm || rn < abs_tol) { converged = 1; break; }
    double rz_new = 0.0;
#pragma omp parallel for reduction(+ : rz_new) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      z[i] = r[i] * dinv[i];
      rz_new += r[i] * z[i];
    }
    const double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  st->iterations = it;
  st->residual_norm = final_res;
  st->relative_residual = final_res / b_norm;
  st->converged = converged;
  st->solve_time = now_s() - t0;
  free(r); free(z); free(p); free(ap); free(dinv);
  return OR_OK;
}

int or_cg_solve(const or_csr *A, const double *b, int64_t max_iter, double tol,
                double abs_tol, int use_precond, double *x, or_stats *st) {
  const int64_t n = A->rows;
  const double t0 = now_s();
  const double b_norm = norm2(b, n);
  memset(st, 0, sizeof(*st));
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  if (b_norm < 1e-25) {
    st->converged = 1;
    st->solve_time = now_s() - t0;
    return OR_OK;
  }
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *z = (double *)malloc((size_t)n * sizeof(double));
  double *p = (double *)malloc((size_t)n * sizeof(double));
  double *ap = (double *)malloc((size_t)n * sizeof(double));
  double *dinv = (double *)malloc((size_t)n * sizeof(double));
  if (!r || !z || !p || !ap || !dinv) {
    free(r); free(z); free(p); free(ap); free(dinv);
    return OR_ERR_OOM;
  }
  if (use_precond) or_jacobi_diag_inv(A, dinv);
  else for (int64_t i = 0; i < n; ++i) dinv[i] = 1.0;
  for (int64_t i = 0; i < n; ++i) r[i] = b[i];
  for (int64_t i = 0; i < n; ++i) z[i] = use_precond ? r[i] * dinv[i] : r[i];
  for (int64_t i = 0; i < n; ++i) p[i] = z[i];
  double rz = 0.0;
  for (int64_t i = 0; i < n; ++i) rz += r[i] * z[i];
  int64_t it = 0;
  int converged = 0;
  double final_res = b_norm;
  while (it < max_iter) {
    or_spmv(A, p, ap);
    double pap = 0.0;
    for (int64_t i = 0; i < n; ++i) pap += p[i] * ap[i];
    if (fabs(pap) < 1e-30) break;
    const double alpha = rz / pap;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * ap[i];
    }
    const double rn = norm2(r, n);
    final_res = rn;
    if (rn < tol * b_norm || rn < abs_tol) { converged = 1; break; }
    for (int64_t i = 0; i < n; ++i) z[i] = use_precond ? r[i] * dinv[i] : r[i];
    double rz_new = 0.0;
    for (int64_t i = 0; i < n; ++i) rz_new += r[i] * z[i];
    const double beta = rz_new / rz;
    rz = rz_new;
    for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  st->iterations = it;
  st->residual_norm = final_res;
  st->relative_residual = b_norm > 1e-20 ? final_res / b_norm : 0.0;
  st->converged = converged;
  st->solve_time = now_s() - t0;
  free(r); free(z); free(p); free(ap); free(dinv);
  return OR_OK;
}

/* src/linalg/iterative.rs:177-280 (BiCGSTAB::solve_with_operator_internal,
 * + :283-291 for the preconditioner choice), statement by statement: the
 * shadow residual is b, the iteration counter excludes a pass that ends on
 * the intermediate residual s, and the stopping tests are absolute
 * thresholds (1e-40) on rho, r_hat.v, t.t and omega as there. */
int or_bicgstab_solve(const or_csr *A, const double *b, int64_t max_iter,
                      double tol, double abs_tol, int use_precond, double *x,
                      or_stats *st) {
  const int64_t n = A->rows;
  const double t0 = now_s();
  const double b_norm = norm2(b, n);
  memset(st, 0, sizeof(*st));
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  if (b_norm < 1e-25) {
    st->converged = 1;
    st->solve_time = now_s() - t0;
    return OR_OK;
  }
  const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(double);
  double *r = (double *)malloc(bytes), *rh = (double *)malloc(bytes);
  double *v = (double *)malloc(bytes), *p = (double *)malloc(bytes);
  double *ph = (double *)malloc(bytes), *s = (double *)malloc(bytes);
  double *sh = (double *)malloc(bytes), *t = (double *)malloc(bytes);
  double *dinv = (double *)malloc(bytes);
  if (!r || !rh || !v || !p || !ph || !s || !sh || !t || !dinv) {
    free(r); free(rh); free(v); free(p); free(ph); free(s); free(sh); free(t);
    free(dinv);
    return OR_ERR_OOM;
  }
  if (use_precond) or_jacobi_diag_inv(A, dinv);
  for (int64_t i = 0; i < n; ++i) {
    r[i] = b[i];
    rh[i] = b[i];
    v[i] = 0.0;
    p[i] = 0.0;
  }
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  int64_t it = 0;
  int converged = 0;
  double final_res = b_norm;
  while (it < max_iter) {
    const double rho_prev = rho;
    rho = 0.0;
    for (int64_t i = 0; i < n; ++i) rho += rh[i] * r[i];
    if (fabs(rho) < 1e-40) break;
    if (it == 0) {
      for (int64_t i = 0; i < n; ++i) p[i] = r[i];
    } else {
      const double beta = (rho / rho_prev) * (alpha / omega);
      for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
    }
    for (int64_t i = 0; i < n; ++i) ph[i] = use_precond ? p[i] * dinv[i] : p[i];
    or_spmv(A, ph, v);
    double rhat_v = 0.0;
    for (int64_t i = 0; i < n; ++i) rhat_v += rh[i] * v[i];
    if (fabs(rhat_v) < 1e-40) break;
    alpha = rho / rhat_v;
    for (int64_t i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    const double s_norm = norm2(s, n);
    if (s_norm < tol * b_norm || s_norm < abs_tol) {
      for (int64_t i = 0; i < n; ++i) x[i] += alpha * ph[i];
      final_res = s_norm;
      converged = 1;
      break;
    }
    for (int64_t i = 0; i < n; ++i) sh[i] = use_precond ? s[i] * dinv[i] : s[i];
    or_spmv(A, sh, t);
    double t_t = 0.0, t_s = 0.0;
    for (int64_t i = 0; i < n; ++i) t_t += t[i] * t[i];
    for (int64_t i = 0; i < n; ++i) t_s += t[i] * s[i];
    if (fabs(t_t) < 1e-40) break;
    omega = t_s / t_t;
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * ph[i] + omega * sh[i];
      r[i] = s[i] - omega * t[i];
    }
    final_res = norm2(r, n);
    ++it;
    if (final_res < tol * b_norm || final_res < abs_tol) {
      converged = 1;
      break;
    }
    if (fabs(omega) < 1e-40) break;
  }
  st->iterations = it;
  st->residual_norm = final_res;
  st->relative_residual = final_res / b_norm;
  st->converged = converged;
  st->solve_time = now_s() - t0;
  free(r); free(rh); free(v); free(p); free(ph); free(s); free(sh); free(t);
  free(dinv);
  return OR_OK;
}

/* src/linalg/iterative.rs:463-473: a Givens rotation zeroing b against a. */
static void givens(double a, double b, double *c, double *s, double *rho) {
  if (fabs(b) < 1e-40) {
    *c = 1.0; *s = 0.0; *rho = a;
  } else if (fabs(b) > fabs(a)) {
    const double tau = a / b;
    *s = 1.0 / sqrt(1.0 + tau * tau);
    *c = *s * tau;
    *rho = b * sqrt(1.0 + tau * tau);
  } else {
    const double tau = b / a;
    *c = 1.0 / sqrt(1.0 + tau * tau);
    *s = *c * tau;
    *rho = a * sqrt(1.0 + tau * tau);
  }
}

/* src/linalg/iterative.rs:333-460 (GMRES::solve_with_operator_internal,
 * + :476-486 for the preconditioner choice), statement by statement:
 * restarted GMRES(m) with right preconditioning, modified Gram-Schmidt,
 * Givens rotations; the iteration counter counts inner steps over all
 * cycles; each cycle starts from the true residual b - A x. */
int or_gmres_solve(const or_csr *A, const double *b, int64_t max_iter,
                   int64_t restart, double tol, double abs_tol, int use_precond,
                   double *x, or_stats *st) {
  const int64_t n = A->rows, m = restart;
  const double t0 = now_s();
  const double b_norm = norm2(b, n);
  memset(st, 0, sizeof(*st));
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  if (b_norm < 1e-25) {
    st->converged = 1;
    st->solve_time = now_s() - t0;
    return OR_OK;
  }
  if (m < 1) return OR_ERR_INVALID_ARG; /* restart 0 never advances (the reference loops) */
  const size_t nb = (size_t)(n > 0 ? n : 1) * sizeof(double);
  double *r = (double *)malloc(nb), *w = (double *)malloc(nb), *z = (double *)malloc(nb);
  double *dy = (double *)malloc(nb), *dinv = (double *)malloc(nb);
  double *V = (double *)malloc(nb * (size_t)(m + 1));
  double *H = (double *)malloc(sizeof(double) * (size_t)((m + 1) * m));
  double *g = (double *)malloc(sizeof(double) * (size_t)(m + 1));
  double *cs = (double *)malloc(sizeof(double) * (size_t)m);
  double *sn = (double *)malloc(sizeof(double) * (size_t)m);
  double *y = (double *)malloc(sizeof(double) * (size_t)m);
  if (!r || !w || !z || !dy || !dinv || !V || !H || !g || !cs || !sn || !y) {
    free(r); free(w); free(z); free(dy); free(dinv); free(V); free(H); free(g);
    free(cs); free(sn); free(y);
    return OR_ERR_OOM;
  }
#define HH(i, j) H[(i) * m + (j)]
  if (use_precond) or_jacobi_diag_inv(A, dinv);
  int64_t total = 0;
  int converged = 0;
  double final_res = b_norm;
  while (total < max_iter) {
    or_spmv(A, x, r); /* compute_residual: r = b - A x */
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
    const double r_norm = norm2(r, n);
    final_res = r_norm;
    if (r_norm < tol * b_norm || r_norm < abs_tol) { converged = 1; break; }
    memset(V, 0, nb * (size_t)(m + 1));
    memset(H, 0, sizeof(double) * (size_t)((m + 1) * m));
    for (int64_t i = 0; i < n; ++i) V[i] = r[i] / r_norm;
    for (int64_t i = 0; i <= m; ++i) g[i] = 0.0;
    g[0] = r_norm;
    for (int64_t i = 0; i < m; ++i) { cs[i] = 0.0; sn[i] = 0.0; }
    int64_t k = 0;
    for (int64_t j = 0; j < m; ++j) {
      if (total >= max_iter) break;
      const double *vj = V + (size_t)j * n;
      for (int64_t i = 0; i < n; ++i) z[i] = use_precond ? vj[i] * dinv[i] : vj[i];
      or_spmv(A, z, w);
      for (int64_t i = 0; i <= j; ++i) {
        const double *vi = V + (size_t)i * n;
        double h = 0.0;
        for (int64_t l = 0; l < n; ++l) h += vi[l] * w[l];
        HH(i, j) = h;
        for (int64_t l = 0; l < n; ++l) w[l] -= h * vi[l];
      }
      HH(j + 1, j) = norm2(w, n);
      if (fabs(HH(j + 1, j)) > 1e-40) {
        double *vn = V + (size_t)(j + 1) * n;
        for (int64_t l = 0; l < n; ++l) vn[l] = w[l] / HH(j + 1, j);
      }
      for (int64_t i = 0; i < j; ++i) {
        const double temp = cs[i] * HH(i, j) + sn[i] * HH(i + 1, j);
        HH(i + 1, j) = -sn[i] * HH(i, j) + cs[i] * HH(i + 1, j);
        HH(i, j) = temp;
      }
      double c, sv, rho;
      givens(HH(j, j), HH(j + 1, j), &c, &sv, &rho);
      cs[j] = c; sn[j] = sv; HH(j, j) = rho; HH(j + 1, j) = 0.0;
      const double temp = cs[j] * g[j];
      g[j + 1] = -sn[j] * g[j];
      g[j] = temp;
      k = j + 1;
      ++total;
      final_res = fabs(g[k]);
      if (final_res < tol * b_norm || final_res < abs_tol) break;
    }
    for (int64_t i = 0; i < k; ++i) y[i] = 0.0;
    for (int64_t i = k - 1; i >= 0; --i) {
      double sum = 0.0;
      for (int64_t j = i + 1; j < k; ++j) sum += HH(i, j) * y[j];
      if (fabs(HH(i, i)) < 1e-40) break;
      y[i] = (g[i] - sum) / HH(i, i);
    }
    for (int64_t i = 0; i < n; ++i) dy[i] = 0.0;
    for (int64_t j = 0; j < k; ++j) {
      const double *vj = V + (size_t)j * n;
      for (int64_t i = 0; i < n; ++i) dy[i] += vj[i] * y[j];
    }
    for (int64_t i = 0; i < n; ++i) x[i] += use_precond ? dy[i] * dinv[i] : dy[i];
    if (final_res < tol * b_norm || final_res < abs_tol) { converged = 1; break; }
  }
#undef HH
  st->iterations = total;
  st->residual_norm = final_res;
  st->relative_residual = final_res / b_norm;
  st->converged = converged;
  st->solve_time = now_s() - t0;
  free(r); free(w); free(z); free(dy); free(dinv); free(V); free(H); free(g);
  free(cs); free(sn); free(y);
  return OR_OK;
}

/* ILUPreconditioner::new (src/linalg/preconditioner.rs:86-108) and
 * factorize_in_place (:126-166). */
int or_ilu0_factor(const or_csr *A, double *factors, int64_t *diag, int64_t *bad,
                   int32_t *kind) {
  const int64_t n = A->rows;
  const int64_t *indptr = A->indptr, *indices = A->indices;
  double *data = factors;
  for (int64_t p = 0; p < A->nnz; ++p) data[p] = A->data[p];
  /* 1. the diagonal position of every row (:93-105) */
  for (int64_t i = 0; i < n; ++i) {
    int found = 0;
    for (int64_t idx = indptr[i]; idx < indptr[i + 1]; ++idx)
      if (indices[idx] == i) {
        diag[i] = idx;
        found = 1;
        break;
      }
    if (!found) {
      *bad = i;
      *kind = 0;
      return OR_ERR_ZERO_PIVOT;
    }
  }
  /* 2. elimination by the earlier rows k < i, merge-style update of row i
   * over row k's upper part (:131-160) */
  for (int64_t i = 0; i < n; ++i) {
    const int64_t row_start = indptr[i], diag_i = diag[i];
    for (int64_t k_idx = row_start; k_idx < diag_i; ++k_idx) {
      const int64_t k = indices[k_idx];
      const int64_t k_diag = diag[k], k_end = indptr[k + 1];
      const double diag_val_k = data[k_diag];
      if (fabs(diag_val_k) < 1e-15) {
        *bad = k;
        *kind = 1;
        return OR_ERR_ZERO_PIVOT;
      }
      const double val_ik = data[k_idx] / diag_val_k;
      data[k_idx] = val_ik;
      int64_t cur = k_idx + 1;
      const int64_t row_i_end = indptr[i + 1];
      for (int64_t kj = k_diag + 1; kj < k_end; ++kj) {
        const int64_t col_j = indices[kj];
        while (cur < row_i_end && indices[cur] < col_j) ++cur;
        if (cur == row_i_end) break;
        if (indices[cur] == col_j) data[cur] -= val_ik * data[kj];
      }
    }
    if (fabs(data[diag_i]) < 1e-15) {
      *bad = i;
      *kind = 2;
      return OR_ERR_ZERO_PIVOT;
    }
  }
  return OR_OK;
}

/* ILUPreconditioner::apply (src/linalg/preconditioner.rs:168-203): forward
 * substitution with the unit lower factor, then backward with the upper. */
void or_ilu0_apply(const or_csr *A, const double *factors, const int64_t *diag,
                   const double *r, double *z) {
  const int64_t n = A->rows;
  const int64_t *indptr = A->indptr, *indices = A->indices;
  for (int64_t i = 0; i < n; ++i) z[i] = r[i];
  for (int64_t i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int64_t idx = indptr[i]; idx < diag[i]; ++idx) sum += factors[idx] * z[indices[idx]];
    z[i] -= sum;
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    double sum = 0.0;
    for (int64_t idx = diag[i] + 1; idx < indptr[i + 1]; ++idx)
      sum += factors[idx] * z[indices[idx]];
    z[i] = (z[i] - sum) / factors[diag[i]];
  }
}

/* ------------------------------------------------------------------------ */
/* Adjoint element backward (no reference counterpart; pinned by FD).        */
/* With theta_e = log k_e (or log E_e), dK_e/dtheta_e = K_e, so              */
/* dL/dtheta_e = -lambda_e^T K_e u_e with u_e including Dirichlet values.    */
/* ------------------------------------------------------------------------ */
int or_thermal_element_backward(int64_t n_nodes, const double *xyz,
                                int64_t n_elems, const int64_t *conn,
                                const double *k_elem, const double *u,
                                const double *lam, double *grad) {
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  int err = OR_OK;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3], K[10][10];
    const int64_t *c = conn + 10 * e;
    gather_nodes(xyz, c, nodes);
    if (or_thermal_stiffness(nodes, k_elem[e], K) != OR_OK) {
#pragma omp atomic write
      err = OR_ERR_SINGULAR_JACOBIAN;
      continue;
    }
    double s = 0.0;
    for (int a = 0; a < 10; ++a) {
      double ka = 0.0;
      for (int b = 0; b < 10; ++b) ka += K[a][b] * u[c[b]];
      s += lam[c[a]] * ka;
    }
    grad[e] = -s;
  }
  return err;
}

int or_elastic_element_backward(int64_t n_nodes, const double *xyz,
                                int64_t n_elems, const int64_t *conn,
                                const double *E_elem, double nu,
                                const double *u, const double *lam,
                                double *grad) {
  if (check_conn(n_nodes, n_elems, conn) != OR_OK) return OR_ERR_INVALID_ARG;
  int err = OR_OK;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elems; ++e) {
    double nodes[10][3], K[30][30];
    const int64_t *c = conn + 10 * e;
    gather_nodes(xyz, c, nodes);
    if (or_elastic_stiffness(nodes, E_elem[e], nu, K) != OR_OK) {
#pragma omp atomic write
      err = OR_ERR_SINGULAR_JACOBIAN;
      continue;
    }
    double s = 0.0;
    for (int a = 0; a < 30; ++a) {
      double ka = 0.0;
      for (int b = 0; b < 30; ++b) ka += K[a][b] * u[3 * c[b / 3] + b % 3];
      s += lam[3 * c[a / 3] + a % 3] * ka;
    }
    grad[e] = -s;
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* AMGPreconditioner (src/linalg/amg.rs).                                    */
/* ------------------------------------------------------------------------ */

static int csr_alloc(or_csr *m, int64_t rows, int64_t cols, int64_t cap) {
  memset(m, 0, sizeof(*m));
  m->rows = rows;
  m->cols = cols;
  m->indptr = (int64_t *)calloc((size_t)rows + 1, sizeof(int64_t));
  m->indices = (int64_t *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(int64_t));
  m->data = (double *)malloc((size_t)(cap > 0 ? cap : 1) * sizeof(double));
  if (!m->indptr || !m->indices || !m->data) {
    or_csr_free(m);
    return OR_ERR_OOM;
  }
  return OR_OK;
}

/* strength_matrix (src/linalg/amg.rs:359-390).  The reference scans every
 * column j of row i through a.get(i, j), which yields exactly the stored
 * entries of the row; the loops below walk those entries in ascending column
 * order.  f64::max ignores a NaN operand, as fmax does. */
int or_amg_strength(const or_csr *A, double theta, or_csr *S) {
  const int64_t n = A->rows;
  if (csr_alloc(S, n, A->cols, A->nnz) != OR_OK) return OR_ERR_OOM;
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    double max_off = 0.0;
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p)
      if (A->indices[p] != i) max_off = fmax(max_off, fabs(A->data[p]));
    const double threshold = theta * max_off;
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p)
      if (A->indices[p] != i && fabs(A->data[p]) >= threshold) {
        S->indices[nnz] = A->indices[p];
        S->data[nnz] = A->data[p];
        ++nnz;
      }
    S->indptr[i + 1] = nnz;
  }
  S->nnz = nnz;
  return OR_OK;
}

static int cmp_i64(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* max-heap of (lambda, node): larger lambda first, then the smaller node */
typedef struct { int64_t lam, node; } lamnode;
static int lam_before(lamnode a, lamnode b) {
  return a.lam > b.lam || (a.lam == b.lam && a.node < b.node);
}
static void heap_push(lamnode *h, int64_t *size, lamnode v) {
  int64_t i = (*size)++;
  h[i] = v;
  while (i > 0) {
    const int64_t up = (i - 1) / 2;
    if (!lam_before(h[i], h[up])) break;
    const lamnode t = h[i]; h[i] = h[up]; h[up] = t;
    i = up;
  }
}
static lamnode heap_pop(lamnode *h, int64_t *size) {
  const lamnode top = h[0];
  h[0] = h[--(*size)];
  int64_t i = 0;
  for (;;) {
    const int64_t l = 2 * i + 1, r = l + 1;
    int64_t m = i;
    if (l < *size && lam_before(h[l], h[m])) m = l;
    if (r < *size && lam_before(h[r], h[m])) m = r;
    if (m == i) break;
    const lamnode t = h[i]; h[i] = h[m]; h[m] = t;
    i = m;
  }
  return top;
}

/* coarsen (src/linalg/amg.rs:286-355).  The reference rescans all nodes for
 * the undecided node of largest lambda (the first such node in index order:
 * its comparison is strict) at every step; here a max-heap ordered by
 * (lambda descending, node ascending) holds every undecided node whose
 * lambda is positive, with stale entries (a node decided since, or whose
 * lambda has grown since) dropped when they surface.  lambda only grows, so
 * the valid top is exactly the node the reference's scan picks.  When no
 * undecided node has a positive lambda the rest become coarse (:312-321). */
int or_amg_coarsen(const or_csr *A, double theta, uint8_t *cf, int64_t *num_coarse) {
  or_csr S;
  int err = or_amg_strength(A, theta, &S);
  if (err != OR_OK) return err;
  const int64_t n = A->rows;
  int64_t *lambda = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  uint8_t *undecided = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
  int64_t cap = n + S.nnz * 8 + 16, size = 0;
  lamnode *heap = (lamnode *)malloc((size_t)cap * sizeof(lamnode));
  if (!lambda || !undecided || !heap) {
    free(lambda); free(undecided); free(heap); or_csr_free(&S);
    return OR_ERR_OOM;
  }
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = 0;
    for (int64_t p = S.indptr[i]; p < S.indptr[i + 1]; ++p) c += S.indices[p] != i;
    lambda[i] = c;
    undecided[i] = 1;
    cf[i] = 0;
    if (c > 0) heap_push(heap, &size, (lamnode){c, i});
  }
  int64_t nc = 0, left = n;
  while (left > 0) {
    int64_t max_node = -1;
    while (size > 0) {
      const lamnode t = heap_pop(heap, &size);
      if (undecided[t.node] && lambda[t.node] == t.lam) {
        max_node = t.node;
        break;
      }
    }
    if (max_node < 0) { /* max_lambda == 0: the rest are coarse */
      for (int64_t i = 0; i < n; ++i)
        if (undecided[i]) {
          cf[i] = 1;
          undecided[i] = 0;
          ++nc;
        }
      break;
    }
    cf[max_node] = 1;
    undecided[max_node] = 0;
    ++nc;
    --left;
    for (int64_t p = S.indptr[max_node]; p < S.indptr[max_node + 1]; ++p) {
      const int64_t j = S.indices[p];
      if (j != max_node && undecided[j]) {
        undecided[j] = 0; /* fine */
        --left;
        for (int64_t q = S.indptr[j]; q < S.indptr[j + 1]; ++q) {
          const int64_t k = S.indices[q];
          if (k != j && undecided[k]) {
            lambda[k] += 1;
            if (size == cap) {
              cap *= 2;
              lamnode *h2 = (lamnode *)realloc(heap, (size_t)cap * sizeof(lamnode));
              if (!h2) {
                free(lambda); free(undecided); free(heap); or_csr_free(&S);
                return OR_ERR_OOM;
              }
              heap = h2;
            }
            heap_push(heap, &size, (lamnode){lambda[k], k});
          }
        }
      }
    }
  }
  *num_coarse = nc;
  free(lambda);
  free(undecided);
  free(heap);
  or_csr_free(&S);
  return OR_OK;
}

/* build_prolongation (src/linalg/amg.rs:396-487): a coarse point keeps its
 * value (weight 1); a fine point takes its strong coarse neighbours (S row
 * order) with weights |s_ij| / sum, or when that sum is at most 1e-14 all
 * coarse entries of its A row the same way, or when that too is at most
 * 1e-14 the first coarse point with weight 1 (if there is one).  Triplets
 * arrive row by row in ascending column, so to_csr keeps them as they are
 * (explicit zeros included). */
int or_amg_prolongation(const or_csr *A, const uint8_t *cf, int64_t num_coarse,
                        double theta, or_csr *P) {
  or_csr S;
  int err = or_amg_strength(A, theta, &S);
  if (err != OR_OK) return err;
  const int64_t n = A->rows;
  int64_t *cidx = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  if (!cidx || csr_alloc(P, n, num_coarse, A->nnz + n) != OR_OK) {
    free(cidx); or_csr_free(&S);
    return OR_ERR_OOM;
  }
  int64_t c = 0;
  for (int64_t i = 0; i < n; ++i) {
    cidx[i] = 0;
    if (cf[i]) cidx[i] = c++;
  }
  int64_t nnz = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (cf[i]) {
      P->indices[nnz] = cidx[i];
      P->data[nnz++] = 1.0;
    } else {
      double sum = 0.0;
      for (int64_t p = S.indptr[i]; p < S.indptr[i + 1]; ++p)
        if (cf[S.indices[p]]) sum += fabs(S.data[p]);
      if (sum > 1e-14) {
        for (int64_t p = S.indptr[i]; p < S.indptr[i + 1]; ++p)
          if (cf[S.indices[p]]) {
            P->indices[nnz] = cidx[S.indices[p]];
            P->data[nnz++] = fabs(S.data[p]) / sum;
          }
      } else {
        double sum_all = 0.0;
        for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p)
          if (cf[A->indices[p]]) sum_all += fabs(A->data[p]);
        if (sum_all > 1e-14) {
          for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p)
            if (cf[A->indices[p]]) {
              P->indices[nnz] = cidx[A->indices[p]];
              P->data[nnz++] = fabs(A->data[p]) / sum_all;
            }
        } else if (num_coarse > 0) {
          P->indices[nnz] = 0;
          P->data[nnz++] = 1.0;
        }
      }
    }
    P->indptr[i + 1] = nnz;
  }
  P->nnz = nnz;
  free(cidx);
  or_csr_free(&S);
  return OR_OK;
}

/* matrix_multiply (src/linalg/amg.rs:195-226).  The reference forms every
 * C[i][j] as 0.0 plus a_ik b_kj over the stored entries k of A's row i in
 * ascending order (b.get(k, j) present), keeping it when |sum| > 1e-14.  A
 * dense accumulator per row receives the same products in the same order
 * (for a fixed j they arrive in ascending k); columns never reached hold
 * the reference's 0.0 and are not kept. */
int or_csr_matmul(const or_csr *A, const or_csr *B, or_csr *C) {
  if (A->cols != B->rows) return OR_ERR_INVALID_ARG;
  const int64_t m = A->rows, nb = B->cols;
  double *acc = (double *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(double));
  int64_t *mark = (int64_t *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(int64_t));
  int64_t *touched = (int64_t *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(int64_t));
  int64_t cap = A->nnz + B->nnz + 16;
  if (!acc || !mark || !touched || csr_alloc(C, m, nb, cap) != OR_OK) {
    free(acc); free(mark); free(touched);
    return OR_ERR_OOM;
  }
  for (int64_t j = 0; j < nb; ++j) mark[j] = -1;
  int64_t nnz = 0;
  for (int64_t i = 0; i < m; ++i) {
    int64_t nt = 0;
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p) {
      const int64_t k = A->indices[p];
      const double a = A->data[p];
      for (int64_t q = B->indptr[k]; q < B->indptr[k + 1]; ++q) {
        const int64_t j = B->indices[q];
        if (mark[j] != i) {
          mark[j] = i;
          acc[j] = 0.0;
          touched[nt++] = j;
        }
        acc[j] += a * B->data[q];
      }
    }
    qsort(touched, (size_t)nt, sizeof(int64_t), cmp_i64); /* ascending columns */
    for (int64_t x = 0; x < nt; ++x) {
      const int64_t j = touched[x];
      if (fabs(acc[j]) > 1e-14) {
        if (nnz == cap) {
          cap *= 2;
          int64_t *i2 = (int64_t *)realloc(C->indices, (size_t)cap * sizeof(int64_t));
          if (i2) C->indices = i2;
          double *d2 = (double *)realloc(C->data, (size_t)cap * sizeof(double));
          if (d2) C->data = d2;
          if (!i2 || !d2) {
            free(acc); free(mark); free(touched); or_csr_free(C);
            return OR_ERR_OOM;
          }
        }
        C->indices[nnz] = j;
        C->data[nnz++] = acc[j];
      }
    }
    C->indptr[i + 1] = nnz;
  }
  C->nnz = nnz;
  free(acc);
  free(mark);
  free(touched);
  return OR_OK;
}

int or_csr_transpose(const or_csr *A, or_csr *T) {
  if (csr_alloc(T, A->cols, A->rows, A->nnz) != OR_OK) return OR_ERR_OOM;
  for (int64_t p = 0; p < A->nnz; ++p) T->indptr[A->indices[p] + 1]++;
  for (int64_t j = 0; j < A->cols; ++j) T->indptr[j + 1] += T->indptr[j];
  int64_t *cur = (int64_t *)malloc((size_t)(A->cols > 0 ? A->cols : 1) * sizeof(int64_t));
  if (!cur) { or_csr_free(T); return OR_ERR_OOM; }
  memcpy(cur, T->indptr, (size_t)A->cols * sizeof(int64_t));
  for (int64_t i = 0; i < A->rows; ++i)
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p) {
      const int64_t q = cur[A->indices[p]]++;
      T->indices[q] = i;
      T->data[q] = A->data[p];
    }
  T->nnz = A->nnz;
  free(cur);
  return OR_OK;
}

/* extract_diag_inv (src/linalg/amg.rs:266-280) */
void or_amg_diag_inv(const or_csr *A, double *dinv) {
  for (int64_t i = 0; i < A->rows; ++i) {
    dinv[i] = 1.0;
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p)
      if (A->indices[p] == i) {
        if (fabs(A->data[p]) > 1e-14) dinv[i] = 1.0 / A->data[p];
        break;
      }
  }
}

static int csr_copy(const or_csr *A, or_csr *B) {
  if (csr_alloc(B, A->rows, A->cols, A->nnz) != OR_OK) return OR_ERR_OOM;
  memcpy(B->indptr, A->indptr, (size_t)(A->rows + 1) * sizeof(int64_t));
  memcpy(B->indices, A->indices, (size_t)A->nnz * sizeof(int64_t));
  memcpy(B->data, A->data, (size_t)A->nnz * sizeof(double));
  B->nnz = A->nnz;
  return OR_OK;
}

void or_amg_free(or_amg *M) {
  for (int64_t l = 0; l < M->nlevels; ++l) {
    or_csr_free(&M->A[l]);
    or_csr_free(&M->P[l]);
    or_csr_free(&M->PT[l]);
    free(M->dinv[l]);
  }
  free(M->A);
  free(M->P);
  free(M->PT);
  free(M->dinv);
  memset(M, 0, sizeof(*M));
}

/* AMGPreconditioner::new (src/linalg/amg.rs:44-129).  A level is kept with
 * its prolongation while the operator has more than coarse_size rows and
 * coarsening leaves between 1 and n - 1 coarse points; the Galerkin operator
 * P^T (A P) becomes the next level's.  When max_levels levels were all
 * coarsened, the last Galerkin operator is not kept (the loop ends), and
 * the last level keeps its P unused. */
int or_amg_create(const or_csr *A, int64_t max_levels, int64_t coarse_size, double theta,
                  or_amg *out) {
  memset(out, 0, sizeof(*out));
  if (max_levels < 1 || A->rows != A->cols) return OR_ERR_INVALID_ARG;
  out->A = (or_csr *)calloc((size_t)max_levels, sizeof(or_csr));
  out->P = (or_csr *)calloc((size_t)max_levels, sizeof(or_csr));
  out->PT = (or_csr *)calloc((size_t)max_levels, sizeof(or_csr));
  out->dinv = (double **)calloc((size_t)max_levels, sizeof(double *));
  if (!out->A || !out->P || !out->PT || !out->dinv) { or_amg_free(out); return OR_ERR_OOM; }
  or_csr cur;
  int err = csr_copy(A, &cur);
  if (err != OR_OK) { or_amg_free(out); return err; }
  for (int64_t level = 0; level < max_levels; ++level) {
    const int64_t n = cur.rows;
    const int64_t l = out->nlevels;
    out->dinv[l] = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    if (!out->dinv[l]) { or_csr_free(&cur); or_amg_free(out); return OR_ERR_OOM; }
    or_amg_diag_inv(&cur, out->dinv[l]);
    out->A[l] = cur;
    out->nlevels = l + 1;
    if (n <= coarse_size) break;
    uint8_t *cf = (uint8_t *)malloc((size_t)(n > 0 ? n : 1));
    int64_t nc = 0;
    if (!cf) { or_amg_free(out); return OR_ERR_OOM; }
    err = or_amg_coarsen(&cur, theta, cf, &nc);
    if (err != OR_OK) { free(cf); or_amg_free(out); return err; }
    if (nc == 0 || nc >= n) { free(cf); break; }
    err = or_amg_prolongation(&cur, cf, nc, theta, &out->P[l]);
    free(cf);
    if (err == OR_OK) err = or_csr_transpose(&out->P[l], &out->PT[l]);
    if (err != OR_OK) { or_amg_free(out); return err; }
    if (level + 1 == max_levels) break; /* the loop's last Galerkin product is dropped */
    or_csr ap, ac;
    err = or_csr_matmul(&out->A[l], &out->P[l], &ap);
    if (err != OR_OK) { or_amg_free(out); return err; }
    err = or_csr_matmul(&out->PT[l], &ap, &ac);
    or_csr_free(&ap);
    if (err != OR_OK) { or_amg_free(out); return err; }
    cur = ac;
  }
  return OR_OK;
}

/* y = A x, each row summed from 0.0 in stored order (matvec, :230-247) */
static void amg_matvec(const or_csr *A, const double *x, double *y) {
  for (int64_t i = 0; i < A->rows; ++i) {
    double sum = 0.0;
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p) sum += A->data[p] * x[A->indices[p]];
    y[i] = sum;
  }
}

/* weighted_jacobi_step (:250-263) with omega 0.67 */
static void amg_jacobi(const or_csr *A, const double *r, double *x, double *xn,
                       const double *dinv) {
  const double omega = 0.67;
  for (int64_t i = 0; i < A->rows; ++i) {
    double ax = 0.0;
    for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p) ax += A->data[p] * x[A->indices[p]];
    xn[i] = x[i] + omega * (r[i] - ax) * dinv[i];
  }
  memcpy(x, xn, (size_t)A->rows * sizeof(double));
}

/* v_cycle (:132-185): x = 0; on the last level ten in-place sweeps
 * x_i += (r_i - (A x)_i) / a_ii (a Gauss-Seidel sweep: rows in order, each
 * reading the entries already updated); otherwise two weighted-Jacobi steps,
 * the residual restricted by P^T, the coarse correction prolonged by P and
 * added, two more steps. */
static int amg_vcycle(const or_amg *M, int64_t level, const double *r, double *x) {
  const or_csr *A = &M->A[level];
  const int64_t n = A->rows;
  const double *dinv = M->dinv[level];
  for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  if (level == M->nlevels - 1) {
    for (int s = 0; s < 10; ++s)
      for (int64_t i = 0; i < n; ++i) {
        double ax = 0.0;
        for (int64_t p = A->indptr[i]; p < A->indptr[i + 1]; ++p)
          ax += A->data[p] * x[A->indices[p]];
        x[i] += (r[i] - ax) * dinv[i];
      }
    return OR_OK;
  }
  const or_csr *P = &M->P[level], *PT = &M->PT[level];
  const int64_t nc = P->cols;
  double *t = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  double *rc = (double *)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(double));
  double *ec = (double *)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(double));
  if (!t || !rc || !ec) { free(t); free(rc); free(ec); return OR_ERR_OOM; }
  for (int k = 0; k < 2; ++k) amg_jacobi(A, r, x, t, dinv);
  amg_matvec(A, x, t);
  for (int64_t i = 0; i < n; ++i) t[i] = r[i] - t[i];
  amg_matvec(PT, t, rc);
  int err = amg_vcycle(M, level + 1, rc, ec);
  if (err == OR_OK) {
    amg_matvec(P, ec, t);
    for (int64_t i = 0; i < n; ++i) x[i] += t[i];
    for (int k = 0; k < 2; ++k) amg_jacobi(A, r, x, t, dinv);
  }
  free(t);
  free(rc);
  free(ec);
  return err;
}

int or_amg_apply(const or_amg *M, const double *r, double *z) {
  if (M->nlevels < 1) return OR_ERR_INVALID_ARG;
  return amg_vcycle(M, 0, r, z);
}
