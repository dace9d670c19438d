// Tet10 element kernels for gfx950.
//
// Arithmetic follows the reference's operation order term for term, and this
// file is compiled without floating-point contraction, so every element
// matrix and load vector is bit-identical to the FP64 restatement of
//   src/fem/element.rs:23-61,117-140       (thermal stiffness / load)
//   src/mechanics/element.rs:33-65         (elastic B^T D B, 4-point rule)
//   src/mechanics/body_force.rs:49-84      (gravity load)
//   src/fem/basis.rs:85-110,196-276        (dN/dL, Jacobian, J^-T)
// The shape-function derivative tables at the quadrature points are
// compile-time constants; the Jacobian is recomputed per quadrature point
// from all ten nodes, so curved (non-affine) Tet10s are handled.
#pragma clang fp contract(off)

#include <type_traits>

#include "internal.hpp"

namespace gs {
namespace {

// src/fem/quadrature.rs:24-38 (literal constants)
constexpr double QA = 0.5854101966249685;
constexpr double QB = 0.1381966011250105;
constexpr double QW = 1.0 / 24.0;

struct Tables {
  double dl[4][10][3];  // (dN_i/dL1 - dN_i/dL0, .../dL2 - ..., .../dL3 - ...)
  double N[4][10];      // N_i at the quadrature points
};

constexpr double qpl(int q, int m) { return q == m ? QA : QB; }

// dN_i/dL_m (src/fem/basis.rs:85-110)
constexpr double dndl(int i, int m, const double L[4]) {
  switch (i) {
  case 0: return m == 0 ? 4.0 * L[0] - 1.0 : 0.0;
  case 1: return m == 1 ? 4.0 * L[1] - 1.0 : 0.0;
  case 2: return m == 2 ? 4.0 * L[2] - 1.0 : 0.0;
  case 3: return m == 3 ? 4.0 * L[3] - 1.0 : 0.0;
  case 4: return m == 0 ? 4.0 * L[1] : (m == 1 ? 4.0 * L[0] : 0.0);
  case 5: return m == 1 ? 4.0 * L[2] : (m == 2 ? 4.0 * L[1] : 0.0);
  case 6: return m == 0 ? 4.0 * L[2] : (m == 2 ? 4.0 * L[0] : 0.0);
  case 7: return m == 0 ? 4.0 * L[3] : (m == 3 ? 4.0 * L[0] : 0.0);
  case 8: return m == 1 ? 4.0 * L[3] : (m == 3 ? 4.0 * L[1] : 0.0);
  default: return m == 2 ? 4.0 * L[3] : (m == 3 ? 4.0 * L[2] : 0.0);
  }
}

// N_i (src/fem/basis.rs:56-73)
constexpr double shape(int i, const double L[4]) {
  switch (i) {
  case 0: return L[0] * (2.0 * L[0] - 1.0);
  case 1: return L[1] * (2.0 * L[1] - 1.0);
  case 2: return L[2] * (2.0 * L[2] - 1.0);
  case 3: return L[3] * (2.0 * L[3] - 1.0);
  case 4: return 4.0 * L[0] * L[1];
  case 5: return 4.0 * L[1] * L[2];
  case 6: return 4.0 * L[2] * L[0];
  case 7: return 4.0 * L[0] * L[3];
  case 8: return 4.0 * L[1] * L[3];
  default: return 4.0 * L[2] * L[3];
  }
}

constexpr Tables make_tables() {
  Tables t{};
  for (int q = 0; q < 4; ++q) {
    const double L[4] = {qpl(q, 0), qpl(q, 1), qpl(q, 2), qpl(q, 3)};
    for (int i = 0; i < 10; ++i) {
      const double d0 = dndl(i, 0, L);
      t.dl[q][i][0] = dndl(i, 1, L) - d0;
      t.dl[q][i][1] = dndl(i, 2, L) - d0;
      t.dl[q][i][2] = dndl(i, 3, L) - d0;
      t.N[q][i] = shape(i, L);
    }
  }
  return t;
}

__device__ constexpr Tables kT = make_tables();

// 5-point rule (src/fem/quadrature.rs:44-59, literal constants) for the
// thermal mass matrix
constexpr double Q5B = 1.0 / 6.0;
struct Tables5 {
  double dl[5][10][3];
  double N[5][10];
  double w[5];
};
constexpr Tables5 make_tables5() {
  Tables5 t{};
  const double P[5][4] = {{0.25, 0.25, 0.25, 0.25},
                          {Q5B, Q5B, Q5B, 0.5},
                          {Q5B, Q5B, 0.5, Q5B},
                          {Q5B, 0.5, Q5B, Q5B},
                          {0.5, Q5B, Q5B, Q5B}};
  const double W[5] = {-2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0};
  for (int q = 0; q < 5; ++q) {
    for (int i = 0; i < 10; ++i) {
      const double d0 = dndl(i, 0, P[q]);
      t.dl[q][i][0] = dndl(i, 1, P[q]) - d0;
      t.dl[q][i][1] = dndl(i, 2, P[q]) - d0;
      t.dl[q][i][2] = dndl(i, 3, P[q]) - d0;
      t.N[q][i] = shape(i, P[q]);
    }
    t.w[q] = W[q];
  }
  return t;
}
__device__ constexpr Tables5 kT5 = make_tables5();

// The four rules of GaussQuadrature (src/fem/quadrature.rs:13-104), their
// literal constants, in barycentric points and weights.
struct QuadRule {
  int n;
  double L[11][4];
  double w[11];
};
constexpr double Q11A2 = 0.0714285714285714, Q11B2 = 0.7857142857142857;
constexpr double Q11A3 = 0.3994035761667992, Q11B3 = 0.1005964238332008;
constexpr double Q11W1 = -0.01315555555555556, Q11W2 = 0.007622222222222222,
                 Q11W3 = 0.024888888888888888;
constexpr QuadRule make_rule(int n) {
  QuadRule r{};
  r.n = n;
  if (n == 1) {
    for (int m = 0; m < 4; ++m) r.L[0][m] = 0.25;
    r.w[0] = 1.0 / 6.0;
  } else if (n == 4) {
    for (int q = 0; q < 4; ++q) {
      for (int m = 0; m < 4; ++m) r.L[q][m] = qpl(q, m);
      r.w[q] = QW;
    }
  } else if (n == 5) {
    const double P[5][4] = {{0.25, 0.25, 0.25, 0.25},
                            {Q5B, Q5B, Q5B, 0.5},
                            {Q5B, Q5B, 0.5, Q5B},
                            {Q5B, 0.5, Q5B, Q5B},
                            {0.5, Q5B, Q5B, Q5B}};
    const double W[5] = {-2.0 / 15.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0, 3.0 / 40.0};
    for (int q = 0; q < 5; ++q) {
      for (int m = 0; m < 4; ++m) r.L[q][m] = P[q][m];
      r.w[q] = W[q];
    }
  } else if (n == 11) {
    const double P[11][4] = {{0.25, 0.25, 0.25, 0.25},
                             {Q11A2, Q11A2, Q11A2, Q11B2},
                             {Q11A2, Q11A2, Q11B2, Q11A2},
                             {Q11A2, Q11B2, Q11A2, Q11A2},
                             {Q11B2, Q11A2, Q11A2, Q11A2},
                             {Q11A3, Q11A3, Q11B3, Q11B3},
                             {Q11A3, Q11B3, Q11A3, Q11B3},
                             {Q11A3, Q11B3, Q11B3, Q11A3},
                             {Q11B3, Q11A3, Q11A3, Q11B3},
                             {Q11B3, Q11A3, Q11B3, Q11A3},
                             {Q11B3, Q11B3, Q11A3, Q11A3}};
    for (int q = 0; q < 11; ++q) {
      for (int m = 0; m < 4; ++m) r.L[q][m] = P[q][m];
      r.w[q] = q == 0 ? Q11W1 : q < 5 ? Q11W2 : Q11W3;
    }
  }
  return r;
}
constexpr QuadRule kRule1 = make_rule(1), kRule4 = make_rule(4), kRule5 = make_rule(5),
                   kRule11 = make_rule(11);

// dN/dL differences of a rule's points (as Tables for the 4-point rule)
template <int NQ> struct RuleTables {
  double dl[NQ][10][3];
  double w[NQ];
};
template <int NQ> constexpr RuleTables<NQ> make_rule_tables(const QuadRule &r) {
  RuleTables<NQ> t{};
  for (int q = 0; q < NQ; ++q) {
    for (int i = 0; i < 10; ++i) {
      const double d0 = dndl(i, 0, r.L[q]);
      t.dl[q][i][0] = dndl(i, 1, r.L[q]) - d0;
      t.dl[q][i][1] = dndl(i, 2, r.L[q]) - d0;
      t.dl[q][i][2] = dndl(i, 3, r.L[q]) - d0;
    }
    t.w[q] = r.w[q];
  }
  return t;
}
__device__ constexpr RuleTables<1> kRT1 = make_rule_tables<1>(kRule1);
__device__ constexpr RuleTables<4> kRT4 = make_rule_tables<4>(kRule4);
__device__ constexpr RuleTables<5> kRT5 = make_rule_tables<5>(kRule5);
__device__ constexpr RuleTables<11> kRT11 = make_rule_tables<11>(kRule11);
// the tables of the NQ-point rule, named in device code (a __device__
// variable's name in host code is a host-side stand-in, not its address)
template <int NQ> __device__ __forceinline__ const RuleTables<NQ> &rule_tables();
template <> __device__ __forceinline__ const RuleTables<1> &rule_tables<1>() { return kRT1; }
template <> __device__ __forceinline__ const RuleTables<4> &rule_tables<4>() { return kRT4; }
template <> __device__ __forceinline__ const RuleTables<5> &rule_tables<5>() { return kRT5; }
template <> __device__ __forceinline__ const RuleTables<11> &rule_tables<11>() { return kRT11; }

enum : int { FLAG_SINGULAR = 1, FLAG_NONFINITE = 2, FLAG_MATERIAL = 4 };

// J (src/fem/basis.rs:196-223), accumulated over the ten nodes in order.
__device__ __forceinline__ void jacobian(int q, const double X[10][3],
                                         double J[3][3]) {
#pragma unroll
  for (int r = 0; r < 3; ++r) J[r][0] = J[r][1] = J[r][2] = 0.0;
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const double d1 = kT.dl[q][i][0], d2 = kT.dl[q][i][1], d3 = kT.dl[q][i][2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      J[r][0] += X[i][r] * d1;
      J[r][1] += X[i][r] * d2;
      J[r][2] += X[i][r] * d3;
    }
  }
}

// det (cofactor expansion on the first row) and adjugate/det inverse; returns
// det.  Same formulas as the FP64 restatement's 3x3 helpers.
__device__ __forceinline__ double inverse3(const double m[3][3], double o[3][3]) {
  const double m12_23 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  const double m11_23 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  const double m11_22 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  const double det = m[0][0] * m12_23 - m[0][1] * m11_23 + m[0][2] * m11_22;
  o[0][0] = m12_23 / det;
  o[0][1] = (m[0][2] * m[2][1] - m[2][2] * m[0][1]) / det;
  o[0][2] = (m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det;
  o[1][0] = -m11_23 / det;
  o[1][1] = (m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det;
  o[1][2] = (m[0][2] * m[1][0] - m[1][2] * m[0][0]) / det;
  o[2][0] = m11_22 / det;
  o[2][1] = (m[0][1] * m[2][0] - m[2][1] * m[0][0]) / det;
  o[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det;
  return det;
}

__device__ __forceinline__ int det_flags(double det) {
  if (!isfinite(det)) return FLAG_NONFINITE;
  return det == 0.0 ? FLAG_SINGULAR : 0;
}

// grad N_i = J^-T (dN/dL1 - dN/dL0, ...) (src/fem/basis.rs:265-272)
__device__ __forceinline__ void gradients(int q, const double Ji[3][3],
                                          double G[10][3]) {
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    const double v0 = kT.dl[q][i][0], v1 = kT.dl[q][i][1], v2 = kT.dl[q][i][2];
#pragma unroll
    for (int r = 0; r < 3; ++r) G[i][r] = Ji[0][r] * v0 + Ji[1][r] * v1 + Ji[2][r] * v2;
  }
}

__device__ __forceinline__ void load_nodes(const double *__restrict__ xyz,
                                           const int *__restrict__ conn,
                                           int64_t e, double X[10][3]) {
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    const int64_t nd = conn[e * 10 + a];
#pragma unroll
    for (int r = 0; r < 3; ++r) X[a][r] = xyz[nd * 3 + r];
  }
}

// index of entry (a, b) of a symmetric 10x10 matrix in its packed upper
// triangle (row-major, a <= b)
__host__ __device__ constexpr int tri_p(int a, int b) {
  return a <= b ? 10 * a - (a * (a - 1)) / 2 + (b - a) : 10 * b - (b * (b - 1)) / 2 + (a - b);
}

// ---------------------------------------------------------------------------
// Thermal: one thread per element (a wavefront handles a 64-element batch).
// Writes the 55 upper-triangle entries of K_e (bitwise symmetric, because
// grad_i . grad_j is evaluated with the same products in either order) in
// element-major layout Ke[e*55 + p(a,b)], and optionally f_e (Fe[e*10+a]).
// ROWS: row a of K_e goes to the incidence slot i = incpos[e*10 + a] of the
// system assembly's row gather instead, Ke[i*12 + b] (b < 10), with f_e[a]
// at i*12 + 10 and a zero at i*12 + 11: 96 bytes, three whole 32-byte
// sectors per row, written as six 16-byte stores.
// ---------------------------------------------------------------------------
template <bool PER_ELEM, bool WITH_LOAD, bool ROWS>
__global__ __launch_bounds__(128) void k_thermal_elements(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    double k, const double *__restrict__ k_elem, double Q,
    double *__restrict__ Ke, double *__restrict__ Fe, const int *__restrict__ incpos,
    int *__restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double X[10][3];
  load_nodes(xyz, conn, e, X);
  const double kk = PER_ELEM ? k_elem[e] : k;
  double K[55];
#pragma unroll
  for (int p = 0; p < 55; ++p) K[p] = 0.0;
  // the points' weights are kept and f_e formed after the loop (the same
  // terms added in the same point order): ten accumulators fewer live across
  // it brings the kernel under 256 VGPRs, two wavefronts per SIMD instead of
  // one (thermal assembly 0.79 -> 0.77 ms)
  double wq[4];
  int bad = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double J[3][3], Ji[3][3], G[10][3];
    jacobian(q, X, J);
    const double det = inverse3(J, Ji);
    bad |= det_flags(det);
    gradients(q, Ji, G);
    const double w = QW * fabs(det);
    wq[q] = w;
    int p = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
      for (int j = i; j < 10; ++j) {
        const double gd = G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2];
        K[p] += kk * gd * w;
        ++p;
      }
  }
  double F[10];
#pragma unroll
  for (int i = 0; i < 10; ++i) {
    F[i] = 0.0;
    if (WITH_LOAD) {
#pragma unroll
      for (int q = 0; q < 4; ++q) F[i] += kT.N[q][i] * Q * wq[q];
    }
  }
  if (bad) atomicOr(flag, bad);
  if (ROWS) {
    typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int a = 0; a < 10; ++a) {
      d2 *o = (d2 *)(Ke + (int64_t)incpos[e * 10 + a] * 12);
#pragma unroll
      for (int b = 0; b < 10; b += 2)
        o[b / 2] = (d2){K[tri_p(a, b)], K[tri_p(a, b + 1)]};
      o[5] = (d2){F[a], 0.0};
    }
    return;
  }
  double *out = Ke + e * 55;
#pragma unroll
  for (int p = 0; p < 55; ++p) out[p] = K[p];
  if (WITH_LOAD) {
#pragma unroll
    for (int i = 0; i < 10; ++i) Fe[e * 10 + i] = F[i];
  }
}

// ElementMatrix::thermal_stiffness (src/fem/element.rs:23-61) with any of
// the four rules: one thread per element, J, det, J^-1 and the gradients at
// each point (the expressions of k_thermal_elements), K_ij += k (grad N_i .
// grad N_j) w over the points in order; all 100 entries, row-major per
// element (Ke[e*100 + 10 i + j]; entry (j, i) carries the same bits).
template <int NQ, bool PER_ELEM>
__global__ __launch_bounds__(128) void k_thermal_rule(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn, double k,
    const double *__restrict__ k_elem, double *__restrict__ Ke, int *__restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  const RuleTables<NQ> *tb = &rule_tables<NQ>();
  double X[10][3];
  load_nodes(xyz, conn, e, X);
  const double kk = PER_ELEM ? k_elem[e] : k;
  double K[55];
#pragma unroll
  for (int p = 0; p < 55; ++p) K[p] = 0.0;
  int bad = 0;
  for (int q = 0; q < NQ; ++q) {
    double J[3][3], Ji[3][3], G[10][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) J[r][0] = J[r][1] = J[r][2] = 0.0;
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        J[r][0] += X[i][r] * tb->dl[q][i][0];
        J[r][1] += X[i][r] * tb->dl[q][i][1];
        J[r][2] += X[i][r] * tb->dl[q][i][2];
      }
    const double det = inverse3(J, Ji);
    bad |= det_flags(det);
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      const double v0 = tb->dl[q][i][0], v1 = tb->dl[q][i][1], v2 = tb->dl[q][i][2];
#pragma unroll
      for (int r = 0; r < 3; ++r) G[i][r] = Ji[0][r] * v0 + Ji[1][r] * v1 + Ji[2][r] * v2;
    }
    const double w = tb->w[q] * fabs(det);
    int p = 0;
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
      for (int j = i; j < 10; ++j) {
        const double gd = G[i][0] * G[j][0] + G[i][1] * G[j][1] + G[i][2] * G[j][2];
        K[p] += kk * gd * w;
        ++p;
      }
  }
  if (bad) atomicOr(flag, bad);
  double *out = Ke + e * 100;
#pragma unroll
  for (int i = 0; i < 10; ++i)
#pragma unroll
    for (int j = 0; j < 10; ++j) out[i * 10 + j] = K[tri_p(i, j)];
}

// ---------------------------------------------------------------------------
// Elasticity: 32 lanes per element, four elements per 128-thread workgroup.
//
//   1. lanes 0..29 stage the ten node coordinates in LDS; lanes 0..3 form J,
//      det and J^-1 at quadrature point q = lane; lanes then form the forty
//      gradients grad N_i(q) (the same expression as `gradients`), all into
//      LDS, so each is computed once per element instead of once per lane;
//   2. lane a < 30 accumulates row a of the 30x30 K_e and f_e[a];
//   3. the rows go through LDS and leave in block-major order
//      Ke[e*900 + (i*10 + j)*9 + ci*3 + cj] for entry (3i+ci, 3j+cj): every
//      store of the workgroup is contiguous, and the gathers read each 3x3
//      node block as 72 contiguous bytes.
//
// Row a = 3i+ca of B^T (D B) only meets the three Voigt rows k where B[k][a]
// is structurally nonzero (ca=0: k=0,3,5; ca=1: 1,3,4; ca=2: 2,4,5) and
// (D B)[k][b] has a single structural term, so the sparse evaluation below
// performs exactly the nonzero products of the dense reference product in
// the same order; the dropped terms are exact zeros.
// ---------------------------------------------------------------------------
constexpr int EL_PER_WG = 4;

// The isotropic Voigt laws share D's pattern: a normal block with Dd on the
// diagonal and Do off it, and Ds on the shear diagonal.
//   MAT_ELASTIC (src/mechanics/constitutive.rs:63-102): factor = E/((1+nu)(1-2nu)),
//     Dd = (1-nu) factor, Do = nu factor, Ds = ((1-2nu)/2) factor;
//   MAT_VISCOUS (src/mechanics/constitutive.rs:227-255): lambda_v = zeta - 2mu/3,
//     Dd = lambda_v + 2mu, Do = lambda_v, Ds = 2mu.
// The element load is none, the body force w N_i (rho g)_c
// (src/mechanics/body_force.rs:49-84), or the stress-history force
// -sum_q w B^T sigma_e (src/mechanics/element.rs:257-295).
struct DLaw {
  double Dd, Do, Ds;
  bool ok;
};

template <int MAT>
__device__ __forceinline__ DLaw d_law(const ElasticKind &k, int64_t e) {
  DLaw d;
  if constexpr (MAT == GS_MAT_ELASTIC) {
    const double Ee = k.e0 ? k.e0[e] : k.p0, nu = k.p1;
    const double factor = Ee / ((1.0 + nu) * (1.0 - 2.0 * nu));
    d.Dd = (1.0 - nu) * factor;
    d.Do = nu * factor;
    d.Ds = ((1.0 - 2.0 * nu) / 2.0) * factor;
    d.ok = Ee > 0.0;
  } else {
    const double mu = k.e0 ? k.e0[e] : k.p0, zeta = k.e1 ? k.e1[e] : k.p1;
    const double f2mu = 2.0 * mu;
    const double lv = zeta - (2.0 * mu / 3.0);
    d.Dd = lv + f2mu;
    d.Do = lv;
    d.Ds = f2mu;
    d.ok = mu > 0.0;
  }
  return d;
}

// d_law from the law's two values (elastic: E, nu; viscous: mu, zeta).
template <int MAT>
__device__ __forceinline__ DLaw d_law_of(double a, double b) {
  ElasticKind k;
  k.p0 = a;
  k.p1 = b;
  return d_law<MAT>(k, 0);
}

template <int MAT, int LOAD>
__global__ __launch_bounds__(128) void k_elastic_elements(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    const ElasticKind kind, double *__restrict__ Ke, double *__restrict__ Fe,
    int *__restrict__ flag) {
  __shared__ double sX[EL_PER_WG][10][3];
  __shared__ double sJi[EL_PER_WG][4][3][3];
  __shared__ double sW[EL_PER_WG][4];
  __shared__ double sG[EL_PER_WG][4][10][3];
  __shared__ double sK[EL_PER_WG][900];
  const int le = threadIdx.x >> 5;  // element within the workgroup
  const int a = threadIdx.x & 31;   // lane within the element
  const int64_t e = (int64_t)blockIdx.x * EL_PER_WG + le;
  const bool live = e < E;
  int bad = 0;
  if (live && a < 30) {
    const int node = a / 3, r = a - 3 * (a / 3);
    sX[le][node][r] = xyz[(int64_t)conn[e * 10 + node] * 3 + r];
  }
  __syncthreads();
  if (live && a < 4) {
    double X[10][3], J[3][3], Ji[3][3];
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
      for (int r = 0; r < 3; ++r) X[i][r] = sX[le][i][r];
    jacobian(a, X, J);
    const double det = inverse3(J, Ji);
    bad |= det_flags(det);
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) sJi[le][a][r][c] = Ji[r][c];
    sW[le][a] = QW * fabs(det);
  }
  __syncthreads();
  if (live) {
    for (int t = a; t < 40; t += 32) {
      const int q = t / 10, i = t - 10 * (t / 10);
      const double v0 = kT.dl[q][i][0], v1 = kT.dl[q][i][1], v2 = kT.dl[q][i][2];
#pragma unroll
      for (int r = 0; r < 3; ++r)
        sG[le][q][i][r] = sJi[le][q][0][r] * v0 + sJi[le][q][1][r] * v1 +
                          sJi[le][q][2][r] * v2;
    }
  }
  __syncthreads();
  if (live && a < 30) {
    const int ia = a / 3, ca = a - 3 * (a / 3);
    const DLaw law = d_law<MAT>(kind, e);
    if (!law.ok) bad |= FLAG_MATERIAL;
    const double Dd = law.Dd, Do = law.Do, Ds = law.Ds;
    double K[30];
#pragma unroll
    for (int b = 0; b < 30; ++b) K[b] = 0.0;
    double f = 0.0;
    const double bfa = LOAD == GS_LOAD_BODY ? kind.bf[ca] : 0.0;
    // the stress components of the three Voigt rows k1 < k2 < k3 where
    // column a of B is structurally nonzero
    double sk1 = 0.0, sk2 = 0.0, sk3 = 0.0;
    if (LOAD == GS_LOAD_STRESS) {
      const double *sg = kind.sig + e * 6;
      sk1 = sg[ca];
      sk2 = sg[ca == 2 ? 4 : 3];
      sk3 = sg[ca == 1 ? 4 : 5];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double w = sW[le][q];
      // B[k][a] for the three structural Voigt rows of column a
      const double gi0 = sG[le][q][ia][0], gi1 = sG[le][q][ia][1], gi2 = sG[le][q][ia][2];
      const double Bk1 = ca == 0 ? gi0 : (ca == 1 ? gi1 : gi2);  // k1 = ca
      const double Bk2 = ca == 0 ? gi1 : (ca == 1 ? gi0 : gi1);  // k2 = 3,3,4
      const double Bk3 = ca == 0 ? gi2 : (ca == 1 ? gi2 : gi0);  // k3 = 5,4,5
#pragma unroll
      for (int j = 0; j < 10; ++j) {
        const double g0 = sG[le][q][j][0], g1 = sG[le][q][j][1], g2 = sG[le][q][j][2];
#pragma unroll
        for (int cb = 0; cb < 3; ++cb) {
          const double gjc = cb == 0 ? g0 : (cb == 1 ? g1 : g2);
          // (D B)[k1][b] = D[ca][cb] * B[cb][b]
          const double db1 = (ca == cb ? Dd : Do) * gjc;
          // (D B)[k][b] = Ds * B[k][b] for shear rows k = 3, 4, 5
          const double B3 = cb == 0 ? g1 : (cb == 1 ? g0 : 0.0);
          const double B4 = cb == 0 ? 0.0 : (cb == 1 ? g2 : g1);
          const double B5 = cb == 0 ? g2 : (cb == 1 ? 0.0 : g0);
          const double db2 = Ds * (ca == 2 ? B4 : B3);
          const double db3 = Ds * (ca == 1 ? B4 : B5);
          const double s = Bk1 * db1 + Bk2 * db2 + Bk3 * db3;
          K[3 * j + cb] += w * s;
        }
      }
      if (LOAD == GS_LOAD_BODY) f += w * kT.N[q][ia] * bfa;
      // (B^T sigma)[a]: the dense product's nonzero terms in ascending Voigt
      // row; f_history -= w * B^T sigma
      if (LOAD == GS_LOAD_STRESS) f -= w * (Bk1 * sk1 + Bk2 * sk2 + Bk3 * sk3);
    }
#pragma unroll
    for (int j = 0; j < 10; ++j)
#pragma unroll
      for (int cb = 0; cb < 3; ++cb) sK[le][(ia * 10 + j) * 9 + ca * 3 + cb] = K[3 * j + cb];
    if (LOAD != GS_LOAD_NONE) Fe[e * 30 + a] = f;
  }
  if (bad) atomicOr(flag, bad);
  __syncthreads();
  if (live) {
    double *out = Ke + e * 900;
    for (int t = a; t < 900; t += 32) out[t] = sK[le][t];
  }
}

// Entry (r, c) of inverse3's adjugate/det inverse and its det, with
// inverse3's expressions (so every lane's det carries the same bits).
__device__ __forceinline__ double inverse_entry(const double m[3][3], int rc, double &det) {
  const double m12_23 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  const double m11_23 = m[1][0] * m[2][2] - m[2][0] * m[1][2];
  const double m11_22 = m[1][0] * m[2][1] - m[2][0] * m[1][1];
  det = m[0][0] * m12_23 - m[0][1] * m11_23 + m[0][2] * m11_22;
  switch (rc) {
  case 0: return m12_23 / det;
  case 1: return (m[0][2] * m[2][1] - m[2][2] * m[0][1]) / det;
  case 2: return (m[0][1] * m[1][2] - m[1][1] * m[0][2]) / det;
  case 3: return -m11_23 / det;
  case 4: return (m[0][0] * m[2][2] - m[2][0] * m[0][2]) / det;
  case 5: return (m[0][2] * m[1][0] - m[1][2] * m[0][0]) / det;
  case 6: return m11_22 / det;
  case 7: return (m[0][1] * m[2][0] - m[2][1] * m[0][0]) / det;
  default: return (m[0][0] * m[1][1] - m[1][0] * m[0][1]) / det;
  }
}

template <int NE, int NT, typename TG>
__device__ __forceinline__ void stage_jacobians(int nel, double (&sX)[NE][10][3],
                                                double (&sJ)[NE][4][3][3],
                                                double (&sW)[NE][4], TG (&sG)[NE][4][10][3],
                                                int &bad, const Tables *tb = &kT);

// The workgroup's element geometry in LDS, one lane per entry so the
// latency-bound part (J from ten nodes, nine FP64 divisions of the inverse)
// runs on many lanes instead of one per element and point: coordinates,
// J and J^-1 per quadrature point (jacobian's and inverse3's expressions),
// w = |det| / 24, and the forty gradients (gradients' expression), rounded
// to TG.  nel live elements from e0; every thread of the workgroup calls it.
template <int NE, int NT, typename TG>
__device__ __forceinline__ void stage_geometry(int nel, int64_t e0, const double *__restrict__ xyz,
                                               const int *__restrict__ conn,
                                               double (&sX)[NE][10][3],
                                               double (&sJ)[NE][4][3][3],
                                               double (&sW)[NE][4], TG (&sG)[NE][4][10][3],
                                               int &bad) {
  static_assert(NE * 36 <= NT, "one lane per J entry");
  const int t = threadIdx.x, nt = NT;
  for (int k = t; k < nel * 30; k += nt) {
    const int le = k / 30, r = k - 30 * (k / 30);
    const int node = r / 3, c = r - 3 * (r / 3);
    sX[le][node][c] = xyz[(int64_t)conn[(e0 + le) * 10 + node] * 3 + c];
  }
  __syncthreads();
  stage_jacobians<NE, NT, TG>(nel, sX, sJ, sW, sG, bad);
}

// The part of stage_geometry after the coordinates are in sX.
template <int NE, int NT, typename TG>
__device__ __forceinline__ void stage_jacobians(int nel, double (&sX)[NE][10][3],
                                                double (&sJ)[NE][4][3][3],
                                                double (&sW)[NE][4], TG (&sG)[NE][4][10][3],
                                                int &bad, const Tables *tb) {
  static_assert(NE * 36 <= NT, "one lane per J entry");
  const int t = threadIdx.x, nt = NT;
  double jv = 0.0;  // J[r][c] of this lane's (element, point, entry)
  if (t < nel * 36) {
    const int k = t;
    const int le = k / 36, r9 = k - 36 * (k / 36);
    const int q = r9 / 9, rc = r9 - 9 * (r9 / 9);
    const int r = rc / 3, c = rc - 3 * (rc / 3);
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < 10; ++i) acc += sX[le][i][r] * tb->dl[q][i][c];
    jv = acc;
  }
  __syncthreads();
  if (t < nel * 36) {
    const int k = t;
    const int le = k / 36, r9 = k - 36 * (k / 36);
    const int q = r9 / 9, rc = r9 - 9 * (r9 / 9);
    sJ[le][q][rc / 3][rc - 3 * (rc / 3)] = jv;
  }
  __syncthreads();
  double iv = 0.0;
  if (t < nel * 36) {
    const int k = t;
    const int le = k / 36, r9 = k - 36 * (k / 36);
    const int q = r9 / 9, rc = r9 - 9 * (r9 / 9);
    double m[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) m[r][c] = sJ[le][q][r][c];
    double det;
    iv = inverse_entry(m, rc, det);
    if (rc == 0) {
      bad |= det_flags(det);
      sW[le][q] = QW * fabs(det);
    }
  }
  __syncthreads();  // every lane has read J
  if (t < nel * 36) {
    const int k = t;
    const int le = k / 36, r9 = k - 36 * (k / 36);
    const int q = r9 / 9, rc = r9 - 9 * (r9 / 9);
    sJ[le][q][rc / 3][rc - 3 * (rc / 3)] = iv;  // J^-1 replaces J
  }
  __syncthreads();
  for (int k = t; k < nel * 120; k += nt) {
    const int le = k / 120, r = k - 120 * (k / 120);
    const int q = r / 30, ic = r - 30 * (r / 30);
    const int i = ic / 3, c = ic - 3 * (ic / 3);
    const double v0 = tb->dl[q][i][0], v1 = tb->dl[q][i][1], v2 = tb->dl[q][i][2];
    sG[le][q][i][c] = (TG)(sJ[le][q][0][c] * v0 + sJ[le][q][1][c] * v1 + sJ[le][q][2][c] * v2);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Elasticity family, one lane per 3x3 node block (i, j) of K_e: 100 lanes per
// element, EB_ELEMS elements per workgroup of EB_THREADS.
//
//   1. the workgroup stages its elements' geometry in LDS (stage_jacobians:
//      J^-1 and w per quadrature point, the forty gradients per element);
//      the law parameters (per element or uniform) and stress history come
//      through LDS too;
//   2. lane (i, j) sums the nine entries (3i+ci, 3j+cb) of B^T (D B) over the
//      four points.  Column 3i+ci of B is structurally nonzero in three
//      Voigt rows (ci=0: 0,3,5; ci=1: 1,3,4; ci=2: 2,4,5) and (D B)[k][3j+cb]
//      has one structural term (D[k][cb] grad_j[cb] on the normal rows, Ds
//      times a gradient entry on the shear rows), so each entry is the dense
//      reference product's nonzero terms in ascending Voigt row, the dropped
//      terms being exact zeros; with ci and cb unrolled no lane selects;
//   3. a wavefront's 64 blocks (64 x 9 values, contiguous in the block-major
//      layout Ke[e*900 + (i*10 + j)*9 + ci*3 + cb]) go through LDS so that
//      every store writes 64 consecutive values.
// T = double is the assembly kernel (bit-identical to the reference order);
// T = float is the FP32 B^T D B variant (geometry in FP64, rounded to FP32
// gradients and weights; products and sums in FP32).
// ---------------------------------------------------------------------------
constexpr int EB_ELEMS = 5;
constexpr int EB_THREADS = 512;

template <typename T, int MAT, int LOAD>
__global__ __launch_bounds__(EB_THREADS) void k_elastic_blocks(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    const ElasticKind kind, T *__restrict__ Ke, T *__restrict__ Fe, int *__restrict__ flag) {
  __shared__ double sX[EB_ELEMS][10][3];
  __shared__ double sJ[EB_ELEMS][4][3][3];
  __shared__ double sW[EB_ELEMS][4];
  __shared__ double sP[EB_ELEMS][2];  // per-element law parameters (e0, e1)
  __shared__ double sS[EB_ELEMS][6];  // per-element stress (GS_LOAD_STRESS)
  __shared__ T sG[EB_ELEMS][4][10][3];
  __shared__ T sOut[EB_THREADS / 64][64 * 9];
  const int t = threadIdx.x;
  const int64_t g = blockIdx.x;
  const int64_t e0 = g * EB_ELEMS;
  const int nel = (int)(E - e0 < EB_ELEMS ? E - e0 : EB_ELEMS);
  int bad = 0;
  // lanes t < 30 nel load one coordinate each, lanes 160.. the law
  // parameters (two per element), lanes 192.. the stress components
  if (t < 30 * nel) {
    const int le = t / 30, r = t - 30 * (t / 30);
    const int node = r / 3, c = r - 3 * (r / 3);
    sX[le][node][c] = xyz[(int64_t)conn[(e0 + le) * 10 + node] * 3 + c];
  }
  if (t >= 160 && t < 160 + 2 * nel) {
    const int le = (t - 160) >> 1;
    const double *src = (t & 1) ? kind.e1 : kind.e0;
    sP[le][t & 1] = src ? src[e0 + le] : ((t & 1) ? kind.p1 : kind.p0);
  }
  if (LOAD == GS_LOAD_STRESS && t >= 192 && t < 192 + 6 * nel)
    sS[(t - 192) / 6][(t - 192) % 6] = kind.sig[e0 * 6 + (t - 192)];
  __syncthreads();
  {
    const Tables *tb = &kT;
    stage_jacobians<EB_ELEMS, EB_THREADS, T>(nel, sX, sJ, sW, sG, bad, tb);
    const int le = t / 100, p = t - 100 * (t / 100);
    T K[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) K[m] = (T)0;
    if (le < nel) {
      const int64_t e = e0 + le;
      const int i = p / 10, j = p - 10 * (p / 10);
      const DLaw law = d_law_of<MAT>(sP[le][0], sP[le][1]);
      if (!law.ok && p == 0) bad |= FLAG_MATERIAL;
      const T Dd = (T)law.Dd, Do = (T)law.Do, Ds = (T)law.Ds;
      T f0 = (T)0, f1 = (T)0, f2 = (T)0;
      T s0 = (T)0, s1 = (T)0, s2 = (T)0, s3 = (T)0, s4 = (T)0, s5 = (T)0;
      if (LOAD == GS_LOAD_STRESS && j == 0) {
        s0 = (T)sS[le][0], s1 = (T)sS[le][1], s2 = (T)sS[le][2];
        s3 = (T)sS[le][3], s4 = (T)sS[le][4], s5 = (T)sS[le][5];
      }
#pragma unroll 1
      for (int q = 0; q < 4; ++q) {
        const T w = (T)sW[le][q];
        const T a0 = sG[le][q][i][0], a1 = sG[le][q][i][1], a2 = sG[le][q][i][2];
        const T b0 = sG[le][q][j][0], b1 = sG[le][q][j][1], b2 = sG[le][q][j][2];
        const T dd0 = Dd * b0, dd1 = Dd * b1, dd2 = Dd * b2;
        const T do0 = Do * b0, do1 = Do * b1, do2 = Do * b2;
        const T ds0 = Ds * b0, ds1 = Ds * b1, ds2 = Ds * b2;
        // ci = 0 (rows 0, 3, 5)
        K[0] += w * (a0 * dd0 + a1 * ds1 + a2 * ds2);
        K[1] += w * (a0 * do1 + a1 * ds0);
        K[2] += w * (a0 * do2 + a2 * ds0);
        // ci = 1 (rows 1, 3, 4)
        K[3] += w * (a1 * do0 + a0 * ds1);
        K[4] += w * (a1 * dd1 + a0 * ds0 + a2 * ds2);
        K[5] += w * (a1 * do2 + a2 * ds1);
        // ci = 2 (rows 2, 4, 5)
        K[6] += w * (a2 * do0 + a0 * ds2);
        K[7] += w * (a2 * do1 + a1 * ds2);
        K[8] += w * (a2 * dd2 + a1 * ds1 + a0 * ds0);
        if (LOAD == GS_LOAD_BODY && j == 0) {
          const T wn = w * (T)tb->N[q][i];
          f0 += wn * (T)kind.bf[0];
          f1 += wn * (T)kind.bf[1];
          f2 += wn * (T)kind.bf[2];
        }
        // f_history -= w * (B^T sigma)[3i+ci], nonzero terms in ascending row
        if (LOAD == GS_LOAD_STRESS && j == 0) {
          f0 -= w * (a0 * s0 + a1 * s3 + a2 * s5);
          f1 -= w * (a1 * s1 + a0 * s3 + a2 * s4);
          f2 -= w * (a2 * s2 + a1 * s4 + a0 * s5);
        }
      }
      if (LOAD != GS_LOAD_NONE && j == 0) {
        T *fo = Fe + e * 30 + 3 * i;
        fo[0] = f0;
        fo[1] = f1;
        fo[2] = f2;
      }
    }
    // this wavefront's blocks t = 64 wv + l are the values
    // [(64 wv + l) * 9, + 9) of the group's contiguous output
    const int wv = t >> 6, l = t & 63;
    T *buf = sOut[wv];
#pragma unroll
    for (int m = 0; m < 9; ++m) buf[l * 9 + m] = K[m];
    __syncthreads();
    const int64_t valid = (int64_t)nel * 900 - (int64_t)wv * 576;  // values of this wave
    T *out = Ke + e0 * 900 + (int64_t)wv * 576;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int idx = k * 64 + l;
      if (idx < valid) out[idx] = buf[idx];
    }
  }
  if (bad) atomicOr(flag, bad);
}

// ---------------------------------------------------------------------------
// Elasticity on the matrix cores (timing variant; not bit-identical).
//
// K_e = sum_q w_q B_q^T (D B_q) is one product of a 30x24 and a 24x30 matrix,
// the 24 inner indices being (q, Voigt row v): A[a][6q+v] = w_q B_q[v][a],
// C[6q+v][b] = (D B_q)[v][b].  A wavefront per element forms it with
// v_mfma_f64_16x16x4f64 on a 2x2 grid of 16x16 output tiles (rows and
// columns 30, 31 are zero padding), six k-steps each.  Lane l supplies
// A[16I + l%16][4s + l/16] and C[4s + l/16][16J + l%16] and holds
// K[16I + 4t + l/16][16J + l%16] in element t of its accumulator.  The
// geometry (J^-1, w, gradients) is shared through LDS as in the VALU kernel;
// the matrix cores sum in their own order, so entries differ from the VALU
// kernel's by rounding only.
// ---------------------------------------------------------------------------
typedef double dbl4 __attribute__((ext_vector_type(4)));

// B_q[v][a] for a = 3i+c (src/mechanics/strain.rs:38-71 Voigt layout)
template <typename T>
__device__ __forceinline__ T b_entry(int v, int c, T g0, T g1, T g2) {
  const T gc = c == 0 ? g0 : (c == 1 ? g1 : g2);
  switch (v) {
  case 0: case 1: case 2: return v == c ? gc : (T)0;
  case 3: return c == 0 ? g1 : (c == 1 ? g0 : (T)0);
  case 4: return c == 0 ? (T)0 : (c == 1 ? g2 : g1);
  default: return c == 0 ? g2 : (c == 1 ? (T)0 : g0);
  }
}

template <typename T, bool PER_ELEM, bool WITH_LOAD>
__global__ __launch_bounds__(256) void k_elastic_elements_mfma(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    double Emod, double nu, const double *__restrict__ E_elem, double bfx,
    double bfy, double bfz, T *__restrict__ Ke, T *__restrict__ Fe,
    int *__restrict__ flag) {
  __shared__ double sX[EL_PER_WG][10][3];
  __shared__ double sJ[EL_PER_WG][4][3][3];
  __shared__ double sW[EL_PER_WG][4];
  __shared__ double sG[EL_PER_WG][4][10][3];
  __shared__ T sK[EL_PER_WG][900];
  const int le = threadIdx.x >> 6;  // element (wavefront) within the workgroup
  const int l = threadIdx.x & 63;   // lane
  const int64_t e0 = (int64_t)blockIdx.x * EL_PER_WG;
  const int64_t e = e0 + le;
  const bool live = e < E;
  const int nel = (int)(E - e0 < EL_PER_WG ? E - e0 : EL_PER_WG);
  int bad = 0;
  stage_geometry<EL_PER_WG, 256, double>(nel, e0, xyz, conn, sX, sJ, sW, sG, bad);
  const double Ee = live ? (PER_ELEM ? E_elem[e] : Emod) : 1.0;
  if (live && l == 0 && !(Ee > 0.0)) bad |= FLAG_MATERIAL;
  const double factor = Ee / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const T Dd = (T)((1.0 - nu) * factor);
  const T Do = (T)(nu * factor);
  const T Ds = (T)(((1.0 - 2.0 * nu) / 2.0) * factor);
  if (live) {
    // operands: a_op[I][s] = A[16I + l%16][4s + l/16], c_op[J][s] = C[4s + l/16][16J + l%16]
    T a_op[2][6], c_op[2][6];
#pragma unroll
    for (int s6 = 0; s6 < 6; ++s6) {
      const int kk = 4 * s6 + (l >> 4);
      const int q = kk / 6, v = kk - 6 * (kk / 6);
      const T w = (T)sW[le][q];
#pragma unroll
      for (int I = 0; I < 2; ++I) {
        const int a = 16 * I + (l & 15);
        T av = (T)0, cv = (T)0;
        if (a < 30) {
          const int i = a / 3, c = a - 3 * (a / 3);
          const T g0 = (T)sG[le][q][i][0], g1 = (T)sG[le][q][i][1], g2 = (T)sG[le][q][i][2];
          const T bv = (T)b_entry(v, c, g0, g1, g2);
          av = w * bv;
          if (v < 3) {
            const T gc = c == 0 ? g0 : (c == 1 ? g1 : g2);
            cv = (v == c ? Dd : Do) * gc;
          } else {
            cv = Ds * bv;
          }
        }
        a_op[I][s6] = av;
        c_op[I][s6] = cv;
      }
    }
    typedef T acc4 __attribute__((ext_vector_type(4)));
    acc4 acc[2][2];
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int J = 0; J < 2; ++J) {
        acc[I][J] = (acc4){(T)0, (T)0, (T)0, (T)0};
#pragma unroll
        for (int s6 = 0; s6 < 6; ++s6) {
          if constexpr (sizeof(T) == 8)
            acc[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(a_op[I][s6], c_op[J][s6],
                                                             acc[I][J], 0, 0, 0);
          else
            acc[I][J] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_op[I][s6], c_op[J][s6],
                                                             acc[I][J], 0, 0, 0);
        }
      }
    // output lane layout (measured on the part): f64 16x16x4 holds row
    // 4t + l/16 in element t, f32 16x16x4 row 4 (l/16) + t; column l % 16
#pragma unroll
    for (int I = 0; I < 2; ++I)
#pragma unroll
      for (int J = 0; J < 2; ++J)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int rr = sizeof(T) == 8 ? 4 * t + (l >> 4) : 4 * (l >> 4) + t;
          const int row = 16 * I + rr, col = 16 * J + (l & 15);
          if (row < 30 && col < 30)
            sK[le][((row / 3) * 10 + col / 3) * 9 + (row % 3) * 3 + col % 3] = acc[I][J][t];
        }
    if (WITH_LOAD && l < 30) {
      const int ia = l / 3, ca = l - 3 * (l / 3);
      const T bfa = (T)(ca == 0 ? bfx : (ca == 1 ? bfy : bfz));
      T f = (T)0;
#pragma unroll
      for (int q = 0; q < 4; ++q) f += (T)sW[le][q] * (T)kT.N[q][ia] * bfa;
      Fe[e * 30 + l] = f;
    }
  }
  if (bad) atomicOr(flag, bad);
  __syncthreads();
  if (live) {
    T *out = Ke + e * 900;
    for (int t = l; t < 900; t += 64) out[t] = sK[le][t];
  }
}

// max |a - b| and max |b| as non-negative double bit patterns (integer max is
// order independent)
template <typename TA>
__global__ void k_maxdiff(int64_t n, const TA *__restrict__ a,
                          const double *__restrict__ b,
                          unsigned long long *__restrict__ out) {
  unsigned long long md = 0, mb = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double d = fabs((double)a[i] - b[i]), v = fabs(b[i]);
    const unsigned long long bd = (unsigned long long)__double_as_longlong(d == d ? d : __builtin_inf());
    const unsigned long long bv = (unsigned long long)__double_as_longlong(v);
    md = bd > md ? bd : md;
    mb = bv > mb ? bv : mb;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long od = __shfl_xor(md, off, 64), ob = __shfl_xor(mb, off, 64);
    md = od > md ? od : md;
    mb = ob > mb ? ob : mb;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(out, md);
    atomicMax(out + 1, mb);
  }
}

// ---------------------------------------------------------------------------
// Thermal mass (src/fem/element.rs:74-105): M_ij = sum_q rho_cp N_i N_j w_q
// |det J_q| over the 5-point rule, all 100 entries in the reference's order
// ((rho_cp N_i) N_j) w, one thread per element, Me[e*100 + i*10 + j].
// The reference does not invert J here, so nothing is flagged.
template <bool PER_ELEM>
__global__ __launch_bounds__(128) void k_thermal_mass(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    double rho_cp, const double *__restrict__ rc_elem, double *__restrict__ Me) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double X[10][3];
  load_nodes(xyz, conn, e, X);
  const double rc = PER_ELEM ? rc_elem[e] : rho_cp;
  double w[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    double J[3][3];
#pragma unroll
    for (int r = 0; r < 3; ++r) J[r][0] = J[r][1] = J[r][2] = 0.0;
#pragma unroll
    for (int i = 0; i < 10; ++i) {
      const double d1 = kT5.dl[q][i][0], d2 = kT5.dl[q][i][1], d3 = kT5.dl[q][i][2];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        J[r][0] += X[i][r] * d1;
        J[r][1] += X[i][r] * d2;
        J[r][2] += X[i][r] * d3;
      }
    }
    const double m12_23 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
    const double m11_23 = J[1][0] * J[2][2] - J[2][0] * J[1][2];
    const double m11_22 = J[1][0] * J[2][1] - J[2][0] * J[1][1];
    const double det = J[0][0] * m12_23 - J[0][1] * m11_23 + J[0][2] * m11_22;
    w[q] = kT5.w[q] * fabs(det);
  }
  double *out = Me + e * 100;
  for (int i = 0; i < 10; ++i)
#pragma unroll
    for (int j = 0; j < 10; ++j) {
      double m = 0.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) m += rc * kT5.N[q][i] * kT5.N[q][j] * w[q];
      out[i * 10 + j] = m;
    }
}

// ---------------------------------------------------------------------------
// Neumann faces (src/fem/boundary.rs:72-253): the 6 nodal contributions of
// a constant flux over a quadratic triangle face, 3-point rule, one thread
// per face; contrib[6k + i] goes to node[6k + i].
__constant__ int kFaceNodes[4][6] = {{1, 2, 3, 5, 9, 8},
                                     {0, 2, 3, 6, 9, 7},
                                     {0, 1, 3, 4, 8, 7},
                                     {0, 1, 2, 4, 5, 6}};

__global__ __launch_bounds__(128) void k_face_flux(
    int64_t F, const double *__restrict__ xyz, const int *__restrict__ conn,
    const int64_t *__restrict__ elem, const int *__restrict__ lface,
    const double *__restrict__ flux, double *__restrict__ contrib,
    int *__restrict__ node) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= F) return;
  const int64_t e = elem[k];
  const int lf = lface[k];
  double c[6][3];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    const int nd = conn[e * 10 + kFaceNodes[lf][i]];
    node[6 * k + i] = nd;
#pragma unroll
    for (int r = 0; r < 3; ++r) c[i][r] = xyz[(int64_t)nd * 3 + r];
  }
  const double qp[3][3] = {{1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0},
                           {2.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0},
                           {1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0}};
  const double fl = flux[k];
  double out[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double l1 = qp[q][0], l2 = qp[q][1], w = qp[q][2];
    const double l3 = 1.0 - l1 - l2;
    const double n[6] = {l1 * (2.0 * l1 - 1.0), l2 * (2.0 * l2 - 1.0), l3 * (2.0 * l3 - 1.0),
                         4.0 * l1 * l2,         4.0 * l2 * l3,         4.0 * l3 * l1};
    const double d[6][2] = {{4.0 * l1 - 1.0, 0.0},
                            {0.0, 4.0 * l2 - 1.0},
                            {-4.0 * l3 + 1.0, -4.0 * l3 + 1.0},
                            {4.0 * l2, 4.0 * l1},
                            {-4.0 * l2, 4.0 * (l3 - l2)},
                            {4.0 * (l3 - l1), -4.0 * l1}};
    double t1[3] = {0.0, 0.0, 0.0}, t2[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        t1[r] += c[i][r] * d[i][0];
        t2[r] += c[i][r] * d[i][1];
      }
    const double x = t1[1] * t2[2] - t1[2] * t2[1];
    const double y = t1[2] * t2[0] - t1[0] * t2[2];
    const double z = t1[0] * t2[1] - t1[1] * t2[0];
    const double det = sqrt(x * x + y * y + z * z);
#pragma unroll
    for (int i = 0; i < 6; ++i) out[i] += w * n[i] * fl * det;
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) contrib[6 * k + i] = out[i];
}

// ---------------------------------------------------------------------------
// Adjoint element backward: grad_e = -lambda_e^T K_e u_e, evaluated per
// quadrature point as -sum_q w_q k (G^T lambda).(G^T u) (thermal) and
// -sum_q w_q (B lambda)^T D (B u) (elastic), one thread per element.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(128) void k_thermal_backward(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    const double *__restrict__ k_elem, const double *__restrict__ u,
    const double *__restrict__ lam, double *__restrict__ grad,
    int *__restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double X[10][3], ue[10], le[10];
  load_nodes(xyz, conn, e, X);
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    const int nd = conn[e * 10 + a];
    ue[a] = u[nd];
    le[a] = lam[nd];
  }
  const double kk = k_elem[e];
  double s = 0.0;
  int bad = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double J[3][3], Ji[3][3], G[10][3];
    jacobian(q, X, J);
    const double det = inverse3(J, Ji);
    bad |= det_flags(det);
    gradients(q, Ji, G);
    const double w = QW * fabs(det);
    double gu[3] = {0.0, 0.0, 0.0}, gl[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        gu[r] += G[i][r] * ue[i];
        gl[r] += G[i][r] * le[i];
      }
    s += kk * (gl[0] * gu[0] + gl[1] * gu[1] + gl[2] * gu[2]) * w;
  }
  if (bad) atomicOr(flag, bad);
  grad[e] = -s;
}

__global__ __launch_bounds__(128) void k_elastic_backward(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn,
    const double *__restrict__ E_elem, double nu, const double *__restrict__ u,
    const double *__restrict__ lam, double *__restrict__ grad,
    int *__restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double X[10][3], ue[10][3], le[10][3];
  load_nodes(xyz, conn, e, X);
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    const int64_t nd = conn[e * 10 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ue[a][c] = u[nd * 3 + c];
      le[a][c] = lam[nd * 3 + c];
    }
  }
  const double Ee = E_elem[e];
  int bad = (Ee > 0.0) ? 0 : FLAG_MATERIAL;
  const double factor = Ee / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double Dd = (1.0 - nu) * factor, Do = nu * factor;
  const double Ds = ((1.0 - 2.0 * nu) / 2.0) * factor;
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double J[3][3], Ji[3][3], G[10][3];
    jacobian(q, X, J);
    const double det = inverse3(J, Ji);
    bad |= det_flags(det);
    gradients(q, Ji, G);
    const double w = QW * fabs(det);
    // displacement gradients H[c][r] = d u_c / d x_r
    double Hu[3][3] = {}, Hl[3][3] = {};
#pragma unroll
    for (int i = 0; i < 10; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          Hu[c][r] += ue[i][c] * G[i][r];
          Hl[c][r] += le[i][c] * G[i][r];
        }
    // Voigt strains [xx, yy, zz, xy, yz, zx] (engineering shear)
    const double eu[6] = {Hu[0][0], Hu[1][1], Hu[2][2], Hu[0][1] + Hu[1][0],
                          Hu[1][2] + Hu[2][1], Hu[2][0] + Hu[0][2]};
    const double el[6] = {Hl[0][0], Hl[1][1], Hl[2][2], Hl[0][1] + Hl[1][0],
                          Hl[1][2] + Hl[2][1], Hl[2][0] + Hl[0][2]};
    const double sig0 = Dd * eu[0] + Do * eu[1] + Do * eu[2];
    const double sig1 = Do * eu[0] + Dd * eu[1] + Do * eu[2];
    const double sig2 = Do * eu[0] + Do * eu[1] + Dd * eu[2];
    const double en = el[0] * sig0 + el[1] * sig1 + el[2] * sig2 +
                      Ds * (el[3] * eu[3] + el[4] * eu[4] + el[5] * eu[5]);
    s += w * en;
  }
  if (bad) atomicOr(flag, bad);
  grad[e] = -s;
}

inline unsigned blocks_for(int64_t work, int bs) {
  return (unsigned)((work + bs - 1) / bs);
}

// update_stresses_maxwell (src/mechanics/stress_update.rs:39-147), one thread
// per element: du = u_next - u_n on the element's 30 DOFs (DOF 3 node + c),
// the strain increment B du averaged over the 4-point rule's points (each
// B du as the dense product's nonzero terms in ascending column), the
// elastic predictor sigma_n + D_el d_eps in the reference's expression form,
// and its deviatoric part relaxed by 1 / (1 + dt / tau_M), tau_M = mu / G.
__global__ __launch_bounds__(128) void k_maxwell_stress(
    int64_t E, const double *__restrict__ xyz, const int *__restrict__ conn, double Emod,
    double nu, double visc, double dt, const double *__restrict__ u_n,
    const double *__restrict__ u_next, const double *__restrict__ sig_in,
    double *__restrict__ sig_out, int *__restrict__ flag) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double X[10][3];
  load_nodes(xyz, conn, e, X);
  double du[30];
#pragma unroll
  for (int a = 0; a < 10; ++a) {
    const int64_t nd = conn[e * 10 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) du[3 * a + c] = u_next[nd * 3 + c] - u_n[nd * 3 + c];
  }
  double avg[6];
#pragma unroll
  for (int v = 0; v < 6; ++v) avg[v] = 0.0;
  int bad = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double J[3][3], Ji[3][3], G[10][3];
    jacobian(q, X, J);
    bad |= det_flags(inverse3(J, Ji));
    gradients(q, Ji, G);
    // Voigt rows [xx, yy, zz, xy, yz, zx]; columns 3i + c (strain.rs:38-71)
    double d0 = G[0][0] * du[0], d1 = G[0][1] * du[1], d2 = G[0][2] * du[2];
    double d3 = G[0][1] * du[0], d4 = G[0][2] * du[1], d5 = G[0][2] * du[0];
    d3 = d3 + G[0][0] * du[1];
    d4 = d4 + G[0][1] * du[2];
    d5 = d5 + G[0][0] * du[2];
#pragma unroll
    for (int i = 1; i < 10; ++i) {
      const double x = du[3 * i], y = du[3 * i + 1], z = du[3 * i + 2];
      d0 = d0 + G[i][0] * x;
      d3 = d3 + G[i][1] * x;
      d5 = d5 + G[i][2] * x;
      d1 = d1 + G[i][1] * y;
      d3 = d3 + G[i][0] * y;
      d4 = d4 + G[i][2] * y;
      d2 = d2 + G[i][2] * z;
      d4 = d4 + G[i][1] * z;
      d5 = d5 + G[i][0] * z;
    }
    avg[0] += d0;
    avg[1] += d1;
    avg[2] += d2;
    avg[3] += d3;
    avg[4] += d4;
    avg[5] += d5;
  }
  if (bad) atomicOr(flag, bad);
#pragma unroll
  for (int v = 0; v < 6; ++v) avg[v] /= 4.0;
  const double Gm = Emod / (2.0 * (1.0 + nu));
  const double tau = visc / (Emod / (2.0 * (1.0 + nu)));
  const double factor = 1.0 / (1.0 + dt / tau);
  const double Kb = Emod / (3.0 * (1.0 - 2.0 * nu));
  const double lambda = Kb - 2.0 * Gm / 3.0;
  const double *sn = sig_in + e * 6;
  const double tr = avg[0] + avg[1] + avg[2];
  const double t0 = sn[0] + lambda * tr + 2.0 * Gm * avg[0];
  const double t1 = sn[1] + lambda * tr + 2.0 * Gm * avg[1];
  const double t2 = sn[2] + lambda * tr + 2.0 * Gm * avg[2];
  const double t3 = sn[3] + Gm * avg[3];
  const double t4 = sn[4] + Gm * avg[4];
  const double t5 = sn[5] + Gm * avg[5];
  const double p = (t0 + t1 + t2) / 3.0;
  double *so = sig_out + e * 6;
  so[0] = (t0 - p) * factor + p;
  so[1] = (t1 - p) * factor + p;
  so[2] = (t2 - p) * factor + p;
  so[3] = t3 * factor;
  so[4] = t4 * factor;
  so[5] = t5 * factor;
}

}  // namespace

void launch_thermal_elements(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, double k, const double *k_elem,
                             double Q, bool with_load, double *Ke, double *Fe,
                             int *flag, const int *incpos) {
  if (E == 0) return;
  using T = std::true_type;
  using F = std::false_type;
  auto go = [&](auto per, auto load) {
    constexpr bool P = decltype(per)::value, L = decltype(load)::value;
    if (incpos)
      hipLaunchKernelGGL((k_thermal_elements<P, L, true>), blocks_for(E, 128), 128, 0, s, E,
                         xyz, conn, k, k_elem, Q, Ke, Fe, incpos, flag);
    else
      hipLaunchKernelGGL((k_thermal_elements<P, L, false>), blocks_for(E, 128), 128, 0, s, E,
                         xyz, conn, k, k_elem, Q, Ke, Fe, incpos, flag);
  };
  auto with_load_flag = [&](auto per) {
    if (with_load) go(per, T{});
    else go(per, F{});
  };
  if (k_elem) with_load_flag(T{});
  else with_load_flag(F{});
  GS_LAUNCH_CHECK("k_thermal_elements");
}

void launch_thermal_rule(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                         double k, const double *k_elem, int npts, double *Ke, int *flag) {
  if (E == 0) return;
  const unsigned g = blocks_for(E, 128);
  auto go = [&](auto per) {
    constexpr bool P = decltype(per)::value;
    switch (npts) {
    case 1:
      hipLaunchKernelGGL((k_thermal_rule<1, P>), g, 128, 0, s, E, xyz, conn, k, k_elem,
                         Ke, flag);
      break;
    case 4:
      hipLaunchKernelGGL((k_thermal_rule<4, P>), g, 128, 0, s, E, xyz, conn, k, k_elem,
                         Ke, flag);
      break;
    case 5:
      hipLaunchKernelGGL((k_thermal_rule<5, P>), g, 128, 0, s, E, xyz, conn, k, k_elem,
                         Ke, flag);
      break;
    default:
      hipLaunchKernelGGL((k_thermal_rule<11, P>), g, 128, 0, s, E, xyz, conn, k, k_elem,
                         Ke, flag);
      break;
    }
  };
  if (k_elem) go(std::true_type{});
  else go(std::false_type{});
  GS_LAUNCH_CHECK("k_thermal_rule");
}

bool quad_rule(int npts, double *points, double *weights) {
  const QuadRule *r = npts == 1 ? &kRule1 : npts == 4 ? &kRule4 : npts == 5 ? &kRule5
                    : npts == 11 ? &kRule11 : nullptr;
  if (!r) return false;
  for (int q = 0; q < npts; ++q) {
    for (int m = 0; m < 4; ++m) points[4 * q + m] = r->L[q][m];
    weights[q] = r->w[q];
  }
  return true;
}

template <typename T, bool ROWS>
static void launch_family(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                          const ElasticKind &k, T *Ke, T *Fe, int *flag) {
  if (E == 0) return;
  auto go = [&](auto mat, auto load) {
    constexpr int M = decltype(mat)::value, L = decltype(load)::value;
    if constexpr (ROWS)
      hipLaunchKernelGGL((k_elastic_elements<M, L>), blocks_for(E, EL_PER_WG), 128, 0, s, E,
                         xyz, conn, k, Ke, Fe, flag);
    else
      hipLaunchKernelGGL((k_elastic_blocks<T, M, L>), blocks_for(E, EB_ELEMS), EB_THREADS, 0, s,
                         E, xyz, conn, k, Ke, Fe, flag);
  };
  using ME = std::integral_constant<int, GS_MAT_ELASTIC>;
  using MV = std::integral_constant<int, GS_MAT_VISCOUS>;
  using LN = std::integral_constant<int, GS_LOAD_NONE>;
  using LB = std::integral_constant<int, GS_LOAD_BODY>;
  using LS = std::integral_constant<int, GS_LOAD_STRESS>;
  if (k.mat == GS_MAT_VISCOUS) {
    if (k.load == GS_LOAD_BODY) go(MV{}, LB{});
    else go(MV{}, LN{});
  } else {
    if (k.load == GS_LOAD_BODY) go(ME{}, LB{});
    else if (k.load == GS_LOAD_STRESS) go(ME{}, LS{});
    else go(ME{}, LN{});
  }
  GS_LAUNCH_CHECK("k_elastic_elements");
}

void launch_elastic_family(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                           const ElasticKind &k, double *Ke, double *Fe, int *flag) {
  launch_family<double, false>(s, E, xyz, conn, k, Ke, Fe, flag);
}

void launch_elastic_family_rows(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                                const ElasticKind &k, double *Ke, double *Fe, int *flag) {
  launch_family<double, true>(s, E, xyz, conn, k, Ke, Fe, flag);
}

void launch_elastic_family_f32(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                               const ElasticKind &k, float *Ke, float *Fe, int *flag) {
  launch_family<float, false>(s, E, xyz, conn, k, Ke, Fe, flag);
}

void launch_elastic_elements(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, double Emod, double nu,
                             const double *E_elem, double rho, const double g3[3],
                             bool with_load, double *Ke, double *Fe, int *flag) {
  ElasticKind k;
  k.mat = GS_MAT_ELASTIC;
  k.p0 = Emod;
  k.p1 = nu;
  k.e0 = E_elem;
  k.load = with_load ? GS_LOAD_BODY : GS_LOAD_NONE;
  for (int c = 0; c < 3; ++c) k.bf[c] = rho * g3[c];
  launch_elastic_family(s, E, xyz, conn, k, Ke, Fe, flag);
}

void launch_maxwell_stress(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                           double Emod, double nu, double visc, double dt, const double *u_n,
                           const double *u_next, const double *sig_in, double *sig_out,
                           int *flag) {
  if (E == 0) return;
  hipLaunchKernelGGL(k_maxwell_stress, blocks_for(E, 128), 128, 0, s, E, xyz, conn, Emod, nu,
                     visc, dt, u_n, u_next, sig_in, sig_out, flag);
  GS_LAUNCH_CHECK("k_maxwell_stress");
}

void launch_thermal_backward(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, const double *k_elem,
                             const double *u, const double *lam, double *grad,
                             int *flag) {
  if (E == 0) return;
  hipLaunchKernelGGL(k_thermal_backward, blocks_for(E, 128), 128, 0, s, E, xyz,
                     conn, k_elem, u, lam, grad, flag);
  GS_LAUNCH_CHECK("k_thermal_backward");
}

void launch_elastic_backward(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, const double *E_elem, double nu,
                             const double *u, const double *lam, double *grad,
                             int *flag) {
  if (E == 0) return;
  hipLaunchKernelGGL(k_elastic_backward, blocks_for(E, 128), 128, 0, s, E, xyz,
                     conn, E_elem, nu, u, lam, grad, flag);
  GS_LAUNCH_CHECK("k_elastic_backward");
}

template <typename T>
static void launch_mfma(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                        double Emod, double nu, const double *E_elem, double rho,
                        const double g3[3], bool with_load, T *Ke, T *Fe, int *flag) {
  if (E == 0) return;
  const unsigned g = blocks_for(E, EL_PER_WG);
  const double bfx = rho * g3[0], bfy = rho * g3[1], bfz = rho * g3[2];
  auto go = [&](auto pe, auto wl) {
    constexpr bool P = decltype(pe)::value, W = decltype(wl)::value;
    hipLaunchKernelGGL((k_elastic_elements_mfma<T, P, W>), g, 256, 0, s, E, xyz, conn, Emod, nu,
                       E_elem, bfx, bfy, bfz, Ke, Fe, flag);
  };
  using Y = std::true_type;
  using N = std::false_type;
  if (E_elem) {
    if (with_load) go(Y{}, Y{});
    else go(Y{}, N{});
  } else {
    if (with_load) go(N{}, Y{});
    else go(N{}, N{});
  }
  GS_LAUNCH_CHECK("k_elastic_elements_mfma");
}

void launch_elastic_elements_mfma(hipStream_t s, int64_t E, const double *xyz,
                                  const int *conn, double Emod, double nu,
                                  const double *E_elem, double rho, const double g3[3],
                                  bool with_load, double *Ke, double *Fe, int *flag) {
  launch_mfma<double>(s, E, xyz, conn, Emod, nu, E_elem, rho, g3, with_load, Ke, Fe, flag);
}

void launch_elastic_elements_mfma_f32(hipStream_t s, int64_t E, const double *xyz,
                                      const int *conn, double Emod, double nu,
                                      const double *E_elem, double rho, const double g3[3],
                                      bool with_load, float *Ke, float *Fe, int *flag) {
  launch_mfma<float>(s, E, xyz, conn, Emod, nu, E_elem, rho, g3, with_load, Ke, Fe, flag);
}

void launch_thermal_mass(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                         double rho_cp, const double *rc_elem, double *Me) {
  if (E == 0) return;
  if (rc_elem)
    hipLaunchKernelGGL((k_thermal_mass<true>), blocks_for(E, 128), 128, 0, s, E, xyz, conn,
                       rho_cp, rc_elem, Me);
  else
    hipLaunchKernelGGL((k_thermal_mass<false>), blocks_for(E, 128), 128, 0, s, E, xyz, conn,
                       rho_cp, rc_elem, Me);
  GS_LAUNCH_CHECK("k_thermal_mass");
}

void launch_face_flux(hipStream_t s, int64_t F, const double *xyz, const int *conn,
                      const int64_t *elem, const int *lface, const double *flux,
                      double *contrib, int *node) {
  if (F == 0) return;
  hipLaunchKernelGGL(k_face_flux, blocks_for(F, 128), 128, 0, s, F, xyz, conn, elem, lface,
                     flux, contrib, node);
  GS_LAUNCH_CHECK("k_face_flux");
}

void launch_maxdiff(hipStream_t s, int64_t n, const double *a, const double *b,
                    unsigned long long *out) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_maxdiff<double>, 2048, 256, 0, s, n, a, b, out);
  GS_LAUNCH_CHECK("k_maxdiff");
}

void launch_maxdiff_f32(hipStream_t s, int64_t n, const float *a, const double *b,
                        unsigned long long *out) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_maxdiff<float>, 2048, 256, 0, s, n, a, b, out);
  GS_LAUNCH_CHECK("k_maxdiff");
}

}  // namespace gs
