/*
 * geofem_oracle.h -- FP64 CPU restatement of geo_simulator's Tet10
 * assemble-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the correctness checker for the
 * HIP implementation in geo_simulator_amd/.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  The product path never links
 * or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Arithmetic is plain IEEE FP64 in the
 * reference's operation order, compiled without floating-point contraction
 * (-ffp-contract=off), because Rust does not fuse multiply-adds.
 *
 * Pinning (what checks each function independently of this file):
 *  - quadrature tables: weights sum to 1/6 and exactness on monomials up to
 *    each rule's degree, against exact rational integrals (tests).
 *  - basis/Jacobian: partition of unity, Kronecker delta, linear-field
 *    reproduction, derivative sums, volumes, curved-element volume (the
 *    reference's own unit tests) and J == vertex-edge matrix for affine tets.
 *  - element matrices/loads: exact rational golden vectors (sympy) for the
 *    reference tet and the six Kuhn tets (tests/golden/).
 *  - global assembly: scipy COO->CSR duplicate summation, dense brute force,
 *    K*1 == 0 (thermal).
 *  - Dirichlet elimination: dense brute force restatement on small meshes,
 *    closed-form nodal solutions (thermal quadratic profile, confined
 *    elastic column).
 *  - CG: the reference's 2x2 KAT, scipy spsolve, iteration-count semantics.
 *  - element backward (adjoint): central finite differences.
 * Bit-exact agreement with nalgebra's 3x3 determinant/inverse rounding and
 * with sprs' duplicate-summation order is parity unpinned (neither crate's
 * source is available); this oracle fixes its own order (ascending element
 * index per CSR slot) and documents it.
 */
#ifndef GEOFEM_ORACLE_H
#define GEOFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OR_OK = 0,
  OR_ERR_INVALID_ARG = 1,
  OR_ERR_SINGULAR_JACOBIAN = 3,
  OR_ERR_BAD_DOFS = 4,
  OR_ERR_INVALID_MATERIAL = 5,
  OR_ERR_OOM = 8,
  OR_ERR_ZERO_PIVOT = 9 /* ILU(0): a row without a diagonal, or a pivot below 1e-15 */
};

/* ---- quadrature (src/fem/quadrature.rs:13-104) ---- */
int or_quad_rule(int npts, const double **pts, const double **weights);

/* ---- Tet10 basis (src/fem/basis.rs) ---- */
void or_shape_functions(const double L[4], double N[10]);
void or_shape_derivs_bary(const double L[4], double dN[10][4]);
void or_jacobian(const double L[4], const double nodes[10][3], double J[3][3]);
double or_det3(const double J[3][3]);
int or_inv3(const double J[3][3], double Jinv[3][3]);
int or_shape_derivs_cart(const double L[4], const double nodes[10][3],
                         double dNdx[10][3]);
double or_element_volume(const double nodes[10][3]);

/* ---- element kernels ---- */
int or_thermal_stiffness(const double nodes[10][3], double k, double Ke[10][10]);
int or_thermal_stiffness_rule(const double nodes[10][3], double k, int npts,
                              double Ke[10][10]);
int or_thermal_load(const double nodes[10][3], double Q, double fe[10]);
int or_constitutive(double E, double nu, double D[6][6]);
void or_b_matrix(const double dNdx[10][3], double B[6][30]);
int or_elastic_stiffness(const double nodes[10][3], double E, double nu,
                         double Ke[30][30]);
int or_gravity_load(const double nodes[10][3], double rho, const double g[3],
                    double fe[30]);
int or_viscous_constitutive(double mu, double zeta, double D[6][6]);
int or_viscosity_stiffness(const double nodes[10][3], double mu, double zeta,
                           double Ke[30][30]);
int or_maxwell_element(const double nodes[10][3], double E, double nu,
                       const double sigma[6], double Ke[30][30], double fe[30]);
int or_maxwell_stress(const double nodes[10][3], double E, double nu, double visc,
                      double dt, const double du[30], const double sn[6], double out[6]);

/* ---- mesh generator (src/mesh_generator_improved.rs:19-169) ---- */
void or_cube_sizes(int64_t nx, int64_t ny, int64_t nz, int64_t *n_nodes,
                   int64_t *n_elems);
int or_generate_cube(int64_t nx, int64_t ny, int64_t nz, double lx, double ly,
                     double lz, double *xyz, int64_t *conn);

/* ---- global assembly (src/fem/assembly.rs) ---- */
typedef struct {
  int64_t rows, cols, nnz;
  int64_t *indptr;  /* rows+1 */
  int64_t *indices; /* nnz, sorted ascending within a row */
  double *data;     /* nnz */
} or_csr;

void or_csr_free(or_csr *m);
int or_csr_from_triplets(int64_t rows, int64_t cols, int64_t ntrip,
                         const int64_t *ti, const int64_t *tj,
                         const double *tv, or_csr *out);
int or_assemble_thermal(int64_t n_nodes, const double *xyz, int64_t n_elems,
                        const int64_t *conn, double k, const double *k_elem,
                        int nthreads, or_csr *out);
int or_assemble_elastic(int64_t n_nodes, const double *xyz, int64_t n_elems,
                        const int64_t *conn, double E, double nu,
                        const double *E_elem, int nthreads, or_csr *out);
int or_assemble_viscosity(int64_t n_nodes, const double *xyz, int64_t n_elems,
                          const int64_t *conn, double mu, double zeta,
                          const double *mu_elem, const double *zeta_elem, int nthreads,
                          or_csr *out);
int or_assemble_maxwell(int64_t n_nodes, const double *xyz, int64_t n_elems,
                        const int64_t *conn, double E, double nu, double visc, double dt,
                        const double *stress, int nthreads, or_csr *out, double *f_hist);
int or_update_stresses_maxwell(int64_t n_nodes, const double *xyz, int64_t n_elems,
                               const int64_t *conn, double E, double nu, double visc,
                               double dt, const double *u_n, const double *u_next,
                               const double *sig_in, double *sig_out);
void or_thermal_mass(const double nodes[10][3], double density, double specific_heat,
                     double M[10][10]);
int or_assemble_thermal_mass(int64_t n_nodes, const double *xyz, int64_t n_elems,
                             const int64_t *conn, double density, double specific_heat,
                             const double *rho_cp_elem, int nthreads, or_csr *out);
int or_assemble_thermal_load(int64_t n_nodes, const double *xyz,
                             int64_t n_elems, const int64_t *conn, double Q,
                             double *f);
int or_assemble_gravity_load(int64_t n_nodes, const double *xyz,
                             int64_t n_elems, const int64_t *conn, double rho,
                             const double g[3], double *f);
int or_apply_dirichlet(const or_csr *K, const double *f, const uint8_t *is_dir,
                       const double *values, or_csr *Kbc, double *fbc);

/* ---- Neumann faces (src/fem/boundary.rs, src/fem/assembly.rs:248-273) ---- */
int or_face_nodes(int local_face, int out[6]);
double or_face_area(const double c[6][3]);
void or_face_flux(const double c[6][3], double flux, double out[6]);
int or_apply_neumann(int64_t n_nodes, const double *xyz, int64_t n_elems,
                     const int64_t *conn, int64_t n_faces, const int64_t *elem_ids,
                     const int32_t *local_faces, const double *fluxes, double *f);

/* ---- linear algebra (src/linalg/) ---- */
typedef struct {
  int64_t iterations;
  double residual_norm;
  double relative_residual;
  int32_t converged;
  int32_t pad_;
  double solve_time;
} or_stats;

void or_spmv(const or_csr *A, const double *x, double *y);
void or_jacobi_diag_inv(const or_csr *A, double *dinv);
/* Timing column only (not a restatement): or_cg_solve's loop on nthreads
 * OpenMP threads, reductions in thread-dependent order. */
int or_cg_solve_omp(const or_csr *A, const double *b, int64_t max_iter, double tol,
                    double abs_tol, int nthreads, double *x, or_stats *st);
int or_cg_solve(const or_csr *A, const double *b, int64_t max_iter, double tol,
                double abs_tol, int use_precond, double *x, or_stats *st);
int or_bicgstab_solve(const or_csr *A, const double *b, int64_t max_iter,
                      double tol, double abs_tol, int use_precond, double *x,
                      or_stats *st);
int or_gmres_solve(const or_csr *A, const double *b, int64_t max_iter,
                   int64_t restart, double tol, double abs_tol, int use_precond,
                   double *x, or_stats *st);

/* ILUPreconditioner::new / update (src/linalg/preconditioner.rs:86-166): the
 * values of A (square, sorted columns) factored in place into factors (nnz),
 * L strictly below the diagonal (unit diagonal implied) and U on and above
 * it, rows in order; diag (rows) receives each row's diagonal position.
 * OR_ERR_ZERO_PIVOT with *bad = the row: a row without a diagonal entry
 * (*kind = 0, checked for all rows first), an eliminating row's pivot below
 * 1e-15 (*kind = 1, *bad = that row k) or a row's own pivot below 1e-15 after
 * its elimination (*kind = 2), the first in the reference's order. */
int or_ilu0_factor(const or_csr *A, double *factors, int64_t *diag, int64_t *bad,
                   int32_t *kind);
/* ILUPreconditioner::apply (src/linalg/preconditioner.rs:168-203): z = U^-1
 * L^-1 r with the factors of or_ilu0_factor on A's pattern. */
void or_ilu0_apply(const or_csr *A, const double *factors, const int64_t *diag,
                   const double *r, double *z);

/* ---- AMGPreconditioner (src/linalg/amg.rs) ----
 * Classical Ruge-Stueben-style AMG as the reference builds it.  Matrices are
 * square with sorted column indices (sprs' CSR).  Efficient loops stand in
 * for the reference's O(n^2) ones where the result is the same by
 * construction (noted at each function). */
/* strength_matrix (:359-390): S[i,j] = A[i,j] for stored j != i with
 * |A[i,j]| >= theta * max_{stored k != i} |A[i,k]|. */
int or_amg_strength(const or_csr *A, double theta, or_csr *S);
/* coarsen (:286-355): cf[i] = 1 for coarse points; returns their count. */
int or_amg_coarsen(const or_csr *A, double theta, uint8_t *cf, int64_t *num_coarse);
/* build_prolongation (:396-487): P (rows x num_coarse). */
int or_amg_prolongation(const or_csr *A, const uint8_t *cf, int64_t num_coarse,
                        double theta, or_csr *P);
/* matrix_multiply (:195-226): C = A B, entries with |sum| > 1e-14 kept. */
int or_csr_matmul(const or_csr *A, const or_csr *B, or_csr *C);
/* B = A^T (sprs transpose_view().to_owned().to_csr()). */
int or_csr_transpose(const or_csr *A, or_csr *T);
/* extract_diag_inv (:266-280). */
void or_amg_diag_inv(const or_csr *A, double *dinv);

typedef struct {
  int64_t nlevels;
  or_csr *A;      /* [nlevels] operators */
  or_csr *P;      /* [nlevels] prolongations; rows == 0 where a level has none */
  or_csr *PT;     /* [nlevels] their transposes */
  double **dinv;  /* [nlevels] extract_diag_inv of each operator */
} or_amg;
/* AMGPreconditioner::new (:44-129): max_levels >= 1. */
int or_amg_create(const or_csr *A, int64_t max_levels, int64_t coarse_size, double theta,
                  or_amg *out);
/* Preconditioner::apply for AMGPreconditioner (:188-192): z = v_cycle(r, 0)
 * (:132-185). */
int or_amg_apply(const or_amg *M, const double *r, double *z);
void or_amg_free(or_amg *M);

/* ---- adjoint element backward (new surface; pinned by central FD) ---- */
int or_thermal_element_backward(int64_t n_nodes, const double *xyz,
                                int64_t n_elems, const int64_t *conn,
                                const double *k_elem, const double *u,
                                const double *lam, double *grad);
int or_elastic_element_backward(int64_t n_nodes, const double *xyz,
                                int64_t n_elems, const int64_t *conn,
                                const double *E_elem, double nu,
                                const double *u, const double *lam,
                                double *grad);

#ifdef __cplusplus
}
#endif
#endif
