/*
 * geofem.h -- C ABI of the MI355X (gfx950) Tet10 assemble-and-solve library
 * (libgeofem.so, built from geo_simulator_amd/csrc/).
 *
 * Each entry point replaces one call of geo_simulator's Rust API; the
 * reference interface is cited beside it (paths relative to the reference
 * repository root).  Conventions:
 *  - Every call returns gs_status; GS_OK == 0.  A condition on which the
 *    reference panics returns a non-zero status instead, and
 *    gs_last_error(ctx) gives the message.  CG non-convergence is not an
 *    error (stats->converged == 0), as in the reference.  After an error
 *    the contents of caller-owned output buffers are unspecified and no
 *    output handle is returned.
 *  - Pointers whose name ends in _dev are device pointers (HBM of the ctx's
 *    device, e.g. from gs_device_alloc or any allocator on that device);
 *    every other pointer is host memory.  Device handles (mesh, dofmap, csr)
 *    are owned by the library and freed with their *_destroy call.
 *  - One HIP stream per context; all work of a context is ordered on it and
 *    every entry point returns after its work is complete.  A context is not
 *    re-entrant; distinct contexts may be used from distinct threads.
 *  - Indices inside device handles are int32: rows, columns and element
 *    counts must stay below 2^31 (GS_ERR_TOO_LARGE otherwise).
 */
#ifndef GEOFEM_H
#define GEOFEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  GS_OK = 0,
  GS_ERR_INVALID_ARG = 1,
  GS_ERR_HIP = 2,
  GS_ERR_SINGULAR_JACOBIAN = 3, /* src/fem/basis.rs:247 "Singular Jacobian" */
  GS_ERR_BAD_DOFS = 4,          /* src/fem/assembly.rs:394 dpn != 3 */
  GS_ERR_INVALID_MATERIAL = 5,  /* src/mechanics/constitutive.rs:31-35 */
  GS_ERR_NONFINITE = 6,         /* non-finite Jacobian determinant */
  GS_ERR_TOO_LARGE = 7,         /* exceeds int32 indexing or a fixed limit */
  GS_ERR_OOM = 8,
  GS_ERR_ZERO_PIVOT = 9         /* ILU(0): src/linalg/preconditioner.rs:103,140,159 */
} gs_status;

typedef enum { GS_PREC_FP64 = 0, GS_PREC_MIXED = 1 } gs_precision;

typedef struct gs_ctx gs_ctx;
typedef struct gs_mesh gs_mesh;
typedef struct gs_dofmap gs_dofmap;
typedef struct gs_csr gs_csr;
typedef struct gs_ilu gs_ilu;
typedef struct gs_amg gs_amg;

/* ---- context / memory (no reference counterpart: device plumbing) ---- */
gs_status gs_ctx_create(int device, gs_ctx **out);
void gs_ctx_destroy(gs_ctx *ctx);
/* Message of the context's last failed call; with ctx NULL, of the calling
 * thread's last failed gs_ctx_create. */
const char *gs_last_error(const gs_ctx *ctx);
gs_status gs_ctx_synchronize(gs_ctx *ctx);
gs_status gs_device_alloc(gs_ctx *ctx, size_t bytes, void **out_dev);
gs_status gs_device_free(gs_ctx *ctx, void *p_dev);
gs_status gs_memcpy_h2d(gs_ctx *ctx, void *dst_dev, const void *src, size_t bytes);
gs_status gs_memcpy_d2h(gs_ctx *ctx, void *dst, const void *src_dev, size_t bytes);

/* ---- mesh (host-side generator + device upload) ----
 * ImprovedMeshGenerator::generate_cube (src/mesh_generator_improved.rs:19-169):
 * same node numbering (vertex grid ix-fastest, then edge midpoints in
 * first-touch order) and element order. */
gs_status gs_generate_cube_sizes(int64_t nx, int64_t ny, int64_t nz,
                                 int64_t *n_nodes, int64_t *n_elems);
gs_status gs_generate_cube(int64_t nx, int64_t ny, int64_t nz, double lx,
                           double ly, double lz, double *xyz /* n_nodes*3 */,
                           int64_t *conn /* n_elems*10 */);
/* Mesh{geometry.nodes, connectivity.tet10_elements}
 * (src/mesh/geometry.rs:8-11,41-50, src/mesh/topology.rs:13-15,55-56). */
gs_status gs_mesh_create(gs_ctx *ctx, const double *xyz, int64_t n_nodes,
                         const int64_t *conn, int64_t n_elems, gs_mesh **out);
void gs_mesh_destroy(gs_mesh *mesh);

/* DofManager::new + set_dirichlet (src/fem/dof.rs:37-50,113-118):
 * n_nodes*dofs_per_node DOFs numbered node*dpn+c; is_dirichlet/values are
 * per DOF (values ignored where is_dirichlet == 0). */
gs_status gs_dofmap_create(gs_ctx *ctx, int64_t n_nodes, int dofs_per_node,
                           const uint8_t *is_dirichlet, const double *values,
                           gs_dofmap **out);
void gs_dofmap_destroy(gs_dofmap *dof);

/* ---- assembly ---- */
/* Assembler::assemble_thermal_stiffness_parallel (src/fem/assembly.rs:69-117).
 * k_per_elem_dev (nullable, n_elems doubles) overrides the uniform k
 * (additive overload for heterogeneous conductivity). */
gs_status gs_assemble_thermal_stiffness(gs_ctx *ctx, gs_mesh *mesh,
                                        const gs_dofmap *dof, double conductivity,
                                        const double *k_per_elem_dev,
                                        gs_csr **out);
/* GaussQuadrature::tet_1point / tet_4point / tet_5point / tet_11point
 * (src/fem/quadrature.rs:13-104): the literal tables the kernels integrate
 * with; n_points in {1, 4, 5, 11}, points n_points x 4 barycentric, weights
 * n_points (host arrays). */
gs_status gs_quadrature_rule(int32_t n_points, double *points, double *weights);
/* ElementMatrix::thermal_stiffness (src/fem/element.rs:23-61) of every element
 * of the mesh (caller element order), integrated with the rule of n_points
 * points (the reference's element routine and assembly use 4; its tests also
 * 1, 5 and 11): Ke_dev receives n_elems x 100 doubles, entry (i, j) of
 * element e at e*100 + 10 i + j.  k_per_elem_dev (nullable) overrides k. */
gs_status gs_thermal_element_matrices(gs_ctx *ctx, gs_mesh *mesh, double conductivity,
                                      const double *k_per_elem_dev, int32_t n_points,
                                      double *Ke_dev);
/* Assembler::assemble_thermal_load (src/fem/assembly.rs:130-154). */
gs_status gs_assemble_thermal_load(gs_ctx *ctx, gs_mesh *mesh,
                                   const gs_dofmap *dof, double source,
                                   double *f_dev);
/* Assembler::assemble_elasticity_stiffness_parallel
 * (src/fem/assembly.rs:389-437); E_per_elem_dev nullable (heterogeneous E). */
gs_status gs_assemble_elasticity_stiffness(gs_ctx *ctx, gs_mesh *mesh,
                                           const gs_dofmap *dof, double youngs,
                                           double poisson,
                                           const double *E_per_elem_dev,
                                           gs_csr **out);
/* ElementMatrix::thermal_mass (src/fem/element.rs:74-105: M_e = int rho c_p
 * N^T N dV with the 5-point rule) for every element, scattered like the
 * stiffness (DOF = node id; the reference has the element routine only).
 * rho_cp_per_elem_dev (nullable) replaces density * specific_heat per element.
 * All 100 entries of each M_e are summed (M is symmetric to rounding, as the
 * reference's). */
gs_status gs_assemble_thermal_mass(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                   double density, double specific_heat,
                                   const double *rho_cp_per_elem_dev, gs_csr **out);
/* Assembler::apply_neumann_bcs (src/fem/assembly.rs:248-273) with
 * Tet10Face::from_element_face / integrate_flux (src/fem/boundary.rs:72-253):
 * f_dev[node] += the constant flux over face local_faces[k] (0..3) of element
 * element_ids[k], 6 quadratic-triangle nodes, 3-point rule.  Face lists are
 * host arrays; contributions to one node are added in face order (the
 * reference iterates a hash map, whose order varies).  An invalid face index
 * (the reference panics) or element id returns GS_ERR_INVALID_ARG. */
gs_status gs_apply_neumann_bcs(gs_ctx *ctx, gs_mesh *mesh, int64_t n_faces,
                               const int64_t *element_ids, const int32_t *local_faces,
                               const double *fluxes, double *f_dev);
/* Assembler::assemble_gravity_load (src/fem/assembly.rs:906-937). */
gs_status gs_assemble_gravity_load(gs_ctx *ctx, gs_mesh *mesh,
                                   const gs_dofmap *dof, double density,
                                   const double gravity[3], double *f_dev);
/* Assembler::apply_dirichlet_bcs (src/fem/assembly.rs:168-230). */
gs_status gs_apply_dirichlet_bcs(gs_ctx *ctx, const gs_csr *K,
                                 const double *f_dev, const gs_dofmap *dof,
                                 gs_csr **K_bc, double *f_bc_dev);

/* Fused assemble_thermal_stiffness_parallel + assemble_thermal_load +
 * apply_dirichlet_bcs, device-resident end to end (the benchmark callers'
 * sequence, benchmarks/thermal_1d_improved_mesh.rs:57-59). */
gs_status gs_assemble_thermal_system(gs_ctx *ctx, gs_mesh *mesh,
                                     const gs_dofmap *dof, double conductivity,
                                     const double *k_per_elem_dev, double source,
                                     gs_csr **K_bc, double *f_bc_dev);
/* Same for elasticity + gravity (benchmarks/elasticity_halfspace_gravity.rs:112-122). */
gs_status gs_assemble_elasticity_system(gs_ctx *ctx, gs_mesh *mesh,
                                        const gs_dofmap *dof, double youngs,
                                        double poisson,
                                        const double *E_per_elem_dev,
                                        double density, const double gravity[3],
                                        gs_csr **K_bc, double *f_bc_dev);

/* ---- values-only re-assembly for iterative callers ----
 * The calls below reuse the mesh's cached CSR pattern and contributor lists
 * (and, for the *_system calls, the cached elimination plan of the DOF map),
 * so each new viscosity field or stress history costs only the element
 * kernel and the numeric gathers. */
/* Assembler::assemble_viscosity_stiffness_parallel (src/fem/assembly.rs:646-694)
 * with NewtonianViscosity{viscosity, bulk_viscosity}
 * (src/mechanics/constitutive.rs:156-255): K = sum_e int B^T D B dV,
 * D = (zeta - 2 mu / 3) on the normal block, plus 2 mu on its diagonal and on
 * the shear diagonal.  mu_per_elem_dev / zeta_per_elem_dev (nullable, n_elems
 * doubles) replace mu / zeta per element (additive overload for Picard-type
 * callers whose effective viscosity varies, src/linalg/picard.rs:262-275).
 * mu <= 0 returns GS_ERR_INVALID_MATERIAL ("Viscosity must be positive"). */
gs_status gs_assemble_viscosity_stiffness(gs_ctx *ctx, gs_mesh *mesh,
                                          const gs_dofmap *dof, double viscosity,
                                          double bulk_viscosity,
                                          const double *mu_per_elem_dev,
                                          const double *zeta_per_elem_dev,
                                          gs_csr **out);
/* assemble_viscosity_stiffness_parallel followed by apply_dirichlet_bcs
 * (Picard's assemble / eliminate pair, src/linalg/picard.rs:264-268),
 * device-resident: f_dev (nullable: zero) is the right-hand side before the
 * elimination, in caller numbering. */
gs_status gs_assemble_viscosity_system(gs_ctx *ctx, gs_mesh *mesh,
                                       const gs_dofmap *dof, double viscosity,
                                       double bulk_viscosity,
                                       const double *mu_per_elem_dev,
                                       const double *zeta_per_elem_dev,
                                       const double *f_dev, gs_csr **K_bc,
                                       double *f_bc_dev);
/* Assembler::assemble_maxwell_viscoelastic_parallel (src/fem/assembly.rs:959-1020)
 * with MaxwellViscoelasticity::new(E, nu, viscosity)
 * (src/mechanics/constitutive.rs:290-325) and ElasticityElement::
 * maxwell_viscoelastic (src/mechanics/element.rs:257-295): K with the elastic
 * D, and f_history = -sum_e sum_q w B^T sigma_e.  stress_dev is the stress
 * history (Mesh::stress_history, src/mesh/state.rs:12-16): n_elems x 6 Voigt
 * [xx, yy, zz, xy, yz, xz] in element order.  dt is unused, as in the
 * reference. */
gs_status gs_assemble_maxwell_viscoelastic(gs_ctx *ctx, gs_mesh *mesh,
                                           const gs_dofmap *dof, double youngs,
                                           double poisson, double viscosity, double dt,
                                           const double *stress_dev, gs_csr **K,
                                           double *f_history_dev);
/* BackwardEuler::step's assembly (src/fem/time_integration.rs:99-110): the
 * Maxwell K and f_history, f = f_ext + f_history, then apply_dirichlet_bcs,
 * device-resident; f_ext_dev in caller numbering. */
gs_status gs_assemble_maxwell_system(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                     double youngs, double poisson, double viscosity,
                                     double dt, const double *stress_dev,
                                     const double *f_ext_dev, gs_csr **K_bc,
                                     double *f_bc_dev);
/* update_stresses_maxwell (src/mechanics/stress_update.rs:39-147):
 * stress_next_dev (n_elems x 6) from stress_dev and the displacements
 * u_n_dev, u_next_dev (DOF 3 node + c, caller numbering). */
gs_status gs_update_stresses_maxwell(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                     double youngs, double poisson, double viscosity,
                                     double dt, const double *u_n_dev,
                                     const double *u_next_dev, const double *stress_dev,
                                     double *stress_next_dev);

/* ---- sparse matrix handle (sprs CsMat<f64> counterpart) ---- */
gs_status gs_csr_info(const gs_csr *A, int64_t *rows, int64_t *cols,
                      int64_t *nnz);
/* Copy to the sprs shape: int64 indptr (rows+1), sorted int64 indices, data. */
gs_status gs_csr_export(gs_ctx *ctx, const gs_csr *A, int64_t *indptr,
                        int64_t *indices, double *data);
/* Upload a host CSR (sorted, duplicate-free rows) for Solver::solve on an
 * arbitrary CsMat. */
gs_status gs_csr_from_host(gs_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz,
                           const int64_t *indptr, const int64_t *indices,
                           const double *data, gs_csr **out);
void gs_csr_destroy(gs_csr *A);
/* impl LinearOperator for CsMat<f64>::apply (src/linalg/solver.rs:62-73). */
gs_status gs_csr_spmv(gs_ctx *ctx, const gs_csr *A, const double *x_dev,
                      double *y_dev);
/* JacobiPreconditioner::new (src/linalg/preconditioner.rs:37-51). */
gs_status gs_jacobi_diag_inv(gs_ctx *ctx, const gs_csr *A, double *dinv_dev);

/* ---- Jacobi-preconditioned CG ----
 * ConjugateGradient::new().with_max_iterations().with_tolerance()
 * .with_abs_tolerance().with_preconditioner() (src/linalg/iterative.rs:16-44)
 * and Solver::solve (src/linalg/solver.rs:108, src/linalg/iterative.rs:123-131). */
typedef struct {
  int64_t max_iterations;    /* default 1000 */
  double tolerance;          /* default 1e-8 (relative to ||b||) */
  double abs_tolerance;      /* default 1e-12 */
  int32_t use_preconditioner;/* default 1 (Jacobi) */
  int32_t precision;         /* gs_precision, default GS_PREC_FP64 */
  double inner_tolerance;    /* mixed: an FP64 flexible-CG step each time the
                                FP32 recursion's residual fell by this
                                factor (default 0: automatic, 1e-3 on the
                                scaled operator below, else 1e-2; on scalar
                                operators with a positive Jacobi diagonal
                                the recursion runs on D^-1/2 A D^-1/2 and
                                the residual is ||D^-1/2 r||) */
  int64_t max_inner_iterations; /* mixed: cap on FP32 passes between FP64
                                   steps (0 = none) */
  int64_t max_outer_iterations; /* mixed: cap on FP64 steps (default 50;
                                   0 = 100) */
  int32_t relative_breakdown; /* FP64: stop on p.Ap == 0 or a non-finite step
                                 instead of the reference's |p.Ap| < 1e-30,
                                 whose meaning depends on the units of b
                                 (no reference counterpart; for reverse
                                 solves); default 0.  Mixed always uses it. */
  int32_t reserved_;          /* 0 */
} gs_cg_opts;

/* SolverStats (src/linalg/solver.rs:5-20) plus mixed-precision counters. */
typedef struct {
  int64_t iterations;        /* FP64: as the reference; mixed: total inner */
  double residual_norm;
  double relative_residual;
  int32_t converged;
  int32_t pad_;
  double solve_time;         /* seconds, host wall clock */
  int64_t outer_iterations;  /* mixed: FP64 flexible-CG steps */
  int64_t spmv_count;        /* SpMV launches issued */
} gs_solver_stats;

gs_status gs_cg_opts_default(gs_cg_opts *opts);
gs_status gs_cg_solve(gs_ctx *ctx, const gs_csr *A, const double *b_dev,
                      const gs_cg_opts *opts, double *x_dev,
                      gs_solver_stats *stats);
/* gs_cg_solve in FP64 without allocating device memory (no reference
 * counterpart: for callers that cannot allocate, such as code under HIP graph
 * capture or tight serving loops).  The solver's vectors, partial sums and
 * state live in a fixed device array of the library (about 96 MiB on every
 * device the library runs on, held until the process ends), so A may have at
 * most 2^21 rows (GS_ERR_TOO_LARGE otherwise).  A must already hold its
 * solver arrays: a matrix returned by gs_assemble_*_system, or one solved
 * once with gs_cg_solve under the same gs_csr_set_spmv_shape flags
 * (GS_ERR_INVALID_ARG otherwise, as for a precision other than FP64).
 * Results equal gs_cg_solve's bit for bit.  Calls on one device share the
 * array: calls that overlap on one device race. */
gs_status gs_cg_solve_noalloc(gs_ctx *ctx, const gs_csr *A, const double *b_dev,
                              const gs_cg_opts *opts, double *x_dev,
                              gs_solver_stats *stats);
/* BiCGSTAB::new().with_max_iterations().with_tolerance().with_abs_tolerance()
 * .with_preconditioner() and Solver::solve (src/linalg/iterative.rs:149-291),
 * for general (nonsymmetric) square matrices: the same loop, stopping tests
 * and iteration count (a pass that ends on the intermediate residual s is
 * not counted).  Uses max_iterations, tolerance, abs_tolerance and
 * use_preconditioner of opts (the reference's defaults: 1000, 1e-8, 1e-12,
 * Jacobi); precision must be GS_PREC_FP64. */
gs_status gs_bicgstab_solve(gs_ctx *ctx, const gs_csr *A, const double *b_dev,
                            const gs_cg_opts *opts, double *x_dev,
                            gs_solver_stats *stats);
/* GMRES::new().with_restart().with_max_iterations().with_tolerance()
 * .with_abs_tolerance().with_preconditioner() and Solver::solve
 * (src/linalg/iterative.rs:302-486): restarted GMRES(restart), right Jacobi
 * preconditioning, modified Gram-Schmidt, Givens rotations, each cycle from
 * the true residual; iterations count inner steps over all cycles.  Uses
 * max_iterations, tolerance, abs_tolerance and use_preconditioner of opts
 * (the reference's defaults: 1000, 1e-8, 1e-12, Jacobi; restart 10);
 * restart in [1, 200] (the reference never advances with 0); FP64. */
gs_status gs_gmres_solve(gs_ctx *ctx, const gs_csr *A, const double *b_dev,
                         const gs_cg_opts *opts, int32_t restart, double *x_dev,
                         gs_solver_stats *stats);

/* ---- ILU(0) preconditioner (src/linalg/preconditioner.rs:78-204) ----
 * ILUPreconditioner::new (:86-108, factorize_in_place :126-166): incomplete
 * LU with zero fill of a square A, in A's caller numbering with columns
 * ascending (the matrix the reference would hold), rows eliminated in
 * order; bit-identical factors.  GS_ERR_ZERO_PIVOT with the reference's
 * message ("Row i has no diag", "Zero diag at k", "Zero diag after update at
 * i") where the reference returns Err. */
gs_status gs_ilu_create(gs_ctx *ctx, const gs_csr *A, gs_ilu **out);
/* ILUPreconditioner::update (:110-124): refactor with A's values; A must
 * have the factored pattern ("Size mismatch" for another size,
 * GS_ERR_INVALID_ARG for another pattern). */
gs_status gs_ilu_update(gs_ctx *ctx, gs_ilu *ilu, const gs_csr *A);
/* Preconditioner::apply for ILUPreconditioner (:168-203): z = U^-1 L^-1 r,
 * forward then backward substitution in the reference's order; device
 * vectors of n doubles in A's caller numbering (r_dev and z_dev distinct). */
gs_status gs_ilu_apply(gs_ctx *ctx, const gs_ilu *ilu, const double *r_dev, double *z_dev);
/* Size, and the dependency levels of the forward and backward substitutions
 * (one launch each); no reference counterpart. */
gs_status gs_ilu_info(const gs_ilu *ilu, int64_t *n, int64_t *nnz, int64_t *levels_lower,
                      int64_t *levels_upper);
/* The factors on A's pattern in caller numbering, sprs shape (host arrays:
 * indptr n+1, indices nnz, factors nnz): the reference's `factors` field. */
gs_status gs_ilu_export(gs_ctx *ctx, const gs_ilu *ilu, int64_t *indptr, int64_t *indices,
                        double *factors);
void gs_ilu_destroy(gs_ilu *ilu);

/* ---- AMG preconditioner (src/linalg/amg.rs) ----
 * AMGPreconditioner::new(&A, max_levels, coarse_size, strength_threshold)
 * (:44-129; the reference's defaults 10, 500, 0.25): the hierarchy of a
 * square A in A's caller numbering with columns ascending, built as the
 * reference builds it (strength of connection, greedy coarsening, direct
 * interpolation, Galerkin operators P^T (A P) with |entries| <= 1e-14
 * dropped); bit-identical levels.  max_levels >= 1 and coarse_size >= 0
 * (GS_ERR_INVALID_ARG otherwise: the reference's V-cycle cannot run
 * without a level). */
gs_status gs_amg_create(gs_ctx *ctx, const gs_csr *A, int64_t max_levels,
                        int64_t coarse_size, double strength_threshold, gs_amg **out);
/* Preconditioner::apply for AMGPreconditioner (:188-192): z = one V-cycle
 * (:132-185) on r, bit-identical to the reference's; device vectors of n
 * doubles in A's caller numbering (r_dev and z_dev distinct). */
gs_status gs_amg_apply(gs_ctx *ctx, const gs_amg *amg, const double *r_dev, double *z_dev);
/* Number of levels (no reference counterpart). */
gs_status gs_amg_levels(const gs_amg *amg, int64_t *levels);
/* Level `level`'s sizes: operator rows and stored entries, the columns and
 * stored entries of its prolongation (0 and 0 where it has none), and the
 * dependency levels of its in-place sweeps (the last level; 0 elsewhere). */
gs_status gs_amg_level_info(const gs_amg *amg, int64_t level, int64_t *rows, int64_t *nnz,
                            int64_t *coarse, int64_t *nnz_p, int64_t *sweep_levels);
/* Level `level`'s operator (which 0) or prolongation (which 1) in sprs shape
 * (host arrays: indptr rows+1, indices and data of its stored entries), the
 * reference's AMGLevel fields a and p; which 2: its Jacobi inverse diagonal
 * into data (diag_inv). */
gs_status gs_amg_export(gs_ctx *ctx, const gs_amg *amg, int64_t level, int32_t which,
                        int64_t *indptr, int64_t *indices, double *data);
void gs_amg_destroy(gs_amg *amg);

/* Solver::solve_with_operator for ConjugateGradient (src/linalg/solver.rs:
 * 112-120, src/linalg/iterative.rs:133-140 and the loop at :46-119): CG on a
 * caller's linear operator with a caller's preconditioner (the traits
 * LinearOperator, src/linalg/solver.rs:29-59, and Preconditioner,
 * src/linalg/preconditioner.rs:7-16).  The operator and the preconditioner
 * are callbacks on two caller-owned device buffers of n doubles: before
 * each call the library writes the vector to transform into in_dev and
 * waits for its own work; the callback writes its result into out_dev and
 * returns 0 once that device work is complete (non-zero aborts the solve
 * with GS_ERR_INVALID_ARG); a callback may call other entry points on the
 * same context.  precond NULL is the IdentityPreconditioner.
 * The loop, thresholds (1e-25, 1e-30) and the stopping tests are the
 * reference's; uses max_iterations, tolerance, abs_tolerance and
 * relative_breakdown of opts; precision must be GS_PREC_FP64 and
 * use_preconditioner is ignored (the callback decides).  b_dev and x_dev in
 * the operator's own numbering. */
typedef int32_t (*gs_operator_fn)(void *user);
gs_status gs_cg_solve_operator(gs_ctx *ctx, int64_t n, gs_operator_fn apply,
                               void *apply_user, gs_operator_fn precond,
                               void *precond_user, double *in_dev, double *out_dev,
                               const double *b_dev, const gs_cg_opts *opts,
                               double *x_dev, gs_solver_stats *stats);

/* ---- adjoint (new surface; no reference counterpart) ----
 * Element backward of theta_e = log k_e (thermal) or log E_e (elastic):
 * grad_e = -lambda_e^T K_e(theta_e) u_e, with u the full solution (Dirichlet
 * values included) and lambda the adjoint solution (zero on Dirichlet DOFs). */
gs_status gs_thermal_element_backward(gs_ctx *ctx, gs_mesh *mesh,
                                      const double *k_per_elem_dev,
                                      const double *u_dev,
                                      const double *lambda_dev,
                                      double *grad_dev);
gs_status gs_elastic_element_backward(gs_ctx *ctx, gs_mesh *mesh,
                                      const double *E_per_elem_dev,
                                      double poisson, const double *u_dev,
                                      const double *lambda_dev,
                                      double *grad_dev);

/* ---- SpMV launch shape (no reference counterpart) ----
 * Override the SpMV engine and launch shape for this matrix.  Default engine
 * (square matrices; the only one CG runs on): sliced ELL, SELL-64 -- a
 * wavefront per 64-row slice, a lane per row, rows length-sorted within
 * windows of 256, 16-bit column offsets where they fit; workgroups 0 = one
 * per 4 slices, else a persistent grid of that many (at most 16384).
 * GS_SPMV_INDEX32 keeps 32-bit column indices.  GS_SPMV_CSR selects the CSR
 * engines for gs_csr_spmv and gs_time_spmv (always used for rectangular
 * matrices): CSR-stream by default (lanes = threads per row in the row sums,
 * 1, 2, 4 or 8), GS_SPMV_VECTOR for lanes-per-row (4, 8, 16 or 32); lanes 0 =
 * chosen from the mean row length.  Operators with 3 DOFs per node (the
 * elasticity assemblies and their Dirichlet eliminations) use the 3x3-block
 * SELL view: slices of 64 node rows, one index per 3x3 block (a block is
 * stored when any of its entries is; the others as zeros).  GS_SPMV_SCALAR
 * keeps the scalar view for them; GS_SPMV_BLOCK3 asks for the block view for
 * any square matrix whose order is a multiple of 3.  Results depend on the
 * engine and shape only through floating-point summation order (each row
 * still sums its entries in ascending column order on both SELL views). */
enum {
  GS_SPMV_VECTOR = 4,
  GS_SPMV_INDEX32 = 8,
  GS_SPMV_CSR = 16,
  GS_SPMV_SCALAR = 32,
  GS_SPMV_BLOCK3 = 64
};
gs_status gs_csr_set_spmv_shape(gs_csr *A, int32_t lanes, int32_t workgroups,
                                int32_t flags);

/* ---- measurement hooks (no reference counterpart) ----
 * Launch the CG's SpMV kernel `reps` times on the ctx stream between two HIP
 * events and return the average duration per launch in milliseconds.
 * GS_PREC_FP64: the FP64 pass's SpMV (q = A p and p.q partials).
 * GS_PREC_MIXED: the FP32 recursion's SpMV -- on a scalar operator whose
 * Jacobi diagonal is positive, the two-launch pass's first kernel (w = A r on
 * the Jacobi-scaled values in int32 fixed point, FP32 vectors, then p and s
 * updated in place and p.s, r.p partials); otherwise the FP32 q = A p. */
gs_status gs_time_spmv(gs_ctx *ctx, const gs_csr *A, int32_t precision,
                       int32_t reps, double *avg_ms);
/* The device's achievable HBM rate: kind 0 streams `bytes` of FP64 values
 * once and sums them (16-byte loads, read-dominant), kind 1 copies them to a
 * second buffer; average milliseconds per pass over `reps` passes, timed with
 * HIP events on the ctx stream (the buffers come from the ctx's pool). */
gs_status gs_time_stream(gs_ctx *ctx, size_t bytes, int32_t kind, int32_t reps,
                         double *avg_ms);
/* Time the elastic element kernel (K_e and the gravity load of every
 * element; variant 0: VALU, the one assembly uses (one lane per 3x3 node
 * block); 1: matrix cores, v_mfma_f64_16x16x4f64; 2: FP32 B^T D B on the
 * VALU (FP32 K_e); 3: FP32 on the matrix cores, v_mfma_f32_16x16x4f32;
 * 4: FP64 VALU with one lane per K_e row) over `reps` launches, and report
 * the largest entry difference against variant 0 relative to the largest
 * entry. */
gs_status gs_time_elastic_elements(gs_ctx *ctx, gs_mesh *mesh, double E,
                                   double poisson, const double *E_per_elem_dev,
                                   int32_t variant, int32_t reps, double *avg_ms,
                                   double *max_rel_diff);

#ifdef __cplusplus
}
#endif
#endif
