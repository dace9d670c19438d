// Internal declarations shared by the libgeofem sources (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/geofem.h"

namespace gs {

// ---------------------------------------------------------------------------
// Errors: C++ exceptions inside the library, gs_status at the C ABI.
// ---------------------------------------------------------------------------
struct Error : std::runtime_error {
  gs_status code;
  Error(gs_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw Error(e == hipErrorOutOfMemory ? GS_ERR_OOM : GS_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}
#define GS_HIP(call) ::gs::hip_check((call), #call)
#define GS_LAUNCH_CHECK(name) ::gs::hip_check(hipGetLastError(), name)

// Shared-ownership device array (hipFree on last release).
template <class T> struct DArr {
  std::shared_ptr<T> p;
  int64_t n = 0;
  T *get() const { return p.get(); }
  explicit operator bool() const { return (bool)p; }
};

template <class T> DArr<T> dalloc(int64_t n) {
  T *ptr = nullptr;
  size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(T);
  hipError_t e = hipMalloc((void **)&ptr, bytes);
  if (e != hipSuccess)
    throw Error(GS_ERR_OOM, "hipMalloc(" + std::to_string(bytes) +
                                " bytes): " + hipGetErrorString(e));
  DArr<T> a;
  a.p = std::shared_ptr<T>(ptr, [](T *q) { (void)hipFree(q); });
  a.n = n;
  return a;
}

// Node-level CSR pattern of a Tet10 mesh plus the slot-sorted contributor
// list that drives the deterministic segmented-sum assembly.
struct NodePattern {
  int64_t nnz = 0;
  DArr<int> row_ptr;     // n_nodes + 1
  DArr<int> col;         // nnz, ascending within a row
  DArr<int> slot_row;    // nnz
  DArr<int> diag_slot;   // n_nodes (-1: node in no element)
  DArr<int> contrib_ptr; // nnz + 1
  DArr<int> contrib;     // 100 * n_elems triplet ids t = e*100 + a*10 + b,
                         // ascending t (hence ascending e) within a slot
};

// DOF-level (3 per node) pattern of the elastic operator, built from the
// node pattern by 3x3 expansion.
struct DofPattern3 {
  int64_t nnz = 0;
  DArr<int> row_ptr; // 3 n_nodes + 1
  DArr<int> col;     // nnz
};

}  // namespace gs

struct gs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void *pinned = nullptr;  // small pinned staging buffer (4 KiB)
  gs::DArr<char> ws;       // solver scratch, grows on demand
  size_t ws_bytes = 0;
};

struct gs_mesh {
  gs_ctx *ctx = nullptr;
  int64_t n_nodes = 0, n_elems = 0;
  gs::DArr<double> xyz; // n_nodes * 3 (AoS, as Point3<f64>), caller numbering
  gs::DArr<int> conn;   // n_elems * 10, caller numbering
  // Internal node numbering used by every device matrix built on this mesh
  // (a locality order, see gs_mesh_create): perm[new] = old, iperm[old] = new.
  gs::DArr<int> perm, iperm;
  gs::DArr<int> perm3;  // DOF-level perm for 3 DOFs per node (lazily)
  bool has_pattern = false;
  gs::NodePattern pat;
  bool has_dof3 = false;
  gs::DofPattern3 dof3;
};

struct gs_dofmap {
  gs_ctx *ctx = nullptr;
  int64_t n_nodes = 0;
  int dpn = 1;
  int64_t n_dofs = 0;
  int64_t n_fixed = 0;
  gs::DArr<uint8_t> is_dir; // n_dofs
  gs::DArr<double> values;  // n_dofs
};

struct gs_csr {
  gs_ctx *ctx = nullptr;
  int64_t rows = 0, cols = 0, nnz = 0;
  gs::DArr<int> row_ptr;
  gs::DArr<int> col;
  gs::DArr<double> val;
  // Row/column numbering: internal index i is caller index dperm[i] (empty:
  // identity).  Vectors cross the C ABI in caller numbering.
  gs::DArr<int> dperm;
  // lazily derived, cached (solver)
  int idx16 = 0;              // 16-bit column offsets: 0 unknown, 1 yes, -1 no
  gs::DArr<short> off16;      // nnz
  gs::DArr<int> anchor;       // one per 1024 nonzeros
  gs::DArr<float> val32;
  gs::DArr<double> dinv;
  gs::DArr<float> dinv32;
  // SpMV launch shape (0 = automatic); see gs_csr_set_spmv_shape
  int spmv_lanes = 0, spmv_grid = 0, spmv_flags = 0;
};

namespace gs {

// ---- elements.hip ----
void launch_thermal_elements(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, double k, const double *k_elem,
                             double Q, bool with_load, double *Ke, double *Fe,
                             int *flag);
void launch_elastic_elements(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, double Emod, double nu,
                             const double *E_elem, double rho, const double g[3],
                             bool with_load, double *Ke, double *Fe, int *flag);
void launch_thermal_backward(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, const double *k_elem,
                             const double *u, const double *lam, double *grad,
                             int *flag);
void launch_elastic_backward(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, const double *E_elem, double nu,
                             const double *u, const double *lam, double *grad,
                             int *flag);

// ---- assembly.hip ----
void build_node_pattern(gs_ctx *ctx, gs_mesh *m);
void build_dof3_pattern(gs_ctx *ctx, gs_mesh *m);
void gather_thermal(hipStream_t s, const NodePattern &p, int64_t E,
                    const double *Ke, double *val);
// Loads land at out_map[r] (caller numbering) when out_map is given, else at
// internal row r.
void gather_thermal_load(hipStream_t s, const NodePattern &p, int64_t n_nodes,
                         const double *Fe, const int *out_map, double *f);
void gather_elastic(hipStream_t s, const NodePattern &p, int64_t n_nodes,
                    const double *Ke, double *val);
void gather_gravity_load(hipStream_t s, const NodePattern &p, int64_t n_nodes,
                         const double *Fe, const int *out_map, double *f);
void build_perm3(gs_ctx *ctx, gs_mesh *m);
// Dirichlet elimination of a CSR matrix in its internal numbering (f, is_dir,
// g, fbc internal too); dperm (nullable) gives each column's caller index,
// which fixes the order of the RHS lift.  Returns the new CSR.
gs_csr *dirichlet_eliminate(gs_ctx *ctx, const gs_csr *K, const double *f,
                            const uint8_t *is_dir, const double *g,
                            double *fbc);
// dst[i] = src[map[i]]  /  dst[map[i]] = src[i]
void permute_gather(hipStream_t s, int64_t n, const double *src, const int *map,
                    double *dst);
void permute_gather_u8(hipStream_t s, int64_t n, const uint8_t *src,
                       const int *map, uint8_t *dst);
void permute_scatter(hipStream_t s, int64_t n, const double *src, const int *map,
                     double *dst);

// ---- solver.hip ----
void jacobi_diag_inv(hipStream_t s, const gs_csr *A, double *dinv);
void csr_spmv(gs_ctx *ctx, const gs_csr *A, const double *x, double *y);
void cg_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
              const gs_cg_opts &o, double *x, gs_solver_stats *st);
double time_spmv(gs_ctx *ctx, const gs_csr *A, int precision, int reps);

int read_flag_and_clear(gs_ctx *ctx, int *flag_dev);

}  // namespace gs
