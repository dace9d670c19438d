// Internal declarations shared by the libgeofem sources (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

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

// Device blocks kept for reuse by exact byte size (one pool per context), so
// repeated assemblies and solves of one problem stop calling hipMalloc.  A
// block returns here when its last DArr releases it; the bytes cached are
// capped, and the rest is freed.
struct Pool {
  std::mutex m;
  std::multimap<size_t, void *> free;
  size_t cached = 0;
  size_t cap = (size_t)16 << 30;
  ~Pool() {
    for (auto &kv : free) (void)hipFree(kv.second);
  }
};

template <class T> DArr<T> palloc(const std::shared_ptr<Pool> &pool, int64_t n) {
  const size_t bytes = (size_t)(n > 0 ? n : 1) * sizeof(T);
  void *ptr = nullptr;
  {
    std::lock_guard<std::mutex> g(pool->m);
    auto it = pool->free.find(bytes);
    if (it != pool->free.end()) {
      ptr = it->second;
      pool->free.erase(it);
      pool->cached -= bytes;
    }
  }
  if (!ptr) {
    hipError_t e = hipMalloc(&ptr, bytes);
    if (e != hipSuccess)
      throw Error(GS_ERR_OOM, "hipMalloc(" + std::to_string(bytes) +
                                  " bytes): " + hipGetErrorString(e));
  }
  std::weak_ptr<Pool> wp = pool;
  DArr<T> a;
  a.p = std::shared_ptr<T>((T *)ptr, [wp, bytes](T *q) {
    if (auto pl = wp.lock()) {
      std::lock_guard<std::mutex> g(pl->m);
      if (pl->cached + bytes <= pl->cap) {
        pl->free.emplace(bytes, (void *)q);
        pl->cached += bytes;
        return;
      }
    }
    (void)hipFree(q);
  });
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

// SELL-64 view of a square matrix (layout in sell.hpp): the solver's operator
// format.  SELL row s is CSR row sperm[s] and caller index vperm[s]; columns
// are SELL rows.  Shared by every matrix with the same pattern and numbering.
// bs = 3: the block variant for 3-DOF-per-node operators -- slices of 64
// block rows (3 DOF rows each), one 3x3 block per lane per chunk, one index
// per block; DOF row 3s+c of the SELL order is block row s, component c.
struct SellPattern {
  int bs = 1;                      // 1: scalar SELL-64; 3: 3x3-block SELL-64
  int64_t n = 0, nnz = 0, nslices = 0, nchunks = 0;
  int64_t entries = 0;             // index slots (scalar: also the value count)
  int64_t nvals = 0;               // values (bs = 3: 9 per index slot)
  bool idx16 = true;
  DArr<int> sperm, siperm, vperm;  // n (DOF rows)
  DArr<int> cb;                    // nslices + 1 chunk offsets
  DArr<short> off;                 // entries (16-bit variant)
  DArr<int> col;                   // entries (32-bit variant)
  DArr<int> pos;                   // nnz: CSR nonzero -> SELL value
};

// Everything about a Dirichlet-eliminated system that depends only on the
// mesh and the DOF map, built once (cached on the mesh): the K_bc pattern in
// the internal numbering, its SELL view, and for every DOF slot of the full K
// where its value goes -- a K_bc nonzero, an entry of the RHS lift
// f_j -= K_ji g_i, or nowhere.  The numeric phase then writes K_bc directly.
struct SystemPlan {
  uint64_t dof_uid = 0;
  int dpn = 1;
  int64_t n = 0, nnz = 0, n_lift = 0;
  DArr<int> row_ptr, col;            // K_bc pattern, internal numbering
  std::shared_ptr<SellPattern> sell; // SELL view of that pattern
  DArr<int> dst;      // full-K DOF slot -> K_bc nonzero (>= 0), lift entry
                      // q (as -2 - q), or -1
  DArr<int> lift_ptr; // n + 1: row r's lift entries, ascending caller column
  DArr<int> lift_col; // n_lift: internal column of each lift entry
  DArr<int> diag;     // n: K_bc nonzero holding row r's diagonal (-1: none)
  DArr<uint8_t> dir;  // n, internal numbering
  DArr<double> g;     // n, internal numbering
  // 3 DOFs per node: the node slot of each block position of the block view
  // (-1: padding), and the node slots with lift destinations
  DArr<int> bslot;
  DArr<unsigned short> bmask;  // kept entries of each block position (bit m)
  // thermal: incidence slots of the row gather.  The incidences (e, a) of
  // node r = conn[e][a], in ascending element order, take the slots
  // (ib[s] + j) * 64 + l, j = 0 .. icnt[64 s + l] - 1, of the node's SELL
  // row 64 s + l: slice s owns ib[s+1] - ib[s] steps of 64 slots, so a
  // wavefront's loads of step j are contiguous.  incpos[e*10 + a] is the
  // slot of (e, a); icode[slot*10 + b] names the accumulator of triplet
  // (e, a, b) (1-byte codes when every row has at most 255 entries, else
  // 2-byte).
  DArr<int> incpos, ib, icnt;
  int64_t n_islots = 0;
  // the row gather's slices, those whose rows and lifts fit the short pass
  // (ROW_LW_SHORT accumulators) first: n_tshort of them
  DArr<int> tslices;
  int64_t n_tshort = 0;
  DArr<uint8_t> icode8;
  DArr<unsigned short> icode16;
  DArr<int> lift_slots;
  int64_t n_lift_slots = 0;
};

}  // namespace gs

struct gs_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void *pinned = nullptr;  // small pinned staging buffer (4 KiB)
  hipEvent_t snap_ev = nullptr;  // marks a solver state snapshot (created with the ctx)
  gs::DArr<char> ws;       // solver scratch, grows on demand
  size_t ws_bytes = 0;
  std::shared_ptr<gs::Pool> pool = std::make_shared<gs::Pool>();
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
  std::vector<std::shared_ptr<gs::SystemPlan>> plans;  // most recent last
};

struct gs_dofmap {
  gs_ctx *ctx = nullptr;
  uint64_t uid = 0;  // unique per created dofmap (keys the system plans)
  int64_t n_nodes = 0;
  int dpn = 1;
  int64_t n_dofs = 0;
  int64_t n_fixed = 0;
  gs::DArr<uint8_t> is_dir; // n_dofs
  gs::DArr<double> values;  // n_dofs
};

// ILU(0) factors of a square matrix in the caller's numbering (ilu.hip):
// pattern, diagonal positions, factors, and the rows of each dependency
// level of the forward (lower) and backward (upper) substitutions.
struct gs_ilu {
  gs_ctx *ctx = nullptr;
  int64_t n = 0, nnz = 0;
  gs::DArr<int> rp, col, diag;
  gs::DArr<double> f;
  gs::DArr<int> ord_lower, ord_upper;
  std::vector<int> lvl_lower, lvl_upper;  // level offsets into ord_*
};

// AMG hierarchy (amg.hip): per level the operator, prolongation and its
// transpose in CSR (caller numbering on level 0), the Jacobi inverse
// diagonal and the V-cycle's vectors; the last level's sweep order.
namespace gs {
struct DevCsr {
  int64_t rows = 0, nnz = 0;
  DArr<int> rp, col;
  DArr<double> val;
  // the same rows sliced for the V-cycle's kernels: slices of 64 rows, entry
  // k of the row at position 64 s + l at (ecb[s] + k) * 64 + l, padded with
  // column -1; erow: the row at each position (rows sorted by length, when
  // that pads less), empty for position = row
  DArr<int> ecb, ecol, erow;
  DArr<double> eval;
  int ew = 4;  // entries per trip of the row sums (8 where rows are long)
  int64_t n_long = 0;  // positions [0, n_long): rows summed a wavefront each
};
struct AmgLevel {
  DevCsr A, P, PT;      // P and PT empty where the level has none
  bool has_p = false;
  bool finite = true;   // every operator entry finite
  int64_t coarse = 0;   // columns of P
  DArr<double> dinv, x, t, r;
  // last level: the rows that read no other row and that no other row reads
  // (swept on their own), and the dependency levels of the others' sweeps
  int64_t n_iso = 0;
  DArr<int> iso;
  int nlev = 0;
  DArr<int> lptr, ord;
  // the coupled rows staged in LDS (sweep_lds bytes, 0: not staged): global
  // row of each local row, local CSR and level order
  size_t sweep_lds = 0;
  int64_t lnnz = 0;
  DArr<int> grow, lrp, lcol, lord;
  DArr<double> lval;
};
}  // namespace gs

struct gs_amg {
  gs_ctx *ctx = nullptr;
  int64_t n = 0;
  std::vector<gs::AmgLevel> lv;
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
  // The CSR values of a fused system (gs_assemble_*_system) are produced from
  // its SELL values on first use (ensure_csr_values); the solvers never need
  // them.
  bool val_lazy = false;
  // lazily derived, cached (solver)
  int idx16 = 0;              // 16-bit CSR column offsets: 0 unknown, 1 yes, -1 no
  gs::DArr<short> off16;      // nnz
  gs::DArr<int> anchor;       // one per 1024 nonzeros
  gs::DArr<double> dinv;      // CSR order
  gs::DArr<float> val32;      // CSR order, FP32 (CSR engine timing only)
  std::shared_ptr<gs::SellPattern> sell;  // SELL view (square matrices)
  bool sell_force32 = false;              // the view was built with int32 forced
  int block = 1;  // 3: rows and columns come in node triples (elasticity)
  gs::DArr<double> sval;                  // SELL values, FP64
  gs::DArr<float> sval32;                 // SELL values, FP32 (mixed precision)
  gs::DArr<double> sdinv;                 // Jacobi inverse diagonal, SELL order
  gs::DArr<float> sdinv32;
  // mixed solve: S = sqrt(D^-1) (SELL order) and fixed-point S A S; sscaled 0
  // unknown, 2 built but not yet checked, 1 usable, -1 not applicable (a
  // non-positive D^-1 entry, flagged in sscale_bad)
  int sscaled = 0;
  gs::DArr<double> sscale;
  gs::DArr<int> svalq;  // S A S in int32 fixed point (value * 2^30)
  gs::DArr<int> sscale_bad;
  // SpMV launch shape (0 = automatic); see gs_csr_set_spmv_shape
  int spmv_lanes = 0, spmv_grid = 0, spmv_flags = 0;
};

namespace gs {

// ---- elements.hip ----
// ElementMatrix::thermal_stiffness with the rule of npts points (1, 4, 5, 11),
// all 100 entries per element
void launch_thermal_rule(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                         double k, const double *k_elem, int npts, double *Ke, int *flag);
// GaussQuadrature::tet_{npts}point tables (false for another npts)
bool quad_rule(int npts, double *points, double *weights);
void launch_thermal_elements(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, double k, const double *k_elem,
                             double Q, bool with_load, double *Ke, double *Fe,
                             int *flag, const int *incpos = nullptr);
// The elastic element family: K_e = sum_q w B^T D B for an isotropic Voigt
// law (elastic or Newtonian viscous) and an optional element load.
enum : int { GS_MAT_ELASTIC = 0, GS_MAT_VISCOUS = 1 };
enum : int { GS_LOAD_NONE = 0, GS_LOAD_BODY = 1, GS_LOAD_STRESS = 2 };
struct ElasticKind {
  int mat = GS_MAT_ELASTIC;
  double p0 = 0.0, p1 = 0.0;      // elastic: E, nu; viscous: mu, zeta
  const double *e0 = nullptr;     // per-element E or mu (nullable, device)
  const double *e1 = nullptr;     // per-element zeta (viscous, nullable)
  int load = GS_LOAD_NONE;
  double bf[3] = {0.0, 0.0, 0.0}; // body force rho g (GS_LOAD_BODY)
  const double *sig = nullptr;    // per-element stress [E][6] (GS_LOAD_STRESS)
};
void launch_elastic_family(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                           const ElasticKind &k, double *Ke, double *Fe, int *flag);
void launch_maxwell_stress(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                           double Emod, double nu, double visc, double dt, const double *u_n,
                           const double *u_next, const double *sig_in, double *sig_out,
                           int *flag);
void launch_elastic_elements(hipStream_t s, int64_t E, const double *xyz,
                             const int *conn, double Emod, double nu,
                             const double *E_elem, double rho, const double g[3],
                             bool with_load, double *Ke, double *Fe, int *flag);
void launch_face_flux(hipStream_t s, int64_t F, const double *xyz, const int *conn,
                      const int64_t *elem, const int *lface, const double *flux,
                      double *contrib, int *node);
void launch_thermal_mass(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                         double rho_cp, const double *rc_elem, double *Me);
// Matrix-core variant of the elastic element kernel (timing; entries equal
// the VALU kernel's up to rounding).
void launch_elastic_elements_mfma(hipStream_t s, int64_t E, const double *xyz,
                                  const int *conn, double Emod, double nu,
                                  const double *E_elem, double rho, const double g[3],
                                  bool with_load, double *Ke, double *Fe, int *flag);
// out[0] = bits of max |a - b|, out[1] = bits of max |b| (out preset by caller)
void launch_maxdiff(hipStream_t s, int64_t n, const double *a, const double *b,
                    unsigned long long *out);
void launch_maxdiff_f32(hipStream_t s, int64_t n, const float *a, const double *b,
                        unsigned long long *out);
void launch_elastic_family_rows(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                                const ElasticKind &k, double *Ke, double *Fe, int *flag);
void launch_elastic_family_f32(hipStream_t s, int64_t E, const double *xyz, const int *conn,
                               const ElasticKind &k, float *Ke, float *Fe, int *flag);
void launch_elastic_elements_mfma_f32(hipStream_t s, int64_t E, const double *xyz,
                                      const int *conn, double Emod, double nu,
                                      const double *E_elem, double rho, const double g3[3],
                                      bool with_load, float *Ke, float *Fe, int *flag);
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
// f += Neumann face fluxes (faces given on the device), deterministic order
void apply_neumann(gs_ctx *ctx, const gs_mesh *m, int64_t F, const int64_t *elem,
                   const int *lface, const double *flux, double *f);
// all-100-entry element matrices (thermal mass)
void gather_thermal_full(hipStream_t s, const NodePattern &p, const double *Me,
                         double *val);
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

// System plans and the fused numeric phase (assembly.hip).
std::shared_ptr<SystemPlan> system_plan(gs_ctx *ctx, gs_mesh *m,
                                        const gs_dofmap *d);
// workgroups of the system gather (entries of its diagonal-maximum array)
int64_t system_gather_blocks(const gs_mesh *m, const SystemPlan &P);
// values of K_bc (SELL, FP64 and FP32), the lift values, and per DOF the
// gathered load (when with_load) and diagonal.  Thermal: Ke holds the
// incidence rows of k_thermal_elements<.., ROWS = true> (the load inside, Fe
// unused, one row per incidence slot of the plan); elastic: the block-major
// K_e and Fe.
void system_gather(hipStream_t s, const gs_mesh *m, const SystemPlan &P, const double *Ke,
                   const double *Fe, bool with_load, double *sval, float *sval32,
                   double *liftv, double *fload, double *dval, unsigned long long *dmax);
void system_finish(gs_ctx *ctx, const gs_mesh *m, const SystemPlan &P, const double *fload,
                   const double *dval, const double *fext, double *sval, float *sval32,
                   const double *liftv, const int *dperm, double *fbc, double *sdinv,
                   float *sdinv32, const unsigned long long *dmax);

// Host <-> device copies on the context's stream, completed before return.
template <class T> DArr<T> to_device(gs_ctx *ctx, const std::vector<T> &h) {
  auto d = dalloc<T>((int64_t)h.size());
  if (!h.empty())
    GS_HIP(hipMemcpyAsync(d.get(), h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice,
                          ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  return d;
}
template <class T> std::vector<T> to_host(gs_ctx *ctx, const T *d, int64_t n) {
  std::vector<T> h((size_t)n);
  if (n > 0)
    GS_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

// ---- ilu.hip ----
// A in caller numbering with sorted columns: row pointers, columns, values
// and, when A is renumbered, the source nonzero of each entry.
struct CallerCsr {
  DArr<int> rp, col, src;
  DArr<double> val;
};
CallerCsr caller_csr(gs_ctx *ctx, const gs_csr *A);
gs_ilu *ilu_create(gs_ctx *ctx, const gs_csr *A);
void ilu_update(gs_ilu *M, const gs_csr *A);
void ilu_apply(const gs_ilu *M, const double *r, double *z);

// ---- amg.hip ----
gs_amg *amg_create(gs_ctx *ctx, const gs_csr *A, int64_t max_levels, int64_t coarse_size,
                   double theta);
void amg_apply(const gs_amg *M, const double *r, double *z);

// ---- operator.hip ----
// block = 3 builds the 3x3-block view (rows must be a multiple of 3; a block
// is stored when any of its 9 entries is, the others as zeros).
std::shared_ptr<SellPattern> build_sell_pattern(gs_ctx *ctx, const gs_csr *A,
                                                bool force32, int block = 1);
// zero values of a block view that no CSR nonzero writes (bs = 1: nothing)
void sell_zero_holes(hipStream_t s, const SellPattern &P, double *sval, float *sval32);
// A->val from A->sval through the SELL pattern's pos, when the values are lazy
void ensure_csr_values(gs_ctx *ctx, const gs_csr *A);
// sval[pos[k]] = val[k], padding zero
template <typename T>
void sell_values(hipStream_t s, const SellPattern &P, const double *val, T *sval);
void sell_jacobi(hipStream_t s, const SellPattern &P, const gs_csr *A, double *d64,
                 float *d32);

// ---- solver.hip ----
void jacobi_diag_inv(hipStream_t s, const gs_csr *A, double *dinv);
void csr_spmv(gs_ctx *ctx, const gs_csr *A, const double *x, double *y);
void cg_solve_operator(gs_ctx *ctx, int64_t n, gs_operator_fn apply, void *apply_user,
                       gs_operator_fn precond, void *precond_user, double *in,
                       double *out, const double *b, const gs_cg_opts &o, double *x,
                       gs_solver_stats *st);
double time_stream(gs_ctx *ctx, size_t bytes, int kind, int reps);
void noalloc_setup(int device);
void cg_solve_noalloc(gs_ctx *ctx, const gs_csr *A, const double *b, const gs_cg_opts &o,
                      double *x, gs_solver_stats *st);
void cg_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
              const gs_cg_opts &o, double *x, gs_solver_stats *st);
void bicgstab_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
                    const gs_cg_opts &o, double *x, gs_solver_stats *st);
void gmres_solve(gs_ctx *ctx, const gs_csr *A, const double *b, const gs_cg_opts &o,
                 int restart, double *x, gs_solver_stats *st);
double time_spmv(gs_ctx *ctx, const gs_csr *A, int precision, int reps);

int read_flag_and_clear(gs_ctx *ctx, int *flag_dev);

}  // namespace gs
