// C ABI of libgeofem (include/geofem.h): handle management, argument
// validation, and the assemble/solve pipelines built from the device kernels.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.hpp"

using namespace gs;

namespace {

template <class F> gs_status guard(gs_ctx *ctx, F &&f) {
  try {
    f();
    // every entry point returns after its work is complete (the header's
    // contract): callers read the outputs on their own streams
    if (ctx && ctx->stream) GS_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx) ctx->err.clear();
    return GS_OK;
  } catch (const Error &e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    if (ctx) ctx->err = "host out of memory";
    return GS_ERR_OOM;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return GS_ERR_INVALID_ARG;
  }
}

void require(bool ok, gs_status code, const char *msg) {
  if (!ok) throw Error(code, msg);
}

void set_device(gs_ctx *ctx) { GS_HIP(hipSetDevice(ctx->device)); }

// Element-kernel failure flags -> reference panic conditions as statuses.
void check_flags(int flags, int mat = GS_MAT_ELASTIC) {
  if (flags & 4)
    throw Error(GS_ERR_INVALID_MATERIAL, mat == GS_MAT_VISCOUS ? "Viscosity must be positive"
                                                                 : "Young's modulus must be positive");
  if (flags & 1) throw Error(GS_ERR_SINGULAR_JACOBIAN, "Singular Jacobian");
  if (flags & 2)
    throw Error(GS_ERR_NONFINITE, "non-finite Jacobian determinant");
}

void check_material(double E, double nu) {
  // src/mechanics/constitutive.rs:30-41
  require(E > 0.0, GS_ERR_INVALID_MATERIAL, "Young's modulus must be positive");
  require(nu > -1.0 && nu < 0.5, GS_ERR_INVALID_MATERIAL,
          "Poisson's ratio must be in (-1, 0.5) for physical stability");
}

gs_csr *new_csr(gs_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz,
                DArr<int> rp, DArr<int> col, DArr<double> val) {
  auto *c = new gs_csr();
  c->ctx = ctx;
  c->rows = rows;
  c->cols = cols;
  c->nnz = nnz;
  c->row_ptr = std::move(rp);
  c->col = std::move(col);
  c->val = std::move(val);
  return c;
}

// thermal K (full structural pattern) [+ element loads Fe]
gs_csr *assemble_thermal_K(gs_ctx *ctx, gs_mesh *m, double k,
                           const double *k_elem, bool with_load, double Q,
                           double *f_dev) {
  build_node_pattern(ctx, m);
  const NodePattern &p = m->pat;
  hipStream_t s = ctx->stream;
  auto Ke = dalloc<double>(m->n_elems * 55);
  DArr<double> Fe;
  if (with_load) Fe = dalloc<double>(m->n_elems * 10);
  auto flag = dalloc<int>(1);
  GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
  launch_thermal_elements(s, m->n_elems, m->xyz.get(), m->conn.get(), k, k_elem,
                          Q, with_load, Ke.get(), with_load ? Fe.get() : nullptr,
                          flag.get());
  check_flags(read_flag_and_clear(ctx, flag.get()));
  auto val = dalloc<double>(p.nnz);
  gather_thermal(s, p, m->n_elems, Ke.get(), val.get());
  if (with_load) gather_thermal_load(s, p, m->n_nodes, Fe.get(), nullptr, f_dev);
  GS_HIP(hipStreamSynchronize(s));
  gs_csr *K = new_csr(ctx, m->n_nodes, m->n_nodes, p.nnz, p.row_ptr, p.col, val);
  K->dperm = m->perm;
  return K;
}

// K of an elastic-family law (full structural pattern, DOF 3 node + c) and,
// when the kind has a load, its assembled load in caller numbering (f_dev).
gs_csr *assemble_dof3_K(gs_ctx *ctx, gs_mesh *m, const ElasticKind &kind, double *f_dev) {
  build_dof3_pattern(ctx, m);
  const NodePattern &p = m->pat;
  hipStream_t s = ctx->stream;
  const bool with_load = kind.load != GS_LOAD_NONE;
  auto Ke = dalloc<double>(m->n_elems * 900);
  DArr<double> Fe;
  if (with_load) Fe = dalloc<double>(m->n_elems * 30);
  auto flag = dalloc<int>(1);
  GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
  launch_elastic_family(s, m->n_elems, m->xyz.get(), m->conn.get(), kind, Ke.get(),
                        with_load ? Fe.get() : nullptr, flag.get());
  check_flags(read_flag_and_clear(ctx, flag.get()), kind.mat);
  auto val = dalloc<double>(m->dof3.nnz);
  gather_elastic(s, p, m->n_nodes, Ke.get(), val.get());
  if (with_load) gather_gravity_load(s, p, m->n_nodes, Fe.get(), m->perm.get(), f_dev);
  build_perm3(ctx, m);
  GS_HIP(hipStreamSynchronize(s));
  gs_csr *K = new_csr(ctx, 3 * m->n_nodes, 3 * m->n_nodes, m->dof3.nnz,
                      m->dof3.row_ptr, m->dof3.col, val);
  K->dperm = m->perm3;
  K->block = 3;
  return K;
}

struct SystemParams {
  bool elastic = false;  // 3 DOFs per node: the elastic family of `ek`
  double k = 0.0, Q = 0.0;
  const double *k_elem = nullptr;
  ElasticKind ek;
  const double *fext = nullptr;  // added to the assembled load (caller numbering)
};

// Element kernel + the cached system plan's numeric phase: K_bc (CSR and its
// SELL view, FP64 and FP32), the Jacobi inverse diagonal and f_bc (caller
// numbering) without materialising the full K.  Results are bit-identical to
// assembling K, the load, and then eliminating.
gs_csr *assemble_system(gs_ctx *ctx, gs_mesh *m, const gs_dofmap *dof,
                        const SystemParams &sp, double *f_bc) {
  auto P = system_plan(ctx, m, dof);
  hipStream_t s = ctx->stream;
  const auto &pool = ctx->pool;
  const int64_t E = m->n_elems;
  const int dpn = sp.elastic ? 3 : 1;
  const bool with_load = !sp.elastic || sp.ek.load != GS_LOAD_NONE;
  // thermal: one 12-double row per incidence slot of the plan, the load inside
  auto Ke = palloc<double>(pool, sp.elastic ? E * 900 : P->n_islots * 12);
  DArr<double> Fe;
  if (with_load && sp.elastic) Fe = palloc<double>(pool, E * 10 * dpn);
  auto flag = palloc<int>(pool, 1);
  GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
  if (sp.elastic)
    launch_elastic_family(s, E, m->xyz.get(), m->conn.get(), sp.ek, Ke.get(), Fe.get(),
                          flag.get());
  else
    launch_thermal_elements(s, E, m->xyz.get(), m->conn.get(), sp.k, sp.k_elem, sp.Q,
                            true, Ke.get(), nullptr, flag.get(), P->incpos.get());
  const SellPattern &S = *P->sell;
  auto sval = palloc<double>(pool, S.nvals);
  // FP32 values: the elastic solves' mixed recursion reads them, so the
  // elastic gather writes them; a thermal solve reads its scaled fixed-point
  // values instead, and the FP32 ones are made when a path needs them
  DArr<float> sval32;
  if (sp.elastic) sval32 = palloc<float>(pool, S.nvals);  // every entry written by the gather
  auto liftv = palloc<double>(pool, P->n_lift);
  auto sdinv = palloc<double>(pool, P->n);
  auto sdinv32 = palloc<float>(pool, P->n);
  const DArr<int> &dperm = sp.elastic ? m->perm3 : m->perm;
  auto dmax = palloc<unsigned long long>(pool, std::max<int64_t>(1, system_gather_blocks(m, *P)));
  auto fload = palloc<double>(pool, P->n);  // per DOF: the gathered load
  auto dval = palloc<double>(pool, P->n);   // per DOF: the gathered diagonal
  // K_bc's CSR values are not written here: ensure_csr_values produces them
  // from the SELL values when a caller asks for the CSR form
  system_gather(s, m, *P, Ke.get(), Fe.get(), with_load, sval.get(), sval32.get(),
                liftv.get(), fload.get(), dval.get(), dmax.get());
  system_finish(ctx, m, *P, with_load ? fload.get() : nullptr, dval.get(), sp.fext,
                sval.get(), sval32.get(), liftv.get(), dperm.get(), f_bc, sdinv.get(),
                sdinv32.get(), dmax.get());
  // the element kernels' status, read once the stream has drained: the
  // gathers index by the plan alone, so values from a rejected element
  // reach only buffers that are discarded here
  check_flags(read_flag_and_clear(ctx, flag.get()), sp.elastic ? sp.ek.mat : GS_MAT_ELASTIC);
  gs_csr *K = new_csr(ctx, P->n, P->n, P->nnz, P->row_ptr, P->col, DArr<double>());
  K->val_lazy = true;
  K->dperm = dperm;
  K->block = dpn;
  K->sell = P->sell;
  K->sval = sval;
  K->sval32 = sval32;
  K->sdinv = sdinv;
  K->sdinv32 = sdinv32;
  return K;
}

void check_mesh_dof(const gs_mesh *m, const gs_dofmap *d) {
  require(m && d, GS_ERR_INVALID_ARG, "null mesh or dofmap");
  require(d->n_nodes == m->n_nodes, GS_ERR_INVALID_ARG,
          "dofmap node count differs from the mesh");
}

// Dirichlet elimination with the caller's numbering at the boundary: f is in
// caller numbering when f_caller, else already internal; f_bc is returned in
// caller numbering.  Masks and values are gathered into K's numbering.
gs_csr *eliminate_caller_rhs(gs_ctx *ctx, const gs_csr *K, const double *f,
                             bool f_caller, const gs_dofmap *dof, double *f_bc) {
  hipStream_t s = ctx->stream;
  const int *dp = K->dperm.get();
  if (!dp) {
    return dirichlet_eliminate(ctx, K, f, dof->is_dir.get(), dof->values.get(),
                               f_bc);
  }
  const int64_t n = K->rows;
  auto dir = dalloc<uint8_t>(n);
  auto g = dalloc<double>(n);
  auto fi = dalloc<double>(n);
  auto fo = dalloc<double>(n);
  permute_gather_u8(s, n, dof->is_dir.get(), dp, dir.get());
  permute_gather(s, n, dof->values.get(), dp, g.get());
  const double *fin = f;
  if (f_caller) {
    permute_gather(s, n, f, dp, fi.get());
    fin = fi.get();
  }
  gs_csr *out = dirichlet_eliminate(ctx, K, fin, dir.get(), g.get(), fo.get());
  permute_scatter(s, n, fo.get(), dp, f_bc);
  GS_HIP(hipStreamSynchronize(s));
  return out;
}

// Locality order of the mesh nodes: lexicographic in the coordinates, the
// longest bounding-box axis slowest, so a row's neighbours sit within a few
// coordinate layers and a window of consecutive rows gathers from a small,
// cache-resident part of the vector.  Ties keep the caller order.  Any
// non-finite coordinate leaves the identity order.
std::vector<int> locality_order(const double *xyz, int64_t n) {
  std::vector<int> order((size_t)n);
  for (int64_t i = 0; i < n; ++i) order[i] = (int)i;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int64_t i = 0; i < 3 * n; ++i)
    if (!std::isfinite(xyz[i])) return order;
  for (int c = 0; c < 3 && n > 0; ++c) {
    lo[c] = hi[c] = xyz[c];
    for (int64_t i = 0; i < n; ++i) {
      lo[c] = std::min(lo[c], xyz[3 * i + c]);
      hi[c] = std::max(hi[c], xyz[3 * i + c]);
    }
  }
  int ax[3] = {0, 1, 2};
  std::stable_sort(ax, ax + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    for (int k = 0; k < 3; ++k) {
      const double va = xyz[3 * (int64_t)a + ax[k]], vb = xyz[3 * (int64_t)b + ax[k]];
      if (va < vb) return true;
      if (va > vb) return false;
    }
    return a < b;
  });
  return order;
}

// ---- host mesh generator (src/mesh_generator_improved.rs:19-169) ----------
struct EdgeTable {
  std::vector<uint64_t> keys;
  std::vector<int64_t> vals;
  uint64_t mask = 0;
  explicit EdgeTable(int64_t expect) {
    uint64_t cap = 16;
    while (cap < (uint64_t)(2 * expect)) cap <<= 1;
    keys.assign(cap, 0);
    vals.assign(cap, -1);
    mask = cap - 1;
  }
  static uint64_t hash(uint64_t x) {
    x ^= x >> 31;
    x *= 0x7fb5d329728ea185ULL;
    x ^= x >> 27;
    x *= 0x81dadef4bc2dd44dULL;
    x ^= x >> 33;
    return x;
  }
  // index of the midpoint node of edge (a,b), created at `next` when new
  int64_t get_or_add(int64_t a, int64_t b, int64_t next, bool &added) {
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    const uint64_t key = (((uint64_t)lo << 32) | (uint64_t)hi) + 1;
    uint64_t h = hash(key) & mask;
    while (keys[h] != 0) {
      if (keys[h] == key) { added = false; return vals[h]; }
      h = (h + 1) & mask;
    }
    keys[h] = key;
    vals[h] = next;
    added = true;
    return next;
  }
};

void check_maxwell(double E, double nu, double visc) {
  // MaxwellViscoelasticity::new (src/mechanics/constitutive.rs:312-325)
  check_material(E, nu);
  require(visc > 0.0, GS_ERR_INVALID_MATERIAL, "Viscosity must be positive");
}

ElasticKind viscous_kind(double mu, double zeta, const double *mu_elem,
                         const double *zeta_elem) {
  // NewtonianViscosity::new (src/mechanics/constitutive.rs:179-186); a
  // per-element viscosity is checked by the element kernel
  if (!mu_elem) require(mu > 0.0, GS_ERR_INVALID_MATERIAL, "Viscosity must be positive");
  ElasticKind k;
  k.mat = GS_MAT_VISCOUS;
  k.p0 = mu;
  k.p1 = zeta;
  k.e0 = mu_elem;
  k.e1 = zeta_elem;
  return k;
}

ElasticKind maxwell_kind(double E, double nu, const double *stress) {
  ElasticKind k;
  k.p0 = E;
  k.p1 = nu;
  k.load = GS_LOAD_STRESS;
  k.sig = stress;
  return k;
}

}  // namespace

extern "C" {

namespace {
// message of the last failed gs_ctx_create on this thread (there is no
// context to hold it)
thread_local std::string create_error;
}  // namespace

gs_status gs_ctx_create(int device, gs_ctx **out) {
  if (!out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  auto *ctx = new gs_ctx();
  ctx->device = device;
  const gs_status st = guard(ctx, [&] {
    int count = 0;
    GS_HIP(hipGetDeviceCount(&count));
    require(device >= 0 && device < count, GS_ERR_INVALID_ARG,
            "device index out of range");
    GS_HIP(hipSetDevice(device));
    GS_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    GS_HIP(hipHostMalloc(&ctx->pinned, 4096, hipHostMallocDefault));
    GS_HIP(hipEventCreateWithFlags(&ctx->snap_ev, hipEventDisableTiming));
    noalloc_setup(device);  // the no-allocation CG's scratch address on this device
  });
  if (st != GS_OK) {
    create_error = "gs_ctx_create: " + ctx->err;
    gs_ctx_destroy(ctx);
    return st;
  }
  create_error.clear();
  *out = ctx;
  return GS_OK;
}

void gs_ctx_destroy(gs_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->ws = DArr<char>();
  ctx->pool.reset();  // frees the cached blocks; live arrays free their own
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->snap_ev) (void)hipEventDestroy(ctx->snap_ev);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *gs_last_error(const gs_ctx *ctx) {
  return ctx ? ctx->err.c_str() : create_error.c_str();
}

gs_status gs_ctx_synchronize(gs_ctx *ctx) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] { GS_HIP(hipStreamSynchronize(ctx->stream)); });
}

gs_status gs_device_alloc(gs_ctx *ctx, size_t bytes, void **out_dev) {
  if (!ctx || !out_dev) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    set_device(ctx);
    GS_HIP(hipMalloc(out_dev, bytes ? bytes : 1));
  });
}

gs_status gs_device_free(gs_ctx *ctx, void *p) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] { GS_HIP(hipFree(p)); });
}

gs_status gs_memcpy_h2d(gs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    GS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

gs_status gs_memcpy_d2h(gs_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    GS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    GS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

gs_status gs_generate_cube_sizes(int64_t nx, int64_t ny, int64_t nz,
                                 int64_t *n_nodes, int64_t *n_elems) {
  if (nx <= 0 || ny <= 0 || nz <= 0 || !n_nodes || !n_elems)
    return GS_ERR_INVALID_ARG;
  *n_nodes = (2 * nx + 1) * (2 * ny + 1) * (2 * nz + 1);
  *n_elems = 6 * nx * ny * nz;
  return GS_OK;
}

gs_status gs_generate_cube(int64_t nx, int64_t ny, int64_t nz, double lx,
                           double ly, double lz, double *xyz, int64_t *conn) {
  int64_t nn = 0, ne = 0;
  if (gs_generate_cube_sizes(nx, ny, nz, &nn, &ne) != GS_OK || !xyz || !conn)
    return GS_ERR_INVALID_ARG;
  try {
    const double dx = lx / (double)nx, dy = ly / (double)ny, dz = lz / (double)nz;
    // vertex grid, ix fastest (src/mesh_generator_improved.rs:29-39)
    int64_t n = 0;
    for (int64_t iz = 0; iz <= nz; ++iz)
      for (int64_t iy = 0; iy <= ny; ++iy)
        for (int64_t ix = 0; ix <= nx; ++ix, ++n) {
          xyz[3 * n] = (double)ix * dx;
          xyz[3 * n + 1] = (double)iy * dy;
          xyz[3 * n + 2] = (double)iz * dz;
        }
    EdgeTable edges(nn);
    const int64_t sy = nx + 1, sz = (nx + 1) * (ny + 1);
    int64_t e = 0;
    auto mid = [&](int64_t a, int64_t b) {
      bool added = false;
      const int64_t id = edges.get_or_add(a, b, n, added);
      if (added) {
        for (int c = 0; c < 3; ++c)
          xyz[3 * n + c] = (xyz[3 * a + c] + xyz[3 * b + c]) / 2.0;
        ++n;
      }
      return id;
    };
    for (int64_t iz = 0; iz < nz; ++iz)
      for (int64_t iy = 0; iy < ny; ++iy)
        for (int64_t ix = 0; ix < nx; ++ix) {
          const int64_t v000 = ix + iy * sy + iz * sz;
          const int64_t v100 = v000 + 1, v010 = v000 + sy, v110 = v000 + 1 + sy;
          const int64_t v001 = v000 + sz, v101 = v001 + 1, v011 = v001 + sy,
                        v111 = v001 + 1 + sy;
          // six tets around the v000-v111 diagonal (:105-118)
          const int64_t T[6][4] = {{v000, v100, v110, v111}, {v000, v100, v111, v101},
                                   {v000, v010, v110, v111}, {v000, v010, v111, v011},
                                   {v000, v001, v111, v101}, {v000, v001, v111, v011}};
          for (const auto &t : T) {
            int64_t v[4] = {t[0], t[1], t[2], t[3]};
            double ed[3][3];
            for (int k = 0; k < 3; ++k)
              for (int c = 0; c < 3; ++c)
                ed[k][c] = xyz[3 * v[k + 1] + c] - xyz[3 * v[0] + c];
            // orientation: e1 . (e2 x e3) < 0 -> swap v2, v3 (:139-155)
            const double det =
                ed[0][0] * (ed[1][1] * ed[2][2] - ed[1][2] * ed[2][1]) +
                ed[0][1] * (ed[1][2] * ed[2][0] - ed[1][0] * ed[2][2]) +
                ed[0][2] * (ed[1][0] * ed[2][1] - ed[1][1] * ed[2][0]);
            if (det < 0.0) std::swap(v[2], v[3]);
            int64_t *c = conn + 10 * e;
            c[0] = v[0]; c[1] = v[1]; c[2] = v[2]; c[3] = v[3];
            // midpoints in the order e01, e12, e20, e03, e13, e23 (:158-163)
            c[4] = mid(v[0], v[1]);
            c[5] = mid(v[1], v[2]);
            c[6] = mid(v[2], v[0]);
            c[7] = mid(v[0], v[3]);
            c[8] = mid(v[1], v[3]);
            c[9] = mid(v[2], v[3]);
            ++e;
          }
        }
    return (n == nn && e == ne) ? GS_OK : GS_ERR_INVALID_ARG;
  } catch (...) {
    return GS_ERR_OOM;
  }
}

gs_status gs_mesh_create(gs_ctx *ctx, const double *xyz, int64_t n_nodes,
                         const int64_t *conn, int64_t n_elems, gs_mesh **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    require(n_nodes >= 0 && n_elems >= 0 && (xyz || n_nodes == 0) &&
                (conn || n_elems == 0),
            GS_ERR_INVALID_ARG, "invalid mesh arrays");
    require(n_nodes < INT32_MAX / 3 && n_elems * 100 < INT32_MAX,
            GS_ERR_TOO_LARGE, "mesh exceeds int32 indexing");
    set_device(ctx);
    std::vector<int> c32((size_t)n_elems * 10);
    for (int64_t i = 0; i < n_elems * 10; ++i) {
      require(conn[i] >= 0 && conn[i] < n_nodes, GS_ERR_INVALID_ARG,
              "connectivity references a node out of range");
      c32[i] = (int)conn[i];
    }
    std::unique_ptr<gs_mesh> m(new gs_mesh());
    m->ctx = ctx;
    m->n_nodes = n_nodes;
    m->n_elems = n_elems;
    m->xyz = dalloc<double>(n_nodes * 3);
    m->conn = dalloc<int>(n_elems * 10);
    if (n_nodes)
      GS_HIP(hipMemcpyAsync(m->xyz.get(), xyz, n_nodes * 3 * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
    if (n_elems)
      GS_HIP(hipMemcpyAsync(m->conn.get(), c32.data(), n_elems * 10 * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
    const std::vector<int> perm = locality_order(xyz, n_nodes);
    std::vector<int> iperm((size_t)n_nodes);
    for (int64_t i = 0; i < n_nodes; ++i) iperm[perm[i]] = (int)i;
    m->perm = dalloc<int>(n_nodes);
    m->iperm = dalloc<int>(n_nodes);
    if (n_nodes) {
      GS_HIP(hipMemcpyAsync(m->perm.get(), perm.data(), n_nodes * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
      GS_HIP(hipMemcpyAsync(m->iperm.get(), iperm.data(), n_nodes * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
    }
    GS_HIP(hipStreamSynchronize(ctx->stream));
    *out = m.release();
  });
}

void gs_mesh_destroy(gs_mesh *m) { delete m; }

gs_status gs_dofmap_create(gs_ctx *ctx, int64_t n_nodes, int dpn,
                           const uint8_t *is_dir, const double *values,
                           gs_dofmap **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    require(n_nodes >= 0 && dpn >= 1, GS_ERR_INVALID_ARG, "invalid dofmap size");
    const int64_t nd = n_nodes * dpn;
    require(nd < INT32_MAX, GS_ERR_TOO_LARGE, "too many DOFs");
    set_device(ctx);
    std::vector<uint8_t> mask((size_t)nd, 0);
    std::vector<double> vals((size_t)nd, 0.0);
    int64_t fixed = 0;
    for (int64_t i = 0; i < nd; ++i) {
      mask[i] = (is_dir && is_dir[i]) ? 1 : 0;
      vals[i] = (mask[i] && values) ? values[i] : 0.0;
      fixed += mask[i];
    }
    static std::atomic<uint64_t> next_uid{1};
    std::unique_ptr<gs_dofmap> d(new gs_dofmap());
    d->ctx = ctx;
    d->uid = next_uid.fetch_add(1);
    d->n_nodes = n_nodes;
    d->dpn = dpn;
    d->n_dofs = nd;
    d->n_fixed = fixed;
    d->is_dir = dalloc<uint8_t>(nd);
    d->values = dalloc<double>(nd);
    if (nd) {
      GS_HIP(hipMemcpyAsync(d->is_dir.get(), mask.data(), nd, hipMemcpyHostToDevice,
                            ctx->stream));
      GS_HIP(hipMemcpyAsync(d->values.get(), vals.data(), nd * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
    }
    GS_HIP(hipStreamSynchronize(ctx->stream));
    *out = d.release();
  });
}

void gs_dofmap_destroy(gs_dofmap *d) { delete d; }

gs_status gs_assemble_thermal_stiffness(gs_ctx *ctx, gs_mesh *mesh,
                                        const gs_dofmap *dof, double k,
                                        const double *k_elem, gs_csr **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    // DOF = node id (src/fem/assembly.rs:95-97); n_dofs must equal n_nodes.
    require(dof->dpn == 1, GS_ERR_BAD_DOFS, "thermal assembly needs 1 DOF per node");
    set_device(ctx);
    *out = assemble_thermal_K(ctx, mesh, k, k_elem, false, 0.0, nullptr);
  });
}

gs_status gs_assemble_thermal_load(gs_ctx *ctx, gs_mesh *mesh,
                                   const gs_dofmap *dof, double Q, double *f) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 1, GS_ERR_BAD_DOFS, "thermal load needs 1 DOF per node");
    require(f || mesh->n_nodes == 0, GS_ERR_INVALID_ARG, "null output");
    set_device(ctx);
    build_node_pattern(ctx, mesh);
    hipStream_t s = ctx->stream;
    auto Ke = dalloc<double>(mesh->n_elems * 55);
    auto Fe = dalloc<double>(mesh->n_elems * 10);
    auto flag = dalloc<int>(1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    launch_thermal_elements(s, mesh->n_elems, mesh->xyz.get(), mesh->conn.get(),
                            1.0, nullptr, Q, true, Ke.get(), Fe.get(), flag.get());
    check_flags(read_flag_and_clear(ctx, flag.get()) & ~1);
    gather_thermal_load(s, mesh->pat, mesh->n_nodes, Fe.get(), mesh->perm.get(), f);
    GS_HIP(hipStreamSynchronize(s));
  });
}

gs_status gs_assemble_elasticity_stiffness(gs_ctx *ctx, gs_mesh *mesh,
                                           const gs_dofmap *dof, double E,
                                           double nu, const double *E_elem,
                                           gs_csr **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Elasticity requires 3 DOF per node");
    if (E_elem) {
      require(nu > -1.0 && nu < 0.5, GS_ERR_INVALID_MATERIAL,
              "Poisson's ratio must be in (-1, 0.5) for physical stability");
    } else {
      check_material(E, nu);
    }
    set_device(ctx);
    ElasticKind k;
    k.p0 = E;
    k.p1 = nu;
    k.e0 = E_elem;
    *out = assemble_dof3_K(ctx, mesh, k, nullptr);
  });
}

gs_status gs_assemble_thermal_mass(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                   double density, double specific_heat,
                                   const double *rho_cp_elem, gs_csr **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    set_device(ctx);
    build_node_pattern(ctx, mesh);
    const NodePattern &p = mesh->pat;
    hipStream_t s = ctx->stream;
    auto Me = dalloc<double>(mesh->n_elems * 100);
    launch_thermal_mass(s, mesh->n_elems, mesh->xyz.get(), mesh->conn.get(),
                        density * specific_heat, rho_cp_elem, Me.get());
    auto val = dalloc<double>(p.nnz);
    gather_thermal_full(s, p, Me.get(), val.get());
    GS_HIP(hipStreamSynchronize(s));
    gs_csr *M = new_csr(ctx, mesh->n_nodes, mesh->n_nodes, p.nnz, p.row_ptr, p.col, val);
    M->dperm = mesh->perm;
    *out = M;
  });
}

gs_status gs_apply_neumann_bcs(gs_ctx *ctx, gs_mesh *mesh, int64_t n_faces,
                               const int64_t *element_ids, const int32_t *local_faces,
                               const double *fluxes, double *f) {
  if (!ctx || !mesh) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(n_faces >= 0, GS_ERR_INVALID_ARG, "negative face count");
    if (n_faces == 0) return;
    require(element_ids && local_faces && fluxes && f, GS_ERR_INVALID_ARG, "null argument");
    for (int64_t k = 0; k < n_faces; ++k) {
      require(local_faces[k] >= 0 && local_faces[k] <= 3, GS_ERR_INVALID_ARG,
              "Invalid face index for tet10");
      require(element_ids[k] >= 0 && element_ids[k] < mesh->n_elems, GS_ERR_INVALID_ARG,
              "face element id out of range");
    }
    set_device(ctx);
    hipStream_t s = ctx->stream;
    auto el = dalloc<int64_t>(n_faces);
    auto lf = dalloc<int>(n_faces);
    auto fl = dalloc<double>(n_faces);
    GS_HIP(hipMemcpyAsync(el.get(), element_ids, n_faces * sizeof(int64_t),
                          hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(lf.get(), local_faces, n_faces * sizeof(int32_t),
                          hipMemcpyHostToDevice, s));
    GS_HIP(hipMemcpyAsync(fl.get(), fluxes, n_faces * sizeof(double), hipMemcpyHostToDevice,
                          s));
    apply_neumann(ctx, mesh, n_faces, el.get(), lf.get(), fl.get(), f);
  });
}

gs_status gs_assemble_gravity_load(gs_ctx *ctx, gs_mesh *mesh,
                                   const gs_dofmap *dof, double rho,
                                   const double g[3], double *f) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Gravity load requires 3 DOF per node");
    require(g && (f || mesh->n_nodes == 0), GS_ERR_INVALID_ARG, "null argument");
    set_device(ctx);
    build_node_pattern(ctx, mesh);
    hipStream_t s = ctx->stream;
    // The 30-entry element load needs no K_e; the elastic kernel computes both,
    // so run it with a valid material and keep only Fe.
    auto Ke = dalloc<double>(mesh->n_elems * 900);
    auto Fe = dalloc<double>(mesh->n_elems * 30);
    auto flag = dalloc<int>(1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    launch_elastic_elements(s, mesh->n_elems, mesh->xyz.get(), mesh->conn.get(),
                            1.0, 0.25, nullptr, rho, g, true, Ke.get(), Fe.get(),
                            flag.get());
    check_flags(read_flag_and_clear(ctx, flag.get()) & ~1);
    gather_gravity_load(s, mesh->pat, mesh->n_nodes, Fe.get(), mesh->perm.get(), f);
    GS_HIP(hipStreamSynchronize(s));
  });
}

gs_status gs_apply_dirichlet_bcs(gs_ctx *ctx, const gs_csr *K, const double *f,
                                 const gs_dofmap *dof, gs_csr **K_bc,
                                 double *f_bc) {
  if (!ctx || !K_bc) return GS_ERR_INVALID_ARG;
  *K_bc = nullptr;
  return guard(ctx, [&] {
    require(K && dof && (f || K->rows == 0) && (f_bc || K->rows == 0),
            GS_ERR_INVALID_ARG, "null argument");
    require(K->rows == dof->n_dofs && K->cols == dof->n_dofs, GS_ERR_INVALID_ARG,
            "matrix size differs from the DOF count");
    set_device(ctx);
    *K_bc = eliminate_caller_rhs(ctx, K, f, true, dof, f_bc);
    (*K_bc)->block = K->block;
  });
}

gs_status gs_assemble_thermal_system(gs_ctx *ctx, gs_mesh *mesh,
                                     const gs_dofmap *dof, double k,
                                     const double *k_elem, double Q,
                                     gs_csr **K_bc, double *f_bc) {
  if (!ctx || !K_bc) return GS_ERR_INVALID_ARG;
  *K_bc = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 1, GS_ERR_BAD_DOFS, "thermal assembly needs 1 DOF per node");
    require(f_bc || mesh->n_nodes == 0, GS_ERR_INVALID_ARG, "null output");
    set_device(ctx);
    SystemParams sp;
    sp.k = k;
    sp.k_elem = k_elem;
    sp.Q = Q;
    *K_bc = assemble_system(ctx, mesh, dof, sp, f_bc);
  });
}

gs_status gs_assemble_elasticity_system(gs_ctx *ctx, gs_mesh *mesh,
                                        const gs_dofmap *dof, double E,
                                        double nu, const double *E_elem,
                                        double rho, const double g[3],
                                        gs_csr **K_bc, double *f_bc) {
  if (!ctx || !K_bc) return GS_ERR_INVALID_ARG;
  *K_bc = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Elasticity requires 3 DOF per node");
    require(g && (f_bc || mesh->n_nodes == 0), GS_ERR_INVALID_ARG, "null argument");
    if (E_elem) {
      require(nu > -1.0 && nu < 0.5, GS_ERR_INVALID_MATERIAL,
              "Poisson's ratio must be in (-1, 0.5) for physical stability");
    } else {
      check_material(E, nu);
    }
    set_device(ctx);
    SystemParams sp;
    sp.elastic = true;
    sp.ek.p0 = E;
    sp.ek.p1 = nu;
    sp.ek.e0 = E_elem;
    sp.ek.load = GS_LOAD_BODY;
    for (int c = 0; c < 3; ++c) sp.ek.bf[c] = rho * g[c];
    *K_bc = assemble_system(ctx, mesh, dof, sp, f_bc);
  });
}

gs_status gs_assemble_viscosity_stiffness(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                          double mu, double zeta, const double *mu_elem,
                                          const double *zeta_elem, gs_csr **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Viscous flow requires 3 DOF per node");
    const ElasticKind k = viscous_kind(mu, zeta, mu_elem, zeta_elem);
    set_device(ctx);
    *out = assemble_dof3_K(ctx, mesh, k, nullptr);
  });
}

gs_status gs_assemble_viscosity_system(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                       double mu, double zeta, const double *mu_elem,
                                       const double *zeta_elem, const double *f,
                                       gs_csr **K_bc, double *f_bc) {
  if (!ctx || !K_bc) return GS_ERR_INVALID_ARG;
  *K_bc = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Viscous flow requires 3 DOF per node");
    require(f_bc || mesh->n_nodes == 0, GS_ERR_INVALID_ARG, "null output");
    SystemParams sp;
    sp.elastic = true;
    sp.ek = viscous_kind(mu, zeta, mu_elem, zeta_elem);
    sp.fext = f;
    set_device(ctx);
    *K_bc = assemble_system(ctx, mesh, dof, sp, f_bc);
  });
}

gs_status gs_assemble_maxwell_viscoelastic(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                           double E, double nu, double visc, double dt,
                                           const double *stress, gs_csr **out,
                                           double *f_history) {
  (void)dt;  // unused by the reference's element routine (element.rs:260)
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Maxwell viscoelasticity requires 3 DOF per node");
    check_maxwell(E, nu, visc);
    require(stress || mesh->n_elems == 0, GS_ERR_INVALID_ARG,
            "Mesh must have stress_history for viscoelasticity");
    require(f_history || mesh->n_nodes == 0, GS_ERR_INVALID_ARG, "null output");
    set_device(ctx);
    *out = assemble_dof3_K(ctx, mesh, maxwell_kind(E, nu, stress), f_history);
  });
}

gs_status gs_assemble_maxwell_system(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                     double E, double nu, double visc, double dt,
                                     const double *stress, const double *f_ext,
                                     gs_csr **K_bc, double *f_bc) {
  (void)dt;
  if (!ctx || !K_bc) return GS_ERR_INVALID_ARG;
  *K_bc = nullptr;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Maxwell viscoelasticity requires 3 DOF per node");
    check_maxwell(E, nu, visc);
    require(stress || mesh->n_elems == 0, GS_ERR_INVALID_ARG,
            "Mesh must have stress_history for viscoelasticity");
    require((f_ext && f_bc) || mesh->n_nodes == 0, GS_ERR_INVALID_ARG, "null argument");
    SystemParams sp;
    sp.elastic = true;
    sp.ek = maxwell_kind(E, nu, stress);
    sp.fext = f_ext;
    set_device(ctx);
    *K_bc = assemble_system(ctx, mesh, dof, sp, f_bc);
  });
}

gs_status gs_update_stresses_maxwell(gs_ctx *ctx, gs_mesh *mesh, const gs_dofmap *dof,
                                     double E, double nu, double visc, double dt,
                                     const double *u_n, const double *u_next,
                                     const double *stress, double *stress_next) {
  if (!ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    check_mesh_dof(mesh, dof);
    require(dof->dpn == 3, GS_ERR_BAD_DOFS, "Maxwell viscoelasticity requires 3 DOF per node");
    check_maxwell(E, nu, visc);
    require((u_n && u_next && stress && stress_next) || mesh->n_elems == 0,
            GS_ERR_INVALID_ARG, "null argument");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    auto flag = palloc<int>(ctx->pool, 1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    launch_maxwell_stress(s, mesh->n_elems, mesh->xyz.get(), mesh->conn.get(), E, nu, visc,
                          dt, u_n, u_next, stress, stress_next, flag.get());
    check_flags(read_flag_and_clear(ctx, flag.get()));
  });
}

gs_status gs_csr_info(const gs_csr *A, int64_t *rows, int64_t *cols,
                      int64_t *nnz) {
  if (!A) return GS_ERR_INVALID_ARG;
  if (rows) *rows = A->rows;
  if (cols) *cols = A->cols;
  if (nnz) *nnz = A->nnz;
  return GS_OK;
}

gs_status gs_csr_export(gs_ctx *ctx, const gs_csr *A, int64_t *indptr,
                        int64_t *indices, double *data) {
  if (!ctx || !A) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    set_device(ctx);
    ensure_csr_values(ctx, A);
    const int64_t rows = A->rows, nnz = A->nnz;
    std::vector<int> rp((size_t)rows + 1), ci((size_t)nnz);
    std::vector<double> dv((size_t)nnz);
    GS_HIP(hipMemcpyAsync(rp.data(), A->row_ptr.get(), (rows + 1) * sizeof(int),
                          hipMemcpyDeviceToHost, ctx->stream));
    if (nnz) {
      GS_HIP(hipMemcpyAsync(ci.data(), A->col.get(), nnz * sizeof(int),
                            hipMemcpyDeviceToHost, ctx->stream));
      GS_HIP(hipMemcpyAsync(dv.data(), A->val.get(), nnz * sizeof(double),
                            hipMemcpyDeviceToHost, ctx->stream));
    }
    if (!A->dperm) {
      GS_HIP(hipStreamSynchronize(ctx->stream));
      for (int64_t i = 0; indptr && i <= rows; ++i) indptr[i] = rp[i];
      for (int64_t i = 0; indices && i < nnz; ++i) indices[i] = ci[i];
      if (data && nnz) memcpy(data, dv.data(), nnz * sizeof(double));
      return;
    }
    // Back to caller numbering: caller row o is internal row inv[o]; its
    // entries get caller columns dperm[c] and are re-sorted by them.
    std::vector<int> dp((size_t)rows), inv((size_t)rows);
    if (rows)
      GS_HIP(hipMemcpyAsync(dp.data(), A->dperm.get(), rows * sizeof(int),
                            hipMemcpyDeviceToHost, ctx->stream));
    GS_HIP(hipStreamSynchronize(ctx->stream));
    for (int64_t i = 0; i < rows; ++i) inv[dp[i]] = (int)i;
    std::vector<std::pair<int, double>> row;
    int64_t pos = 0;
    if (indptr) indptr[0] = 0;
    for (int64_t o = 0; o < rows; ++o) {
      const int r = inv[o];
      row.clear();
      for (int k = rp[r]; k < rp[r + 1]; ++k) row.emplace_back(dp[ci[k]], dv[k]);
      std::sort(row.begin(), row.end(),
                [](const std::pair<int, double> &a, const std::pair<int, double> &b) {
                  return a.first < b.first;
                });
      for (const auto &e : row) {
        if (indices) indices[pos] = e.first;
        if (data) data[pos] = e.second;
        ++pos;
      }
      if (indptr) indptr[o + 1] = pos;
    }
  });
}

gs_status gs_csr_from_host(gs_ctx *ctx, int64_t rows, int64_t cols, int64_t nnz,
                           const int64_t *indptr, const int64_t *indices,
                           const double *data, gs_csr **out) {
  if (!ctx || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    require(rows >= 0 && cols >= 0 && nnz >= 0 && indptr, GS_ERR_INVALID_ARG,
            "invalid CSR arguments");
    require(rows < INT32_MAX && cols < INT32_MAX && nnz < INT32_MAX,
            GS_ERR_TOO_LARGE, "CSR exceeds int32 indexing");
    require(indptr[0] == 0 && indptr[rows] == nnz, GS_ERR_INVALID_ARG,
            "indptr does not span nnz");
    std::vector<int> rp((size_t)rows + 1), ci((size_t)nnz);
    for (int64_t r = 0; r < rows; ++r) {
      require(indptr[r + 1] >= indptr[r], GS_ERR_INVALID_ARG, "indptr not monotone");
      for (int64_t k = indptr[r]; k < indptr[r + 1]; ++k) {
        require(indices[k] >= 0 && indices[k] < cols, GS_ERR_INVALID_ARG,
                "column index out of range");
        require(k == indptr[r] || indices[k] > indices[k - 1], GS_ERR_INVALID_ARG,
                "columns must be sorted and unique within a row");
        ci[k] = (int)indices[k];
      }
    }
    for (int64_t r = 0; r <= rows; ++r) rp[r] = (int)indptr[r];
    set_device(ctx);
    auto drp = dalloc<int>(rows + 1);
    auto dci = dalloc<int>(nnz);
    auto dv = dalloc<double>(nnz);
    GS_HIP(hipMemcpyAsync(drp.get(), rp.data(), (rows + 1) * sizeof(int),
                          hipMemcpyHostToDevice, ctx->stream));
    if (nnz) {
      GS_HIP(hipMemcpyAsync(dci.get(), ci.data(), nnz * sizeof(int),
                            hipMemcpyHostToDevice, ctx->stream));
      GS_HIP(hipMemcpyAsync(dv.get(), data, nnz * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
    }
    GS_HIP(hipStreamSynchronize(ctx->stream));
    *out = new_csr(ctx, rows, cols, nnz, drp, dci, dv);
  });
}

void gs_csr_destroy(gs_csr *A) {
  if (!A) return;
  if (A->ctx) (void)hipStreamSynchronize(A->ctx->stream);
  delete A;
}

gs_status gs_csr_spmv(gs_ctx *ctx, const gs_csr *A, const double *x, double *y) {
  if (!ctx || !A) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((x && y) || A->rows == 0, GS_ERR_INVALID_ARG, "null vector");
    set_device(ctx);
    csr_spmv(ctx, A, x, y);
  });
}

gs_status gs_quadrature_rule(int32_t n_points, double *points, double *weights) {
  if (!points || !weights) return GS_ERR_INVALID_ARG;
  return quad_rule(n_points, points, weights) ? GS_OK : GS_ERR_INVALID_ARG;
}

gs_status gs_thermal_element_matrices(gs_ctx *ctx, gs_mesh *mesh, double conductivity,
                                      const double *k_per_elem, int32_t n_points,
                                      double *Ke) {
  if (!ctx || !mesh) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(n_points == 1 || n_points == 4 || n_points == 5 || n_points == 11,
            GS_ERR_INVALID_ARG, "quadrature rules have 1, 4, 5 or 11 points");
    require(Ke || mesh->n_elems == 0, GS_ERR_INVALID_ARG, "null output");
    set_device(ctx);
    hipStream_t s = ctx->stream;
    auto flag = palloc<int>(ctx->pool, 1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    launch_thermal_rule(s, mesh->n_elems, mesh->xyz.get(), mesh->conn.get(), conductivity,
                        k_per_elem, n_points, Ke, flag.get());
    check_flags(read_flag_and_clear(ctx, flag.get()));
  });
}

gs_status gs_ilu_create(gs_ctx *ctx, const gs_csr *A, gs_ilu **out) {
  if (!ctx || !A || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    set_device(ctx);
    *out = ilu_create(ctx, A);
  });
}

gs_status gs_ilu_update(gs_ctx *ctx, gs_ilu *ilu, const gs_csr *A) {
  if (!ctx || !ilu || !A || ilu->ctx != ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    set_device(ctx);
    ilu_update(ilu, A);
  });
}

gs_status gs_ilu_apply(gs_ctx *ctx, const gs_ilu *ilu, const double *r, double *z) {
  if (!ctx || !ilu || ilu->ctx != ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((r && z) || ilu->n == 0, GS_ERR_INVALID_ARG, "null vector");
    require(r != z || ilu->n == 0, GS_ERR_INVALID_ARG, "r and z must be distinct buffers");
    set_device(ctx);
    ilu_apply(ilu, r, z);
  });
}

gs_status gs_ilu_info(const gs_ilu *ilu, int64_t *n, int64_t *nnz, int64_t *levels_lower,
                      int64_t *levels_upper) {
  if (!ilu) return GS_ERR_INVALID_ARG;
  if (n) *n = ilu->n;
  if (nnz) *nnz = ilu->nnz;
  if (levels_lower) *levels_lower = (int64_t)ilu->lvl_lower.size() - 1;
  if (levels_upper) *levels_upper = (int64_t)ilu->lvl_upper.size() - 1;
  return GS_OK;
}

gs_status gs_ilu_export(gs_ctx *ctx, const gs_ilu *ilu, int64_t *indptr, int64_t *indices,
                        double *factors) {
  if (!ctx || !ilu || !indptr || ilu->ctx != ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((indices && factors) || ilu->nnz == 0, GS_ERR_INVALID_ARG, "null output");
    set_device(ctx);
    std::vector<int> rp((size_t)ilu->n + 1), col((size_t)ilu->nnz);
    GS_HIP(hipMemcpyAsync(rp.data(), ilu->rp.get(), rp.size() * sizeof(int),
                          hipMemcpyDeviceToHost, ctx->stream));
    if (ilu->nnz > 0) {
      GS_HIP(hipMemcpyAsync(col.data(), ilu->col.get(), col.size() * sizeof(int),
                            hipMemcpyDeviceToHost, ctx->stream));
      GS_HIP(hipMemcpyAsync(factors, ilu->f.get(), ilu->nnz * sizeof(double),
                            hipMemcpyDeviceToHost, ctx->stream));
    }
    GS_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < rp.size(); ++i) indptr[i] = rp[i];
    for (size_t k = 0; k < col.size(); ++k) indices[k] = col[k];
  });
}

void gs_ilu_destroy(gs_ilu *ilu) {
  if (!ilu) return;
  if (ilu->ctx) (void)hipStreamSynchronize(ilu->ctx->stream);
  delete ilu;
}

gs_status gs_amg_create(gs_ctx *ctx, const gs_csr *A, int64_t max_levels,
                        int64_t coarse_size, double strength_threshold, gs_amg **out) {
  if (!ctx || !A || !out) return GS_ERR_INVALID_ARG;
  *out = nullptr;
  return guard(ctx, [&] {
    set_device(ctx);
    *out = amg_create(ctx, A, max_levels, coarse_size, strength_threshold);
  });
}

gs_status gs_amg_apply(gs_ctx *ctx, const gs_amg *amg, const double *r, double *z) {
  if (!ctx || !amg || amg->ctx != ctx) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((r && z) || amg->n == 0, GS_ERR_INVALID_ARG, "null vector");
    require(r != z || amg->n == 0, GS_ERR_INVALID_ARG, "r and z must be distinct buffers");
    set_device(ctx);
    amg_apply(amg, r, z);
  });
}

gs_status gs_amg_levels(const gs_amg *amg, int64_t *levels) {
  if (!amg || !levels) return GS_ERR_INVALID_ARG;
  *levels = (int64_t)amg->lv.size();
  return GS_OK;
}

gs_status gs_amg_level_info(const gs_amg *amg, int64_t level, int64_t *rows, int64_t *nnz,
                            int64_t *coarse, int64_t *nnz_p, int64_t *sweep_levels) {
  if (!amg || level < 0 || level >= (int64_t)amg->lv.size()) return GS_ERR_INVALID_ARG;
  const AmgLevel &L = amg->lv[(size_t)level];
  if (rows) *rows = L.A.rows;
  if (nnz) *nnz = L.A.nnz;
  if (coarse) *coarse = L.has_p ? L.coarse : 0;
  if (nnz_p) *nnz_p = L.has_p ? L.P.nnz : 0;
  if (sweep_levels) *sweep_levels = L.nlev + (L.n_iso > 0 ? 1 : 0);
  return GS_OK;
}

gs_status gs_amg_export(gs_ctx *ctx, const gs_amg *amg, int64_t level, int32_t which,
                        int64_t *indptr, int64_t *indices, double *data) {
  if (!ctx || !amg || amg->ctx != ctx || level < 0 || level >= (int64_t)amg->lv.size() ||
      which < 0 || which > 2)
    return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    const AmgLevel &L = amg->lv[(size_t)level];
    set_device(ctx);
    if (which == 2) {
      require(data || L.A.rows == 0, GS_ERR_INVALID_ARG, "null output");
      const std::vector<double> d = to_host(ctx, L.dinv.get(), L.A.rows);
      std::copy(d.begin(), d.end(), data);
      return;
    }
    const DevCsr &M = which == 0 ? L.A : L.P;
    require(which == 0 || L.has_p, GS_ERR_INVALID_ARG, "this level has no prolongation");
    require(indptr && ((indices && data) || M.nnz == 0), GS_ERR_INVALID_ARG, "null output");
    const std::vector<int> rp = to_host(ctx, M.rp.get(), M.rows + 1);
    const std::vector<int> col = to_host(ctx, M.col.get(), M.nnz);
    const std::vector<double> val = to_host(ctx, M.val.get(), M.nnz);
    for (size_t i = 0; i < rp.size(); ++i) indptr[i] = rp[i];
    for (size_t k = 0; k < col.size(); ++k) {
      indices[k] = col[k];
      data[k] = val[k];
    }
  });
}

void gs_amg_destroy(gs_amg *amg) {
  if (!amg) return;
  if (amg->ctx) (void)hipStreamSynchronize(amg->ctx->stream);
  delete amg;
}

gs_status gs_jacobi_diag_inv(gs_ctx *ctx, const gs_csr *A, double *dinv) {
  if (!ctx || !A) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(dinv || A->rows == 0, GS_ERR_INVALID_ARG, "null output");
    set_device(ctx);
    ensure_csr_values(ctx, A);
    if (A->dperm) {
      auto d = dalloc<double>(A->rows);
      jacobi_diag_inv(ctx->stream, A, d.get());
      permute_scatter(ctx->stream, A->rows, d.get(), A->dperm.get(), dinv);
    } else {
      jacobi_diag_inv(ctx->stream, A, dinv);
    }
    GS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

gs_status gs_cg_opts_default(gs_cg_opts *o) {
  if (!o) return GS_ERR_INVALID_ARG;
  // src/linalg/iterative.rs:16-24
  o->max_iterations = 1000;
  o->tolerance = 1e-8;
  o->abs_tolerance = 1e-12;
  o->use_preconditioner = 1;
  o->precision = GS_PREC_FP64;
  o->inner_tolerance = 0.0;
  o->max_inner_iterations = 0;
  o->max_outer_iterations = 50;
  o->relative_breakdown = 0;
  o->reserved_ = 0;
  return GS_OK;
}

gs_status gs_cg_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
                      const gs_cg_opts *opts, double *x, gs_solver_stats *st) {
  if (!ctx || !A || !st) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((b && x) || A->rows == 0, GS_ERR_INVALID_ARG, "null vector");
    gs_cg_opts o;
    gs_cg_opts_default(&o);
    if (opts) o = *opts;
    require(o.precision == GS_PREC_FP64 || o.precision == GS_PREC_MIXED,
            GS_ERR_INVALID_ARG, "unknown precision");
    require(o.precision == GS_PREC_FP64 || o.inner_tolerance >= 0.0,
            GS_ERR_INVALID_ARG, "mixed precision needs inner_tolerance >= 0");
    set_device(ctx);
    cg_solve(ctx, A, b, o, x, st);
  });
}

gs_status gs_cg_solve_noalloc(gs_ctx *ctx, const gs_csr *A, const double *b,
                              const gs_cg_opts *opts, double *x, gs_solver_stats *st) {
  if (!ctx || !A || !st) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((b && x) || A->rows == 0, GS_ERR_INVALID_ARG, "null vector");
    gs_cg_opts o;
    gs_cg_opts_default(&o);
    if (opts) o = *opts;
    set_device(ctx);
    cg_solve_noalloc(ctx, A, b, o, x, st);
  });
}

gs_status gs_cg_solve_operator(gs_ctx *ctx, int64_t n, gs_operator_fn apply,
                               void *apply_user, gs_operator_fn precond,
                               void *precond_user, double *in, double *out,
                               const double *b, const gs_cg_opts *opts, double *x,
                               gs_solver_stats *st) {
  if (!ctx || !st) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(n >= 0 && n < INT32_MAX, GS_ERR_INVALID_ARG, "invalid size");
    require(apply || n == 0, GS_ERR_INVALID_ARG, "null operator");
    require((b && x && in && out) || n == 0, GS_ERR_INVALID_ARG, "null vector");
    gs_cg_opts o;
    gs_cg_opts_default(&o);
    if (opts) o = *opts;
    require(o.precision == GS_PREC_FP64, GS_ERR_INVALID_ARG,
            "a caller's operator is solved in FP64");
    set_device(ctx);
    if (n == 0) {  // ||b|| = 0 < 1e-25: converged at once
      *st = gs_solver_stats{};
      st->converged = 1;
      return;
    }
    cg_solve_operator(ctx, n, apply, apply_user, precond, precond_user, in, out, b, o, x,
                      st);
  });
}

gs_status gs_bicgstab_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
                            const gs_cg_opts *opts, double *x, gs_solver_stats *st) {
  if (!ctx || !A || !st) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((b && x) || A->rows == 0, GS_ERR_INVALID_ARG, "null vector");
    gs_cg_opts o;
    gs_cg_opts_default(&o);
    if (opts) o = *opts;
    set_device(ctx);
    bicgstab_solve(ctx, A, b, o, x, st);
  });
}

gs_status gs_gmres_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
                         const gs_cg_opts *opts, int32_t restart, double *x,
                         gs_solver_stats *st) {
  if (!ctx || !A || !st) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((b && x) || A->rows == 0, GS_ERR_INVALID_ARG, "null vector");
    gs_cg_opts o;
    gs_cg_opts_default(&o);
    if (opts) o = *opts;
    set_device(ctx);
    gmres_solve(ctx, A, b, o, restart, x, st);
  });
}

gs_status gs_thermal_element_backward(gs_ctx *ctx, gs_mesh *mesh,
                                      const double *k_elem, const double *u,
                                      const double *lam, double *grad) {
  if (!ctx || !mesh) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((k_elem && u && lam && grad) || mesh->n_elems == 0,
            GS_ERR_INVALID_ARG, "null argument");
    set_device(ctx);
    auto flag = dalloc<int>(1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), ctx->stream));
    launch_thermal_backward(ctx->stream, mesh->n_elems, mesh->xyz.get(),
                            mesh->conn.get(), k_elem, u, lam, grad, flag.get());
    check_flags(read_flag_and_clear(ctx, flag.get()));
    GS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

gs_status gs_elastic_element_backward(gs_ctx *ctx, gs_mesh *mesh,
                                      const double *E_elem, double nu,
                                      const double *u, const double *lam,
                                      double *grad) {
  if (!ctx || !mesh) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require((E_elem && u && lam && grad) || mesh->n_elems == 0,
            GS_ERR_INVALID_ARG, "null argument");
    require(nu > -1.0 && nu < 0.5, GS_ERR_INVALID_MATERIAL,
            "Poisson's ratio must be in (-1, 0.5) for physical stability");
    set_device(ctx);
    auto flag = dalloc<int>(1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), ctx->stream));
    launch_elastic_backward(ctx->stream, mesh->n_elems, mesh->xyz.get(),
                            mesh->conn.get(), E_elem, nu, u, lam, grad, flag.get());
    check_flags(read_flag_and_clear(ctx, flag.get()));
    GS_HIP(hipStreamSynchronize(ctx->stream));
  });
}

gs_status gs_csr_set_spmv_shape(gs_csr *A, int32_t lanes, int32_t workgroups,
                                int32_t flags) {
  if (!A) return GS_ERR_INVALID_ARG;
  if (flags & ~(GS_SPMV_VECTOR | GS_SPMV_INDEX32 | GS_SPMV_CSR | GS_SPMV_SCALAR |
                GS_SPMV_BLOCK3))
    return GS_ERR_INVALID_ARG;
  if ((flags & GS_SPMV_SCALAR) && (flags & GS_SPMV_BLOCK3)) return GS_ERR_INVALID_ARG;
  const bool csr = (flags & GS_SPMV_CSR) != 0;
  const bool vec = (flags & GS_SPMV_VECTOR) != 0;
  if (!csr && (vec || lanes != 0)) return GS_ERR_INVALID_ARG;  // SELL: no lanes
  const bool ok = lanes == 0 ||
                  (vec ? (lanes == 4 || lanes == 8 || lanes == 16 || lanes == 32)
                       : (lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8));
  if (!ok) return GS_ERR_INVALID_ARG;
  if (workgroups < 0 || workgroups > 16384) return GS_ERR_INVALID_ARG;
  A->spmv_lanes = lanes;
  A->spmv_grid = workgroups;
  A->spmv_flags = flags;
  return GS_OK;
}

gs_status gs_time_elastic_elements(gs_ctx *ctx, gs_mesh *mesh, double E,
                                   double nu, const double *E_elem,
                                   int32_t variant, int32_t reps, double *avg_ms,
                                   double *max_rel_diff) {
  if (!ctx || !mesh || !avg_ms || !max_rel_diff) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(variant >= 0 && variant <= 4, GS_ERR_INVALID_ARG, "unknown kernel variant");
    require(reps >= 1, GS_ERR_INVALID_ARG, "reps must be positive");
    if (E_elem) {
      require(nu > -1.0 && nu < 0.5, GS_ERR_INVALID_MATERIAL,
              "Poisson's ratio must be in (-1, 0.5) for physical stability");
    } else {
      check_material(E, nu);
    }
    set_device(ctx);
    hipStream_t s = ctx->stream;
    const int64_t ne = mesh->n_elems;
    const auto &pool = ctx->pool;
    const bool f32 = variant == 2 || variant == 3;
    auto ref = palloc<double>(pool, ne * 900);
    auto Ke = palloc<double>(pool, f32 ? 1 : ne * 900);
    auto Fe = palloc<double>(pool, ne * 30);
    auto Ke32 = palloc<float>(pool, f32 ? ne * 900 : 1);
    auto Fe32 = palloc<float>(pool, f32 ? ne * 30 : 1);
    auto flag = palloc<int>(pool, 1);
    auto bits = palloc<unsigned long long>(pool, 2);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    GS_HIP(hipMemsetAsync(bits.get(), 0, 2 * sizeof(unsigned long long), s));
    const double g[3] = {0.0, 0.0, -9.81};
    ElasticKind k;
    k.p0 = E;
    k.p1 = nu;
    k.e0 = E_elem;
    k.load = GS_LOAD_BODY;
    for (int c = 0; c < 3; ++c) k.bf[c] = 3300.0 * g[c];
    auto run = [&](int v, double *out) {
      switch (v) {
      case 0: launch_elastic_family(s, ne, mesh->xyz.get(), mesh->conn.get(), k, out,
                                    Fe.get(), flag.get()); break;
      case 1: launch_elastic_elements_mfma(s, ne, mesh->xyz.get(), mesh->conn.get(), E, nu,
                                           E_elem, 3300.0, g, true, out, Fe.get(), flag.get());
              break;
      case 2: launch_elastic_family_f32(s, ne, mesh->xyz.get(), mesh->conn.get(), k,
                                        Ke32.get(), Fe32.get(), flag.get()); break;
      case 3: launch_elastic_elements_mfma_f32(s, ne, mesh->xyz.get(), mesh->conn.get(), E,
                                               nu, E_elem, 3300.0, g, true, Ke32.get(),
                                               Fe32.get(), flag.get()); break;
      default: launch_elastic_family_rows(s, ne, mesh->xyz.get(), mesh->conn.get(), k, out,
                                          Fe.get(), flag.get());
      }
    };
    run(0, ref.get());
    run(variant, Ke.get());
    hipEvent_t e0, e1;
    GS_HIP(hipEventCreate(&e0));
    GS_HIP(hipEventCreate(&e1));
    GS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) run(variant, Ke.get());
    GS_HIP(hipEventRecord(e1, s));
    GS_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    check_flags(read_flag_and_clear(ctx, flag.get()));
    if (f32)
      launch_maxdiff_f32(s, ne * 900, Ke32.get(), ref.get(), bits.get());
    else
      launch_maxdiff(s, ne * 900, Ke.get(), ref.get(), bits.get());
    unsigned long long hb[2] = {0, 0};
    GS_HIP(hipMemcpyAsync(ctx->pinned, bits.get(), sizeof(hb), hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    memcpy(hb, ctx->pinned, sizeof(hb));
    double md, mb;
    memcpy(&md, &hb[0], sizeof(double));
    memcpy(&mb, &hb[1], sizeof(double));
    *avg_ms = (double)ms / reps;
    *max_rel_diff = mb > 0.0 ? md / mb : md;
  });
}

gs_status gs_time_stream(gs_ctx *ctx, size_t bytes, int32_t kind, int32_t reps,
                         double *avg_ms) {
  if (!ctx || !avg_ms) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(kind == 0 || kind == 1, GS_ERR_INVALID_ARG, "unknown stream kind");
    require(reps >= 1, GS_ERR_INVALID_ARG, "reps must be positive");
    set_device(ctx);
    *avg_ms = time_stream(ctx, bytes, kind, reps);
  });
}

gs_status gs_time_spmv(gs_ctx *ctx, const gs_csr *A, int32_t precision,
                       int32_t reps, double *avg_ms) {
  if (!ctx || !A || !avg_ms) return GS_ERR_INVALID_ARG;
  return guard(ctx, [&] {
    require(precision == GS_PREC_FP64 || precision == GS_PREC_MIXED,
            GS_ERR_INVALID_ARG, "unknown precision");
    set_device(ctx);
    *avg_ms = time_spmv(ctx, A, precision, reps);
  });
}

}  // extern "C"
