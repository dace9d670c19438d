// Jacobi-preconditioned CG on gfx950 (FP64, and FP32-inner mixed precision).
//
// Control flow is the reference's ConjugateGradient (src/linalg/
// iterative.rs:46-119) statement for statement:
//   x0 = 0, r = b, z = D^-1 r, p = z, rz = r.z
//   loop while it < max_it:
//     q = A p; pq = p.q; break if |pq| < 1e-30; alpha = rz/pq
//     x += alpha p; r -= alpha q; stop if ||r|| < tol ||b|| or ||r|| < abs_tol
//     z = D^-1 r; beta = rz'/rz; p = z + beta p; it += 1
// One kernel per line group: K1 q = A p (+ per-workgroup p.q partials),
// K2 the x/r/z update (+ r.r, r.z partials), K3 the direction update.
//
// Reductions use no atomics and no device-scope fences (on this part a
// device-scope release writes back the XCD's whole L2, which is ruinous when
// every workgroup does it).  A producer kernel writes one partial per
// workgroup; every workgroup of the consumer kernel sums all partials in
// workgroup order -- the same deterministic value everywhere -- and only
// workgroup 0 records scalars in the state.  r.z is double-buffered by pass
// parity so no kernel reads a state word that it also writes.  Kernels return
// at once when the state says stop, which lets the host enqueue passes in
// batches and read the state back only between batches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "spmv.hpp"

namespace gs {
namespace {

constexpr int BS = 256;
constexpr int GMAX = 2048;       // workgroups of the grid-stride vector kernels
constexpr int GRID_MAX = 16384;  // workgroups of any SpMV kernel

struct CgState {
  double rz[2];  // r.z at the start of pass k lives in rz[k & 1]
  double alpha, beta, pq;
  double rnorm, bnorm, tol_abs, abs_tol;
  long long iteration, max_iter;
  int stop, converged;
};

// ---- reductions -----------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sums of NV values, returned in every thread (fixed order).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[BS / 64][NV];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  __syncthreads();
  if (l == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[w][i] = v[i];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = sh[0][i];
#pragma unroll
    for (int k = 1; k < BS / 64; ++k) s += sh[k][i];
    v[i] = s;
  }
}

// Write this workgroup's NV partial sums (after a block reduction).
template <int NV>
__device__ __forceinline__ void put_partials(double (&v)[NV], double *part) {
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) part[blockIdx.x * NV + i] = v[i];
  }
}

// Sum of the producer's partials (np workgroups x NV), in every thread.
template <int NV>
__device__ __forceinline__ void get_total(const double *__restrict__ part,
                                          int np, double (&tot)[NV]) {
#pragma unroll
  for (int i = 0; i < NV; ++i) tot[i] = 0.0;
  for (int b = threadIdx.x; b < np; b += BS) {
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] += part[b * NV + i];
  }
  block_sum<NV>(tot);
}

__device__ __forceinline__ bool leader() {
  return blockIdx.x == 0 && threadIdx.x == 0;
}

// ---- SpMV kernels (engines in spmv.hpp) ------------------------------------
// y = A x
template <int ENG, int W, class C, typename T>
__global__ __launch_bounds__(BS) void k_spmv(int n, const int *__restrict__ rp,
                                             const C cols,
                                             const T *__restrict__ val,
                                             const T *__restrict__ x,
                                             T *__restrict__ y) {
  auto epi = [&](int row, T s) { y[row] = s; };
  spmv_rows<ENG, W, T>(n, rp, cols, val, x, epi);
}

// K1: q = A p; per-workgroup partials of p.q.
template <int ENG, int W, class C, typename T>
__global__ __launch_bounds__(BS) void k_cg_spmv(
    int n, const int *__restrict__ rp, const C cols, const T *__restrict__ val,
    const T *__restrict__ p, T *__restrict__ q, const CgState *__restrict__ st,
    double *__restrict__ part) {
  if (st->stop) return;
  double acc[1] = {0.0};
  auto epi = [&](int row, T s) {
    q[row] = s;
    acc[0] += (double)p[row] * (double)s;
  };
  spmv_rows<ENG, W, T>(n, rp, cols, val, p, epi);
  put_partials<1>(acc, part);
}

// K2 (pass `it`): pq = sum of K1 partials; break if |pq| < 1e-30;
// alpha = rz/pq; x += alpha p; r -= alpha q; z = D^-1 r; partials r.r, r.z.
template <typename T, bool PREC>
__global__ __launch_bounds__(BS) void k_cg_update(
    int n, long long it, T *__restrict__ x, T *__restrict__ r,
    T *__restrict__ z, const T *__restrict__ p, const T *__restrict__ q,
    const T *__restrict__ dinv, CgState *__restrict__ st,
    const double *__restrict__ part_in, int np_in, double *__restrict__ part) {
  if (st->stop) return;
  double pq[1];
  get_total<1>(part_in, np_in, pq);
  if (fabs(pq[0]) < 1e-30) {  // src/linalg/iterative.rs:85
    if (leader()) {
      st->pq = pq[0];
      st->stop = 1;
    }
    return;
  }
  const double alpha_d = st->rz[it & 1] / pq[0];
  if (leader()) {
    st->pq = pq[0];
    st->alpha = alpha_d;
  }
  const T alpha = (T)alpha_d;
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = PREC ? ri * dinv[i] : ri;
    z[i] = zi;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)ri * (double)zi;
  }
  put_partials<2>(acc, part);
}

// K3 (pass `it`): ||r|| test; beta = rz'/rz; p = z + beta p; it += 1.
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_direction(
    int n, long long it, const T *__restrict__ z, T *__restrict__ p,
    CgState *__restrict__ st, const double *__restrict__ part_in, int np_in) {
  if (st->stop) return;
  double tot[2];
  get_total<2>(part_in, np_in, tot);
  const double rn = sqrt(tot[0]);
  if (rn < st->tol_abs || rn < st->abs_tol) {  // src/linalg/iterative.rs:95-98
    if (leader()) {
      st->rnorm = rn;
      st->converged = 1;
      st->stop = 1;
    }
    return;
  }
  const double beta_d = tot[1] / st->rz[it & 1];
  if (leader()) {
    st->rnorm = rn;
    st->beta = beta_d;
    st->rz[(it + 1) & 1] = tot[1];
    st->iteration = it + 1;
    if (it + 1 >= st->max_iter) st->stop = 1;
  }
  const T beta = (T)beta_d;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    p[i] = z[i] + beta * p[i];
}

// Set-up: x = 0, r = b, z = D^-1 r, p = z; partials b.b, r.z.
template <typename T, bool PREC>
__global__ __launch_bounds__(BS) void k_cg_init(
    int n, const T *__restrict__ b, T *__restrict__ x, T *__restrict__ r,
    T *__restrict__ z, T *__restrict__ p, const T *__restrict__ dinv,
    double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const T bi = b[i];
    x[i] = (T)0;
    r[i] = bi;
    const T zi = PREC ? bi * dinv[i] : bi;
    z[i] = zi;
    p[i] = zi;
    acc[0] += (double)bi * (double)bi;
    acc[1] += (double)bi * (double)zi;
  }
  put_partials<2>(acc, part);
}

// Set-up, single workgroup: ||b||, rz, stopping thresholds, early exits.
__global__ __launch_bounds__(BS) void k_cg_init_final(
    CgState *__restrict__ st, const double *__restrict__ part_in, int np_in,
    long long max_iter, double tol, double abs_tol) {
  double tot[2];
  get_total<2>(part_in, np_in, tot);
  if (threadIdx.x != 0) return;
  CgState s{};
  const double bn = sqrt(tot[0]);
  s.bnorm = bn;
  s.rnorm = bn;
  s.rz[0] = tot[1];
  s.tol_abs = tol * bn;
  s.abs_tol = abs_tol;
  s.max_iter = max_iter;
  if (bn < 1e-25) {  // src/linalg/iterative.rs:60-68
    s.stop = 1;
    s.converged = 1;
    s.rnorm = 0.0;
  } else if (max_iter <= 0) {
    s.stop = 1;
  }
  *st = s;
}

// ---- mixed-precision helpers ----------------------------------------------
// r = b - A x (FP64); per-workgroup partials of r.r
template <int ENG, int W, class C>
__global__ __launch_bounds__(BS) void k_residual(
    int n, const int *__restrict__ rp, const C cols,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, double *__restrict__ r,
    double *__restrict__ part) {
  double acc[1] = {0.0};
  auto epi = [&](int row, double s) {
    const double ri = b[row] - s;
    r[row] = ri;
    acc[0] += ri * ri;
  };
  spmv_rows<ENG, W, double>(n, rp, cols, val, x, epi);
  put_partials<1>(acc, part);
}

__global__ __launch_bounds__(BS) void k_norm_final(
    CgState *__restrict__ st, const double *__restrict__ part_in, int np_in) {
  double tot[1];
  get_total<1>(part_in, np_in, tot);
  if (threadIdx.x == 0) st->rnorm = sqrt(tot[0]);
}

__global__ void k_scale_to_f32(int n, const double *__restrict__ r,
                               const CgState *__restrict__ st,
                               float *__restrict__ r32) {
  const double inv = 1.0 / st->rnorm;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    r32[i] = (float)(r[i] * inv);
}

__global__ void k_axpy_f32(int n, const float *__restrict__ d,
                           const CgState *__restrict__ st,
                           double *__restrict__ x) {
  const double s = st->rnorm;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    x[i] += s * (double)d[i];
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a,
                         float *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BS)
    b[i] = (float)a[i];
}

__global__ void k_zero_f64(int n, double *__restrict__ x) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) x[i] = 0.0;
}

__global__ void k_fill_pattern(int n, double *__restrict__ x) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    x[i] = 1.0 + 0.5 * sin(0.001 * (double)i);
}

__global__ void k_fill_pattern32(int n, float *__restrict__ x) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    x[i] = 1.0f + 0.5f * sinf(0.001f * (float)i);
}

// JacobiPreconditioner::new (src/linalg/preconditioner.rs:37-51)
__global__ void k_jacobi(int n, const int *__restrict__ rp,
                         const int *__restrict__ ci,
                         const double *__restrict__ val,
                         double *__restrict__ dinv) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  double d = 1.0;
  for (int k = rp[r]; k < rp[r + 1]; ++k)
    if (ci[k] == r) {
      if (fabs(val[k]) > 1e-14) d = 1.0 / val[k];
      break;
    }
  dinv[r] = d;
}

__global__ void k_state_reset(CgState *st) { *st = CgState{}; }

inline unsigned grid_vec(int64_t n) {
  const int64_t need = (n + BS - 1) / BS;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, GMAX));
}

// anchor[g] = row holding nonzero g*1024 (binary search in the row pointer)
__global__ void k_anchor(int n, int ng, const int *__restrict__ rp,
                         int *__restrict__ anchor) {
  const int g = blockIdx.x * BS + threadIdx.x;
  if (g >= ng) return;
  const int k = g << ANCHOR_SHIFT;
  int lo = 0, hi = n;  // rp[lo] <= k < rp[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (rp[mid] <= k) lo = mid;
    else hi = mid;
  }
  anchor[g] = lo;
}

__global__ void k_offsets16(int64_t nnz, const int *__restrict__ ci,
                            const int *__restrict__ anchor,
                            short *__restrict__ off, int *__restrict__ bad) {
  for (int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * BS) {
    const int d = ci[k] - anchor[k >> ANCHOR_SHIFT];
    if (d < -32768 || d > 32767) *bad = 1;
    off[k] = (short)d;
  }
}

Shape shape_of(const gs_csr *A, bool fp32) {
  Shape sh;
  const int64_t n = A->rows;
  const double avg = n > 0 ? (double)A->nnz / (double)n : 0.0;
  const bool forced = A->spmv_lanes != 0 || (A->spmv_flags & GS_SPMV_VECTOR);
  if (forced) {
    sh.eng = (A->spmv_flags & GS_SPMV_VECTOR) ? 1 : 0;
  } else {
    // measured on MI355X: the stream engine wins in FP64, the vector engine
    // in FP32
    sh.eng = fp32 ? 1 : 0;
  }
  // Defaults from launch-shape sweeps on MI355X at 1.08M rows x 27 nnz/row:
  // FP64 stream engine, 4 threads/row, 2048 workgroups; FP32 vector engine,
  // 4 lanes/row, 8192 workgroups.
  if (sh.eng == 0) {
    // rows per workgroup R = 256 / W: a few thousand nonzeros per row block
    sh.W = A->spmv_lanes ? A->spmv_lanes : (avg <= 40.0 ? 4 : 8);
  } else {
    sh.W = A->spmv_lanes ? A->spmv_lanes
                         : (avg <= 40.0 ? 4 : (avg <= 100.0 ? 8 : 16));
  }
  const int64_t rows_per_pass = BS / sh.W;
  const int64_t want = A->spmv_grid > 0 ? A->spmv_grid : (fp32 ? 8192 : 2048);
  const int64_t cap = std::max<int64_t>(1, (n + rows_per_pass - 1) / rows_per_pass);
  sh.grid = (int)std::min<int64_t>(std::min<int64_t>(want, cap), GRID_MAX);
  sh.idx16 = A->idx16 > 0 && !(A->spmv_flags & GS_SPMV_INDEX32);
  return sh;
}

Cols32 cols32(const gs_csr *A) { return Cols32{A->col.get()}; }
Cols16 cols16(const gs_csr *A) { return Cols16{A->off16.get(), A->anchor.get()}; }

template <typename T>
void launch_cg_spmv(hipStream_t s, const Shape &sh, const gs_csr *A,
                    const T *val, const T *p, T *q, const CgState *st,
                    double *part) {
  const int n = (int)A->rows;
  dispatch_shape(sh, [&](auto eng, auto w) {
    constexpr int E = decltype(eng)::value, Wv = decltype(w)::value;
    if (sh.idx16)
      hipLaunchKernelGGL((k_cg_spmv<E, Wv, Cols16, T>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), cols16(A), val, p, q, st, part);
    else
      hipLaunchKernelGGL((k_cg_spmv<E, Wv, Cols32, T>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), cols32(A), val, p, q, st, part);
  });
}

template <typename T>
void launch_spmv(hipStream_t s, const Shape &sh, const gs_csr *A, const T *val,
                 const T *x, T *y) {
  const int n = (int)A->rows;
  dispatch_shape(sh, [&](auto eng, auto w) {
    constexpr int E = decltype(eng)::value, Wv = decltype(w)::value;
    if (sh.idx16)
      hipLaunchKernelGGL((k_spmv<E, Wv, Cols16, T>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), cols16(A), val, x, y);
    else
      hipLaunchKernelGGL((k_spmv<E, Wv, Cols32, T>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), cols32(A), val, x, y);
  });
}

// r = b - A x and ||r|| -> st->rnorm
void launch_residual(hipStream_t s, const Shape &sh, const gs_csr *A,
                     const double *x, const double *b, double *r, CgState *st,
                     double *part) {
  const int n = (int)A->rows;
  dispatch_shape(sh, [&](auto eng, auto w) {
    constexpr int E = decltype(eng)::value, Wv = decltype(w)::value;
    if (sh.idx16)
      hipLaunchKernelGGL((k_residual<E, Wv, Cols16>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), cols16(A), A->val.get(), x, b, r, part);
    else
      hipLaunchKernelGGL((k_residual<E, Wv, Cols32>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), cols32(A), A->val.get(), x, b, r, part);
  });
  hipLaunchKernelGGL(k_norm_final, 1, BS, 0, s, st, part, sh.grid);
}

// Solver scratch carved from one per-context buffer that only grows, so
// repeated solves on the same problem size allocate nothing.
struct Work {
  double *part1 = nullptr;  // one partial per SpMV workgroup
  double *part2 = nullptr;  // two partials per vector-kernel workgroup
  CgState *st = nullptr;
  void *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
};

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

size_t work_bytes(int64_t n, size_t esz) {
  return 2 * align256(GRID_MAX * sizeof(double)) + align256(sizeof(CgState)) +
         4 * align256((size_t)std::max<int64_t>(n, 1) * esz);
}

char *carve_work(char *base, int64_t n, size_t esz, Work &w) {
  w.part1 = (double *)base;
  base += align256(GRID_MAX * sizeof(double));
  w.part2 = (double *)base;
  base += align256(GRID_MAX * sizeof(double));
  w.st = (CgState *)base;
  base += align256(sizeof(CgState));
  void **v[4] = {&w.r, &w.z, &w.p, &w.q};
  for (void **x : v) {
    *x = base;
    base += align256((size_t)std::max<int64_t>(n, 1) * esz);
  }
  return base;
}

char *ctx_scratch(gs_ctx *ctx, size_t bytes) {
  if (ctx->ws_bytes < bytes) {
    ctx->ws = DArr<char>();
    ctx->ws = dalloc<char>((int64_t)bytes);
    ctx->ws_bytes = bytes;
  }
  return ctx->ws.get();
}

CgState read_state(gs_ctx *ctx, const CgState *st) {
  CgState h;
  GS_HIP(hipMemcpyAsync(ctx->pinned, st, sizeof(CgState), hipMemcpyDeviceToHost,
                        ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(&h, ctx->pinned, sizeof(CgState));
  return h;
}

struct PcgResult {
  long long iterations = 0;
  double rnorm = 0.0, bnorm = 0.0;
  bool converged = false;
  long long spmv = 0;
};

// Device-resident PCG in one precision.  Passes are enqueued in batches sized
// from the observed residual decay; the state is read once per batch.
template <typename T>
PcgResult pcg(gs_ctx *ctx, const Shape &sh, const gs_csr *A, const T *val,
              const T *b, T *x, const T *dinv, long long max_iter, double tol,
              double abs_tol, Work &w) {
  hipStream_t s = ctx->stream;
  const int n = (int)A->rows;
  T *r = (T *)w.r, *z = (T *)w.z, *p = (T *)w.p, *q = (T *)w.q;
  CgState *st = w.st;
  const unsigned gv = grid_vec(n);
  if (dinv)
    hipLaunchKernelGGL((k_cg_init<T, true>), gv, BS, 0, s, n, b, x, r, z, p,
                       dinv, w.part2);
  else
    hipLaunchKernelGGL((k_cg_init<T, false>), gv, BS, 0, s, n, b, x, r, z, p,
                       dinv, w.part2);
  hipLaunchKernelGGL(k_cg_init_final, 1, BS, 0, s, st, w.part2, (int)gv,
                     max_iter, tol, abs_tol);
  GS_LAUNCH_CHECK("k_cg_init");
  PcgResult res;
  long long launched = 0;  // pass index of the next launch
  long long batch = 8;
  double prev_rn = -1.0;
  long long prev_it = 0;
  for (;;) {
    for (long long i = 0; i < batch; ++i, ++launched) {
      launch_cg_spmv<T>(s, sh, A, val, p, q, st, w.part1);
      if (dinv)
        hipLaunchKernelGGL((k_cg_update<T, true>), gv, BS, 0, s, n, launched, x,
                           r, z, p, q, dinv, st, w.part1, sh.grid, w.part2);
      else
        hipLaunchKernelGGL((k_cg_update<T, false>), gv, BS, 0, s, n, launched, x,
                           r, z, p, q, dinv, st, w.part1, sh.grid, w.part2);
      hipLaunchKernelGGL((k_cg_direction<T>), gv, BS, 0, s, n, launched, z, p, st,
                         w.part2, (int)gv);
    }
    GS_LAUNCH_CHECK("cg batch");
    const CgState h = read_state(ctx, st);
    res.iterations = h.iteration;
    res.rnorm = h.rnorm;
    res.bnorm = h.bnorm;
    res.converged = h.converged != 0;
    if (h.stop) break;
    // next batch: passes to reach the target at the observed decay rate
    long long next = 16;
    if (prev_rn > 0.0 && h.iteration > prev_it && h.rnorm > 0.0) {
      const double rate =
          std::pow(h.rnorm / prev_rn, 1.0 / (double)(h.iteration - prev_it));
      const double target = std::max(h.tol_abs, h.abs_tol);
      if (rate < 0.999 && target > 0.0) {
        const double need = std::log(target / h.rnorm) / std::log(rate);
        next = (long long)std::ceil(0.9 * need);
      }
    }
    prev_rn = h.rnorm;
    prev_it = h.iteration;
    batch = std::max(4ll, std::min(64ll, next));
    batch = std::min(batch, std::max(1ll, max_iter - h.iteration));
  }
  // SpMV launches that did work: one per pass, the converging pass included
  res.spmv = res.iterations + (res.converged && res.bnorm >= 1e-25 ? 1 : 0);
  return res;
}

}  // namespace

// ---------------------------------------------------------------------------
void jacobi_diag_inv(hipStream_t s, const gs_csr *A, double *dinv) {
  if (A->rows == 0) return;
  hipLaunchKernelGGL(k_jacobi, (unsigned)((A->rows + BS - 1) / BS), BS, 0, s,
                     (int)A->rows, A->row_ptr.get(), A->col.get(), A->val.get(),
                     dinv);
  GS_LAUNCH_CHECK("k_jacobi");
}

void csr_spmv(gs_ctx *ctx, const gs_csr *A, const double *x, double *y) {
  if (A->rows == 0) return;
  hipStream_t s = ctx->stream;
  const int *dp = A->dperm.get();
  if (!dp) {
    launch_spmv<double>(s, shape_of(A, false), A, A->val.get(), x, y);
  } else {
    const size_t vec64 = align256((size_t)A->rows * sizeof(double));
    char *base = ctx_scratch(ctx, 2 * vec64);
    double *xi = (double *)base, *yi = (double *)(base + vec64);
    permute_gather(s, A->rows, x, dp, xi);
    launch_spmv<double>(s, shape_of(A, false), A, A->val.get(), xi, yi);
    permute_scatter(s, A->rows, yi, dp, y);
  }
  GS_LAUNCH_CHECK("k_spmv");
  GS_HIP(hipStreamSynchronize(s));
}

static void ensure_dinv(gs_ctx *ctx, gs_csr *A) {
  if (!A->dinv) {
    A->dinv = dalloc<double>(A->rows);
    jacobi_diag_inv(ctx->stream, A, A->dinv.get());
  }
}

// 16-bit anchored column offsets (see spmv.hpp), built once per matrix;
// A->idx16 = 1 when every offset fits, -1 when the matrix keeps int32.
static void ensure_idx16(gs_ctx *ctx, gs_csr *A) {
  if (A->idx16 != 0) return;
  if (A->nnz == 0 || A->rows == 0) {
    A->idx16 = -1;
    return;
  }
  hipStream_t s = ctx->stream;
  const int ng = (int)((A->nnz + (1 << ANCHOR_SHIFT) - 1) >> ANCHOR_SHIFT);
  auto anchor = dalloc<int>(ng);
  auto off = dalloc<short>(A->nnz);
  auto bad = dalloc<int>(1);
  GS_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
  hipLaunchKernelGGL(k_anchor, (unsigned)((ng + BS - 1) / BS), BS, 0, s,
                     (int)A->rows, ng, A->row_ptr.get(), anchor.get());
  hipLaunchKernelGGL(k_offsets16, (unsigned)GMAX, BS, 0, s, A->nnz, A->col.get(),
                     anchor.get(), off.get(), bad.get());
  GS_LAUNCH_CHECK("k_offsets16");
  int h = 0;
  GS_HIP(hipMemcpyAsync(ctx->pinned, bad.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  GS_HIP(hipStreamSynchronize(s));
  memcpy(&h, ctx->pinned, sizeof(int));
  if (h) {
    A->idx16 = -1;
  } else {
    A->anchor = anchor;
    A->off16 = off;
    A->idx16 = 1;
  }
}

static void ensure_f32(gs_ctx *ctx, gs_csr *A) {
  if (!A->val32) {
    A->val32 = dalloc<float>(A->nnz);
    if (A->nnz > 0)
      hipLaunchKernelGGL(k_to_f32, (unsigned)GMAX, BS, 0, ctx->stream, A->nnz,
                         A->val.get(), A->val32.get());
  }
  if (!A->dinv32) {
    ensure_dinv(ctx, A);
    A->dinv32 = dalloc<float>(A->rows);
    if (A->rows > 0)
      hipLaunchKernelGGL(k_to_f32, (unsigned)GMAX, BS, 0, ctx->stream, A->rows,
                         A->dinv.get(), A->dinv32.get());
  }
  GS_LAUNCH_CHECK("k_to_f32");
}

void cg_solve(gs_ctx *ctx, const gs_csr *Ac, const double *b,
              const gs_cg_opts &o, double *x, gs_solver_stats *st) {
  gs_csr *A = const_cast<gs_csr *>(Ac);  // derived caches only
  const auto t0 = std::chrono::steady_clock::now();
  memset(st, 0, sizeof(*st));
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "CG needs a square matrix");
  const int n = (int)A->rows;
  hipStream_t s = ctx->stream;
  if (o.use_preconditioner) ensure_dinv(ctx, A);
  ensure_idx16(ctx, A);
  const Shape sh = shape_of(A, false);
  const Shape sh32 = shape_of(A, true);
  const bool mixed = !(o.precision == GS_PREC_FP64 || n == 0);
  const int *dp = A->dperm.get();
  const size_t vec64 = align256((size_t)std::max(n, 1) * sizeof(double));
  // scratch: FP64 work block [+ b, x in internal numbering]
  //          [+ FP32 inner work block + r64, r32, d32]
  size_t bytes = work_bytes(n, sizeof(double)) + (dp ? 2 * vec64 : 0);
  if (mixed)
    bytes += work_bytes(n, sizeof(float)) + vec64 +
             2 * align256((size_t)std::max(n, 1) * sizeof(float));
  char *base = ctx_scratch(ctx, bytes);
  Work w;
  base = carve_work(base, n, sizeof(double), w);
  // vectors cross the ABI in caller numbering; the matrix is internal
  const double *b_user = b;
  double *x_user = x;
  if (dp) {
    double *bi = (double *)base;
    double *xi = (double *)(base + vec64);
    base += 2 * vec64;
    permute_gather(s, n, b_user, dp, bi);
    b = bi;
    x = xi;
  }
  if (!mixed) {
    const PcgResult r = pcg<double>(ctx, sh, A, A->val.get(), b, x,
                                    o.use_preconditioner ? A->dinv.get() : nullptr,
                                    o.max_iterations, o.tolerance,
                                    o.abs_tolerance, w);
    st->iterations = r.iterations;
    st->residual_norm = r.rnorm;
    st->relative_residual = r.bnorm > 1e-20 ? r.rnorm / r.bnorm : 0.0;
    st->converged = r.converged;
    st->outer_iterations = 0;
    st->spmv_count = r.spmv;
  } else {
    // FP64 defect correction around an FP32 Jacobi-PCG:
    //   r = b - A x (FP64); solve A d = r/||r|| in FP32 (FP64 dot products);
    //   x += ||r|| d (FP64); until ||r|| < tol ||b|| or < abs_tol.
    ensure_f32(ctx, A);
    Work wi;
    base = carve_work(base, n, sizeof(float), wi);
    double *r64 = (double *)base;
    base += align256((size_t)std::max(n, 1) * sizeof(double));
    float *r32 = (float *)base;
    base += align256((size_t)std::max(n, 1) * sizeof(float));
    float *d32 = (float *)base;
    CgState *S = w.st;
    const unsigned gv = grid_vec(n);
    hipLaunchKernelGGL(k_state_reset, 1, 1, 0, s, S);
    hipLaunchKernelGGL(k_zero_f64, gv, BS, 0, s, n, x);
    // r = b - A*0 = b, through the residual kernel for its norm
    launch_residual(s, sh, A, x, b, r64, S, w.part1);
    GS_LAUNCH_CHECK("k_residual");
    CgState h = read_state(ctx, S);
    const double bnorm = h.rnorm;
    double rn = bnorm;
    long long inner_total = 0, outer = 0, spmv = 1;
    bool converged = false;
    if (bnorm < 1e-25) {
      converged = true;
      rn = 0.0;
    } else {
      const float *dinv32 = o.use_preconditioner ? A->dinv32.get() : nullptr;
      const long long max_outer =
          o.max_outer_iterations > 0 ? o.max_outer_iterations : 100;
      while (outer < max_outer) {
        if (rn < o.tolerance * bnorm || rn < o.abs_tolerance) {
          converged = true;
          break;
        }
        if (inner_total >= o.max_iterations) break;
        hipLaunchKernelGGL(k_scale_to_f32, gv, BS, 0, s, n, r64, S, r32);
        const long long cap = std::min<long long>(
            o.max_inner_iterations > 0 ? o.max_inner_iterations : o.max_iterations,
            o.max_iterations - inner_total);
        const PcgResult ri = pcg<float>(ctx, sh32, A, A->val32.get(), r32, d32,
                                        dinv32, cap, o.inner_tolerance, 0.0, wi);
        inner_total += ri.iterations + (ri.converged ? 1 : 0);
        spmv += ri.spmv;
        hipLaunchKernelGGL(k_axpy_f32, gv, BS, 0, s, n, d32, S, x);
        launch_residual(s, sh, A, x, b, r64, S, w.part1);
        GS_LAUNCH_CHECK("defect correction");
        ++spmv;
        ++outer;
        h = read_state(ctx, S);
        rn = h.rnorm;
      }
      if (!converged && (rn < o.tolerance * bnorm || rn < o.abs_tolerance))
        converged = true;
    }
    st->iterations = inner_total;
    st->residual_norm = rn;
    st->relative_residual = bnorm > 1e-20 ? rn / bnorm : 0.0;
    st->converged = converged;
    st->outer_iterations = outer;
    st->spmv_count = spmv;
  }
  if (dp) permute_scatter(s, n, x, dp, x_user);
  (void)b_user;
  GS_HIP(hipStreamSynchronize(s));
  st->solve_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double time_spmv(gs_ctx *ctx, const gs_csr *Ac, int precision, int reps) {
  gs_csr *A = const_cast<gs_csr *>(Ac);
  const int n = (int)A->rows;
  hipStream_t s = ctx->stream;
  if (n == 0 || reps <= 0) return 0.0;
  ensure_idx16(ctx, A);
  const bool fp32 = precision != GS_PREC_FP64;
  const Shape sh = shape_of(A, fp32);
  Work w;
  carve_work(ctx_scratch(ctx, work_bytes(n, sizeof(double))), n, sizeof(double), w);
  hipLaunchKernelGGL(k_state_reset, 1, 1, 0, s, w.st);
  hipEvent_t e0, e1;
  GS_HIP(hipEventCreate(&e0));
  GS_HIP(hipEventCreate(&e1));
  float ms = 0.f;
  if (!fp32) {
    double *p = (double *)w.p, *q = (double *)w.q;
    hipLaunchKernelGGL(k_fill_pattern, grid_vec(n), BS, 0, s, n, p);
    launch_cg_spmv<double>(s, sh, A, A->val.get(), p, q, w.st, w.part1);
    GS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i)
      launch_cg_spmv<double>(s, sh, A, A->val.get(), p, q, w.st, w.part1);
    GS_HIP(hipEventRecord(e1, s));
  } else {
    ensure_f32(ctx, A);
    float *p = (float *)w.p, *q = (float *)w.q;
    hipLaunchKernelGGL(k_fill_pattern32, grid_vec(n), BS, 0, s, n, p);
    launch_cg_spmv<float>(s, sh, A, A->val32.get(), p, q, w.st, w.part1);
    GS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i)
      launch_cg_spmv<float>(s, sh, A, A->val32.get(), p, q, w.st, w.part1);
    GS_HIP(hipEventRecord(e1, s));
  }
  GS_LAUNCH_CHECK("time_spmv");
  GS_HIP(hipEventSynchronize(e1));
  GS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

}  // namespace gs
