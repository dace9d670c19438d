// Jacobi-preconditioned CG on gfx950 (FP64, and FP32-inner mixed precision).
//
// Control flow is the reference's ConjugateGradient (src/linalg/
// iterative.rs:46-119) statement for statement:
//   x0 = 0, r = b, z = D^-1 r, p = z, rz = r.z
//   loop while it < max_it:
//     q = A p; pq = p.q; break if |pq| < 1e-30; alpha = rz/pq
//     x += alpha p; r -= alpha q; stop if ||r|| < tol ||b|| or ||r|| < abs_tol
//     z = D^-1 r; beta = rz'/rz; p = z + beta p; it += 1
// Each line group is one kernel; every scalar stays in device memory
// (CgState) and each reduction finishes in the last workgroup to arrive, in a
// fixed order, so results are deterministic run to run.  Kernels return at
// once when the state says stop, which lets the host enqueue iterations in
// batches and read the state back only between batches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.hpp"

namespace gs {
namespace {

constexpr int BS = 256;
constexpr int GMAX = 2048;  // max workgroups of the grid-stride CG kernels

struct CgState {
  double rz, alpha, beta, pq;
  double rnorm, bnorm, tol_abs, abs_tol;
  long long iteration, max_iter;
  int stop, converged;
  unsigned int cnt[4];
};

// ---- reductions -----------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of NV values; the result is valid in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[BS / 64][NV];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = wave_sum(v[i]);
  if (l == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sh[w][i] = v[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double s = sh[0][i];
#pragma unroll
      for (int k = 1; k < BS / 64; ++k) s += sh[k][i];
      v[i] = s;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ void store_agent(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(const double *p) {
  return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

// Publish this block's NV partial sums; returns true in every thread of the
// last block to finish, which then holds the grid totals in tot[] (thread 0
// after the block reduction; broadcast through shared memory to all).
template <int NV>
__device__ __forceinline__ bool grid_reduce(double (&v)[NV], double *partials,
                                            unsigned int *counter,
                                            double (&tot)[NV]) {
  __shared__ bool amlast;
  __shared__ double shtot[NV];
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) store_agent(&partials[blockIdx.x * NV + i], v[i]);
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    amlast = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!amlast) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS) {
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] += load_agent(&partials[b * NV + i]);
  }
  block_sum<NV>(acc);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) shtot[i] = acc[i];
    *counter = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < NV; ++i) tot[i] = shtot[i];
  return true;
}

// ---- SpMV row kernel: L lanes per row, wave64 -------------------------------
template <int L, typename T>
__device__ __forceinline__ T row_dot(const int *__restrict__ rp,
                                     const int *__restrict__ ci,
                                     const T *__restrict__ val,
                                     const T *__restrict__ x, int row, int lane) {
  const int b = rp[row], e = rp[row + 1];
  T s = (T)0;
  for (int k = b + lane; k < e; k += L) s += val[k] * x[ci[k]];
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
  return s;
}

// y = A x
template <int L, typename T>
__global__ __launch_bounds__(BS) void k_spmv(int n, const int *__restrict__ rp,
                                             const int *__restrict__ ci,
                                             const T *__restrict__ val,
                                             const T *__restrict__ x,
                                             T *__restrict__ y) {
  const int lane = threadIdx.x & (L - 1);
  const int ng = (gridDim.x * BS) / L;
  for (int row = (blockIdx.x * BS + threadIdx.x) / L; row < n; row += ng) {
    const T s = row_dot<L, T>(rp, ci, val, x, row, lane);
    if (lane == 0) y[row] = s;
  }
}

// CG step 1: q = A p, pq = p.q, alpha = rz / pq (breakdown if |pq| < 1e-30).
template <int L, typename T>
__global__ __launch_bounds__(BS) void k_cg_spmv(
    int n, const int *__restrict__ rp, const int *__restrict__ ci,
    const T *__restrict__ val, const T *__restrict__ p, T *__restrict__ q,
    CgState *__restrict__ st, double *__restrict__ partials) {
  if (st->stop) return;
  const int lane = threadIdx.x & (L - 1);
  const int ng = (gridDim.x * BS) / L;
  double acc[1] = {0.0};
  for (int row = (blockIdx.x * BS + threadIdx.x) / L; row < n; row += ng) {
    const T s = row_dot<L, T>(rp, ci, val, p, row, lane);
    if (lane == 0) {
      q[row] = s;
      acc[0] += (double)p[row] * (double)s;
    }
  }
  double tot[1];
  if (grid_reduce<1>(acc, partials, &st->cnt[0], tot) && threadIdx.x == 0) {
    st->pq = tot[0];
    if (fabs(tot[0]) < 1e-30) st->stop = 1;
    else st->alpha = st->rz / tot[0];
  }
}

// CG step 2: x += alpha p; r -= alpha q; convergence test on ||r||;
// z = D^-1 r; beta = rz_new / rz; it += 1.
template <typename T, bool PREC>
__global__ __launch_bounds__(BS) void k_cg_update(
    int n, T *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
    const T *__restrict__ p, const T *__restrict__ q,
    const T *__restrict__ dinv, CgState *__restrict__ st,
    double *__restrict__ partials) {
  if (st->stop) return;
  const T alpha = (T)st->alpha;
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
  double tot[2];
  if (grid_reduce<2>(acc, partials, &st->cnt[1], tot) && threadIdx.x == 0) {
    const double rn = sqrt(tot[0]);
    st->rnorm = rn;
    if (rn < st->tol_abs || rn < st->abs_tol) {
      st->converged = 1;
      st->stop = 1;
    } else {
      st->beta = tot[1] / st->rz;
      st->rz = tot[1];
      st->iteration += 1;
      if (st->iteration >= st->max_iter) st->stop = 1;
    }
  }
}

// CG step 3: p = z + beta p.
template <typename T>
__global__ __launch_bounds__(BS) void k_cg_direction(
    int n, const T *__restrict__ z, T *__restrict__ p,
    const CgState *__restrict__ st) {
  if (st->stop) return;
  const T beta = (T)st->beta;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    p[i] = z[i] + beta * p[i];
}

// CG set-up: x = 0, r = b, z = D^-1 r, p = z, rz = r.z, ||b||.
template <typename T, bool PREC>
__global__ __launch_bounds__(BS) void k_cg_init(
    int n, const T *__restrict__ b, T *__restrict__ x, T *__restrict__ r,
    T *__restrict__ z, T *__restrict__ p, const T *__restrict__ dinv,
    CgState *__restrict__ st, double *__restrict__ partials, long long max_iter,
    double tol, double abs_tol) {
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
  double tot[2];
  if (grid_reduce<2>(acc, partials, &st->cnt[2], tot) && threadIdx.x == 0) {
    const double bn = sqrt(tot[0]);
    st->bnorm = bn;
    st->rnorm = bn;
    st->rz = tot[1];
    st->tol_abs = tol * bn;
    st->abs_tol = abs_tol;
    st->iteration = 0;
    st->max_iter = max_iter;
    st->converged = 0;
    st->stop = 0;
    st->alpha = st->beta = st->pq = 0.0;
    if (bn < 1e-25) {  // src/linalg/iterative.rs:60-68
      st->stop = 1;
      st->converged = 1;
      st->rnorm = 0.0;
    } else if (max_iter <= 0) {
      st->stop = 1;
    }
  }
}

// ---- mixed-precision helpers ----------------------------------------------
template <int L>
__global__ __launch_bounds__(BS) void k_residual(
    int n, const int *__restrict__ rp, const int *__restrict__ ci,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, double *__restrict__ r,
    CgState *__restrict__ st, double *__restrict__ partials) {
  const int lane = threadIdx.x & (L - 1);
  const int ng = (gridDim.x * BS) / L;
  double acc[1] = {0.0};
  for (int row = (blockIdx.x * BS + threadIdx.x) / L; row < n; row += ng) {
    const double s = row_dot<L, double>(rp, ci, val, x, row, lane);
    if (lane == 0) {
      const double ri = b[row] - s;
      r[row] = ri;
      acc[0] += ri * ri;
    }
  }
  double tot[1];
  if (grid_reduce<1>(acc, partials, &st->cnt[3], tot) && threadIdx.x == 0)
    st->rnorm = sqrt(tot[0]);
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

__global__ void k_copy_f64(int n, const double *__restrict__ a, double *__restrict__ b) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) b[i] = a[i];
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

inline unsigned grid_rows(int64_t n, int L) {
  const int64_t need = (n * L + BS - 1) / BS;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, GMAX));
}
inline unsigned grid_vec(int64_t n) {
  const int64_t need = (n + BS - 1) / BS;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, GMAX));
}

int choose_lanes(int64_t rows, int64_t nnz) {
  const double avg = rows > 0 ? (double)nnz / (double)rows : 0.0;
  if (avg <= 12.0) return 4;
  if (avg <= 40.0) return 8;
  if (avg <= 100.0) return 16;
  return 32;
}

// ---- per-precision dispatch ---------------------------------------------------
template <typename T>
void launch_cg_spmv(hipStream_t s, int L, int n, const int *rp, const int *ci,
                    const T *val, const T *p, T *q, CgState *st, double *part) {
  const unsigned g = grid_rows(n, L);
  switch (L) {
  case 4: hipLaunchKernelGGL((k_cg_spmv<4, T>), g, BS, 0, s, n, rp, ci, val, p, q, st, part); break;
  case 8: hipLaunchKernelGGL((k_cg_spmv<8, T>), g, BS, 0, s, n, rp, ci, val, p, q, st, part); break;
  case 16: hipLaunchKernelGGL((k_cg_spmv<16, T>), g, BS, 0, s, n, rp, ci, val, p, q, st, part); break;
  default: hipLaunchKernelGGL((k_cg_spmv<32, T>), g, BS, 0, s, n, rp, ci, val, p, q, st, part); break;
  }
}

template <typename T>
void launch_spmv(hipStream_t s, int L, int n, const int *rp, const int *ci,
                 const T *val, const T *x, T *y) {
  const unsigned g = grid_rows(n, L);
  switch (L) {
  case 4: hipLaunchKernelGGL((k_spmv<4, T>), g, BS, 0, s, n, rp, ci, val, x, y); break;
  case 8: hipLaunchKernelGGL((k_spmv<8, T>), g, BS, 0, s, n, rp, ci, val, x, y); break;
  case 16: hipLaunchKernelGGL((k_spmv<16, T>), g, BS, 0, s, n, rp, ci, val, x, y); break;
  default: hipLaunchKernelGGL((k_spmv<32, T>), g, BS, 0, s, n, rp, ci, val, x, y); break;
  }
}

void launch_residual(hipStream_t s, int L, int n, const int *rp, const int *ci,
                     const double *val, const double *x, const double *b,
                     double *r, CgState *st, double *part) {
  const unsigned g = grid_rows(n, L);
  switch (L) {
  case 4: hipLaunchKernelGGL((k_residual<4>), g, BS, 0, s, n, rp, ci, val, x, b, r, st, part); break;
  case 8: hipLaunchKernelGGL((k_residual<8>), g, BS, 0, s, n, rp, ci, val, x, b, r, st, part); break;
  case 16: hipLaunchKernelGGL((k_residual<16>), g, BS, 0, s, n, rp, ci, val, x, b, r, st, part); break;
  default: hipLaunchKernelGGL((k_residual<32>), g, BS, 0, s, n, rp, ci, val, x, b, r, st, part); break;
  }
}

// Solver scratch carved from one per-context buffer that only grows, so
// repeated solves on the same problem size allocate nothing.
struct Work {
  double *partials = nullptr;
  CgState *st = nullptr;
  void *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
};

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// bytes for `k` Work blocks with vectors of `n` elements of `esz` bytes
size_t work_bytes(int64_t n, size_t esz) {
  return align256(2 * GMAX * sizeof(double)) + align256(sizeof(CgState)) +
         4 * align256((size_t)std::max<int64_t>(n, 1) * esz);
}

char *carve_work(char *base, int64_t n, size_t esz, Work &w) {
  w.partials = (double *)base;
  base += align256(2 * GMAX * sizeof(double));
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

// Device-resident PCG on one precision.  Iterations are enqueued in batches
// sized from the observed residual decay; the state is read once per batch.
template <typename T>
PcgResult pcg(gs_ctx *ctx, int L, int n, const int *rp, const int *ci,
              const T *val, const T *b, T *x, const T *dinv, long long max_iter,
              double tol, double abs_tol, Work &w) {
  hipStream_t s = ctx->stream;
  struct VecView { T *v; T *get() const { return v; } };
  const VecView r{(T *)w.r}, z{(T *)w.z}, p{(T *)w.p}, q{(T *)w.q};
  CgState *st = w.st;
  double *part = w.partials;
  hipLaunchKernelGGL(k_state_reset, 1, 1, 0, s, st);
  const unsigned gv = grid_vec(n);
  if (dinv)
    hipLaunchKernelGGL((k_cg_init<T, true>), gv, BS, 0, s, n, b, x, r.get(),
                       z.get(), p.get(), dinv, st, part, max_iter, tol, abs_tol);
  else
    hipLaunchKernelGGL((k_cg_init<T, false>), gv, BS, 0, s, n, b, x, r.get(),
                       z.get(), p.get(), dinv, st, part, max_iter, tol, abs_tol);
  GS_LAUNCH_CHECK("k_cg_init");
  PcgResult res;
  long long batch = 8;
  double prev_rn = -1.0;
  long long prev_it = 0;
  for (;;) {
    for (long long i = 0; i < batch; ++i) {
      launch_cg_spmv<T>(s, L, n, rp, ci, val, p.get(), q.get(), st, part);
      if (dinv)
        hipLaunchKernelGGL((k_cg_update<T, true>), gv, BS, 0, s, n, x, r.get(),
                           z.get(), p.get(), q.get(), dinv, st, part);
      else
        hipLaunchKernelGGL((k_cg_update<T, false>), gv, BS, 0, s, n, x, r.get(),
                           z.get(), p.get(), q.get(), dinv, st, part);
      hipLaunchKernelGGL((k_cg_direction<T>), gv, BS, 0, s, n, z.get(), p.get(),
                         st);
    }
    GS_LAUNCH_CHECK("cg batch");
    const CgState h = read_state(ctx, st);
    res.iterations = h.iteration;
    res.rnorm = h.rnorm;
    res.bnorm = h.bnorm;
    res.converged = h.converged != 0;
    if (h.stop) break;
    // next batch: iterations to reach the target at the observed decay rate
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
  // SpMV launches that did work: one per pass (the converging pass included)
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

static int lanes_of(const gs_csr *A) {
  return A->lanes ? A->lanes : choose_lanes(A->rows, A->nnz);
}

void csr_spmv(gs_ctx *ctx, const gs_csr *A, const double *x, double *y) {
  if (A->rows == 0) return;
  launch_spmv<double>(ctx->stream, lanes_of(A), (int)A->rows, A->row_ptr.get(),
                      A->col.get(), A->val.get(), x, y);
  GS_LAUNCH_CHECK("k_spmv");
  GS_HIP(hipStreamSynchronize(ctx->stream));
}

static void ensure_dinv(gs_ctx *ctx, gs_csr *A) {
  if (!A->dinv) {
    A->dinv = dalloc<double>(A->rows);
    jacobi_diag_inv(ctx->stream, A, A->dinv.get());
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
  gs_csr *A = const_cast<gs_csr *>(Ac);  // caches only
  const auto t0 = std::chrono::steady_clock::now();
  memset(st, 0, sizeof(*st));
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "CG needs a square matrix");
  const int n = (int)A->rows;
  const int L = lanes_of(A);
  hipStream_t s = ctx->stream;
  if (o.use_preconditioner) ensure_dinv(ctx, A);
  const bool mixed = !(o.precision == GS_PREC_FP64 || n == 0);
  // scratch: FP64 work block [+ FP32 inner work block + r64, r32, d32]
  size_t bytes = work_bytes(n, sizeof(double));
  if (mixed)
    bytes += work_bytes(n, sizeof(float)) +
             align256((size_t)std::max(n, 1) * sizeof(double)) +
             2 * align256((size_t)std::max(n, 1) * sizeof(float));
  char *base = ctx_scratch(ctx, bytes);
  Work w;
  base = carve_work(base, n, sizeof(double), w);
  if (!mixed) {
    const PcgResult r = pcg<double>(ctx, L, n, A->row_ptr.get(), A->col.get(),
                                    A->val.get(), b, x,
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
    struct { double *p; double *get() const { return p; } } r64{(double *)base};
    base += align256((size_t)std::max(n, 1) * sizeof(double));
    struct { float *p; float *get() const { return p; } } r32{(float *)base};
    base += align256((size_t)std::max(n, 1) * sizeof(float));
    struct { float *p; float *get() const { return p; } } d32{(float *)base};
    CgState *S = w.st;
    const unsigned gv = grid_vec(n);
    hipLaunchKernelGGL(k_state_reset, 1, 1, 0, s, S);
    hipLaunchKernelGGL(k_zero_f64, gv, BS, 0, s, n, x);
    // r = b - A*0 = b, computed by the residual kernel for its norm
    launch_residual(s, L, n, A->row_ptr.get(), A->col.get(), A->val.get(), x,
                    b, r64.get(), S, w.partials);
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
        hipLaunchKernelGGL(k_scale_to_f32, gv, BS, 0, s, n, r64.get(), S,
                           r32.get());
        const long long cap = std::min<long long>(
            o.max_inner_iterations > 0 ? o.max_inner_iterations : o.max_iterations,
            o.max_iterations - inner_total);
        const PcgResult ri = pcg<float>(
            ctx, L, n, A->row_ptr.get(), A->col.get(), A->val32.get(), r32.get(),
            d32.get(), dinv32, cap, o.inner_tolerance, 0.0, wi);
        inner_total += ri.iterations + (ri.converged ? 1 : 0);
        spmv += ri.spmv;
        hipLaunchKernelGGL(k_axpy_f32, gv, BS, 0, s, n, d32.get(), S, x);
        launch_residual(s, L, n, A->row_ptr.get(), A->col.get(), A->val.get(),
                        x, b, r64.get(), S, w.partials);
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
  GS_HIP(hipStreamSynchronize(s));
  st->solve_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double time_spmv(gs_ctx *ctx, const gs_csr *Ac, int precision, int reps) {
  gs_csr *A = const_cast<gs_csr *>(Ac);
  const int n = (int)A->rows;
  const int L = lanes_of(A);
  hipStream_t s = ctx->stream;
  if (n == 0 || reps <= 0) return 0.0;
  Work w;
  carve_work(ctx_scratch(ctx, work_bytes(n, sizeof(double))), n, sizeof(double), w);
  CgState *st = w.st;
  hipLaunchKernelGGL(k_state_reset, 1, 1, 0, s, st);
  hipEvent_t e0, e1;
  GS_HIP(hipEventCreate(&e0));
  GS_HIP(hipEventCreate(&e1));
  float ms = 0.f;
  if (precision == GS_PREC_FP64) {
    double *p = (double *)w.p, *q = (double *)w.q;
    hipLaunchKernelGGL(k_fill_pattern, grid_vec(n), BS, 0, s, n, p);
    launch_cg_spmv<double>(s, L, n, A->row_ptr.get(), A->col.get(),
                           A->val.get(), p, q, st, w.partials);
    GS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i)
      launch_cg_spmv<double>(s, L, n, A->row_ptr.get(), A->col.get(),
                             A->val.get(), p, q, st, w.partials);
    GS_HIP(hipEventRecord(e1, s));
    GS_HIP(hipEventSynchronize(e1));
    GS_HIP(hipEventElapsedTime(&ms, e0, e1));
  } else {
    ensure_f32(ctx, A);
    float *p = (float *)w.p, *q = (float *)w.q;
    hipLaunchKernelGGL(k_fill_pattern32, grid_vec(n), BS, 0, s, n, p);
    launch_cg_spmv<float>(s, L, n, A->row_ptr.get(), A->col.get(),
                          A->val32.get(), p, q, st, w.partials);
    GS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i)
      launch_cg_spmv<float>(s, L, n, A->row_ptr.get(), A->col.get(),
                            A->val32.get(), p, q, st, w.partials);
    GS_HIP(hipEventRecord(e1, s));
    GS_HIP(hipEventSynchronize(e1));
    GS_HIP(hipEventElapsedTime(&ms, e0, e1));
  }
  GS_LAUNCH_CHECK("time_spmv");
  const CgState h = read_state(ctx, st);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (h.stop) throw Error(GS_ERR_INVALID_ARG, "SpMV timing probe hit |p.Ap| < 1e-30");
  return (double)ms / reps;
}

}  // namespace gs
