// Jacobi-preconditioned CG on gfx950 (FP64, and FP32 with FP64 residual
// replacement).
//
// Control flow is the reference's ConjugateGradient (src/linalg/
// iterative.rs:46-119) statement for statement:
//   x0 = 0, r = b, z = D^-1 r, p = z, rz = r.z
//   loop while it < max_it:
//     q = A p; pq = p.q; break if |pq| < 1e-30; alpha = rz/pq
//     x += alpha p; r -= alpha q; stop if ||r|| < tol ||b|| or ||r|| < abs_tol
//     z = D^-1 r; beta = rz'/rz; p = z + beta p; it += 1
// One pass = three launches: K1 q = A p (SELL SpMV, per-workgroup partials of
// p.q); K2 alpha and the x/r/z update (+ r.r, r.z partials); K3 the stopping
// test, beta and the direction update.  K2/K3 run on a few hundred large
// workgroups, each of which sums the producer's partials itself in a fixed
// order, so every result is deterministic and no kernel needs atomics or
// device-scope fences (on this part a device-scope release writes back the
// XCD's whole L2).
//
// Passes are enqueued in batches without host round trips.  Every kernel of
// pass `it` checks the state first and returns at once unless the solve is
// at pass `it` and neither stopped nor paused; the host reads the state back
// only between batches.
//
// Vectors live in the SELL order of the operator (sell.hpp); the C ABI's
// vectors are gathered into it and scattered back.
//
// Mixed precision: the passes run in FP32 (SELL values, vectors and SpMV;
// dot products accumulated in FP64) on the system scaled by 1/||b||.  When
// the recursive FP32 residual has dropped by inner_tolerance since the last
// replacement, or reaches the target, F2 pauses the solve; the host folds the
// FP32 correction into the FP64 solution, recomputes r = b - A x in FP64
// (the stopping test uses this true residual), and continues the FP32
// recursion from it while keeping the search direction (residual
// replacement), so Krylov information is not thrown away as in plain
// defect correction.  One cycle can gain at most about eps32 * cond(A) on
// the true residual, so replacements come every inner_tolerance (1e-2 by
// default) and the direction is reset when the two residuals disagree.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "internal.hpp"
#include "sell.hpp"
#include "spmv.hpp"

namespace gs {
namespace {

constexpr int BS = 256;
constexpr int FBS = 1024;        // finalize kernels: one workgroup
constexpr int GMAX = 2048;       // workgroups of the grid-stride vector kernels
constexpr int VGMAX = 256;       // workgroups (of FBS threads) of a pass's vector kernels
constexpr int GRID_MAX = 16384;  // workgroups of any SpMV kernel

struct CgState {
  double rz[2];  // r.z at the start of pass k lives in rz[k & 1]
  double alpha, beta, pq;
  double rnorm, bnorm, tol_abs, abs_tol, tol;
  double rrec;   // mixed: recursive FP32 residual norm (scaled by 1/||b||)
  double rbase;  // mixed: scaled true residual at the last replacement
  double rtol;   // mixed: replacement factor (inner_tolerance)
  double scale;  // mixed: 1/||b||
  long long iteration, max_iter;
  int stop, converged, pause, mixed, breakdown;
};

__device__ __forceinline__ bool active(const CgState *st, long long it) {
  return !st->stop && !st->pause && st->iteration == it;
}

// ---- reductions -----------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums of NV values over a workgroup of NT threads, returned in every thread
// (fixed order).
template <int NV, int NT>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  __shared__ double sh[NT / 64][NV];
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
    for (int k = 1; k < NT / 64; ++k) s += sh[k][i];
    v[i] = s;
  }
}

// This workgroup's NV partial sums -> part[blockIdx.x * NV + i].
template <int NV>
__device__ __forceinline__ void put_partials(double (&v)[NV], double *part) {
  block_sum<NV, BS>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) part[blockIdx.x * NV + i] = v[i];
  }
}

// Finalize kernels: sum of np x NV partials, in every thread.
template <int NV>
__device__ __forceinline__ void sum_partials(const double *__restrict__ part,
                                             int np, double (&tot)[NV]) {
#pragma unroll
  for (int i = 0; i < NV; ++i) tot[i] = 0.0;
  for (int b = threadIdx.x; b < np; b += FBS) {
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] += part[b * NV + i];
  }
  block_sum<NV, FBS>(tot);
}

// ---- SpMV kernels ------------------------------------------------------------
// y = A x (SELL)
template <typename T, class IDX>
__global__ __launch_bounds__(BS) void k_sell_spmv(int n, int ns,
                                                  const int *__restrict__ cb,
                                                  const IDX idx,
                                                  const T *__restrict__ val,
                                                  const T *__restrict__ x,
                                                  T *__restrict__ y) {
  auto epi = [&](int row, T s) { y[row] = s; };
  sell_rows<T>(n, ns, cb, idx, val, x, epi);
}

// K1 (pass it): q = A p; per-workgroup partials of p.q.
template <typename T, class IDX>
__global__ __launch_bounds__(BS) void k_cg_q(int n, int ns,
                                             const int *__restrict__ cb,
                                             const IDX idx,
                                             const T *__restrict__ val,
                                             const T *__restrict__ p,
                                             T *__restrict__ q,
                                             const CgState *__restrict__ st,
                                             long long it,
                                             double *__restrict__ part) {
  if (!active(st, it)) return;
  double acc[1] = {0.0};
  auto epi = [&](int row, T s) {
    q[row] = s;
    acc[0] += (double)p[row] * (double)s;
  };
  sell_rows<T>(n, ns, cb, idx, val, p, epi);
  put_partials<1>(acc, part);
}

// r = b - A x (FP64, SELL); partials of r.r
template <class IDX>
__global__ __launch_bounds__(BS) void k_sell_residual(
    int n, int ns, const int *__restrict__ cb, const IDX idx,
    const double *__restrict__ val, const double *__restrict__ x,
    const double *__restrict__ b, double *__restrict__ r,
    double *__restrict__ part) {
  double acc[1] = {0.0};
  auto epi = [&](int row, double s) {
    const double ri = b[row] - s;
    r[row] = ri;
    acc[0] += ri * ri;
  };
  sell_rows<double>(n, ns, cb, idx, val, x, epi);
  put_partials<1>(acc, part);
}

// CSR engines (rectangular matrices, and comparison timing): y = A x
template <int ENG, int W, class C, typename T>
__global__ __launch_bounds__(BS) void k_spmv(int n, const int *__restrict__ rp,
                                             const C cols,
                                             const T *__restrict__ val,
                                             const T *__restrict__ x,
                                             T *__restrict__ y) {
  auto epi = [&](int row, T s) { y[row] = s; };
  spmv_rows<ENG, W, T>(n, rp, cols, val, x, epi);
}

// ---- pass kernels ------------------------------------------------------------
// The two vector kernels of a pass run on at most VGMAX workgroups of FBS
// threads.  Each workgroup first sums the previous kernel's partials itself
// (same order everywhere, so every workgroup takes the same decisions) and
// workgroup 0 records the scalars.  No workgroup reads a state word that the
// leader of the same kernel writes, except `iteration`, which the leader only
// advances when the pass completes -- so seeing it+1 also means "live".

// K2 (pass it): pq; FP64 breaks if |pq| < 1e-30 (src/linalg/iterative.rs:85);
// alpha = rz/pq; x += alpha p; r -= alpha q; z = D^-1 r; partials r.r, r.z.
template <typename T, bool PREC>
__global__ __launch_bounds__(FBS) void k_cg_alpha(
    int n, long long it, T *__restrict__ x, T *__restrict__ r,
    T *__restrict__ z, const T *__restrict__ p, const T *__restrict__ q,
    const T *__restrict__ dinv, CgState *__restrict__ st,
    const double *__restrict__ part_in, int np_in, double *__restrict__ part) {
  if (!active(st, it)) return;
  double pq[1];
  sum_partials<1>(part_in, np_in, pq);
  const double alpha_d = st->rz[it & 1] / pq[0];
  // The scaled FP32 system has its own units, where the reference's absolute
  // 1e-30 threshold means nothing; it breaks down only on a zero or
  // non-finite step.
  const bool brk = st->mixed ? (pq[0] == 0.0 || !isfinite(alpha_d))
                             : fabs(pq[0]) < 1e-30;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq[0];
    if (brk) {
      if (st->mixed) {
        st->pause = 1;
        st->breakdown = 1;
      } else {
        st->stop = 1;
      }
    } else {
      st->alpha = alpha_d;
    }
  }
  if (brk) return;
  const T alpha = (T)alpha_d;
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    r[i] = ri;
    const T zi = PREC ? ri * dinv[i] : ri;
    z[i] = zi;
    acc[0] += (double)ri * (double)ri;
    acc[1] += (double)ri * (double)zi;
  }
  block_sum<2, FBS>(acc);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// K3 (pass it).  FP64: the reference's stopping test
// (src/linalg/iterative.rs:95-98; the converging pass is not counted), then
// beta = rz'/rz, it += 1 (stop at max_iter), p = z + beta p.  Mixed: always
// completes the pass, and pauses for a residual replacement when the
// recursive residual reached the target, dropped by rtol since the last
// replacement, or the pass budget is spent.
template <typename T>
__global__ __launch_bounds__(FBS) void k_cg_beta(
    int n, long long it, const T *__restrict__ z, T *__restrict__ p,
    CgState *__restrict__ st, const double *__restrict__ part_in, int np_in) {
  const long long at = st->iteration;
  if (!((at == it && !st->stop && !st->pause) || at == it + 1)) return;
  double tot[2];
  sum_partials<2>(part_in, np_in, tot);
  const double rn = sqrt(tot[0]);
  const bool mixed = st->mixed != 0;
  const bool conv = !mixed && (rn < st->tol_abs || rn < st->abs_tol);
  const double beta_d = tot[1] / st->rz[it & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (conv) {
      st->rnorm = rn;
      st->converged = 1;
      st->stop = 1;
    } else {
      if (mixed) st->rrec = rn;
      else st->rnorm = rn;
      st->beta = beta_d;
      st->rz[(it + 1) & 1] = tot[1];
      st->iteration = it + 1;
      if (!mixed) {
        if (it + 1 >= st->max_iter) st->stop = 1;
      } else if (rn < st->tol || rn * st->bnorm < st->abs_tol ||
                 rn < st->rtol * st->rbase || it + 1 >= st->max_iter) {
        st->pause = 1;
      }
    }
  }
  if (conv) return;
  const T beta = (T)beta_d;
  for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS)
    p[i] = z[i] + beta * p[i];
}

// FP64 set-up: x = 0, r = b, z = D^-1 r, p = z; partials b.b, r.z.
template <bool PREC>
__global__ __launch_bounds__(BS) void k_cg_init(
    int n, const double *__restrict__ b, double *__restrict__ x,
    double *__restrict__ r, double *__restrict__ z, double *__restrict__ p,
    const double *__restrict__ dinv, double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    const double zi = PREC ? bi * dinv[i] : bi;
    z[i] = zi;
    p[i] = zi;
    acc[0] += bi * bi;
    acc[1] += bi * zi;
  }
  put_partials<2>(acc, part);
}

__global__ void k_state_setup(CgState *__restrict__ st, long long max_iter,
                              double tol, double abs_tol, int mixed, double rtol) {
  CgState s{};
  s.max_iter = max_iter;
  s.tol = tol;
  s.abs_tol = abs_tol;
  s.mixed = mixed;
  s.rtol = rtol;
  *st = s;
}

// FP64 set-up, single workgroup: ||b||, rz, thresholds, early exits
// (src/linalg/iterative.rs:60-68).
__global__ __launch_bounds__(FBS) void k_fin_init(CgState *__restrict__ st,
                                                  const double *__restrict__ part,
                                                  int np) {
  double tot[2];
  sum_partials<2>(part, np, tot);
  if (threadIdx.x != 0) return;
  const double bn = sqrt(tot[0]);
  st->bnorm = bn;
  st->rnorm = bn;
  st->rz[0] = tot[1];
  st->tol_abs = st->tol * bn;
  if (bn < 1e-25) {
    st->stop = 1;
    st->converged = 1;
    st->rnorm = 0.0;
  } else if (st->max_iter <= 0) {
    st->stop = 1;
  }
}

// ---- mixed-precision kernels -------------------------------------------------
// x = 0, r = b; partials b.b
__global__ __launch_bounds__(BS) void k_mx_start(int n, const double *__restrict__ b,
                                                 double *__restrict__ x,
                                                 double *__restrict__ r,
                                                 double *__restrict__ part) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    acc[0] += bi * bi;
  }
  put_partials<1>(acc, part);
}

// ||r|| from partials; on the first call also ||b||, 1/||b||, thresholds
__global__ __launch_bounds__(FBS) void k_fin_norm(CgState *__restrict__ st,
                                                  const double *__restrict__ part,
                                                  int np, int first) {
  double tot[1];
  sum_partials<1>(part, np, tot);
  if (threadIdx.x != 0) return;
  const double rn = sqrt(tot[0]);
  st->rnorm = rn;
  if (first) {
    st->bnorm = rn;
    st->scale = rn > 0.0 ? 1.0 / rn : 0.0;
    st->tol_abs = st->tol * rn;
  }
}

// FP32 recursion (re)started from the FP64 residual: rs = r/||b||,
// zs = D^-1 rs [and p = zs, xs = 0 on a restart]; partials rs.zs
template <bool PREC, bool RESTART>
__global__ __launch_bounds__(BS) void k_mx_rescale(
    int n, const double *__restrict__ r, const float *__restrict__ dinv,
    float *__restrict__ rs, float *__restrict__ zs, float *__restrict__ p,
    float *__restrict__ xs, const CgState *__restrict__ st,
    double *__restrict__ part) {
  const double sc = st->scale;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const float ri = (float)(r[i] * sc);
    rs[i] = ri;
    const float zi = PREC ? ri * dinv[i] : ri;
    zs[i] = zi;
    if (RESTART) {
      p[i] = zi;
      xs[i] = 0.0f;
    }
    acc[0] += (double)ri * (double)zi;
  }
  put_partials<1>(acc, part);
}

__global__ __launch_bounds__(FBS) void k_fin_rescale(CgState *__restrict__ st,
                                                     const double *__restrict__ part,
                                                     int np) {
  double tot[1];
  sum_partials<1>(part, np, tot);
  if (threadIdx.x != 0) return;
  st->rz[st->iteration & 1] = tot[0];
  st->rbase = st->rnorm * st->scale;
  st->rrec = st->rbase;
  st->pause = 0;
  st->breakdown = 0;
}

// x += ||b|| xs (FP64); xs = 0
__global__ __launch_bounds__(BS) void k_mx_fold(int n, float *__restrict__ xs,
                                                double *__restrict__ x,
                                                const CgState *__restrict__ st) {
  const double bn = st->bnorm;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    x[i] += bn * (double)xs[i];
    xs[i] = 0.0f;
  }
}

// ---- misc --------------------------------------------------------------------
template <typename T>
__global__ void k_fill_pattern(int n, T *__restrict__ x) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    x[i] = (T)(1.0 + 0.5 * sin(0.001 * (double)i));
}

__global__ void k_to_f32(int64_t n, const double *__restrict__ a,
                         float *__restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BS)
    b[i] = (float)a[i];
}

// JacobiPreconditioner::new (src/linalg/preconditioner.rs:37-51), CSR order
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

// ---- SELL launch helpers -------------------------------------------------------
struct SellOp {
  const SellPattern *P = nullptr;
  int grid = 1;
  int n() const { return (int)P->n; }
  int ns() const { return (int)P->nslices; }
};

SellOp sell_op(const gs_csr *A) {
  SellOp o;
  o.P = A->sell.get();
  const int64_t nb = std::max<int64_t>(1, (o.P->nslices + SELL_WPB - 1) / SELL_WPB);
  const int64_t want = A->spmv_grid > 0 ? A->spmv_grid : nb;
  o.grid = (int)std::min<int64_t>(std::min<int64_t>(want, nb), GRID_MAX);
  return o;
}

// Calls f(idx) with the pattern's index reader.
template <class F> void with_idx(const SellPattern &P, F &&f) {
  if (P.idx16) f(SellIdx16{P.off.get()});
  else f(SellIdx32{P.col.get()});
}

template <typename T>
void launch_cg_q(hipStream_t s, const SellOp &o, const T *val, const T *p, T *q,
                 const CgState *st, long long it, double *part) {
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_cg_q<T, decltype(idx)>), o.grid, BS, 0, s, o.n(), o.ns(),
                       o.P->cb.get(), idx, val, p, q, st, it, part);
  });
}

template <typename T>
void launch_sell_spmv(hipStream_t s, const SellOp &o, const T *val, const T *x,
                      T *y) {
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_sell_spmv<T, decltype(idx)>), o.grid, BS, 0, s, o.n(),
                       o.ns(), o.P->cb.get(), idx, val, x, y);
  });
}

void launch_residual(hipStream_t s, const SellOp &o, const double *val,
                     const double *x, const double *b, double *r, double *part) {
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_sell_residual<decltype(idx)>), o.grid, BS, 0, s, o.n(),
                       o.ns(), o.P->cb.get(), idx, val, x, b, r, part);
  });
}

// ---- CSR engine launch shape (rectangular matrices, comparison timing) ------
Shape csr_shape(const gs_csr *A, bool fp32) {
  Shape sh;
  const int64_t n = A->rows;
  const double avg = n > 0 ? (double)A->nnz / (double)n : 0.0;
  const bool forced = A->spmv_lanes != 0 || (A->spmv_flags & GS_SPMV_VECTOR);
  sh.eng = forced ? ((A->spmv_flags & GS_SPMV_VECTOR) ? 1 : 0) : (fp32 ? 1 : 0);
  if (sh.eng == 0) {
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

template <typename T>
void launch_csr_spmv(hipStream_t s, const Shape &sh, const gs_csr *A, const T *val,
                     const T *x, T *y) {
  const int n = (int)A->rows;
  dispatch_shape(sh, [&](auto eng, auto w) {
    constexpr int E = decltype(eng)::value, Wv = decltype(w)::value;
    if (sh.idx16)
      hipLaunchKernelGGL((k_spmv<E, Wv, Cols16, T>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), Cols16{A->off16.get(), A->anchor.get()},
                         val, x, y);
    else
      hipLaunchKernelGGL((k_spmv<E, Wv, Cols32, T>), sh.grid, BS, 0, s, n,
                         A->row_ptr.get(), Cols32{A->col.get()}, val, x, y);
  });
}

// ---- scratch -------------------------------------------------------------------
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of consecutive 256-byte-aligned pieces of one buffer.
struct Layout {
  size_t bytes = 0;
  template <typename T> size_t add(int64_t count) {
    const size_t o = bytes;
    bytes += align256((size_t)std::max<int64_t>(count, 1) * sizeof(T));
    return o;
  }
};

// One per-context buffer that only grows, so repeated solves of one size
// allocate nothing.
char *ctx_scratch(gs_ctx *ctx, size_t bytes) {
  if (ctx->ws_bytes < bytes) {
    GS_HIP(hipStreamSynchronize(ctx->stream));
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

// Passes for the next batch: enough to reach `target` at the decay rate seen
// since the previous read, within [4, 64].
struct Batcher {
  double prev = -1.0;
  long long prev_it = 0;
  long long next(double now, long long it, double target, long long left) {
    long long b = 16;
    if (prev > 0.0 && it > prev_it && now > 0.0 && target > 0.0 && now > target) {
      const double rate = std::pow(now / prev, 1.0 / (double)(it - prev_it));
      if (rate < 0.999)
        b = (long long)std::ceil(0.9 * std::log(target / now) / std::log(rate));
    }
    prev = now;
    prev_it = it;
    b = std::max(4ll, std::min(64ll, b));
    return std::min(b, std::max(1ll, left));
  }
};

// Workgroups of the pass's vector kernels (FBS threads each).
inline unsigned grid_fat(int64_t n) {
  const int64_t need = (n + FBS - 1) / FBS;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(need, VGMAX));
}

// One pass: K1 q = A p, K2 alpha + update, K3 beta + direction.
template <typename T>
void enqueue_pass(hipStream_t s, const SellOp &o, const T *val, T *x, T *r, T *z,
                  T *p, T *q, const T *dinv, CgState *st, long long it,
                  double *part1, double *part2) {
  const int n = o.n();
  const unsigned gf = grid_fat(n);
  launch_cg_q<T>(s, o, val, p, q, st, it, part1);
  if (dinv)
    hipLaunchKernelGGL((k_cg_alpha<T, true>), gf, FBS, 0, s, n, it, x, r, z, p, q,
                       dinv, st, part1, o.grid, part2);
  else
    hipLaunchKernelGGL((k_cg_alpha<T, false>), gf, FBS, 0, s, n, it, x, r, z, p, q,
                       dinv, st, part1, o.grid, part2);
  hipLaunchKernelGGL((k_cg_beta<T>), gf, FBS, 0, s, n, it, z, p, st, part2, (int)gf);
}

struct PcgResult {
  long long iterations = 0, outer = 0, spmv = 0, restarts = 0;
  double rnorm = 0.0, bnorm = 0.0;
  bool converged = false;
};

// FP64 Jacobi-PCG with the reference's control flow; b, x in SELL order.
PcgResult pcg64(gs_ctx *ctx, const gs_csr *A, const double *b, double *x,
                bool precond, long long max_iter, double tol, double abs_tol,
                char *base, Layout &L) {
  hipStream_t s = ctx->stream;
  const SellOp o = sell_op(A);
  const int n = o.n();
  double *part1 = (double *)(base + L.add<double>(GRID_MAX));
  double *part2 = (double *)(base + L.add<double>(2 * GMAX));
  CgState *st = (CgState *)(base + L.add<CgState>(1));
  double *r = (double *)(base + L.add<double>(n));
  double *z = (double *)(base + L.add<double>(n));
  double *p = (double *)(base + L.add<double>(n));
  double *q = (double *)(base + L.add<double>(n));
  const double *dinv = precond ? A->sdinv.get() : nullptr;
  const double *val = A->sval.get();
  const unsigned gv = grid_vec(n);
  hipLaunchKernelGGL(k_state_setup, 1, 1, 0, s, st, max_iter, tol, abs_tol, 0, 0.0);
  if (dinv)
    hipLaunchKernelGGL((k_cg_init<true>), gv, BS, 0, s, n, b, x, r, z, p, dinv, part2);
  else
    hipLaunchKernelGGL((k_cg_init<false>), gv, BS, 0, s, n, b, x, r, z, p, dinv, part2);
  hipLaunchKernelGGL(k_fin_init, 1, FBS, 0, s, st, part2, (int)gv);
  GS_LAUNCH_CHECK("k_cg_init");
  PcgResult res;
  Batcher bt;
  for (;;) {
    const CgState h = read_state(ctx, st);
    res.iterations = h.iteration;
    res.rnorm = h.rnorm;
    res.bnorm = h.bnorm;
    res.converged = h.converged != 0;
    if (h.stop) break;
    const long long batch = bt.next(h.rnorm, h.iteration,
                                    std::max(h.tol_abs, h.abs_tol),
                                    max_iter - h.iteration);
    for (long long i = 0; i < batch; ++i)
      enqueue_pass<double>(s, o, val, x, r, z, p, q, dinv, st, h.iteration + i,
                           part1, part2);
    GS_LAUNCH_CHECK("cg batch");
  }
  // one SpMV per pass, the converging (uncounted) pass included
  res.spmv = res.iterations + (res.converged && res.bnorm >= 1e-25 ? 1 : 0);
  return res;
}

// FP32 Jacobi-PCG with FP64 residual replacement; b, x in SELL order.
PcgResult pcg_mixed(gs_ctx *ctx, const gs_csr *A, const double *b, double *x,
                    bool precond, const gs_cg_opts &opt, char *base, Layout &L) {
  hipStream_t s = ctx->stream;
  const SellOp o = sell_op(A);
  const int n = o.n();
  double *part1 = (double *)(base + L.add<double>(GRID_MAX));
  double *part2 = (double *)(base + L.add<double>(2 * GMAX));
  CgState *st = (CgState *)(base + L.add<CgState>(1));
  double *r64 = (double *)(base + L.add<double>(n));
  float *xs = (float *)(base + L.add<float>(n));
  float *rs = (float *)(base + L.add<float>(n));
  float *zs = (float *)(base + L.add<float>(n));
  float *p = (float *)(base + L.add<float>(n));
  float *q = (float *)(base + L.add<float>(n));
  const float *dinv = precond ? A->sdinv32.get() : nullptr;
  const float *val32 = A->sval32.get();
  const double *val = A->sval.get();
  const unsigned gv = grid_vec(n);
  const long long max_iter = opt.max_iterations;
  const long long max_outer =
      opt.max_outer_iterations > 0 ? opt.max_outer_iterations : 100;
  const long long inner_cap = opt.max_inner_iterations;
  hipLaunchKernelGGL(k_state_setup, 1, 1, 0, s, st, max_iter, opt.tolerance,
                     opt.abs_tolerance, 1, opt.inner_tolerance);
  hipLaunchKernelGGL(k_mx_start, gv, BS, 0, s, n, b, x, r64, part1);
  hipLaunchKernelGGL(k_fin_norm, 1, FBS, 0, s, st, part1, (int)gv, 1);
  GS_LAUNCH_CHECK("k_mx_start");
  CgState h = read_state(ctx, st);
  PcgResult res;
  res.bnorm = h.bnorm;
  res.rnorm = h.bnorm;
  if (h.bnorm < 1e-25) {  // src/linalg/iterative.rs:60-68
    res.converged = true;
    res.rnorm = 0.0;
    return res;
  }
  if (max_iter <= 0) return res;
  auto rescale = [&](bool restart) {
    if (dinv) {
      if (restart)
        hipLaunchKernelGGL((k_mx_rescale<true, true>), gv, BS, 0, s, n, r64, dinv,
                           rs, zs, p, xs, st, part2);
      else
        hipLaunchKernelGGL((k_mx_rescale<true, false>), gv, BS, 0, s, n, r64, dinv,
                           rs, zs, p, xs, st, part2);
    } else {
      if (restart)
        hipLaunchKernelGGL((k_mx_rescale<false, true>), gv, BS, 0, s, n, r64, dinv,
                           rs, zs, p, xs, st, part2);
      else
        hipLaunchKernelGGL((k_mx_rescale<false, false>), gv, BS, 0, s, n, r64, dinv,
                           rs, zs, p, xs, st, part2);
    }
    hipLaunchKernelGGL(k_fin_rescale, 1, FBS, 0, s, st, part2, (int)gv);
    GS_LAUNCH_CHECK("k_mx_rescale");
  };
  rescale(true);
  h = read_state(ctx, st);
  Batcher bt;
  long long last_rep = 0;  // pass index of the last replacement
  int breakdowns = 0;
  for (;;) {
    bool replace = h.pause != 0;
    if (!replace && inner_cap > 0 && h.iteration - last_rep >= inner_cap) replace = true;
    if (replace) {
      hipLaunchKernelGGL(k_mx_fold, gv, BS, 0, s, n, xs, x, st);
      launch_residual(s, o, val, x, b, r64, part1);
      hipLaunchKernelGGL(k_fin_norm, 1, FBS, 0, s, st, part1, o.grid, 0);
      GS_LAUNCH_CHECK("residual replacement");
      h = read_state(ctx, st);
      ++res.outer;
      ++res.spmv;
      res.rnorm = h.rnorm;
      res.iterations = h.iteration;
      if (h.rnorm < h.tol_abs || h.rnorm < h.abs_tol) {
        res.converged = true;
        break;
      }
      if (h.iteration >= max_iter || res.outer >= max_outer) break;
      if (h.breakdown) {
        if (++breakdowns >= 2) break;
      } else {
        breakdowns = 0;
      }
      // Keep the search direction only while the true residual agrees with
      // the recursive one; a large gap (the FP32 operator's rounding limits
      // what one cycle can gain) would make the next step length
      // inconsistent with p, so the recursion restarts from p = z instead.
      const double gap = h.rnorm * h.scale / std::max(h.rrec, 1e-300);
      rescale(h.breakdown != 0 || gap > 2.0);
      if (gap > 2.0) ++res.restarts;
      last_rep = h.iteration;
      h = read_state(ctx, st);
    }
    long long left = max_iter - h.iteration;
    if (inner_cap > 0) left = std::min(left, inner_cap - (h.iteration - last_rep));
    const double target = std::max({h.tol, h.rtol * h.rbase, h.abs_tol / h.bnorm});
    const long long batch = bt.next(h.rrec, h.iteration, target, left);
    for (long long i = 0; i < batch; ++i)
      enqueue_pass<float>(s, o, val32, xs, rs, zs, p, q, dinv, st, h.iteration + i,
                          part1, part2);
    GS_LAUNCH_CHECK("cg batch");
    const long long before = h.iteration;
    h = read_state(ctx, st);
    res.spmv += h.iteration - before + (h.breakdown ? 1 : 0);
  }
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

// The SELL view and its value/diagonal arrays, built once per matrix.
static void ensure_sell(gs_ctx *ctx, gs_csr *A, bool fp32, bool precond) {
  const bool force32 = (A->spmv_flags & GS_SPMV_INDEX32) != 0;
  hipStream_t s = ctx->stream;
  if (!A->sell || A->sell_force32 != force32) {
    A->sell = build_sell_pattern(ctx, A, force32);
    A->sell_force32 = force32;
    A->sval = DArr<double>();
    A->sval32 = DArr<float>();
  }
  const SellPattern &P = *A->sell;
  if (!A->sval) {
    A->sval = dalloc<double>(P.entries);
    sell_values<double>(s, P, A->val.get(), A->sval.get());
  }
  if (fp32 && !A->sval32) {
    A->sval32 = dalloc<float>(P.entries);
    sell_values<float>(s, P, A->val.get(), A->sval32.get());
  }
  if (precond && (!A->sdinv || A->sdinv.n != P.n)) {
    A->sdinv = dalloc<double>(P.n);
    A->sdinv32 = dalloc<float>(P.n);
    sell_jacobi(s, P, A, A->sdinv.get(), A->sdinv32.get());
  }
}

// 16-bit anchored column offsets for the CSR engines, built once per matrix;
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

static bool use_sell(const gs_csr *A) {
  return A->rows == A->cols && !(A->spmv_flags & GS_SPMV_CSR);
}

void csr_spmv(gs_ctx *ctx, const gs_csr *Ac, const double *x, double *y) {
  gs_csr *A = const_cast<gs_csr *>(Ac);  // derived caches only
  if (A->rows == 0) return;
  hipStream_t s = ctx->stream;
  const int64_t n = A->rows;
  if (use_sell(A)) {
    ensure_sell(ctx, A, false, false);
    Layout L;
    const size_t ox = L.add<double>(n), oy = L.add<double>(n);
    char *base = ctx_scratch(ctx, L.bytes);
    double *xi = (double *)(base + ox), *yi = (double *)(base + oy);
    const int *vp = A->sell->vperm.get();
    permute_gather(s, n, x, vp, xi);
    launch_sell_spmv<double>(s, sell_op(A), A->sval.get(), xi, yi);
    permute_scatter(s, n, yi, vp, y);
  } else {
    ensure_idx16(ctx, A);
    const Shape sh = csr_shape(A, false);
    const int *dp = A->dperm.get();
    if (!dp) {
      launch_csr_spmv<double>(s, sh, A, A->val.get(), x, y);
    } else {
      Layout L;
      const size_t ox = L.add<double>(n), oy = L.add<double>(n);
      char *base = ctx_scratch(ctx, L.bytes);
      double *xi = (double *)(base + ox), *yi = (double *)(base + oy);
      permute_gather(s, n, x, dp, xi);
      launch_csr_spmv<double>(s, sh, A, A->val.get(), xi, yi);
      permute_scatter(s, n, yi, dp, y);
    }
  }
  GS_LAUNCH_CHECK("spmv");
  GS_HIP(hipStreamSynchronize(s));
}

void cg_solve(gs_ctx *ctx, const gs_csr *Ac, const double *b,
              const gs_cg_opts &o, double *x, gs_solver_stats *st) {
  gs_csr *A = const_cast<gs_csr *>(Ac);  // derived caches only
  const auto t0 = std::chrono::steady_clock::now();
  memset(st, 0, sizeof(*st));
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "CG needs a square matrix");
  const int64_t n = A->rows;
  if (n == 0) {  // ||b|| = 0 < 1e-25: converged at once
    st->converged = 1;
    return;
  }
  hipStream_t s = ctx->stream;
  const bool mixed = o.precision != GS_PREC_FP64;
  const bool precond = o.use_preconditioner != 0;
  if (A->spmv_flags & GS_SPMV_CSR)
    throw Error(GS_ERR_INVALID_ARG, "CG runs on the SELL engine only");
  ensure_sell(ctx, A, mixed, precond);
  Layout L;
  const size_t ob = L.add<double>(n), ox = L.add<double>(n);
  // solver scratch follows; size it with a dry layout of the same pieces
  Layout dry = L;
  dry.add<double>(GRID_MAX);
  dry.add<double>(2 * GMAX);
  dry.add<CgState>(1);
  if (mixed) {
    dry.add<double>(n);
    for (int i = 0; i < 5; ++i) dry.add<float>(n);
  } else {
    for (int i = 0; i < 4; ++i) dry.add<double>(n);
  }
  char *base = ctx_scratch(ctx, dry.bytes);
  double *bi = (double *)(base + ob), *xi = (double *)(base + ox);
  const int *vp = A->sell->vperm.get();
  permute_gather(s, n, b, vp, bi);
  PcgResult r;
  if (!mixed)
    r = pcg64(ctx, A, bi, xi, precond, o.max_iterations, o.tolerance,
              o.abs_tolerance, base, L);
  else
    r = pcg_mixed(ctx, A, bi, xi, precond, o, base, L);
  permute_scatter(s, n, xi, vp, x);
  GS_LAUNCH_CHECK("cg_solve");
  GS_HIP(hipStreamSynchronize(s));
  st->iterations = r.iterations;
  st->residual_norm = r.rnorm;
  st->relative_residual = r.bnorm > 1e-20 ? r.rnorm / r.bnorm : 0.0;
  st->converged = r.converged;
  st->outer_iterations = r.outer;
  st->spmv_count = r.spmv;
  st->solve_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

double time_spmv(gs_ctx *ctx, const gs_csr *Ac, int precision, int reps) {
  gs_csr *A = const_cast<gs_csr *>(Ac);
  const int64_t n = A->rows;
  hipStream_t s = ctx->stream;
  if (n == 0 || reps <= 0) return 0.0;
  const bool fp32 = precision != GS_PREC_FP64;
  const bool sell = use_sell(A);
  if (sell) ensure_sell(ctx, A, fp32, false);
  else ensure_idx16(ctx, A);
  if (!sell && fp32 && !A->val32) {
    A->val32 = dalloc<float>(A->nnz);
    hipLaunchKernelGGL(k_to_f32, (unsigned)GMAX, BS, 0, s, A->nnz, A->val.get(),
                       A->val32.get());
  }
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  Layout L;
  const size_t opart = L.add<double>(GRID_MAX), ost = L.add<CgState>(1);
  const size_t op = L.add<char>((int64_t)(n * esz)), oq = L.add<char>((int64_t)(n * esz));
  char *base = ctx_scratch(ctx, L.bytes);
  double *part = (double *)(base + opart);
  CgState *st = (CgState *)(base + ost);
  void *p = base + op, *q = base + oq;
  hipLaunchKernelGGL(k_state_setup, 1, 1, 0, s, st, 1ll << 62, 0.0, 0.0, 0, 0.0);
  hipEvent_t e0, e1;
  GS_HIP(hipEventCreate(&e0));
  GS_HIP(hipEventCreate(&e1));
  auto run = [&](auto tag) {
    using T = decltype(tag);
    const T *val = fp32 ? (const T *)(sell ? (const void *)A->sval32.get()
                                           : (const void *)A->val32.get())
                        : (const T *)(sell ? (const void *)A->sval.get()
                                           : (const void *)A->val.get());
    hipLaunchKernelGGL((k_fill_pattern<T>), grid_vec(n), BS, 0, s, (int)n, (T *)p);
    auto once = [&] {
      if (sell)
        launch_cg_q<T>(s, sell_op(A), val, (const T *)p, (T *)q, st, 0, part);
      else
        launch_csr_spmv<T>(s, csr_shape(A, fp32), A, val, (const T *)p, (T *)q);
    };
    once();
    GS_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) once();
    GS_HIP(hipEventRecord(e1, s));
  };
  if (fp32) run(float{});
  else run(double{});
  GS_LAUNCH_CHECK("time_spmv");
  float ms = 0.f;
  GS_HIP(hipEventSynchronize(e1));
  GS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

}  // namespace gs
