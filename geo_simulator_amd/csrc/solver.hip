// Jacobi-preconditioned CG on gfx950: FP64, and mixed precision (FP64
// flexible CG around an FP32 recursion).
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
// Mixed precision: FP64 flexible CG (Notay's FCG with one previous
// direction kept) whose preconditioner is a fresh FP32 Jacobi-PCG solve of the
// current residual (FP32 SELL values, vectors and SpMV; dot products
// accumulated in FP64) to relative tolerance inner_tolerance.  Every outer
// step is a Galerkin step with the exact FP64 operator: residual, solution
// and search directions stay in FP64, the A-norm error never grows, and the
// stopping test is the reference's on the FP64 residual (checked against
// b - A x before returning).  Unlike plain defect correction (x += A32^-1 r),
// this converges also when eps32 * cond(A) is not small, as for the
// heterogeneous elastic operators.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "internal.hpp"
#include "sell.hpp"
#include "spmv.hpp"

namespace gs {
// the SELL view's arrays of A (values in FP64, and FP32 with fp32; the Jacobi
// diagonal with precond), made where missing
static void ensure_sell(gs_ctx *ctx, gs_csr *A, bool fp32, bool precond);

namespace {

constexpr int BS = 256;
constexpr int FBS = 1024;        // finalize kernels: one workgroup
constexpr int GMAX = 2048;       // workgroups of the grid-stride vector kernels
constexpr int VGMAX = 256;       // workgroups (of FBS threads) of a pass's vector kernels
constexpr int GRID_MAX = 16384;  // workgroups of any SpMV kernel
constexpr int UNR = 4;           // rows per thread per trip in a pass's vector kernels

struct CgState {
  double rz[2];  // r.z at the start of pass k lives in rz[k & 1]
  double alpha, beta, pq;
  double rnorm, bnorm, tol_abs, abs_tol, tol;
  double rbase;  // mixed inner: ||r|| (scaled) when the recursion (re)started
  double rtol;   // mixed inner: pause once ||r|| < rtol * rbase
  double gam[2];       // scaled recursion: r.r at the start of pass k in gam[k & 1]
  long long fresh_it;  // scaled recursion: the pass that starts a new direction
  long long alpha_it;  // the last pass whose K2 made its update (written by K2 only)
  long long iteration, max_iter;
  int stop, converged, pause, mixed;
  int brk_rel;   // break down only on p.Ap == 0 or a non-finite step
  // flexible CG (mixed precision), outer steps decided on the device
  // (k_fcg_decide, k_fcg_decide2): the outer residual norm after the last
  // step, the residual the recursion predicted for it, the steps taken and
  // the true-residual checks run, the next recursion's restart flag, a
  // pending check, and how the solve ended (1 converged, 2 no descent left;
  // stop is set with it).  Only those kernels set stop in a mixed solve.
  double outer_rn, outer_pred;
  long long outer_steps, outer_checks;
  int outer_restart, outer_check, outer_end;
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

// Finalize kernels (workgroups of NT threads): sum of np x NV partials, in
// every thread.
template <int NV, int NT = FBS>
__device__ __forceinline__ void sum_partials(const double *__restrict__ part,
                                             int np, double (&tot)[NV]) {
#pragma unroll
  for (int i = 0; i < NV; ++i) tot[i] = 0.0;
  for (int b = threadIdx.x; b < np; b += NT) {
#pragma unroll
    for (int i = 0; i < NV; ++i) tot[i] += part[b * NV + i];
  }
  block_sum<NV, NT>(tot);
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
    double *__restrict__ part, const int *__restrict__ gate) {
  if (gate && *gate == 0) return;  // a check the device did not ask for
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
// workgroup 0 records the scalars.  A kernel's entry test reads only words
// that an earlier kernel wrote, or words whose change by its own leader leads
// every workgroup to the same decision anyway: K3 runs when K2 of the same
// pass recorded its update in alpha_it, a word K3 never writes (K3's leader
// advances `iteration` and may set `pause` while other workgroups of K3 are
// still starting, so those two cannot gate K3).

// K2 (pass it): pq; break if |pq| < 1e-30 (src/linalg/iterative.rs:85);
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
  // The mixed inner solve works on a scaled FP32 system whose units make the
  // reference's absolute 1e-30 meaningless; it breaks down only on a zero or
  // non-finite step, and then pauses for the outer step.  A solve asked for
  // the same scale-free test (brk_rel) stops there.
  const bool brk = (st->mixed || st->brk_rel) ? (pq[0] == 0.0 || !isfinite(alpha_d))
                                              : fabs(pq[0]) < 1e-30;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq[0];
    if (brk) {
      if (st->mixed) st->pause = 1;
      else st->stop = 1;
    } else {
      st->alpha = alpha_d;
      st->alpha_it = it;
    }
  }
  if (brk) return;
  const T alpha = (T)alpha_d;
  double acc[2] = {0.0, 0.0};
  // UNR rows per thread per trip, all loads issued before the first use
  // (rows, and so the partial sums, in the same order as a plain loop)
  const long long stride = (long long)gridDim.x * FBS;
  for (long long i0 = (long long)blockIdx.x * FBS + threadIdx.x; i0 < n;
       i0 += UNR * stride) {
    T pv[UNR], qv[UNR], xv[UNR], rv[UNR], dv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        pv[u] = p[i];
        qv[u] = q[i];
        xv[u] = x[i];
        rv[u] = r[i];
        if (PREC) dv[u] = dinv[i];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        x[i] = xv[u] + alpha * pv[u];
        const T ri = rv[u] - alpha * qv[u];
        r[i] = ri;
        const T zi = PREC ? ri * dv[u] : ri;
        z[i] = zi;
        acc[0] += (double)ri * (double)ri;
        acc[1] += (double)ri * (double)zi;
      }
    }
  }
  block_sum<2, FBS>(acc);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

// K3 (pass it): the reference's stopping test (src/linalg/iterative.rs:95-98;
// the converging pass is not counted), then beta = rz'/rz, it += 1 (stop at
// max_iter), p = z + beta p.  As the inner solver of the mixed path it never
// stops by itself: it completes the pass and pauses for the outer FP64 step
// once ||r|| fell below rtol * rbase, below the outer target tol, or the pass
// budget is spent.
template <typename T>
__global__ __launch_bounds__(FBS) void k_cg_beta(
    int n, long long it, const T *__restrict__ z, T *__restrict__ p,
    CgState *__restrict__ st, const double *__restrict__ part_in, int np_in) {
  if (st->alpha_it != it) return;
  double tot[2];
  sum_partials<2>(part_in, np_in, tot);
  const double rn = sqrt(tot[0]);
  const bool mixed = st->mixed != 0;
  const bool conv = !mixed && (rn < st->tol_abs || rn < st->abs_tol);
  const double beta_d = tot[1] / st->rz[it & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rnorm = rn;
    if (conv) {
      st->converged = 1;
      st->stop = 1;
    } else {
      st->beta = beta_d;
      st->rz[(it + 1) & 1] = tot[1];
      st->iteration = it + 1;
      if (!mixed) {
        if (it + 1 >= st->max_iter) st->stop = 1;
      } else if (rn < st->rtol * st->rbase || rn < st->tol ||
                 it + 1 >= st->max_iter) {
        st->pause = 1;
      }
    }
  }
  if (conv) return;
  const T beta = (T)beta_d;
  const long long stride = (long long)gridDim.x * FBS;
  for (long long i0 = (long long)blockIdx.x * FBS + threadIdx.x; i0 < n;
       i0 += UNR * stride) {
    T zv[UNR], pv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        zv[u] = z[i];
        pv[u] = p[i];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) p[i] = zv[u] + beta * pv[u];
    }
  }
}

// ---- two-kernel pass of the scaled recursion (FP32 vectors) -------------------
// CG on S A S (S = D^-1/2: Jacobi folded into the operator, so z = r; its
// values in int32 fixed point, 2^-30 resolution, FP32 vectors) with
// the direction update moved into the SpMV kernel.  A p_k is not formed from
// p_k (whose entries other lanes are still writing) but by linearity,
// s_k = A p_k = A r_k + beta_k s_{k-1}, the recurrence of Chronopoulos and
// Gear (J. Comput. Appl. Math. 25, 1989); the step length uses the exact
// products of the pass, alpha = (r, p) / (p, s), an exact line search along
// p.  Per pass:
//   KA  gamma = r.r (the previous pass's partials); pause test on ||r||;
//       beta = gamma / gamma_prev (0 on a fresh start);
//       w = A r; p = r + beta p; s = w + beta s; partials p.s, r.p
//   KB  alpha = r.p / p.s; x += alpha p; r -= alpha s; partials r.r
// Two launches and two reductions per pass instead of three of each.
template <typename T, class IDX, typename TV>
__global__ __launch_bounds__(BS) void k_cg2_dir(int n, int ns, const int *__restrict__ cb,
                                                const IDX idx, const TV *__restrict__ val,
                                                const T *__restrict__ r,
                                                T *__restrict__ p, T *__restrict__ sv,
                                                CgState *__restrict__ st, long long it,
                                                const double *__restrict__ part_r,
                                                int np_r, double *__restrict__ part) {
  if (!active(st, it)) return;
  double g[1];
  sum_partials<1, BS>(part_r, np_r, g);
  const double rn = sqrt(g[0]);
  const bool pause = rn < st->rtol * st->rbase || rn < st->tol || it >= st->max_iter;
  const bool fresh = st->fresh_it == it;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rnorm = rn;
    st->gam[it & 1] = g[0];
    if (pause) st->pause = 1;
  }
  if (pause) return;
  double beta_d = fresh ? 0.0 : g[0] / st->gam[(it - 1) & 1];
  if (!isfinite(beta_d)) beta_d = 0.0;
  const T beta = (T)beta_d;
  double acc[2] = {0.0, 0.0};
  auto epi = [&](int row, T w) {
    const T ri = r[row];
    T pi = ri, si = w;
    if (!fresh) {
      pi = ri + beta * p[row];
      si = w + beta * sv[row];
    }
    p[row] = pi;
    sv[row] = si;
    acc[0] += (double)pi * (double)si;
    acc[1] += (double)ri * (double)pi;
  };
  sell_rows<T>(n, ns, cb, idx, val, r, epi);
  put_partials<2>(acc, part);
}

template <typename T>
__global__ __launch_bounds__(FBS) void k_cg2_step(
    int n, long long it, T *__restrict__ x, T *__restrict__ r, const T *__restrict__ p,
    const T *__restrict__ sv, CgState *__restrict__ st,
    const double *__restrict__ part_in, int np_in, double *__restrict__ part) {
  const long long at = st->iteration;
  if (!((at == it && !st->stop && !st->pause) || at == it + 1)) return;
  double t[2];
  sum_partials<2>(part_in, np_in, t);
  const double alpha_d = t[1] / t[0];
  // no positive curvature left along p, or a non-finite step: pause for the
  // outer step (which then restarts the recursion)
  const bool brk = !(t[0] > 0.0) || !isfinite(alpha_d);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (brk) {
      st->pq = 0.0;
      st->pause = 1;
    } else {
      st->pq = t[0];
      st->alpha = alpha_d;
      st->iteration = it + 1;
    }
  }
  if (brk) return;
  const T alpha = (T)alpha_d;
  double acc[1] = {0.0};
  const long long stride = (long long)gridDim.x * FBS;
  for (long long i0 = (long long)blockIdx.x * FBS + threadIdx.x; i0 < n;
       i0 += UNR * stride) {
    T pv[UNR], sq[UNR], xv[UNR], rv[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        pv[u] = p[i];
        sq[u] = sv[i];
        xv[u] = x[i];
        rv[u] = r[i];
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) {
        x[i] = xv[u] + alpha * pv[u];
        const T ri = rv[u] - alpha * sq[u];
        r[i] = ri;
        acc[0] += (double)ri * (double)ri;
      }
    }
  }
  block_sum<1, FBS>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
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

__global__ void k_state_setup(CgState *__restrict__ st, long long max_iter,
                              double tol, double abs_tol, int mixed, double rtol,
                              int brk_rel) {
  CgState s{};
  s.alpha_it = -1;
  s.brk_rel = brk_rel;
  s.max_iter = max_iter;
  s.tol = tol;
  s.abs_tol = abs_tol;
  s.mixed = mixed;
  s.rtol = rtol;
  *st = s;
}

// Set-up, single workgroup: ||b||, rz, thresholds, early exits
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

// ---- flexible-CG (mixed precision) kernels ----------------------------------
// Sums of NV partials -> out[0..NV) (single workgroup).
template <int NV>
__global__ __launch_bounds__(FBS) void k_fin_sum(const double *__restrict__ part,
                                                 int np, double *__restrict__ out) {
  double tot[NV];
  sum_partials<NV>(part, np, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) out[i] = tot[i];
  }
}

// x = 0, r = b; partials b.b
__global__ __launch_bounds__(BS) void k_fcg_start(int n, const double *__restrict__ b,
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

// Inner FP32 recursion (re)started from the FP64 residual: rs = r/||b||,
// zs = D^-1 rs [and p = zs on a restart]; partials rs.zs
// With dev (the device's outer-step decision), nothing once the solve has
// stopped, and the restart flag from the state instead of RESTART.
template <bool PREC, bool RESTART>
__global__ __launch_bounds__(BS) void k_fcg_inner(
    int n, const double *__restrict__ r, double scale,
    const float *__restrict__ dinv, float *__restrict__ rs, float *__restrict__ zs,
    float *__restrict__ p, double *__restrict__ part, const CgState *__restrict__ dev) {
  if (dev && dev->stop) return;
  const bool restart = dev ? dev->outer_restart != 0 : RESTART;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const float ri = (float)(r[i] * scale);
    rs[i] = ri;
    const float zi = PREC ? ri * dinv[i] : ri;
    zs[i] = zi;
    if (restart) p[i] = zi;
    acc[0] += (double)ri * (double)zi;
  }
  put_partials<1>(acc, part);
}

// The scaled recursion (re)started: rs = S r/||b||; partials rs.rs, one per
// workgroup of FBS threads (the layout of k_cg2_step's partials).
__global__ __launch_bounds__(FBS) void k_fcg_inner_scaled(
    int n, const double *__restrict__ r, double scale, const double *__restrict__ ssc,
    float *__restrict__ rs, double *__restrict__ part, const CgState *__restrict__ dev) {
  if (dev && dev->stop) return;  // the device ended the solve
  double acc[1] = {0.0};
  for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS) {
    const float ri = (float)(ssc[i] * r[i] * scale);
    rs[i] = ri;
    acc[0] += (double)ri * (double)ri;
  }
  block_sum<1, FBS>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// Inner recursion state after a (re)start: rz, rbase; resume passes.  With
// rbase < 0 the recursion's residual is its own r.z (the scaled recursion,
// z = r): rbase = sqrt(r.z), and the pause threshold becomes tol_ratio
// times it.
__global__ __launch_bounds__(FBS) void k_fin_inner(CgState *__restrict__ st,
                                                   const double *__restrict__ part,
                                                   int np, double rbase, double tol_ratio,
                                                   int restart, double outer_rn) {
  double tot[1];
  sum_partials<1>(part, np, tot);
  if (threadIdx.x != 0) return;
  st->outer_rn = outer_rn;
  st->rz[st->iteration & 1] = tot[0];
  if (rbase < 0.0) {
    rbase = sqrt(tot[0]);
    st->tol = rbase * tol_ratio;
  }
  st->fresh_it = restart ? st->iteration : -1;
  st->rbase = rbase;
  st->rnorm = rbase;
  st->pause = 0;
}

// k_fin_inner on the device's outer-step decision: the outer residual and
// the restart flag come from the state (rbase = rn scale, or the scaled
// recursion's own sqrt(r.r) with the pause threshold target / rn of it);
// nothing once the solve has stopped.
__global__ __launch_bounds__(FBS) void k_fin_inner_dev(CgState *__restrict__ st,
                                                       const double *__restrict__ part, int np,
                                                       int scaled, double scale,
                                                       double target) {
  if (st->stop) return;
  double tot[1];
  sum_partials<1>(part, np, tot);
  if (threadIdx.x != 0) return;
  const double rn = st->outer_rn;
  st->rz[st->iteration & 1] = tot[0];
  double rbase = rn * scale;
  if (scaled) {
    rbase = sqrt(tot[0]);
    st->tol = rbase * (target / rn);
  }
  st->fresh_it = st->outer_restart ? st->iteration : -1;
  st->rbase = rbase;
  st->rnorm = rbase;
  st->pause = 0;
}

// The outer step's decision on the device, first half (one thread), after
// the step's scalars sc (gamma, alpha, p.Ap, taken, r.r): no descent left
// ends the solve; else the new outer residual, the residual the recursion
// predicted for it (its own reduction applied to the previous outer
// residual), and whether to confirm convergence on b - A x.
__global__ void k_fcg_decide(CgState *__restrict__ st, const double *__restrict__ sc,
                             double target) {
  if (sc[3] == 0.0) {
    st->outer_end = 2;
    st->outer_check = 0;
    st->stop = 1;
    return;
  }
  st->outer_pred = st->rbase > 0.0 ? st->outer_rn * (st->rnorm / st->rbase) : 0.0;
  const double rn = sqrt(sc[4]);
  st->outer_rn = rn;
  st->outer_steps += 1;
  st->outer_check = rn < target ? 1 : 0;
}

// Second half, after the gated check (dots[0] = |b - A x|^2 when it ran):
// converged ends the solve; else the recursion keeps its direction when the
// step agreed with it (alpha near 1, the new residual near the predicted one,
// a positive curvature) and the host saw no stall, and restarts otherwise.
__global__ void k_fcg_decide2(CgState *__restrict__ st, const double *__restrict__ sc,
                              const double *__restrict__ dots, double target, int stalled) {
  if (st->stop) return;
  double rn = st->outer_rn;
  if (st->outer_check) {
    st->outer_checks += 1;
    rn = sqrt(dots[0]);
    st->outer_rn = rn;
    if (rn < target) {
      st->outer_end = 1;
      st->stop = 1;
      return;
    }
  }
  const bool agree = !stalled && fabs(sc[1] - 1.0) < 0.25 && rn < 2.0 * st->outer_pred &&
                     !(st->pq == 0.0);
  st->outer_restart = agree ? 0 : 1;
}

// d = ||b|| xs (the inner correction in FP64 and in units of b; S xs for
// the scaled recursion); xs = 0
__global__ __launch_bounds__(BS) void k_fcg_expand(int n, float *__restrict__ xs,
                                                   double scale,
                                                   const double *__restrict__ ssc,
                                                   double *__restrict__ d) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    d[i] = ssc ? scale * (ssc[i] * (double)xs[i]) : scale * (double)xs[i];
    xs[i] = 0.0f;
  }
}

// Fixed-point values of S A S, S = diag(ssc) (scalar SELL order), from the
// FP64 values: entry (i, j) becomes round(s_i a_ij s_j 2^30) as int32; padding
// 0.  With a positive Jacobi diagonal the scaled entries of a symmetric
// positive definite operator lie in [-1, 1] (the diagonal is 1); an entry
// outside (-2, 2) sets *bad, and the solve then keeps the unscaled operator.
// One wavefront per slice, blocks of SELL_WPB slices in XCD-contiguous
// ranges as the SpMV, so the gathered s_j stay in the XCD's L2.
template <class IDX>
__global__ __launch_bounds__(BS) void k_scale_sell(int n, int ns, const int *__restrict__ cb,
                                                   const IDX idx,
                                                   const double *__restrict__ val,
                                                   const double *__restrict__ ssc,
                                                   int *__restrict__ out, int *bad) {
  static_assert(IDX::bs == 1, "scalar SELL only");
  typedef double D4 __attribute__((ext_vector_type(4)));
  typedef int Q4 __attribute__((ext_vector_type(4)));
  const int nb = (ns + SELL_WPB - 1) / SELL_WPB;
  const int G = gridDim.x, gi = blockIdx.x;
  int first = gi, last = nb, stride = G;
  if (G >= 8) {
    const int xcd = gi & 7;
    first = (int)((int64_t)nb * xcd / 8) + (gi >> 3);
    last = (int)((int64_t)nb * (xcd + 1) / 8);
    stride = (G >> 3) + (xcd < (G & 7) ? 1 : 0);
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  bool out_of_range = false;
  for (int b = first; b < last; b += stride) {
    const int sl = b * SELL_WPB + w;
    if (sl >= ns) continue;
    const int row = sl * SELL_C + lane;
    const bool live = row < n;
    const double si = live ? ssc[row] : 0.0;
    for (int c = cb[sl]; c < cb[sl + 1]; ++c) {
      const size_t e0 = (size_t)c * SELL_CHUNK + lane * SELL_CW;
      const D4 v = *(const D4 *)(val + e0);
      int cc[4];
      bool ok[4];
      if constexpr (IDX::wide) {
        typedef int I4 __attribute__((ext_vector_type(4)));
        const I4 i4 = *(const I4 *)(idx.col + e0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ok[j] = i4[j] >= 0 && live;
          cc[j] = i4[j];
        }
      } else {
        typedef short S4 __attribute__((ext_vector_type(4)));
        const S4 i4 = *(const S4 *)(idx.off + e0);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          ok[j] = i4[j] != SELL_PAD16 && live;
          cc[j] = sl * SELL_C + i4[j];
        }
      }
      Q4 o;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double sj = ok[j] ? ssc[cc[j]] : 0.0;
        const double t = ok[j] ? si * v[j] * sj * 0x1p30 : 0.0;
        const bool fits = t > -0x1.fffp30 && t < 0x1.fffp30;  // false for NaN
        out_of_range |= !fits;
        o[j] = fits ? (int)rint(t) : 0;
      }
      *(Q4 *)(out + e0) = o;
    }
  }
  if (out_of_range) *bad = 1;
}

// Timing set-up of the two-launch pass: a pass that continues a direction.
__global__ void k_state_continue(CgState *__restrict__ st) { st->fresh_it = -1; }

// Bad flag (an int) as a double, beside the dots the host reads next.
__global__ void k_flag_to_dot(const int *__restrict__ flag, double *__restrict__ out) {
  *out = (double)*flag;
}

// ssc = sqrt(d) when every d > 0 and finite; flag = 1 otherwise
__global__ __launch_bounds__(BS) void k_sqrt_dinv(int n, const double *__restrict__ d,
                                                  double *__restrict__ ssc,
                                                  int *__restrict__ flag) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const double v = d[i];
    if (!(v > 0.0) || !isfinite(v)) *flag = 1;
    ssc[i] = sqrt(v);
  }
}

// w = A d (FP64 SELL); partials d.Ap_prev
template <class IDX>
__global__ __launch_bounds__(BS) void k_fcg_spmv(
    int n, int ns, const int *__restrict__ cb, const IDX idx,
    const double *__restrict__ val, const double *__restrict__ d,
    double *__restrict__ w, const double *__restrict__ ap_prev,
    double *__restrict__ part) {
  double acc[1] = {0.0};
  auto epi = [&](int row, double s) {
    w[row] = s;
    if (ap_prev) acc[0] += d[row] * ap_prev[row];
  };
  sell_rows<double>(n, ns, cb, idx, val, d, epi);
  put_partials<1>(acc, part);
}

// p = d - gamma p; ap = w - gamma ap (first step: p = d, ap = w);
// partials p.r, p.ap
// Outer-step scalars kept on the device (one read by the host per step):
// sc[0] gamma, sc[1] alpha, sc[2] p.Ap of the last direction, sc[3] 1 when
// the step is taken (p.Ap > 0 and a finite alpha), sc[4] r.r after it.
// gamma = (d, A p_prev) / (p_prev, A p_prev); 0 on the first step.
__global__ __launch_bounds__(FBS) void k_fcg_gamma(const double *__restrict__ part, int np,
                                                   int first, double *__restrict__ sc) {
  if (first) {
    if (threadIdx.x == 0) sc[0] = 0.0;
    return;
  }
  double t[1];
  sum_partials<1>(part, np, t);
  if (threadIdx.x == 0) sc[0] = t[0] / sc[2];
}

// alpha = (p, r) / (p, Ap) when p.Ap > 0 and the ratio is finite, else the
// step is not taken (no descent left)
__global__ __launch_bounds__(FBS) void k_fcg_alpha(const double *__restrict__ part, int np,
                                                   double *__restrict__ sc) {
  double t[2];
  sum_partials<2>(part, np, t);
  if (threadIdx.x != 0) return;
  const bool ok = t[1] > 0.0 && isfinite(t[0] / t[1]);
  sc[1] = ok ? t[0] / t[1] : 0.0;
  if (ok) sc[2] = t[1];
  sc[3] = ok ? 1.0 : 0.0;
}

__global__ __launch_bounds__(BS) void k_fcg_dir(int n, const double *__restrict__ d,
                                                const double *__restrict__ w,
                                                const double *__restrict__ r,
                                                const double *__restrict__ sc, int first,
                                                double *__restrict__ p,
                                                double *__restrict__ ap,
                                                double *__restrict__ part) {
  const double gamma = sc[0];
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const double pi = first ? d[i] : d[i] - gamma * p[i];
    const double ai = first ? w[i] : w[i] - gamma * ap[i];
    p[i] = pi;
    ap[i] = ai;
    acc[0] += pi * r[i];
    acc[1] += pi * ai;
  }
  put_partials<2>(acc, part);
}

// x += alpha p; r -= alpha ap; partials r.r (nothing when the step is
// not taken)
__global__ __launch_bounds__(BS) void k_fcg_update(int n, const double *__restrict__ sc,
                                                   const double *__restrict__ p,
                                                   const double *__restrict__ ap,
                                                   double *__restrict__ x,
                                                   double *__restrict__ r,
                                                   double *__restrict__ part) {
  if (sc[3] == 0.0) return;
  const double alpha = sc[1];
  double acc[1] = {0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    acc[0] += ri * ri;
  }
  put_partials<1>(acc, part);
}

// ---- misc --------------------------------------------------------------------
template <typename T>
__global__ void k_zero(int n, T *__restrict__ x) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) x[i] = (T)0;
}

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
// Calls f with the pattern's index view; vsize is the byte size of the
// values the kernel reads, which decides whether they stream (load_vals).
template <class F> void with_idx(const SellPattern &P, F &&f, int vsize) {
  const int64_t bytes = P.nvals * vsize + P.entries * (P.idx16 ? 2 : 4);
  const bool stream = bytes > SELL_RESIDENT_BYTES;
  auto pick = [&](auto tag) {
    constexpr bool S = decltype(tag)::value;
    if (P.bs == 3) {
      if (P.idx16) f(BSellIdx16<S>{P.off.get()});
      else f(BSellIdx32<S>{P.col.get()});
    } else {
      if (P.idx16) f(SellIdx16<S>{P.off.get()});
      else f(SellIdx32<S>{P.col.get()});
    }
  };
  if (stream) pick(std::true_type{});
  else pick(std::false_type{});
}

template <typename T>
void launch_cg_q(hipStream_t s, const SellOp &o, const T *val, const T *p, T *q,
                 const CgState *st, long long it, double *part) {
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_cg_q<T, decltype(idx)>), o.grid, BS, 0, s, o.n(), o.ns(),
                       o.P->cb.get(), idx, val, p, q, st, it, part);
  }, sizeof(T));
}

template <typename T>
void launch_sell_spmv(hipStream_t s, const SellOp &o, const T *val, const T *x,
                      T *y) {
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_sell_spmv<T, decltype(idx)>), o.grid, BS, 0, s, o.n(),
                       o.ns(), o.P->cb.get(), idx, val, x, y);
  }, sizeof(T));
}

// r = b - A x and partials r.r; with gate, only when *gate != 0 (a device
// decision; the partials are then left as they were)
void launch_residual(hipStream_t s, const SellOp &o, const double *val,
                     const double *x, const double *b, double *r, double *part,
                     const int *gate = nullptr) {
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_sell_residual<decltype(idx)>), o.grid, BS, 0, s, o.n(),
                       o.ns(), o.P->cb.get(), idx, val, x, b, r, part, gate);
  }, 8);
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

// State snapshots: a copy of the solver state into the context's pinned
// buffer (past the part read_state and read_dots use), enqueued between two
// passes and followed by an event, so the host reads the state of that
// point while the passes after it keep the GPU busy.
constexpr size_t SNAP_OFFSET = 2048;
constexpr long long SNAP_TAIL = 6;  // passes enqueued after each snapshot
static_assert(SNAP_OFFSET + sizeof(CgState) <= 4096, "snapshot exceeds the pinned buffer");

void snap_enqueue(gs_ctx *ctx, const CgState *st) {
  GS_HIP(hipMemcpyAsync((char *)ctx->pinned + SNAP_OFFSET, st, sizeof(CgState),
                        hipMemcpyDeviceToHost, ctx->stream));
  GS_HIP(hipEventRecord(ctx->snap_ev, ctx->stream));
}

CgState snap_wait(gs_ctx *ctx) {
  GS_HIP(hipEventSynchronize(ctx->snap_ev));
  CgState h;
  memcpy(&h, (const char *)ctx->pinned + SNAP_OFFSET, sizeof(CgState));
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

// One pass of the scaled recursion: KA (SpMV + direction), KB (step).
// part_r holds the r.r partials of the pass's residual (grid_fat(n) of them).
template <typename T, typename TV>
void enqueue_pass2(hipStream_t s, const SellOp &o, const TV *val, T *x, T *r, T *p,
                   T *sv, CgState *st, long long it, double *part1, double *part_r) {
  const int n = o.n();
  const unsigned gf = grid_fat(n);
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_cg2_dir<T, decltype(idx), TV>), o.grid, BS, 0, s, n, o.ns(),
                       o.P->cb.get(), idx, val, r, p, sv, st, it, part_r, (int)gf, part1);
  }, sizeof(TV));
  hipLaunchKernelGGL((k_cg2_step<T>), gf, FBS, 0, s, n, it, x, r, p, sv, st, part1,
                     o.grid, part_r);
}

struct PcgResult {
  long long iterations = 0, outer = 0, spmv = 0;
  double rnorm = 0.0, bnorm = 0.0;
  bool converged = false;
};

template <typename T> struct PcgWork {
  double *part1 = nullptr, *part2 = nullptr;
  CgState *st = nullptr;
  T *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
};

template <typename T>
PcgWork<T> carve_pcg(char *base, Layout &L, int64_t n) {
  PcgWork<T> w;
  w.part1 = (double *)(base + L.add<double>(2 * GRID_MAX));
  w.part2 = (double *)(base + L.add<double>(2 * GMAX));
  w.st = (CgState *)(base + L.add<CgState>(1));
  w.r = (T *)(base + L.add<T>(n));
  w.z = (T *)(base + L.add<T>(n));
  w.p = (T *)(base + L.add<T>(n));
  w.q = (T *)(base + L.add<T>(n));
  return w;
}

// Jacobi-PCG with the reference's control flow, in precision T (dot products
// in FP64); b, x in SELL order.
template <typename T>
PcgResult pcg_core(gs_ctx *ctx, const SellOp &o, const T *val, const T *dinv,
                   const T *b, T *x, long long max_iter, double tol,
                   double abs_tol, bool brk_rel, const PcgWork<T> &w) {
  hipStream_t s = ctx->stream;
  const int n = o.n();
  const unsigned gv = grid_vec(n);
  hipLaunchKernelGGL(k_state_setup, 1, 1, 0, s, w.st, max_iter, tol, abs_tol, 0, 0.0,
                     brk_rel ? 1 : 0);
  if (dinv)
    hipLaunchKernelGGL((k_cg_init<T, true>), gv, BS, 0, s, n, b, x, w.r, w.z, w.p,
                       dinv, w.part2);
  else
    hipLaunchKernelGGL((k_cg_init<T, false>), gv, BS, 0, s, n, b, x, w.r, w.z, w.p,
                       dinv, w.part2);
  hipLaunchKernelGGL(k_fin_init, 1, FBS, 0, s, w.st, w.part2, (int)gv);
  GS_LAUNCH_CHECK("k_cg_init");
  PcgResult res;
  Batcher bt;
  // Batches are pipelined: the state is snapshotted SNAP_TAIL passes before
  // a batch ends and the next batch is enqueued from that snapshot, so the
  // GPU does not drain between batches; passes enqueued past the stop
  // return at once.  When the passes in flight should already reach the
  // target, the host drains the stream and reads the state instead.
  CgState h = read_state(ctx, w.st);
  long long next_it = h.iteration;  // the next pass to enqueue
  while (!h.stop) {
    const long long need = bt.next(h.rnorm, h.iteration, std::max(h.tol_abs, h.abs_tol),
                                   max_iter - h.iteration);
    const long long ahead = next_it - h.iteration;
    if (ahead >= need || next_it >= max_iter) {
      h = read_state(ctx, w.st);
      continue;
    }
    const long long batch = std::min(need - ahead, max_iter - next_it);
    const long long snap_after = batch > SNAP_TAIL ? batch - SNAP_TAIL : batch;
    for (long long i = 0; i < batch; ++i) {
      enqueue_pass<T>(s, o, val, x, w.r, w.z, w.p, w.q, dinv, w.st, next_it + i,
                      w.part1, w.part2);
      if (i + 1 == snap_after) snap_enqueue(ctx, w.st);
    }
    GS_LAUNCH_CHECK("cg batch");
    next_it += batch;
    h = snap_wait(ctx);
  }
  // passes after the stop returned at once: the state is final
  res.iterations = h.iteration;
  res.rnorm = h.rnorm;
  res.bnorm = h.bnorm;
  res.converged = h.converged != 0;
  // one SpMV per pass, the converging (uncounted) pass included
  res.spmv = res.iterations + (res.converged && res.bnorm >= 1e-25 ? 1 : 0);
  return res;
}

// Reads n doubles written by the finalize kernels.
void read_dots(gs_ctx *ctx, const double *dev, double *out, int n) {
  GS_HIP(hipMemcpyAsync(ctx->pinned, dev, n * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(out, ctx->pinned, n * sizeof(double));
}

// Mixed precision: FP64 flexible CG, FCG(1) (Notay, SIAM J. Sci. Comput. 22,
// 2000), whose preconditioner is an FP32 Jacobi-PCG recursion on the system
// scaled by 1/||b||.
//   r = b, x = 0; start the FP32 recursion at rs = r/||b||
//   repeat until ||r|| < tol ||b|| or < abs_tol:
//     run FP32 passes until the recursion's residual fell by inner_tolerance
//     d     = ||b|| xs,  xs = 0                     (the FP32 correction)
//     w     = A d                                   (FP64)
//     gamma = (d, A p) / (p, A p)      [0 on the first step]
//     p     = d - gamma p,  Ap = w - gamma Ap
//     alpha = (p, r) / (p, Ap)
//     x    += alpha p,      r -= alpha Ap           (FP64)
//     continue the FP32 recursion from rs = r/||b|| keeping its direction
//     when the outer step agreed with it (alpha near 1, the new residual near
//     the recursion's own); otherwise restart it (p = z)
// Keeping the direction preserves the Krylov information of the FP32 solve
// across outer steps; the FP64 line search keeps every step safe when the
// FP32 operator's rounding (eps32 * cond(A)) makes the recursion drift.  The
// loop ends on the true residual b - A x.
PcgResult pcg_fcg(gs_ctx *ctx, gs_csr *A, const double *b, double *x,
                  bool precond, const gs_cg_opts &opt, char *base, Layout &L) {
  hipStream_t s = ctx->stream;
  const SellOp o = sell_op(A);
  const int n = o.n();
  const unsigned gv = grid_vec(n);
  double *r = (double *)(base + L.add<double>(n));
  double *d = (double *)(base + L.add<double>(n));
  double *wv = (double *)(base + L.add<double>(n));
  double *p = (double *)(base + L.add<double>(n));
  double *ap = (double *)(base + L.add<double>(n));
  double *dots = (double *)(base + L.add<double>(8));
  float *xs = (float *)(base + L.add<float>(n));
  const PcgWork<float> iw = carve_pcg<float>(base, L, n);
  PcgResult res;
  double h[2];
  hipLaunchKernelGGL(k_fcg_start, gv, BS, 0, s, n, b, x, r, iw.part2);
  hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, iw.part2, (int)gv, dots);
  GS_LAUNCH_CHECK("k_fcg_start");
  const bool pending = precond && A->sscaled == 2;
  if (pending) hipLaunchKernelGGL(k_flag_to_dot, 1, 1, 0, s, A->sscale_bad.get(), dots + 1);
  read_dots(ctx, dots, h, pending ? 2 : 1);
  if (pending) {
    A->sscaled = h[1] == 0.0 ? 1 : -1;
    A->sscale_bad = DArr<int>();
    if (A->sscaled < 0) {
      A->sscale = DArr<double>();
      A->svalq = DArr<int>();
    }
  }
  // Jacobi folded into the recursion's operator (S A S, S = D^-1/2, fixed
  // point) when every D^-1 entry is positive and every scaled entry fits:
  // the recursion then needs no z vector
  const bool scaled = precond && A->sscaled == 1;
  // the FP32 values, for the recursion on A itself (made here when the
  // matrix's assembly did not write them)
  if (!scaled) ensure_sell(ctx, A, true, precond);
  const double *ssc = scaled ? A->sscale.get() : nullptr;
  const float *dinv32 = precond && !scaled ? A->sdinv32.get() : nullptr;
  const float *val32 = A->sval32.get();
  const int *valq = A->svalq.get();
  const double *val = A->sval.get();
  const double bnorm = std::sqrt(h[0]);
  res.bnorm = bnorm;
  res.rnorm = bnorm;
  if (bnorm < 1e-25) {  // src/linalg/iterative.rs:60-68
    res.converged = true;
    res.rnorm = 0.0;
    return res;
  }
  const long long max_iter = opt.max_iterations;
  if (max_iter <= 0) return res;
  const long long max_outer =
      opt.max_outer_iterations > 0 ? opt.max_outer_iterations : 100;
  const double target = std::max(opt.tolerance * bnorm, opt.abs_tolerance);
  const double scale = 1.0 / bnorm;
  // inner_tolerance 0: the reduction measured best per recursion (thermal
  // Tet10, 1.08M DOFs, to rel-res 1e-8 and 1e-12): 1e-3 on the fixed-point
  // scaled operator, 1e-2 on the FP32 ones (the 3x3-block elastic operator
  // lost 12% / 100% at 1e-4)
  const double itol = opt.inner_tolerance > 0.0 ? opt.inner_tolerance : scaled ? 1e-3 : 1e-2;
  hipLaunchKernelGGL(k_state_setup, 1, 1, 0, s, iw.st, max_iter, target * scale,
                     0.0, 1, itol, 0);
  hipLaunchKernelGGL((k_zero<float>), gv, BS, 0, s, n, xs);
  const unsigned gf = grid_fat(n);
  // (Re)start the recursion from the outer residual r.  Host path: restart
  // and the outer residual rn given; device path (dev): both taken from the
  // state, where k_fcg_decide2 left them, and nothing runs once the device
  // has ended the solve.
  auto start_inner = [&](bool restart, double rn, bool dev) {
    const CgState *gate = dev ? iw.st : nullptr;
    if (scaled) {
      hipLaunchKernelGGL(k_fcg_inner_scaled, gf, FBS, 0, s, n, r, scale, ssc, iw.r,
                         iw.part2, gate);
    } else if (dinv32) {
      if (restart && !dev)
        hipLaunchKernelGGL((k_fcg_inner<true, true>), gv, BS, 0, s, n, r, scale,
                           dinv32, iw.r, iw.z, iw.p, iw.part2, gate);
      else
        hipLaunchKernelGGL((k_fcg_inner<true, false>), gv, BS, 0, s, n, r, scale,
                           dinv32, iw.r, iw.z, iw.p, iw.part2, gate);
    } else {
      if (restart && !dev)
        hipLaunchKernelGGL((k_fcg_inner<false, true>), gv, BS, 0, s, n, r, scale,
                           dinv32, iw.r, iw.z, iw.p, iw.part2, gate);
      else
        hipLaunchKernelGGL((k_fcg_inner<false, false>), gv, BS, 0, s, n, r, scale,
                           dinv32, iw.r, iw.z, iw.p, iw.part2, gate);
    }
    // the scaled recursion measures its residual in its own norm ||S r||:
    // its base is its r.r and its target the same fraction of it as the
    // outer target is of ||r||
    if (dev)
      hipLaunchKernelGGL(k_fin_inner_dev, 1, FBS, 0, s, iw.st, iw.part2,
                         (int)(scaled ? gf : gv), scaled ? 1 : 0, scale, target);
    else
      hipLaunchKernelGGL(k_fin_inner, 1, FBS, 0, s, iw.st, iw.part2,
                         (int)(scaled ? gf : gv), scaled ? -1.0 : rn * scale, target / rn,
                         restart ? 1 : 0, rn);
    GS_LAUNCH_CHECK("k_fcg_inner");
  };
  start_inner(true, bnorm, false);
  double rn = bnorm;
  double *const sc = dots + 2;  // the outer step's device scalars
  bool first = true, checked = false;
  Batcher bt;
  CgState hs{};  // pass 0, running: the state k_fin_inner leaves
  // A cycle whose recursion has not halved its best residual within
  // max(200, 2 x the longest cycle so far) passes has stalled on a stale
  // direction: it is cut short and the recursion restarts.
  long long longest = 0, cycle_start = 0;
  // the cycle was (re)started without the host reading its state (the
  // first one, and those the device's outer-step decision restarted): its
  // first batch is enqueued blind, at the Batcher's first size
  bool blind = true;
  long long checks_seen = 0;
  while (res.outer < max_outer) {
    // FP32 passes until the recursion pauses, in pipelined batches (as in
    // pcg_core); the passes enqueued never go past the cycle's limits (the
    // caller's cap, the stall window, max_iterations), and the loop leaves
    // on a drained, current state
    bool stalled = false;
    double best = hs.rnorm;
    long long best_at = hs.iteration;
    const long long it_begin = hs.iteration;
    long long next_it = hs.iteration;
    auto enqueue_batch = [&](long long batch) {
      const long long snap_after = batch > SNAP_TAIL ? batch - SNAP_TAIL : batch;
      for (long long i = 0; i < batch; ++i) {
        if (scaled)
          enqueue_pass2<float, int>(s, o, valq, xs, iw.r, iw.p, iw.q, iw.st, next_it + i,
                                    iw.part1, iw.part2);
        else
          enqueue_pass<float>(s, o, val32, xs, iw.r, iw.z, iw.p, iw.q, dinv32, iw.st,
                              next_it + i, iw.part1, iw.part2);
        if (i + 1 == snap_after) snap_enqueue(ctx, iw.st);
      }
      GS_LAUNCH_CHECK("cg batch");
      next_it += batch;
      hs = snap_wait(ctx);
    };
    if (blind) {
      blind = false;
      long long first_batch = std::min(16ll, max_iter + 1 - next_it);
      if (opt.max_inner_iterations > 0)
        first_batch = std::min(first_batch, cycle_start + opt.max_inner_iterations - next_it);
      if (first_batch > 0) {
        enqueue_batch(first_batch);
        // the cycle's starting residual (k_fin_inner_dev set rnorm = rbase),
        // also the Batcher's first point
        best = hs.rbase;
        best_at = it_begin;
        bt.next(hs.rbase, it_begin, 0.0, 1);
      }
    }
    while (!hs.pause && !hs.stop) {
      const long long window = std::max(200ll, 2 * longest);
      const bool fresh = next_it == hs.iteration;
      if (fresh && opt.max_inner_iterations > 0 &&
          hs.iteration - cycle_start >= opt.max_inner_iterations)
        break;  // the caller's cap on passes per cycle
      if (hs.rnorm < 0.5 * best) {
        best = hs.rnorm;
        best_at = hs.iteration;
      } else if (fresh && hs.iteration - best_at >= window) {
        stalled = true;
        break;
      }
      const double tgt = std::max(hs.rtol * hs.rbase, hs.tol);
      // (pass max_iter itself may be enqueued: the scaled recursion pauses
      // at the start of that pass, the three-launch one already returns)
      long long limit = std::min(max_iter + 1, best_at + window);
      if (opt.max_inner_iterations > 0)
        limit = std::min(limit, cycle_start + opt.max_inner_iterations);
      const long long need = bt.next(hs.rnorm, hs.iteration, tgt, limit - hs.iteration);
      const long long ahead = next_it - hs.iteration;
      if (ahead >= need || next_it >= limit) {
        hs = read_state(ctx, iw.st);
        next_it = std::max(next_it, hs.iteration);
        continue;
      }
      enqueue_batch(std::min(need - ahead, limit - next_it));
    }
    if (next_it > hs.iteration) hs = read_state(ctx, iw.st);  // passes still in flight
    res.spmv += hs.iteration - it_begin;
    res.spmv += hs.outer_checks - checks_seen;  // true-residual checks the device ran
    checks_seen = hs.outer_checks;
    longest = std::max(longest, hs.iteration - cycle_start);
    res.iterations = hs.iteration;
    if (hs.outer_end != 0) {
      // the device ended the solve at its last outer step (taken on its
      // decision): converged on the true residual, or no descent left
      rn = hs.outer_rn;
      if (hs.outer_end == 1) {
        res.converged = true;
        checked = true;
      } else {
        --res.outer;  // the step that found no descent was counted
        checked = false;
      }
      break;
    }
    rn = hs.outer_rn;  // the outer residual this cycle started from
    // the residual the recursion predicts: its own reduction applied to the
    // outer residual it started from
    const double predicted = hs.rbase > 0.0 ? rn * (hs.rnorm / hs.rbase) : 0.0;
    // outer FP64 step along the FP32 correction
    hipLaunchKernelGGL(k_fcg_expand, gv, BS, 0, s, n, xs, bnorm, ssc, d);
    with_idx(*o.P, [&](auto idx) {
      hipLaunchKernelGGL((k_fcg_spmv<decltype(idx)>), o.grid, BS, 0, s, n, o.ns(),
                         o.P->cb.get(), idx, val, d, wv, first ? nullptr : ap,
                         iw.part1);
    }, 8);
    ++res.spmv;
    hipLaunchKernelGGL(k_fcg_gamma, 1, FBS, 0, s, iw.part1, o.grid, first ? 1 : 0, sc);
    hipLaunchKernelGGL(k_fcg_dir, gv, BS, 0, s, n, d, wv, r, sc, first ? 1 : 0, p, ap,
                       iw.part2);
    hipLaunchKernelGGL(k_fcg_alpha, 1, FBS, 0, s, iw.part2, (int)gv, sc);
    hipLaunchKernelGGL(k_fcg_update, gv, BS, 0, s, n, sc, p, ap, x, r, iw.part2);
    hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, iw.part2, (int)gv, sc + 4);
    GS_LAUNCH_CHECK("k_fcg_update");
    // A step the recursion predicts will not reach the target, with passes
    // and outer steps left, is decided on the device and the next cycle
    // starts without a host round trip; the host learns the outcome from the
    // next cycle's first snapshot.  A step that may converge or end the
    // solve is decided here, on the scalars read back.
    const bool dev = !(predicted < 4.0 * target) && hs.iteration < max_iter &&
                     res.outer + 1 < max_outer;
    if (dev) {
      hipLaunchKernelGGL(k_fcg_decide, 1, 1, 0, s, iw.st, sc, target);
      launch_residual(s, o, val, x, b, r, iw.part1, &iw.st->outer_check);
      hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, iw.part1, o.grid, dots);
      hipLaunchKernelGGL(k_fcg_decide2, 1, 1, 0, s, iw.st, sc, dots, target,
                         stalled ? 1 : 0);
      GS_LAUNCH_CHECK("k_fcg_decide");
      start_inner(false, 0.0, true);
      first = false;
      ++res.outer;
      checked = false;
      bt = Batcher{};
      cycle_start = hs.iteration;
      hs.pause = 0;  // as k_fin_inner_dev leaves it, unless the solve ended
      blind = true;
      continue;
    }
    double hs4[4];  // alpha, p.Ap, step taken, r.r
    read_dots(ctx, sc + 1, hs4, 4);
    if (hs4[2] == 0.0) break;  // no descent left
    const double alpha = hs4[0];
    first = false;
    rn = std::sqrt(hs4[3]);
    ++res.outer;
    checked = false;
    if (rn < target) {
      // confirm on the true residual r = b - A x
      launch_residual(s, o, val, x, b, r, iw.part1);
      hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, iw.part1, o.grid, dots);
      GS_LAUNCH_CHECK("k_sell_residual");
      read_dots(ctx, dots, h, 1);
      ++res.spmv;
      rn = std::sqrt(h[0]);
      checked = true;
      if (rn < target) {
        res.converged = true;
        break;
      }
    }
    if (hs.iteration >= max_iter) break;
    const bool agree = !stalled && std::fabs(alpha - 1.0) < 0.25 &&
                       rn < 2.0 * predicted && !(hs.pq == 0.0);
    start_inner(!agree, rn, false);
    bt = Batcher{};
    hs = read_state(ctx, iw.st);
    cycle_start = hs.iteration;
  }
  if (!checked) {  // report the true residual
    launch_residual(s, o, val, x, b, r, iw.part1);
    hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, iw.part1, o.grid, dots);
    GS_LAUNCH_CHECK("k_sell_residual");
    read_dots(ctx, dots, h, 1);
    ++res.spmv;
    rn = std::sqrt(h[0]);
    res.converged = rn < target;
  }
  res.rnorm = rn;
  return res;
}

// ---- BiCGSTAB (src/linalg/iterative.rs:177-291) -------------------------------
// The reference's loop, one pass (iteration) per five launches:
//   KP  end of the previous pass (final_res = ||r||, it += 1, stop on
//       ||r|| < tol ||b|| or < abs_tol, on |omega| < 1e-40) and the loop top
//       (it < max_iter, rho = r^.r, |rho| < 1e-40); p = r + beta (p - omega v)
//       (p = r on the first pass), p^ = D^-1 p
//   KV  v = A p^ (SELL), partials r^.v
//   KS  alpha = rho / r^.v (|r^.v| < 1e-40 stops), s = r - alpha v, s^ = D^-1 s,
//       partials s.s
//   KT  t = A s^ (SELL), partials t.t, t.s
//   KX  stop on ||s|| < tol ||b|| or < abs_tol after x += alpha p^ (the pass is
//       not counted); |t.t| < 1e-40 stops; omega = t.s / t.t;
//       x += alpha p^ + omega s^, r = s - omega t; partials r.r, r^.r
// Every workgroup of a kernel takes the decisions from the same partials;
// workgroup 0 records them.  A kernel runs only when its predecessor of the
// same pass completed without stopping, which that predecessor's leader
// records in its own word mark[k] (read by later kernels only).
struct BgState {
  double rho[2];  // rho of pass it in rho[it & 1]
  double alpha, omega, bnorm, tol_abs, abs_tol, final_res;
  long long iteration, max_iter;
  long long mark[5];  // KP: mark[0], KS: mark[2], KX: mark[4]
  int stop, converged;
  int on_s;           // the solve ended on the intermediate residual s
};

__global__ void k_bg_setup(BgState *__restrict__ st, long long max_iter, double tol,
                           double abs_tol) {
  BgState s{};
  s.rho[0] = s.rho[1] = 1.0;
  s.alpha = s.omega = 1.0;
  s.max_iter = max_iter;
  s.tol_abs = tol;  // times ||b|| in k_bg_fin_init
  s.abs_tol = abs_tol;
  for (int k = 0; k < 5; ++k) s.mark[k] = -2;
  *st = s;
}

// x = 0, r = r^ = b, v = p = 0; partials b.b (twice: ||b||^2 and rho)
__global__ __launch_bounds__(BS) void k_bg_init(int n, const double *__restrict__ b,
                                                double *__restrict__ x, double *__restrict__ r,
                                                double *__restrict__ rh, double *__restrict__ v,
                                                double *__restrict__ p,
                                                double *__restrict__ part) {
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    const double bi = b[i];
    x[i] = 0.0;
    r[i] = bi;
    rh[i] = bi;
    v[i] = 0.0;
    p[i] = 0.0;
    acc[0] += bi * bi;
    acc[1] += bi * bi;
  }
  put_partials<2>(acc, part);
}

// ||b|| and the early exit for b = 0 (src/linalg/iterative.rs:185-189)
__global__ __launch_bounds__(FBS) void k_bg_fin_init(BgState *__restrict__ st,
                                                     const double *__restrict__ part,
                                                     int np) {
  double tot[2];
  sum_partials<2>(part, np, tot);
  if (threadIdx.x != 0) return;
  const double bn = sqrt(tot[0]);
  st->bnorm = bn;
  st->final_res = bn;
  st->tol_abs *= bn;
  if (bn < 1e-25) {
    st->stop = 1;
    st->converged = 1;
    st->final_res = 0.0;
  } else {
    st->mark[4] = -1;  // pass 0 may start
  }
}

template <bool PREC>
__global__ __launch_bounds__(FBS) void k_bg_dir(
    int n, long long it, const double *__restrict__ r, double *__restrict__ p,
    const double *__restrict__ v, const double *__restrict__ dinv,
    double *__restrict__ ph, BgState *__restrict__ st,
    const double *__restrict__ part_in, int np_in) {
  if (st->mark[4] != it - 1) return;
  double tot[2];  // (r.r, r^.r)
  sum_partials<2>(part_in, np_in, tot);
  bool conv = false, stop = false;
  const double res = sqrt(tot[0]);
  if (it > 0) {
    conv = res < st->tol_abs || res < st->abs_tol;
    stop = conv || fabs(st->omega) < 1e-40;
  }
  stop = stop || it >= st->max_iter;
  const double rho = tot[1];
  stop = stop || fabs(rho) < 1e-40;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (it > 0) st->final_res = res;
    st->iteration = it;
    if (conv) st->converged = 1;
    if (stop) {
      st->stop = 1;
    } else {
      st->rho[it & 1] = rho;
      st->mark[0] = it;
    }
  }
  if (stop) return;
  const double omega = st->omega;
  const double beta = it == 0 ? 0.0 : (rho / st->rho[(it - 1) & 1]) * (st->alpha / omega);
  for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS) {
    const double pi = it == 0 ? r[i] : r[i] + beta * (p[i] - omega * v[i]);
    p[i] = pi;
    ph[i] = PREC ? pi * dinv[i] : pi;
  }
}

// KV / KT: y = A x (SELL) with partials of (w.y) [and y.y when NV == 2:
// partials (y.y, y.w)]
template <int NV, class IDX>
__global__ __launch_bounds__(BS) void k_bg_spmv(
    int n, int ns, const int *__restrict__ cb, const IDX idx, const double *__restrict__ val,
    const double *__restrict__ xin, double *__restrict__ y, const double *__restrict__ w,
    const BgState *__restrict__ st, int mk, long long it, double *__restrict__ part) {
  if (st->mark[mk] != it) return;
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  auto epi = [&](int row, double sv) {
    y[row] = sv;
    if (NV == 1) {
      acc[0] += w[row] * sv;
    } else {
      acc[0] += sv * sv;
      acc[NV - 1] += sv * w[row];
    }
  };
  sell_rows<double>(n, ns, cb, idx, val, xin, epi);
  put_partials<NV>(acc, part);
}

template <bool PREC>
__global__ __launch_bounds__(FBS) void k_bg_s(
    int n, long long it, const double *__restrict__ r, const double *__restrict__ v,
    const double *__restrict__ dinv, double *__restrict__ sv, double *__restrict__ sh,
    BgState *__restrict__ st, const double *__restrict__ part_in, int np_in,
    double *__restrict__ part) {
  if (st->mark[0] != it) return;
  double tot[1];
  sum_partials<1>(part_in, np_in, tot);
  const bool stop = fabs(tot[0]) < 1e-40;
  const double alpha = st->rho[it & 1] / tot[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (stop) {
      st->stop = 1;
    } else {
      st->alpha = alpha;
      st->mark[2] = it;
    }
  }
  if (stop) return;
  double acc[1] = {0.0};
  for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS) {
    const double si = r[i] - alpha * v[i];
    sv[i] = si;
    sh[i] = PREC ? si * dinv[i] : si;
    acc[0] += si * si;
  }
  block_sum<1, FBS>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

__global__ __launch_bounds__(FBS) void k_bg_x(
    int n, long long it, double *__restrict__ x, double *__restrict__ r,
    const double *__restrict__ rh, const double *__restrict__ ph,
    const double *__restrict__ sv, const double *__restrict__ sh,
    const double *__restrict__ t, BgState *__restrict__ st,
    const double *__restrict__ part_s, int np_s, const double *__restrict__ part_t,
    int np_t, double *__restrict__ part) {
  if (st->mark[2] != it) return;
  double ss[1], tt[2];  // s.s; (t.t, t.s)
  sum_partials<1>(part_s, np_s, ss);
  sum_partials<2>(part_t, np_t, tt);
  const double alpha = st->alpha;
  const double s_norm = sqrt(ss[0]);
  const bool sconv = s_norm < st->tol_abs || s_norm < st->abs_tol;
  const bool brk = !sconv && fabs(tt[0]) < 1e-40;
  const double omega = tt[1] / tt[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (sconv) {
      st->final_res = s_norm;
      st->converged = 1;
      st->on_s = 1;
      st->stop = 1;
    } else if (brk) {
      st->stop = 1;
    } else {
      st->omega = omega;
      st->mark[4] = it;
    }
  }
  if (brk) return;
  if (sconv) {
    for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS)
      x[i] += alpha * ph[i];
    return;
  }
  double acc[2] = {0.0, 0.0};
  for (int i = blockIdx.x * FBS + threadIdx.x; i < n; i += gridDim.x * FBS) {
    x[i] += alpha * ph[i] + omega * sh[i];
    const double ri = sv[i] - omega * t[i];
    r[i] = ri;
    acc[0] += ri * ri;
    acc[1] += rh[i] * ri;
  }
  block_sum<2, FBS>(acc);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc[0];
    part[2 * blockIdx.x + 1] = acc[1];
  }
}

BgState read_bg_state(gs_ctx *ctx, const BgState *st) {
  static_assert(sizeof(BgState) <= 4096, "state exceeds the pinned staging buffer");
  BgState h;
  GS_HIP(hipMemcpyAsync(ctx->pinned, st, sizeof(BgState), hipMemcpyDeviceToHost,
                        ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(&h, ctx->pinned, sizeof(BgState));
  return h;
}

struct BgWork {
  double *r, *rh, *v, *p, *ph, *s, *sh, *t, *part_a, *part_b, *part_c;
  BgState *st;
};

BgWork carve_bg(char *base, Layout &L, int64_t n) {
  BgWork w;
  w.r = (double *)(base + L.add<double>(n));
  w.rh = (double *)(base + L.add<double>(n));
  w.v = (double *)(base + L.add<double>(n));
  w.p = (double *)(base + L.add<double>(n));
  w.ph = (double *)(base + L.add<double>(n));
  w.s = (double *)(base + L.add<double>(n));
  w.sh = (double *)(base + L.add<double>(n));
  w.t = (double *)(base + L.add<double>(n));
  w.part_a = (double *)(base + L.add<double>(2 * GRID_MAX));  // SpMV partials
  w.part_b = (double *)(base + L.add<double>(2 * GMAX));      // KS partials
  w.part_c = (double *)(base + L.add<double>(2 * GMAX));      // init / KX partials
  w.st = (BgState *)(base + L.add<BgState>(1));
  return w;
}

// KP (and, unless only_dir, KV KS KT KX) of pass it.
void enqueue_bg_pass(hipStream_t s, const SellOp &o, const double *val,
                     const double *dinv, double *x, const BgWork &w, long long it,
                     int np_first, bool only_dir) {
  const int n = o.n();
  const unsigned gf = grid_fat(n);
  const int np_c = it == 0 ? np_first : (int)gf;
  if (dinv)
    hipLaunchKernelGGL((k_bg_dir<true>), gf, FBS, 0, s, n, it, w.r, w.p, w.v, dinv,
                       w.ph, w.st, w.part_c, np_c);
  else
    hipLaunchKernelGGL((k_bg_dir<false>), gf, FBS, 0, s, n, it, w.r, w.p, w.v, dinv,
                       w.ph, w.st, w.part_c, np_c);
  if (only_dir) return;
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_bg_spmv<1, decltype(idx)>), o.grid, BS, 0, s, n, o.ns(),
                       o.P->cb.get(), idx, val, (const double *)w.ph, w.v,
                       (const double *)w.rh, (const BgState *)w.st, 0, it, w.part_a);
  }, 8);
  if (dinv)
    hipLaunchKernelGGL((k_bg_s<true>), gf, FBS, 0, s, n, it, w.r, w.v, dinv, w.s, w.sh,
                       w.st, w.part_a, o.grid, w.part_b);
  else
    hipLaunchKernelGGL((k_bg_s<false>), gf, FBS, 0, s, n, it, w.r, w.v, dinv, w.s, w.sh,
                       w.st, w.part_a, o.grid, w.part_b);
  with_idx(*o.P, [&](auto idx) {
    hipLaunchKernelGGL((k_bg_spmv<2, decltype(idx)>), o.grid, BS, 0, s, n, o.ns(),
                       o.P->cb.get(), idx, val, (const double *)w.sh, w.t,
                       (const double *)w.s, (const BgState *)w.st, 2, it, w.part_a);
  }, 8);
  hipLaunchKernelGGL(k_bg_x, gf, FBS, 0, s, n, it, x, w.r, w.rh, w.ph, w.s, w.sh, w.t,
                     w.st, w.part_b, (int)gf, w.part_a, o.grid, w.part_c);
}

// BiCGSTAB on the SELL operator; b, x in SELL order.
PcgResult bicgstab_core(gs_ctx *ctx, const SellOp &o, const double *val,
                        const double *dinv, const double *b, double *x,
                        long long max_iter, double tol, double abs_tol,
                        const BgWork &w) {
  hipStream_t s = ctx->stream;
  const int n = o.n();
  const unsigned gv = grid_vec(n);
  hipLaunchKernelGGL(k_bg_setup, 1, 1, 0, s, w.st, max_iter, tol, abs_tol);
  hipLaunchKernelGGL(k_bg_init, gv, BS, 0, s, n, b, x, w.r, w.rh, w.v, w.p, w.part_c);
  hipLaunchKernelGGL(k_bg_fin_init, 1, FBS, 0, s, w.st, w.part_c, (int)gv);
  GS_LAUNCH_CHECK("k_bg_init");
  PcgResult res;
  BgState h = read_bg_state(ctx, w.st);
  res.bnorm = h.bnorm;
  long long it = 0;
  Batcher bt;
  while (!h.stop) {
    if (it >= max_iter) {  // the last pass's end and the loop-top test
      enqueue_bg_pass(s, o, val, dinv, x, w, it, (int)gv, true);
    } else {
      const long long batch = bt.next(h.final_res, it, std::max(h.tol_abs, h.abs_tol),
                                      max_iter - it);
      for (long long i = 0; i < batch; ++i)
        enqueue_bg_pass(s, o, val, dinv, x, w, it + i, (int)gv, false);
      it += batch;
    }
    GS_LAUNCH_CHECK("bicgstab batch");
    h = read_bg_state(ctx, w.st);
  }
  res.iterations = h.iteration;
  res.rnorm = h.final_res;
  res.converged = h.converged != 0;
  res.spmv = 2 * h.iteration + (h.on_s ? 2 : 0);
  return res;
}

// ---- GMRES (src/linalg/iterative.rs:313-486) ----------------------------------
// Restarted GMRES(m), right-preconditioned, modified Gram-Schmidt, Givens
// rotations, entirely on the device: per inner step z = M v_j, w = A z (SELL,
// partials v_0.w), one kernel per Gram-Schmidt projection (sums the previous
// partials itself: h_ij, w -= h_ij v_i, partials of the next projection), a
// normalisation kernel (h_{j+1,j}, v_{j+1}) and a one-thread kernel that
// applies the earlier rotations to the new Hessenberg column, forms the next
// rotation, updates g and the residual estimate |g_{j+1}| and decides
// whether the cycle ends -- the reference's arithmetic in its order.  Every
// kernel returns at once once the cycle or the solve has ended (GmState),
// so the host enqueues whole cycles and reads the state once per cycle.

struct GmState {
  double rnorm;      // residual norm at the start of the cycle
  double final_res;  // the reference's final_res
  long long total;   // inner steps over all cycles
  long long spmv;
  int k;             // steps of the current cycle
  int stop;          // 0 running, 1 cycle ended, 2 solve ended
  int converged;
  int pad_;
};

// y = A x (SELL); partials of u.y
template <class IDX>
__global__ __launch_bounds__(BS) void k_spmv_dot(int n, int ns, const int *__restrict__ cb,
                                                 const IDX idx, const double *__restrict__ val,
                                                 const double *__restrict__ x,
                                                 double *__restrict__ y,
                                                 const double *__restrict__ u,
                                                 double *__restrict__ part,
                                                 const GmState *__restrict__ st) {
  if (st->stop) return;
  double acc[1] = {0.0};
  auto epi = [&](int row, double sv) {
    y[row] = sv;
    acc[0] += u[row] * sv;
  };
  sell_rows<double>(n, ns, cb, idx, val, x, epi);
  put_partials<1>(acc, part);
}

// v_0 = r / ||r|| (the reference divides)
__global__ __launch_bounds__(BS) void k_gm_first(int n, const double *__restrict__ r,
                                                 double *__restrict__ v,
                                                 const GmState *__restrict__ st) {
  if (st->stop) return;
  const double rn = st->rnorm;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) v[i] = r[i] / rn;
}

// z = D^-1 v
__global__ __launch_bounds__(BS) void k_gm_precond(int n, const double *__restrict__ v,
                                                   const double *__restrict__ dinv,
                                                   double *__restrict__ z,
                                                   const GmState *__restrict__ st) {
  if (st->stop) return;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) z[i] = v[i] * dinv[i];
}

// Projection i of step j: h = sum of the partials of v_i.w; H[i] = h;
// w -= h v_i; partials of next.w (next = v_{i+1}, or w itself after the last
// projection, for ||w||).
__global__ __launch_bounds__(FBS) void k_gm_project(int n, const double *__restrict__ vi,
                                                    const double *__restrict__ next,
                                                    double *__restrict__ w,
                                                    double *__restrict__ hcol, int i,
                                                    const double *__restrict__ part_in,
                                                    int np_in, double *__restrict__ part,
                                                    const GmState *__restrict__ st) {
  if (st->stop) return;
  double tot[1];
  sum_partials<1>(part_in, np_in, tot);
  const double h = tot[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) hcol[i] = h;
  double acc[1] = {0.0};
  for (int l = blockIdx.x * FBS + threadIdx.x; l < n; l += gridDim.x * FBS) {
    const double wl = w[l] - h * vi[l];
    w[l] = wl;
    acc[0] += (next ? next[l] : wl) * wl;
  }
  block_sum<1, FBS>(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc[0];
}

// H[j+1][j] = ||w||; v_{j+1} = w / H[j+1][j] when above 1e-40, else zero
// (the reference's fresh v_{j+1} of the cycle)
__global__ __launch_bounds__(FBS) void k_gm_normalize(int n, const double *__restrict__ w,
                                                      double *__restrict__ vnext,
                                                      double *__restrict__ hcol, int j,
                                                      const double *__restrict__ part_in,
                                                      int np_in, const GmState *__restrict__ st) {
  if (st->stop) return;
  double tot[1];
  sum_partials<1>(part_in, np_in, tot);
  const double hn = sqrt(tot[0]);
  if (blockIdx.x == 0 && threadIdx.x == 0) hcol[j + 1] = hn;
  const bool ok = fabs(hn) > 1e-40;
  for (int l = blockIdx.x * FBS + threadIdx.x; l < n; l += gridDim.x * FBS)
    vnext[l] = ok ? w[l] / hn : 0.0;
}

// The Hessenberg, rotations and right-hand side of a cycle, on the device
// (H row-major (m+1) x m, g m+1, cs and sn m, y m).
struct GmArrays {
  double *H, *g, *cs, *sn, *y;
};

// Start of a cycle (src/linalg/iterative.rs:356-381): the solve ends when
// max_iterations is reached or the true residual meets the tolerance;
// otherwise H, g, cs, sn are cleared and g[0] = ||r||.
__global__ __launch_bounds__(FBS) void k_gm_begin(GmState *st, const double *__restrict__ dots,
                                                  GmArrays a, int m, double target,
                                                  double abs_tol, long long max_iter) {
  __shared__ int go;
  if (threadIdx.x == 0) {
    go = 0;
    if (st->stop != 2) {
      if (st->total >= max_iter) {
        st->stop = 2;
      } else {
        const double r_norm = sqrt(dots[0]);
        st->final_res = r_norm;
        st->spmv += 1;
        if (r_norm < target || r_norm < abs_tol) {
          st->converged = 1;
          st->stop = 2;
        } else {
          st->rnorm = r_norm;
          st->k = 0;
          st->stop = 0;
          go = 1;
        }
      }
    }
  }
  __syncthreads();
  if (!go) return;
  for (int t = threadIdx.x; t < (m + 1) * m; t += FBS) a.H[t] = 0.0;
  for (int t = threadIdx.x; t <= m; t += FBS) a.g[t] = t == 0 ? st->rnorm : 0.0;
  for (int t = threadIdx.x; t < m; t += FBS) {
    a.cs[t] = 0.0;
    a.sn[t] = 0.0;
  }
}

// src/linalg/iterative.rs:463-473
__device__ void gm_givens(double a, double b, double &c, double &s, double &rho) {
#pragma clang fp contract(off)
  if (fabs(b) < 1e-40) {
    c = 1.0; s = 0.0; rho = a;
  } else if (fabs(b) > fabs(a)) {
    const double tau = a / b;
    s = 1.0 / sqrt(1.0 + tau * tau);
    c = s * tau;
    rho = b * sqrt(1.0 + tau * tau);
  } else {
    const double tau = b / a;
    c = 1.0 / sqrt(1.0 + tau * tau);
    s = c * tau;
    rho = a * sqrt(1.0 + tau * tau);
  }
}

// End of inner step j (src/linalg/iterative.rs:393-414), one thread.
__global__ void k_gm_rotate(GmState *st, const double *__restrict__ hcol, GmArrays a, int m,
                            int j, double target, double abs_tol, long long max_iter) {
#pragma clang fp contract(off)
  if (st->stop) return;
  double *H = a.H;
  auto HH = [&](int i, int c) -> double & { return H[(size_t)i * m + c]; };
  for (int i = 0; i <= j + 1; ++i) HH(i, j) = hcol[i];
  for (int i = 0; i < j; ++i) {
    const double temp = a.cs[i] * HH(i, j) + a.sn[i] * HH(i + 1, j);
    HH(i + 1, j) = -a.sn[i] * HH(i, j) + a.cs[i] * HH(i + 1, j);
    HH(i, j) = temp;
  }
  double c, s, rho;
  gm_givens(HH(j, j), HH(j + 1, j), c, s, rho);
  a.cs[j] = c;
  a.sn[j] = s;
  HH(j, j) = rho;
  HH(j + 1, j) = 0.0;
  const double temp = a.cs[j] * a.g[j];
  a.g[j + 1] = -a.sn[j] * a.g[j];
  a.g[j] = temp;
  st->k = j + 1;
  st->total += 1;
  st->spmv += 1;
  const double fr = fabs(a.g[j + 1]);
  st->final_res = fr;
  // the cycle ends on the estimate, at max_iterations (the next step's
  // check), or after its m-th step
  if (fr < target || fr < abs_tol || st->total >= max_iter || j + 1 == m) st->stop = 1;
}

// End of a cycle (src/linalg/iterative.rs:416-433): back substitution with
// its 1e-40 break (one thread), then the solve ends if the estimate met
// the tolerance.
__global__ void k_gm_backsub(GmState *st, GmArrays a, int m, double target, double abs_tol) {
#pragma clang fp contract(off)
  if (st->stop == 2) return;
  const int k = st->k;
  const double *H = a.H;
  for (int i = 0; i < k; ++i) a.y[i] = 0.0;
  for (int i = k - 1; i >= 0; --i) {
    double sum = 0.0;
    for (int c = i + 1; c < k; ++c) sum += H[(size_t)i * m + c] * a.y[c];
    if (fabs(H[(size_t)i * m + i]) < 1e-40) break;
    a.y[i] = (a.g[i] - sum) / H[(size_t)i * m + i];
  }
}

// x += M dy, dy = sum_j V_j y_j (accumulated from 0 in j order)
template <bool PREC>
__global__ __launch_bounds__(BS) void k_gm_update(int n, const double *__restrict__ V,
                                                  const double *__restrict__ y,
                                                  const double *__restrict__ dinv,
                                                  double *__restrict__ x,
                                                  const GmState *__restrict__ st) {
  if (st->stop == 2) return;
  const int k = st->k;
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    double dy = 0.0;
    for (int j = 0; j < k; ++j) dy += V[(size_t)j * n + i] * y[j];
    x[i] += PREC ? dy * dinv[i] : dy;
  }
}

__global__ void k_gm_end(GmState *st, double target, double abs_tol) {
  if (st->stop == 2) return;
  const double fr = st->final_res;
  if (fr < target || fr < abs_tol) {
    st->converged = 1;
    st->stop = 2;
  }
}

__global__ void k_gm_init(GmState *st, double b_norm) {
  st->rnorm = 0.0;
  st->final_res = b_norm;
  st->total = 0;
  st->spmv = 0;
  st->k = 0;
  st->stop = 0;
  st->converged = 0;
  st->pad_ = 0;
}

PcgResult gmres_core(gs_ctx *ctx, const SellOp &o, const double *val, const double *dinv,
                     const double *b, double *x, long long max_iter, int m, double tol,
                     double abs_tol, char *base, Layout &L) {
  hipStream_t s = ctx->stream;
  const int n = o.n();
  const unsigned gv = grid_vec(n), gf = grid_fat(n);
  double *r = (double *)(base + L.add<double>(n));
  double *w = (double *)(base + L.add<double>(n));
  double *z = (double *)(base + L.add<double>(n));
  double *V = (double *)(base + L.add<double>((int64_t)n * (m + 1)));
  double *hcol = (double *)(base + L.add<double>(m + 2));
  GmArrays a;
  a.H = (double *)(base + L.add<double>((int64_t)(m + 1) * m));
  a.g = (double *)(base + L.add<double>(m + 1));
  a.cs = (double *)(base + L.add<double>(m));
  a.sn = (double *)(base + L.add<double>(m));
  a.y = (double *)(base + L.add<double>(m));
  double *part1 = (double *)(base + L.add<double>(GRID_MAX));
  double *part2 = (double *)(base + L.add<double>(GMAX));
  double *dots = (double *)(base + L.add<double>(8));
  GmState *st = (GmState *)(base + L.add<GmState>(1));
  PcgResult res;
  double h1[1];
  hipLaunchKernelGGL(k_fcg_start, gv, BS, 0, s, n, b, x, r, part2);  // x = 0, ||b||
  hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, part2, (int)gv, dots);
  GS_LAUNCH_CHECK("k_fcg_start");
  read_dots(ctx, dots, h1, 1);
  const double b_norm = std::sqrt(h1[0]);
  hipLaunchKernelGGL(k_gm_init, 1, 1, 0, s, st, b_norm);
  res.bnorm = b_norm;
  res.rnorm = b_norm;
  if (b_norm < 1e-25) {  // src/linalg/iterative.rs:346-348
    res.converged = true;
    res.rnorm = 0.0;
    return res;
  }
  const double target = tol * b_norm;
  GmState h{};
  for (;;) {
    launch_residual(s, o, val, x, b, r, part1);  // r = b - A x
    hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, part1, o.grid, dots);
    hipLaunchKernelGGL(k_gm_begin, 1, FBS, 0, s, st, (const double *)dots, a, m, target,
                       abs_tol, max_iter);
    hipLaunchKernelGGL(k_gm_first, gv, BS, 0, s, n, (const double *)r, V, (const GmState *)st);
    for (int j = 0; j < m; ++j) {
      const double *vj = V + (size_t)j * n;
      const double *zin = vj;
      if (dinv) {
        hipLaunchKernelGGL(k_gm_precond, gv, BS, 0, s, n, vj, dinv, z, (const GmState *)st);
        zin = z;
      }
      with_idx(*o.P, [&](auto idx) {
        hipLaunchKernelGGL((k_spmv_dot<decltype(idx)>), o.grid, BS, 0, s, n, o.ns(),
                           o.P->cb.get(), idx, val, zin, w, (const double *)V, part1,
                           (const GmState *)st);
      }, 8);
      // projections i = 0..j; the first sums the SpMV's partials
      for (int i = 0; i <= j; ++i) {
        const double *nxt = i < j ? V + (size_t)(i + 1) * n : nullptr;
        const double *pin = i == 0 ? part1 : part2 + (i & 1 ? 0 : GMAX / 2);
        const int npin = i == 0 ? o.grid : (int)gf;
        double *pout = part2 + (i & 1 ? GMAX / 2 : 0);
        hipLaunchKernelGGL(k_gm_project, gf, FBS, 0, s, n, V + (size_t)i * n, nxt, w, hcol,
                           i, pin, npin, pout, (const GmState *)st);
      }
      hipLaunchKernelGGL(k_gm_normalize, gf, FBS, 0, s, n, (const double *)w,
                         V + (size_t)(j + 1) * n, hcol, j,
                         (const double *)(part2 + (j & 1 ? GMAX / 2 : 0)), (int)gf,
                         (const GmState *)st);
      hipLaunchKernelGGL(k_gm_rotate, 1, 1, 0, s, st, (const double *)hcol, a, m, j, target,
                         abs_tol, max_iter);
      GS_LAUNCH_CHECK("gmres step");
    }
    hipLaunchKernelGGL(k_gm_backsub, 1, 1, 0, s, st, a, m, target, abs_tol);
    if (dinv)
      hipLaunchKernelGGL((k_gm_update<true>), gv, BS, 0, s, n, (const double *)V,
                         (const double *)a.y, dinv, x, (const GmState *)st);
    else
      hipLaunchKernelGGL((k_gm_update<false>), gv, BS, 0, s, n, (const double *)V,
                         (const double *)a.y, dinv, x, (const GmState *)st);
    hipLaunchKernelGGL(k_gm_end, 1, 1, 0, s, st, target, abs_tol);
    GS_LAUNCH_CHECK("gmres cycle");
    GS_HIP(hipMemcpyAsync(ctx->pinned, st, sizeof(GmState), hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    memcpy(&h, ctx->pinned, sizeof(GmState));
    if (h.stop == 2) break;
  }
  res.iterations = h.total;
  res.rnorm = h.final_res;
  res.converged = h.converged != 0;
  res.spmv = h.spmv;
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
// Block size of the SELL view: the 3x3-block view for 3-DOF-per-node
// operators (or when asked for), the scalar one otherwise or when forced.
static int sell_block(const gs_csr *A) {
  if (A->spmv_flags & GS_SPMV_SCALAR) return 1;
  const bool want = A->block == 3 || (A->spmv_flags & GS_SPMV_BLOCK3);
  return want && A->rows > 0 && A->rows % 3 == 0 ? 3 : 1;
}

static void ensure_sell(gs_ctx *ctx, gs_csr *A, bool fp32, bool precond) {
  const bool force32 = (A->spmv_flags & GS_SPMV_INDEX32) != 0;
  const int block = sell_block(A);
  hipStream_t s = ctx->stream;
  if (!A->sell || A->sell_force32 != force32 || A->sell->bs != block) {
    ensure_csr_values(ctx, A);  // from the old view's values, before they go
    const bool reorder = !A->sell || A->sell->bs != block;
    A->sell = build_sell_pattern(ctx, A, force32, block);
    A->sell_force32 = force32;
    A->sval = DArr<double>();
    A->sval32 = DArr<float>();
    A->svalq = DArr<int>();
    A->sscale = DArr<double>();
    A->sscale_bad = DArr<int>();
    A->sscaled = 0;
    if (reorder) {  // the SELL order of the rows changed
      A->sdinv = DArr<double>();
      A->sdinv32 = DArr<float>();
    }
  }
  const SellPattern &P = *A->sell;
  if (!A->sval) {
    ensure_csr_values(ctx, A);
    A->sval = dalloc<double>(P.nvals);
    sell_values<double>(s, P, A->val.get(), A->sval.get());
  }
  if (fp32 && !A->sval32) {
    ensure_csr_values(ctx, A);
    A->sval32 = dalloc<float>(P.nvals);
    sell_values<float>(s, P, A->val.get(), A->sval32.get());
  }
  if (precond && (!A->sdinv || A->sdinv.n != P.n)) {
    ensure_csr_values(ctx, A);
    A->sdinv = dalloc<double>(P.n);
    A->sdinv32 = dalloc<float>(P.n);
    sell_jacobi(s, P, A, A->sdinv.get(), A->sdinv32.get());
  }
}

// Fixed-point values of S A S, S = diag(sqrt(D^-1)), for the mixed solve's
// two-launch recursion on scalar operators (built once per matrix, after the
// Jacobi diagonal).  The scaled entries lie in [-1, 1], where int32 at 2^-30
// resolution holds them 2^7 times finer than FP32's 24-bit significand at 1
// in the same 4 bytes: the recursion's operator rounding, which delays its
// convergence, drops accordingly.  Nothing is waited for here: whether every
// D^-1 entry was positive and every scaled entry fit (sscaled 2, pending) is
// read with the solve's first dot product, and when not the recursion
// applies D^-1 to its own z on the FP32 operator instead (sscaled -1).
static void ensure_scaled(gs_ctx *ctx, gs_csr *A) {
  if (A->sscaled != 0) return;
  hipStream_t s = ctx->stream;
  const SellPattern &P = *A->sell;
  if (P.bs != 1) {
    A->sscaled = -1;
    return;
  }
  const int n = (int)P.n;
  auto ssc = palloc<double>(ctx->pool, P.n);
  auto bad = palloc<int>(ctx->pool, 1);
  auto v = palloc<int>(ctx->pool, P.nvals);
  hipLaunchKernelGGL((k_zero<int>), 1, BS, 0, s, 1, bad.get());
  hipLaunchKernelGGL(k_sqrt_dinv, grid_vec(n), BS, 0, s, n, A->sdinv.get(), ssc.get(),
                     bad.get());
  GS_LAUNCH_CHECK("k_sqrt_dinv");
  const unsigned g = (unsigned)std::max<int64_t>(
      1, std::min<int64_t>((P.nslices + SELL_WPB - 1) / SELL_WPB, GRID_MAX));
  with_idx(P, [&](auto idx) {
    if constexpr (decltype(idx)::bs == 1)
      hipLaunchKernelGGL((k_scale_sell<decltype(idx)>), g, BS, 0, s, n, (int)P.nslices,
                         P.cb.get(), idx, A->sval.get(), ssc.get(), v.get(), bad.get());
  }, 4);
  GS_LAUNCH_CHECK("k_scale_sell");
  A->sscale = std::move(ssc);
  A->svalq = std::move(v);
  A->sscale_bad = std::move(bad);
  A->sscaled = 2;
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
    ensure_csr_values(ctx, A);
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

// Stats of a finished solve.
static void fill_stats(const PcgResult &r, std::chrono::steady_clock::time_point t0,
                       gs_solver_stats *st) {
  st->iterations = r.iterations;
  st->residual_norm = r.rnorm;
  st->relative_residual = r.bnorm > 1e-20 ? r.rnorm / r.bnorm : 0.0;
  st->converged = r.converged;
  st->outer_iterations = r.outer;
  st->spmv_count = r.spmv;
  st->solve_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// The FP64 solve on a prepared SELL view, all scratch carved from base (the
// layout of fp64_layout): b and x gathered into the SELL order, the PCG, x
// scattered back.
static Layout fp64_layout(int64_t n) {
  Layout L;
  L.add<double>(n);
  L.add<double>(n);
  carve_pcg<double>(nullptr, L, n);
  return L;
}

static PcgResult fp64_solve(gs_ctx *ctx, const gs_csr *A, const double *b,
                            const gs_cg_opts &o, double *x, char *base, bool precond) {
  hipStream_t s = ctx->stream;
  const int64_t n = A->rows;
  Layout L;
  const size_t ob = L.add<double>(n), ox = L.add<double>(n);
  double *bi = (double *)(base + ob), *xi = (double *)(base + ox);
  const PcgWork<double> w = carve_pcg<double>(base, L, n);
  const int *vp = A->sell->vperm.get();
  permute_gather(s, n, b, vp, bi);
  const PcgResult r = pcg_core<double>(ctx, sell_op(A), A->sval.get(),
                                       precond ? A->sdinv.get() : nullptr, bi, xi,
                                       o.max_iterations, o.tolerance, o.abs_tolerance,
                                       o.relative_breakdown != 0, w);
  permute_scatter(s, n, xi, vp, x);
  GS_LAUNCH_CHECK("cg_solve");
  GS_HIP(hipStreamSynchronize(s));
  return r;
}

// ---- CG without allocation ------------------------------------------------------
// gs_cg_solve_noalloc's scratch: the FP64 solve's vectors (b and x in SELL
// order, r, z, p, q), its partial sums and its state, for operators of up to
// NOALLOC_MAX_ROWS rows, in one file-scope device array (about 96 MiB, held
// by every device the library runs on until the process ends).  Host code
// never uses the array's name as an address: noalloc_setup looks up each
// device's copy once, with that device current, and the solve passes that
// address to its kernels.  Every call on one device shares that device's
// copy, so calls that overlap on one device race.
constexpr int64_t NOALLOC_MAX_ROWS = int64_t(1) << 21;
constexpr size_t noalloc_bytes() {
  return 6 * align256((size_t)NOALLOC_MAX_ROWS * sizeof(double)) +
         align256(2 * GRID_MAX * sizeof(double)) + align256(2 * GMAX * sizeof(double)) +
         align256(sizeof(CgState));
}
static __device__ __align__(256) char scratch[noalloc_bytes()];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex noalloc_mu;
static char *noalloc_ptr[NOALLOC_MAX_DEVICES] = {};

void noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
  std::lock_guard<std::mutex> g(noalloc_mu);
  if (noalloc_ptr[device]) return;
  void *ptr = nullptr;
  GS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  noalloc_ptr[device] = (char *)ptr;
}

void cg_solve_noalloc(gs_ctx *ctx, const gs_csr *A, const double *b, const gs_cg_opts &o,
                      double *x, gs_solver_stats *st) {
  const auto t0 = std::chrono::steady_clock::now();
  memset(st, 0, sizeof(*st));
  if (o.precision != GS_PREC_FP64)
    throw Error(GS_ERR_INVALID_ARG, "the no-allocation CG runs in FP64 only");
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "CG needs a square matrix");
  const int64_t n = A->rows;
  if (n > NOALLOC_MAX_ROWS)
    throw Error(GS_ERR_TOO_LARGE, "the no-allocation CG takes at most 2^21 rows");
  if (n == 0) {  // ||b|| = 0 < 1e-25: converged at once
    st->converged = 1;
    return;
  }
  const bool precond = o.use_preconditioner != 0;
  const bool force32 = (A->spmv_flags & GS_SPMV_INDEX32) != 0;
  if ((A->spmv_flags & GS_SPMV_CSR) || !A->sell || A->sell_force32 != force32 ||
      A->sell->bs != sell_block(A) || !A->sval || (precond && !A->sdinv))
    throw Error(GS_ERR_INVALID_ARG,
                "the no-allocation CG needs a matrix whose solver arrays exist: one "
                "returned by a gs_assemble_*_system call or solved once with gs_cg_solve");
  char *base = ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES
                   ? noalloc_ptr[ctx->device]
                   : nullptr;
  if (!base) throw Error(GS_ERR_HIP, "no-allocation scratch not set up on this device");
  if (fp64_layout(n).bytes > noalloc_bytes())
    throw Error(GS_ERR_TOO_LARGE, "the no-allocation CG's scratch is too small");
  fill_stats(fp64_solve(ctx, A, b, o, x, base, precond), t0, st);
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
  // FP64 values (and Jacobi diagonal); the mixed solve makes its FP32 values
  // only if its recursion runs on A itself (not on the scaled operator)
  ensure_sell(ctx, A, mixed && sell_block(A) == 3, precond);
  if (!mixed) {
    char *base = ctx_scratch(ctx, fp64_layout(n).bytes);
    fill_stats(fp64_solve(ctx, A, b, o, x, base, precond), t0, st);
    return;
  }
  if (precond) ensure_scaled(ctx, A);
  Layout L;
  const size_t ob = L.add<double>(n), ox = L.add<double>(n);
  // size the scratch with a dry run of the same layout
  Layout dry = L;
  for (int i = 0; i < 5; ++i) dry.add<double>(n);
  dry.add<double>(8);
  dry.add<float>(n);
  carve_pcg<float>(nullptr, dry, n);
  char *base = ctx_scratch(ctx, dry.bytes);
  double *bi = (double *)(base + ob), *xi = (double *)(base + ox);
  const int *vp = A->sell->vperm.get();
  permute_gather(s, n, b, vp, bi);
  const PcgResult r = pcg_fcg(ctx, A, bi, xi, precond, o, base, L);
  permute_scatter(s, n, xi, vp, x);
  GS_LAUNCH_CHECK("cg_solve");
  GS_HIP(hipStreamSynchronize(s));
  fill_stats(r, t0, st);
}

// ---- CG on a caller's operator (src/linalg/iterative.rs:46-119) ------------
// The vector work of the reference's loop on the device, one kernel per
// statement; the dot products as fixed-order partial sums read by the host,
// which also calls the operator and the preconditioner between them.
__global__ __launch_bounds__(BS) void k_op_dot(int n, const double *__restrict__ a,
                                               const double *__restrict__ b,
                                               double *__restrict__ part) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    acc[0] += a[i] * b[i];
  put_partials<1>(acc, part);
}

// x += alpha p; r -= alpha ap; partials r.r
__global__ __launch_bounds__(BS) void k_op_xr(int n, double alpha, const double *__restrict__ p,
                                              const double *__restrict__ ap,
                                              double *__restrict__ x, double *__restrict__ r,
                                              double *__restrict__ part) {
  double acc[1] = {0.0};
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
    x[i] += alpha * p[i];
    const double ri = r[i] - alpha * ap[i];
    r[i] = ri;
    acc[0] += ri * ri;
  }
  put_partials<1>(acc, part);
}

// p = z + beta p
__global__ __launch_bounds__(BS) void k_op_dir(int n, double beta, const double *__restrict__ z,
                                               double *__restrict__ p) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS)
    p[i] = z[i] + beta * p[i];
}

__global__ __launch_bounds__(BS) void k_op_copy(int n, const double *__restrict__ a,
                                                double *__restrict__ b) {
  for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) b[i] = a[i];
}

void cg_solve_operator(gs_ctx *ctx, int64_t n64, gs_operator_fn apply, void *apply_user,
                       gs_operator_fn precond, void *precond_user, double *in,
                       double *out, const double *b, const gs_cg_opts &o, double *x,
                       gs_solver_stats *st) {
  const auto t0 = std::chrono::steady_clock::now();
  *st = gs_solver_stats{};
  auto finish = [&] {
    st->solve_time =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  };
  hipStream_t s = ctx->stream;
  const int n = (int)n64;
  const unsigned gv = grid_vec(n);
  // the loop's vectors come from the pool, not the context's scratch buffer:
  // the callbacks may call other entry points on this context (a CsrMatrix
  // operator does), which use that buffer
  auto rv = palloc<double>(ctx->pool, n), pv = palloc<double>(ctx->pool, n);
  auto zv = palloc<double>(ctx->pool, n), partv = palloc<double>(ctx->pool, GMAX);
  auto dotv = palloc<double>(ctx->pool, 2);
  double *r = rv.get(), *p = pv.get(), *zc = zv.get(), *part = partv.get(), *dot = dotv.get();
  auto dotp = [&](const double *u, const double *v) {
    hipLaunchKernelGGL(k_op_dot, gv, BS, 0, s, n, u, v, part);
    hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, part, (int)gv, dot);
    GS_LAUNCH_CHECK("k_op_dot");
    double h;
    read_dots(ctx, dot, &h, 1);
    return h;
  };
  // y = F(v) through the caller's buffers; returns the buffer holding y
  auto call = [&](gs_operator_fn f, void *user, const double *v) -> double * {
    hipLaunchKernelGGL(k_op_copy, gv, BS, 0, s, n, v, in);
    GS_LAUNCH_CHECK("k_op_copy");
    GS_HIP(hipStreamSynchronize(s));
    if (f(user) != 0) throw Error(GS_ERR_INVALID_ARG, "operator callback failed");
    return out;
  };
  // the preconditioner: the caller's, or the identity (a copy)
  auto prec = [&](const double *v) -> const double * {
    if (!precond) {
      hipLaunchKernelGGL(k_op_copy, gv, BS, 0, s, n, v, zc);
      return zc;
    }
    return call(precond, precond_user, v);
  };
  const double bnorm = std::sqrt(dotp(b, b));
  if (bnorm < 1e-25) {  // src/linalg/iterative.rs:60-68
    hipLaunchKernelGGL((k_zero<double>), gv, BS, 0, s, n, x);
    GS_LAUNCH_CHECK("k_zero");
    GS_HIP(hipStreamSynchronize(s));
    st->converged = 1;
    finish();
    return;
  }
  hipLaunchKernelGGL((k_zero<double>), gv, BS, 0, s, n, x);
  hipLaunchKernelGGL(k_op_copy, gv, BS, 0, s, n, b, r);
  const double *z = prec(r);
  hipLaunchKernelGGL(k_op_copy, gv, BS, 0, s, n, z, p);
  double rz = dotp(r, z);
  long long it = 0;
  bool converged = false;
  double final_res = bnorm;
  const bool brk_rel = o.relative_breakdown != 0;
  while (it < o.max_iterations) {
    const double *ap = call(apply, apply_user, p);
    ++st->spmv_count;
    const double pap = dotp(p, ap);
    if (brk_rel ? (pap == 0.0 || !std::isfinite(rz / pap)) : std::fabs(pap) < 1e-30) break;
    const double alpha = rz / pap;
    hipLaunchKernelGGL(k_op_xr, gv, BS, 0, s, n, alpha, p, ap, x, r, part);
    hipLaunchKernelGGL((k_fin_sum<1>), 1, FBS, 0, s, part, (int)gv, dot);
    GS_LAUNCH_CHECK("k_op_xr");
    double rr;
    read_dots(ctx, dot, &rr, 1);
    const double rn = std::sqrt(rr);
    final_res = rn;
    if (rn < o.tolerance * bnorm || rn < o.abs_tolerance) {
      converged = true;
      break;
    }
    z = prec(r);
    const double rz_new = dotp(r, z);
    const double beta = rz_new / rz;
    rz = rz_new;
    hipLaunchKernelGGL(k_op_dir, gv, BS, 0, s, n, beta, z, p);
    GS_LAUNCH_CHECK("k_op_dir");
    ++it;
  }
  GS_HIP(hipStreamSynchronize(s));
  st->iterations = it;
  st->residual_norm = final_res;
  st->relative_residual = bnorm > 1e-20 ? final_res / bnorm : 0.0;
  st->converged = converged;
  finish();
}

// BiCGSTAB::solve (src/linalg/iterative.rs:283-291): Jacobi when
// use_preconditioner, FP64 only.
void bicgstab_solve(gs_ctx *ctx, const gs_csr *Ac, const double *b,
                    const gs_cg_opts &o, double *x, gs_solver_stats *st) {
  gs_csr *A = const_cast<gs_csr *>(Ac);  // derived caches only
  const auto t0 = std::chrono::steady_clock::now();
  memset(st, 0, sizeof(*st));
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "BiCGSTAB needs a square matrix");
  if (o.precision != GS_PREC_FP64)
    throw Error(GS_ERR_INVALID_ARG, "BiCGSTAB runs in FP64 only");
  if (A->spmv_flags & GS_SPMV_CSR)
    throw Error(GS_ERR_INVALID_ARG, "BiCGSTAB runs on the SELL engine only");
  const int64_t n = A->rows;
  if (n == 0) {
    st->converged = 1;
    return;
  }
  hipStream_t s = ctx->stream;
  const bool precond = o.use_preconditioner != 0;
  ensure_sell(ctx, A, false, precond);
  Layout L;
  const size_t ob = L.add<double>(n), ox = L.add<double>(n);
  Layout dry = L;
  carve_bg(nullptr, dry, n);
  char *base = ctx_scratch(ctx, dry.bytes);
  double *bi = (double *)(base + ob), *xi = (double *)(base + ox);
  const BgWork w = carve_bg(base, L, n);
  const int *vp = A->sell->vperm.get();
  permute_gather(s, n, b, vp, bi);
  const PcgResult r = bicgstab_core(ctx, sell_op(A), A->sval.get(),
                                    precond ? A->sdinv.get() : nullptr, bi, xi,
                                    o.max_iterations, o.tolerance, o.abs_tolerance, w);
  permute_scatter(s, n, xi, vp, x);
  GS_LAUNCH_CHECK("bicgstab_solve");
  GS_HIP(hipStreamSynchronize(s));
  st->iterations = r.iterations;
  st->residual_norm = r.rnorm;
  st->relative_residual = r.bnorm > 0.0 ? r.rnorm / r.bnorm : 0.0;
  st->converged = r.converged;
  st->spmv_count = r.spmv;
  st->solve_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// GMRES::solve (src/linalg/iterative.rs:476-486): Jacobi when
// use_preconditioner, FP64 only; restart m in [1, 200].
void gmres_solve(gs_ctx *ctx, const gs_csr *Ac, const double *b, const gs_cg_opts &o,
                 int restart, double *x, gs_solver_stats *st) {
  gs_csr *A = const_cast<gs_csr *>(Ac);  // derived caches only
  const auto t0 = std::chrono::steady_clock::now();
  memset(st, 0, sizeof(*st));
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "GMRES needs a square matrix");
  if (o.precision != GS_PREC_FP64) throw Error(GS_ERR_INVALID_ARG, "GMRES runs in FP64 only");
  if (restart < 1 || restart > 200)
    throw Error(GS_ERR_INVALID_ARG, "GMRES restart must be in [1, 200]");
  if (A->spmv_flags & GS_SPMV_CSR)
    throw Error(GS_ERR_INVALID_ARG, "GMRES runs on the SELL engine only");
  const int64_t n = A->rows;
  if (n == 0) {
    st->converged = 1;
    return;
  }
  hipStream_t s = ctx->stream;
  const bool precond = o.use_preconditioner != 0;
  ensure_sell(ctx, A, false, precond);
  Layout L;
  const size_t ob = L.add<double>(n), ox = L.add<double>(n);
  Layout dry = L;
  for (int i = 0; i < 3; ++i) dry.add<double>(n);
  dry.add<double>(n * (int64_t)(restart + 1));
  dry.add<double>(restart + 2);
  dry.add<double>((int64_t)(restart + 1) * restart);
  dry.add<double>(restart + 1);
  for (int i = 0; i < 3; ++i) dry.add<double>(restart);
  dry.add<double>(GRID_MAX);
  dry.add<double>(GMAX);
  dry.add<double>(8);
  dry.add<GmState>(1);
  char *base = ctx_scratch(ctx, dry.bytes);
  double *bi = (double *)(base + ob), *xi = (double *)(base + ox);
  const int *vp = A->sell->vperm.get();
  permute_gather(s, n, b, vp, bi);
  const PcgResult r = gmres_core(ctx, sell_op(A), A->sval.get(),
                                 precond ? A->sdinv.get() : nullptr, bi, xi,
                                 o.max_iterations, restart, o.tolerance, o.abs_tolerance,
                                 base, L);
  permute_scatter(s, n, xi, vp, x);
  GS_LAUNCH_CHECK("gmres_solve");
  GS_HIP(hipStreamSynchronize(s));
  st->iterations = r.iterations;
  st->residual_norm = r.rnorm;
  st->relative_residual = r.bnorm > 0.0 ? r.rnorm / r.bnorm : 0.0;
  st->converged = r.converged;
  st->spmv_count = r.spmv;
  st->solve_time =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Resolves a pending scaled operator (sscaled 2) with a synchronous read.
static void resolve_scaled(gs_ctx *ctx, gs_csr *A) {
  if (A->sscaled != 2) return;
  int bad = 0;
  GS_HIP(hipMemcpyAsync(&bad, A->sscale_bad.get(), sizeof(int), hipMemcpyDeviceToHost,
                        ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  A->sscaled = bad ? -1 : 1;
  A->sscale_bad = DArr<int>();
  if (bad) {
    A->sscale = DArr<double>();
    A->svalq = DArr<int>();
  }
}

double time_spmv(gs_ctx *ctx, const gs_csr *Ac, int precision, int reps) {
  gs_csr *A = const_cast<gs_csr *>(Ac);
  const int64_t n = A->rows;
  hipStream_t s = ctx->stream;
  if (n == 0 || reps <= 0) return 0.0;
  const bool fp32 = precision != GS_PREC_FP64;
  const bool sell = use_sell(A);
  if (sell) {
    ensure_sell(ctx, A, fp32, fp32);
  } else {
    ensure_idx16(ctx, A);
    ensure_csr_values(ctx, A);
  }
  // the mixed solve's SpMV on a scalar operator with a positive diagonal is
  // the two-launch pass's KA on the scaled values
  bool cg2 = false;
  if (sell && fp32 && A->sell->bs == 1) {
    ensure_scaled(ctx, A);
    resolve_scaled(ctx, A);
    cg2 = A->sscaled == 1;
  }
  if (!sell && fp32 && !A->val32) {
    A->val32 = dalloc<float>(A->nnz);
    hipLaunchKernelGGL(k_to_f32, (unsigned)GMAX, BS, 0, s, A->nnz, A->val.get(),
                       A->val32.get());
  }
  const size_t esz = fp32 ? sizeof(float) : sizeof(double);
  Layout L;
  const size_t opart = L.add<double>(2 * GRID_MAX), ost = L.add<CgState>(1);
  const size_t opr = L.add<double>(VGMAX);
  const size_t op = L.add<char>((int64_t)(n * esz)), oq = L.add<char>((int64_t)(n * esz));
  const size_t ow = L.add<char>((int64_t)(n * esz));
  char *base = ctx_scratch(ctx, L.bytes);
  double *part = (double *)(base + opart);
  double *part_r = (double *)(base + opr);
  CgState *st = (CgState *)(base + ost);
  void *p = base + op, *q = base + oq, *wv = base + ow;
  hipLaunchKernelGGL(k_state_setup, 1, 1, 0, s, st, 1ll << 62, 0.0, 0.0, 0, 0.0, 0);
  if (cg2) {
    hipLaunchKernelGGL(k_state_continue, 1, 1, 0, s, st);
    hipLaunchKernelGGL((k_zero<double>), 1, BS, 0, s, VGMAX, part_r);
  }
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
    hipLaunchKernelGGL((k_fill_pattern<T>), grid_vec(n), BS, 0, s, (int)n, (T *)q);
    hipLaunchKernelGGL((k_fill_pattern<T>), grid_vec(n), BS, 0, s, (int)n, (T *)wv);
    const SellOp o = sell ? sell_op(A) : SellOp{};
    auto once = [&] {
      if (cg2) {
        // p and s are rewritten in place: beta = 0 (gamma_prev = 0) keeps
        // them bounded across repetitions
        with_idx(*o.P, [&](auto idx) {
          hipLaunchKernelGGL((k_cg2_dir<T, decltype(idx), int>), o.grid, BS, 0, s, (int)n,
                             o.ns(), o.P->cb.get(), idx, (const int *)A->svalq.get(),
                             (const T *)p, (T *)q, (T *)wv, st, 0, part_r,
                             (int)grid_fat(n), part);
        }, sizeof(T));
      } else if (sell)
        launch_cg_q<T>(s, o, val, (const T *)p, (T *)q, st, 0, part);
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
