// Algebraic multigrid preconditioner (AMGPreconditioner, src/linalg/amg.rs).
//
// Set-up on the host, V-cycle on the device.
//
// Set-up (AMGPreconditioner::new, :44-129) is the reference's: strength of
// connection (|a_ij| >= theta max_k |a_ik|), greedy coarsening (repeatedly
// the undecided point with the most strong connections, its strong
// neighbours made fine), direct interpolation from strong coarse neighbours,
// and the Galerkin operator P^T (A P) with products of magnitude <= 1e-14
// dropped.  The coarsening is sequential by definition (each choice changes
// the counts the next one reads), so the hierarchy is built on the host
// from A in the caller's numbering, with loops that give the reference's
// result at O(nnz log n) instead of its O(n^2) scans: a max-heap keyed by
// (count, smallest index) with stale entries skipped stands in for the scan
// for the largest count, and a sparse accumulator for the dense column loop
// of matrix_multiply (every product reaches its sum in the same order).
//
// V-cycle (:132-185) on the device, one thread per row: weighted Jacobi
// (omega 0.67) twice, the residual, restriction by P^T, the coarser cycle,
// prolongation by P, weighted Jacobi twice.  The last level runs the
// reference's ten in-place sweeps (Gauss-Seidel: rows in order, each reading
// the entries already updated) in one workgroup, rows grouped in dependency
// levels: row i comes after every row j < i that it reads (its new value)
// and after every row k < i that reads x_i (the old value), so rows of one
// level neither read nor overwrite each other's entries and run at once.
// Every row sum starts at 0.0 and adds in ascending column order, compiled
// without contraction, so the V-cycle equals the reference's bit for bit.
#include <algorithm>
#include <cmath>
#include <queue>
#include <string>
#include <vector>

#include "internal.hpp"

namespace gs {
namespace {

constexpr int BS = 256;
constexpr int GS_NT = 1024;  // threads of the last level's sweep kernel
constexpr double AMG_OMEGA = 0.67;
constexpr int AMG_SMOOTH = 2;     // num_smoothing_steps (:126)
constexpr int AMG_SWEEPS = 10;    // the last level's sweeps (:138)

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + BS - 1) / BS); }

// ---- host set-up ----------------------------------------------------------
struct HCsr {
  int64_t rows = 0, cols = 0;
  std::vector<int64_t> rp{0};
  std::vector<int> col;
  std::vector<double> val;
  int64_t nnz() const { return (int64_t)col.size(); }
};

// strength_matrix (:359-390): the stored off-diagonal entries of row i with
// |a_ij| >= theta * max_{stored k != i} |a_ik| (f64::max skips NaN, as fmax)
HCsr strength(const HCsr &A, double theta) {
  HCsr S;
  S.rows = A.rows;
  S.cols = A.cols;
  S.rp.assign((size_t)A.rows + 1, 0);
  for (int64_t i = 0; i < A.rows; ++i) {
    double mx = 0.0;
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
      if (A.col[p] != i) mx = std::fmax(mx, std::fabs(A.val[p]));
    const double thr = theta * mx;
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
      if (A.col[p] != i && std::fabs(A.val[p]) >= thr) {
        S.col.push_back(A.col[p]);
        S.val.push_back(A.val[p]);
      }
    S.rp[i + 1] = S.nnz();
  }
  return S;
}

// coarsen (:286-355): 1 marks a coarse point
std::vector<unsigned char> coarsen(const HCsr &A, double theta, int64_t &num_coarse) {
  const HCsr S = strength(A, theta);
  const int64_t n = A.rows;
  std::vector<int64_t> lam((size_t)n);
  std::vector<unsigned char> cf((size_t)n, 0), open((size_t)n, 1);
  // top: the largest count, then the smallest index
  auto later = [](const std::pair<int64_t, int64_t> &a, const std::pair<int64_t, int64_t> &b) {
    return a.first < b.first || (a.first == b.first && a.second > b.second);
  };
  std::priority_queue<std::pair<int64_t, int64_t>, std::vector<std::pair<int64_t, int64_t>>,
                      decltype(later)>
      heap(later);
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = 0;
    for (int64_t p = S.rp[i]; p < S.rp[i + 1]; ++p) c += S.col[p] != i;
    lam[i] = c;
    if (c > 0) heap.emplace(c, i);
  }
  int64_t nc = 0, left = n;
  while (left > 0) {
    int64_t pick = -1;
    while (!heap.empty()) {
      const auto t = heap.top();
      heap.pop();
      if (open[t.second] && lam[t.second] == t.first) {
        pick = t.second;
        break;
      }
    }
    if (pick < 0) {  // no undecided point has a strong connection left
      for (int64_t i = 0; i < n; ++i)
        if (open[i]) {
          cf[i] = 1;
          open[i] = 0;
          ++nc;
        }
      break;
    }
    cf[pick] = 1;
    open[pick] = 0;
    ++nc;
    --left;
    for (int64_t p = S.rp[pick]; p < S.rp[pick + 1]; ++p) {
      const int64_t j = S.col[p];
      if (j == pick || !open[j]) continue;
      open[j] = 0;  // fine
      --left;
      for (int64_t q = S.rp[j]; q < S.rp[j + 1]; ++q) {
        const int64_t k = S.col[q];
        if (k != j && open[k]) heap.emplace(++lam[k], k);
      }
    }
  }
  num_coarse = nc;
  return cf;
}

// build_prolongation (:396-487)
HCsr prolongation(const HCsr &A, const std::vector<unsigned char> &cf, int64_t nc,
                  double theta) {
  const HCsr S = strength(A, theta);
  const int64_t n = A.rows;
  std::vector<int> cidx((size_t)n, 0);
  for (int64_t i = 0, c = 0; i < n; ++i)
    if (cf[i]) cidx[i] = (int)c++;
  HCsr P;
  P.rows = n;
  P.cols = nc;
  P.rp.assign((size_t)n + 1, 0);
  auto take = [&](const HCsr &M, int64_t i) {  // the coarse entries of row i, |a| / sum
    double sum = 0.0;
    for (int64_t p = M.rp[i]; p < M.rp[i + 1]; ++p)
      if (cf[M.col[p]]) sum += std::fabs(M.val[p]);
    if (!(sum > 1e-14)) return false;
    for (int64_t p = M.rp[i]; p < M.rp[i + 1]; ++p)
      if (cf[M.col[p]]) {
        P.col.push_back(cidx[M.col[p]]);
        P.val.push_back(std::fabs(M.val[p]) / sum);
      }
    return true;
  };
  for (int64_t i = 0; i < n; ++i) {
    if (cf[i]) {
      P.col.push_back(cidx[i]);
      P.val.push_back(1.0);
    } else if (!take(S, i) && !take(A, i) && nc > 0) {
      P.col.push_back(0);  // no coarse neighbour at all: the first coarse point
      P.val.push_back(1.0);
    }
    P.rp[i + 1] = P.nnz();
  }
  return P;
}

HCsr transpose(const HCsr &A) {
  HCsr T;
  T.rows = A.cols;
  T.cols = A.rows;
  T.rp.assign((size_t)A.cols + 1, 0);
  for (int c : A.col) ++T.rp[(size_t)c + 1];
  for (int64_t j = 0; j < A.cols; ++j) T.rp[j + 1] += T.rp[j];
  T.col.resize(A.col.size());
  T.val.resize(A.val.size());
  std::vector<int64_t> at(T.rp.begin(), T.rp.end() - 1);
  for (int64_t i = 0; i < A.rows; ++i)
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
      const int64_t q = at[A.col[p]]++;
      T.col[q] = (int)i;
      T.val[q] = A.val[p];
    }
  return T;
}

// matrix_multiply (:195-226): C = A B, |c_ij| > 1e-14 kept
HCsr matmul(const HCsr &A, const HCsr &B) {
  HCsr C;
  C.rows = A.rows;
  C.cols = B.cols;
  C.rp.assign((size_t)A.rows + 1, 0);
  std::vector<double> acc((size_t)std::max<int64_t>(1, B.cols));
  std::vector<int64_t> seen((size_t)std::max<int64_t>(1, B.cols), -1);
  std::vector<int> cols;
  for (int64_t i = 0; i < A.rows; ++i) {
    cols.clear();
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
      const double a = A.val[p];
      const int64_t k = A.col[p];
      for (int64_t q = B.rp[k]; q < B.rp[k + 1]; ++q) {
        const int j = B.col[q];
        if (seen[j] != i) {
          seen[j] = i;
          acc[j] = 0.0;
          cols.push_back(j);
        }
        acc[j] += a * B.val[q];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int j : cols)
      if (std::fabs(acc[j]) > 1e-14) {
        C.col.push_back(j);
        C.val.push_back(acc[j]);
      }
    C.rp[i + 1] = C.nnz();
  }
  return C;
}

// extract_diag_inv (:266-280)
std::vector<double> diag_inv(const HCsr &A) {
  std::vector<double> d((size_t)A.rows, 1.0);
  for (int64_t i = 0; i < A.rows; ++i)
    for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
      if (A.col[p] == i) {
        if (std::fabs(A.val[p]) > 1e-14) d[i] = 1.0 / A.val[p];
        break;
      }
  return d;
}

// Dependency levels of the in-place sweep: row i after the rows j < i it
// reads and the rows k < i that read it.
void sweep_levels(const HCsr &A, std::vector<int> &ptr, std::vector<int> &ord) {
  const int64_t n = A.rows;
  std::vector<int> lv((size_t)n, 0);
  const HCsr T = transpose(A);
  int nlev = n > 0 ? 1 : 0;
  for (int64_t i = 0; i < n; ++i) {
    int v = 0;
    for (int64_t p = A.rp[i]; p < A.rp[i + 1] && A.col[p] < i; ++p) v = std::max(v, lv[A.col[p]] + 1);
    for (int64_t p = T.rp[i]; p < T.rp[i + 1] && T.col[p] < i; ++p) v = std::max(v, lv[T.col[p]] + 1);
    lv[i] = v;
    nlev = std::max(nlev, v + 1);
  }
  ptr.assign((size_t)nlev + 1, 0);
  for (int v : lv) ++ptr[(size_t)v + 1];
  for (int l = 0; l < nlev; ++l) ptr[(size_t)l + 1] += ptr[(size_t)l];
  ord.resize((size_t)n);
  std::vector<int> at(ptr.begin(), ptr.end() - 1);
  for (int64_t i = 0; i < n; ++i) ord[(size_t)at[(size_t)lv[i]]++] = (int)i;
}

DevCsr upload(gs_ctx *ctx, const HCsr &A) {
  if (A.nnz() >= INT32_MAX)
    throw Error(GS_ERR_TOO_LARGE, "AMG level with 2^31 or more stored entries");
  DevCsr D;
  D.rows = A.rows;
  D.nnz = A.nnz();
  std::vector<int> rp(A.rp.begin(), A.rp.end());
  D.rp = to_device(ctx, rp);
  D.col = to_device(ctx, A.col);
  D.val = to_device(ctx, A.val);
  return D;
}

// ---- V-cycle kernels (one thread per row; sums from 0.0 in stored order) --
__device__ __forceinline__ double row_sum(int b, int e, const int *__restrict__ col,
                                          const double *__restrict__ val, const double *x) {
  double sum = 0.0;
  int p = b;
  for (; p + 8 <= e; p += 8) {  // eight entries' loads together, additions in order
    int c[8];
    double a[8], v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c[u] = col[p + u];
      a[u] = val[p + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) sum += a[u] * v[u];
  }
  for (; p < e; ++p) sum += val[p] * x[col[p]];
  return sum;
}

__global__ void k_amg_zero(int64_t n, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// weighted_jacobi_step (:250-263): xo_i = xi_i + omega (r_i - (A xi)_i) dinv_i
__global__ void k_amg_jacobi(int64_t n, const int *__restrict__ rp, const int *__restrict__ col,
                             const double *__restrict__ val, const double *__restrict__ dinv,
                             const double *__restrict__ r, const double *__restrict__ xi,
                             double *__restrict__ xo) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double ax = row_sum(rp[i], rp[i + 1], col, val, xi);
  const double t = AMG_OMEGA * (r[i] - ax);
  xo[i] = xi[i] + t * dinv[i];
}

// res_i = r_i - (A x)_i
__global__ void k_amg_residual(int64_t n, const int *__restrict__ rp, const int *__restrict__ col,
                               const double *__restrict__ val, const double *__restrict__ r,
                               const double *__restrict__ x, double *__restrict__ res) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  res[i] = r[i] - row_sum(rp[i], rp[i + 1], col, val, x);
}

// y = M x (matvec, :230-247)
__global__ void k_amg_spmv(int64_t n, const int *__restrict__ rp, const int *__restrict__ col,
                           const double *__restrict__ val, const double *__restrict__ x,
                           double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = row_sum(rp[i], rp[i + 1], col, val, x);
}

// x_i += (P e)_i  (the correction formed first, then added, :168-175)
__global__ void k_amg_prolong(int64_t n, const int *__restrict__ rp, const int *__restrict__ col,
                              const double *__restrict__ val, const double *__restrict__ e,
                              double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double c = row_sum(rp[i], rp[i + 1], col, val, e);
  x[i] = x[i] + c;
}

// The last level's sweeps (:136-146), one workgroup: levels in order, the
// rows of a level at once, a barrier (with its workgroup-scope fence)
// between levels and sweeps.
__global__ __launch_bounds__(GS_NT) void k_amg_sweeps(
    int nlev, const int *__restrict__ lptr, const int *__restrict__ ord,
    const int *__restrict__ rp, const int *__restrict__ col, const double *__restrict__ val,
    const double *__restrict__ dinv, const double *__restrict__ r, double *x) {
  for (int s = 0; s < AMG_SWEEPS; ++s) {
    for (int l = 0; l < nlev; ++l) {
      for (int t = lptr[l] + threadIdx.x; t < lptr[l + 1]; t += GS_NT) {
        const int i = ord[t];
        const double ax = row_sum(rp[i], rp[i + 1], col, val, x);
        const double d = (r[i] - ax) * dinv[i];
        x[i] = x[i] + d;
      }
      __syncthreads();
    }
  }
}

}  // namespace

gs_amg *amg_create(gs_ctx *ctx, const gs_csr *A, int64_t max_levels, int64_t coarse_size,
                   double theta) {
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "AMG needs a square matrix");
  if (max_levels < 1) throw Error(GS_ERR_INVALID_ARG, "AMG needs max_levels >= 1");
  if (coarse_size < 0) throw Error(GS_ERR_INVALID_ARG, "AMG needs coarse_size >= 0");
  ensure_csr_values(ctx, A);
  std::unique_ptr<gs_amg> M(new gs_amg());
  M->ctx = ctx;
  M->n = A->rows;
  // the operator in the caller's numbering, on the host
  HCsr cur;
  {
    CallerCsr C = caller_csr(ctx, A);
    const std::vector<int> rp = to_host(ctx, C.rp.get(), A->rows + 1);
    cur.rows = cur.cols = A->rows;
    cur.rp.assign(rp.begin(), rp.end());
    cur.col = to_host(ctx, C.col.get(), A->nnz);
    cur.val = to_host(ctx, C.val.get(), A->nnz);
  }
  for (int64_t level = 0; level < max_levels; ++level) {
    AmgLevel L;
    const int64_t n = cur.rows;
    L.dinv = to_device(ctx, diag_inv(cur));
    L.x = dalloc<double>(n);
    L.t = dalloc<double>(n);
    L.r = dalloc<double>(n);
    bool last = n <= coarse_size;  // kept without P (:57-66, :74-81)
    HCsr P;
    int64_t nc = 0;
    if (!last) {
      const std::vector<unsigned char> cf = coarsen(cur, theta, nc);
      last = nc == 0 || nc >= n;
      if (!last) P = prolongation(cur, cf, nc, theta);
    }
    // the V-cycle ends here when this level has no P or the loop ends
    const bool end = last || level + 1 == max_levels;
    L.A = upload(ctx, cur);
    if (end) {
      std::vector<int> ptr, ord;
      sweep_levels(cur, ptr, ord);
      L.nlev = (int)ptr.size() - 1;
      L.lptr = to_device(ctx, ptr);
      L.ord = to_device(ctx, ord);
    }
    if (!last) {
      const HCsr PT = transpose(P);
      L.P = upload(ctx, P);
      L.PT = upload(ctx, PT);
      L.has_p = true;
      L.coarse = nc;
      if (!end) cur = matmul(PT, matmul(cur, P));  // A_c = P^T (A P) (:105-113)
    }
    M->lv.push_back(std::move(L));
    if (end) break;
  }
  return M.release();
}

static void vcycle(const gs_amg *M, size_t l, const double *r, double *x) {
  hipStream_t s = M->ctx->stream;
  const AmgLevel &L = M->lv[l];
  const int64_t n = L.A.rows;
  if (n == 0) return;
  const unsigned g = nblk(n);
  hipLaunchKernelGGL(k_amg_zero, g, BS, 0, s, n, x);
  const int *rp = L.A.rp.get(), *col = L.A.col.get();
  const double *val = L.A.val.get(), *dinv = L.dinv.get();
  if (l + 1 == M->lv.size()) {
    hipLaunchKernelGGL(k_amg_sweeps, 1, GS_NT, 0, s, L.nlev, L.lptr.get(), L.ord.get(), rp,
                       col, val, dinv, r, x);
    GS_LAUNCH_CHECK("k_amg_sweeps");
    return;
  }
  double *t = L.t.get();
  auto smooth = [&] {
    for (int k = 0; k < AMG_SMOOTH; ++k) {
      const double *xi = (k & 1) ? t : x;
      double *xo = (k & 1) ? x : t;
      hipLaunchKernelGGL(k_amg_jacobi, g, BS, 0, s, n, rp, col, val, dinv, r, xi, xo);
    }
  };
  static_assert(AMG_SMOOTH % 2 == 0, "the smoothing steps end in x");
  smooth();
  hipLaunchKernelGGL(k_amg_residual, g, BS, 0, s, n, rp, col, val, r, x, t);
  const AmgLevel &C = M->lv[l + 1];
  hipLaunchKernelGGL(k_amg_spmv, nblk(L.coarse), BS, 0, s, L.coarse, L.PT.rp.get(),
                     L.PT.col.get(), L.PT.val.get(), t, C.r.get());
  GS_LAUNCH_CHECK("k_amg_spmv");
  vcycle(M, l + 1, C.r.get(), C.x.get());
  hipLaunchKernelGGL(k_amg_prolong, g, BS, 0, s, n, L.P.rp.get(), L.P.col.get(),
                     L.P.val.get(), C.x.get(), x);
  smooth();
  GS_LAUNCH_CHECK("k_amg_jacobi");
}

void amg_apply(const gs_amg *M, const double *r, double *z) { vcycle(M, 0, r, z); }

}  // namespace gs
