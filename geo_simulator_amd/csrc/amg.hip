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
// result in about O(nnz) instead of its O(n^2) scans: a two-level bitmap of
// point indices per count stands in for the scan for the largest count
// (smallest index first), and a sparse accumulator for the
// dense column loop of matrix_multiply (every product reaches its sum in the
// same order).  Rows of the strength, prolongation and Galerkin matrices are
// built on host threads, each row by one thread.
//
// V-cycle (:132-185) on the device, one thread per row (operators and
// prolongations sliced by 64 rows, rows sorted by length where that saves
// padding, a slice's lanes loading their rows' entries side by side; the few
// rows longer than 64 entries take a wavefront each, whose lanes form the
// products side by side for one lane to add in order): weighted Jacobi
// (omega 0.67) twice, the first step from x = 0 without the operator, the
// residual, restriction by P^T, the coarser cycle, prolongation by P,
// weighted Jacobi twice.  The last level runs the reference's ten in-place
// sweeps (Gauss-Seidel: rows in order, each reading the entries already
// updated): rows that read no other row and that no other row reads take
// their ten updates at once, one thread each; the others run in one
// workgroup, grouped in dependency levels: row i comes after every row
// j < i that it reads (its new value) and after every row k < i that reads
// x_i (the old value), so rows of one level neither read nor overwrite each
// other's entries and run at once; their rows, vectors and schedule are
// staged in LDS when they fit.
// Every row sum starts at 0.0 and adds in ascending column order, compiled
// without contraction, so the V-cycle equals the reference's bit for bit.
#include <algorithm>
#include <cmath>
#include <functional>
#include <memory>
#include <string>
#include <thread>
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

// Host threads for the row-parallel set-up loops (each row is computed by
// one thread in the reference's order, so the result does not depend on the
// thread count).
inline int host_threads(int64_t rows) {
  const unsigned h = std::thread::hardware_concurrency();
  const int64_t t = std::min<int64_t>(h ? h : 1, 16);
  return (int)std::max<int64_t>(1, std::min<int64_t>(t, rows / 4096));
}

// A CSR matrix built row by row: make() gives each thread its row function
// row(i, col, val), which appends row i's entries in order.
template <class Make>
HCsr build_rows(int64_t rows, int64_t cols, Make make) {
  const int T = host_threads(rows);
  std::vector<std::vector<int>> tc((size_t)T);
  std::vector<std::vector<double>> tv((size_t)T);
  HCsr C;
  C.rows = rows;
  C.cols = cols;
  C.rp.assign((size_t)rows + 1, 0);
  auto work = [&](int t) {
    auto row = make();
    const int64_t b = rows * t / T, e = rows * (t + 1) / T;
    for (int64_t i = b; i < e; ++i) {
      const size_t before = tc[t].size();
      row(i, tc[t], tv[t]);
      C.rp[i + 1] = (int64_t)(tc[t].size() - before);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
  work(0);
  for (auto &th : pool) th.join();
  for (int64_t i = 0; i < rows; ++i) C.rp[i + 1] += C.rp[i];
  C.col.resize((size_t)C.rp[rows]);
  C.val.resize((size_t)C.rp[rows]);
  for (int t = 0; t < T; ++t) {
    const int64_t at = C.rp[rows * t / T];
    std::copy(tc[t].begin(), tc[t].end(), C.col.begin() + at);
    std::copy(tv[t].begin(), tv[t].end(), C.val.begin() + at);
  }
  return C;
}

// strength_matrix (:359-390): the stored off-diagonal entries of row i with
// |a_ij| >= theta * max_{stored k != i} |a_ik| (f64::max skips NaN, as fmax)
HCsr strength(const HCsr &A, double theta) {
  return build_rows(A.rows, A.cols, [&] {
    return [&](int64_t i, std::vector<int> &col, std::vector<double> &val) {
      double mx = 0.0;
      for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
        if (A.col[p] != i) mx = std::fmax(mx, std::fabs(A.val[p]));
      const double thr = theta * mx;
      for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p)
        if (A.col[p] != i && std::fabs(A.val[p]) >= thr) {
          col.push_back(A.col[p]);
          val.push_back(A.val[p]);
        }
    };
  });
}

// A set of point indices as a two-level bitmap: insert and erase in O(1),
// the smallest member by scanning one summary word per 4096 points.  The
// bitmap is kept in pages of 4096 points, allocated when a member first
// falls in them, so a set holds memory in proportion to its members.
struct IndexSet {
  std::vector<uint64_t> summary;                     // bit w: page s, word w nonzero
  std::vector<std::unique_ptr<uint64_t[]>> page;     // 64 words each, or none
  int64_t count = 0;
  explicit IndexSet(int64_t n)
      : summary((size_t)((n + 4095) / 4096), 0), page((size_t)((n + 4095) / 4096)) {}
  void insert(int64_t i) {
    const int64_t s = i >> 12, w = (i >> 6) & 63;
    auto &pg = page[(size_t)s];
    if (!pg) pg.reset(new uint64_t[64]());
    pg[w] |= 1ull << (i & 63);
    summary[(size_t)s] |= 1ull << w;
    ++count;
  }
  void erase(int64_t i) {
    const int64_t s = i >> 12, w = (i >> 6) & 63;
    uint64_t *pg = page[(size_t)s].get();
    pg[w] &= ~(1ull << (i & 63));
    if (!pg[w]) summary[(size_t)s] &= ~(1ull << w);
    --count;
  }
  int64_t smallest() const {
    for (size_t s = 0; s < summary.size(); ++s)
      if (summary[s]) {
        const int w = __builtin_ctzll(summary[s]);
        return (int64_t)s * 4096 + w * 64 + __builtin_ctzll(page[s][w]);
      }
    return -1;
  }
};

// coarsen (:286-355) on the strength matrix S: 1 marks a coarse point.  The
// undecided points with a positive count sit in the set of their count; the
// top set's smallest index is the point the reference's scan picks (the
// largest count, the first index among equals).
std::vector<unsigned char> coarsen(const HCsr &S, int64_t &num_coarse) {
  const int64_t n = S.rows;
  std::vector<int64_t> lam((size_t)n);
  std::vector<unsigned char> cf((size_t)n, 0), open((size_t)n, 1);
  std::vector<std::unique_ptr<IndexSet>> bucket(1);
  int64_t top = 0;
  auto set_of = [&](int64_t l) -> IndexSet & {
    if (l >= (int64_t)bucket.size()) bucket.resize((size_t)l + 1);
    if (!bucket[(size_t)l]) bucket[(size_t)l].reset(new IndexSet(n));
    return *bucket[(size_t)l];
  };
  for (int64_t i = 0; i < n; ++i) {
    int64_t c = 0;
    for (int64_t p = S.rp[i]; p < S.rp[i + 1]; ++p) c += S.col[p] != i;
    lam[i] = c;
    if (c > 0) {
      set_of(c).insert(i);
      top = std::max(top, c);
    }
  }
  auto drop = [&](int64_t i) {  // i leaves the set of its count; an empty set is freed
    auto &b = bucket[(size_t)lam[i]];
    b->erase(i);
    if (b->count == 0) b.reset();
  };
  auto decide = [&](int64_t i) {  // i leaves the undecided points
    open[i] = 0;
    if (lam[i] > 0) drop(i);
  };
  int64_t nc = 0, left = n;
  while (left > 0) {
    while (top > 0 && (!bucket[(size_t)top] || bucket[(size_t)top]->count == 0)) --top;
    if (top == 0) {  // no undecided point has a strong connection left
      for (int64_t i = 0; i < n; ++i)
        if (open[i]) {
          cf[i] = 1;
          open[i] = 0;
          ++nc;
        }
      break;
    }
    const int64_t pick = bucket[(size_t)top]->smallest();
    cf[pick] = 1;
    decide(pick);
    ++nc;
    --left;
    for (int64_t p = S.rp[pick]; p < S.rp[pick + 1]; ++p) {
      const int64_t j = S.col[p];
      if (j == pick || !open[j]) continue;
      decide(j);  // fine
      --left;
      for (int64_t q = S.rp[j]; q < S.rp[j + 1]; ++q) {
        const int64_t k = S.col[q];
        if (k != j && open[k]) {
          if (lam[k] > 0) drop(k);
          ++lam[k];
          set_of(lam[k]).insert(k);
          top = std::max(top, lam[k]);
        }
      }
    }
  }
  num_coarse = nc;
  return cf;
}

// build_prolongation (:396-487), with the level's strength matrix S
HCsr prolongation(const HCsr &A, const HCsr &S, const std::vector<unsigned char> &cf,
                  int64_t nc) {
  const int64_t n = A.rows;
  std::vector<int> cidx((size_t)n, 0);
  for (int64_t i = 0, c = 0; i < n; ++i)
    if (cf[i]) cidx[i] = (int)c++;
  return build_rows(n, nc, [&] {
    return [&](int64_t i, std::vector<int> &col, std::vector<double> &val) {
      // the coarse entries of row i of M, |m_ij| / their sum, unless the sum
      // is at most 1e-14
      auto take = [&](const HCsr &M) {
        double sum = 0.0;
        for (int64_t p = M.rp[i]; p < M.rp[i + 1]; ++p)
          if (cf[M.col[p]]) sum += std::fabs(M.val[p]);
        if (!(sum > 1e-14)) return false;
        for (int64_t p = M.rp[i]; p < M.rp[i + 1]; ++p)
          if (cf[M.col[p]]) {
            col.push_back(cidx[M.col[p]]);
            val.push_back(std::fabs(M.val[p]) / sum);
          }
        return true;
      };
      if (cf[i]) {
        col.push_back(cidx[i]);
        val.push_back(1.0);
      } else if (!take(S) && !take(A) && nc > 0) {
        col.push_back(0);  // no coarse neighbour at all: the first coarse point
        val.push_back(1.0);
      }
    };
  });
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
  const int64_t nb = std::max<int64_t>(1, B.cols);
  return build_rows(A.rows, B.cols, [&] {
    // per thread: a dense accumulator over B's columns, the row that last
    // touched each column, and the row's columns
    auto acc = std::make_shared<std::vector<double>>((size_t)nb);
    auto seen = std::make_shared<std::vector<int>>((size_t)nb, -1);
    auto cols = std::make_shared<std::vector<int>>();
    return [&, acc, seen, cols](int64_t i, std::vector<int> &col, std::vector<double> &val) {
      double *const ac = acc->data();
      int *const sn = seen->data();
      const int64_t *const brp = B.rp.data();
      const int *const bcol = B.col.data();
      const double *const bval = B.val.data();
      const int row = (int)i;
      std::vector<int> &cs = *cols;
      cs.clear();
      for (int64_t p = A.rp[i]; p < A.rp[i + 1]; ++p) {
        const double a = A.val[p];
        const int64_t k = A.col[p];
        const int64_t q1 = brp[k + 1];
        for (int64_t q = brp[k]; q < q1; ++q) {
          const int j = bcol[q];
          if (sn[j] != row) {
            sn[j] = row;
            ac[j] = 0.0;
            cs.push_back(j);
          }
          ac[j] += a * bval[q];
        }
      }
      std::sort(cs.begin(), cs.end());
      for (int j : cs)
        if (std::fabs(ac[j]) > 1e-14) {
          col.push_back(j);
          val.push_back(ac[j]);
        }
    };
  });
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

// The in-place sweeps' schedule: the isolated rows (every stored column of
// the row is i, and column i is stored in no other row), whose ten updates
// touch no other row's entry; and the dependency levels of the others: row i
// after the rows j < i it reads and the rows k < i that read it.
void sweep_levels(const HCsr &A, std::vector<int> &iso, std::vector<int> &ptr,
                  std::vector<int> &ord) {
  const int64_t n = A.rows;
  std::vector<int> lv((size_t)n, -1);
  const HCsr T = transpose(A);
  auto alone = [](const HCsr &M, int64_t i) {
    for (int64_t p = M.rp[i]; p < M.rp[i + 1]; ++p)
      if (M.col[p] != i) return false;
    return true;
  };
  iso.clear();
  int nlev = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (alone(A, i) && alone(T, i)) {
      iso.push_back((int)i);
      continue;
    }
    int v = 0;
    for (int64_t p = A.rp[i]; p < A.rp[i + 1] && A.col[p] < i; ++p) v = std::max(v, lv[A.col[p]] + 1);
    for (int64_t p = T.rp[i]; p < T.rp[i + 1] && T.col[p] < i; ++p) v = std::max(v, lv[T.col[p]] + 1);
    lv[i] = v;
    nlev = std::max(nlev, v + 1);
  }
  ptr.assign((size_t)nlev + 1, 0);
  for (int v : lv)
    if (v >= 0) ++ptr[(size_t)v + 1];
  for (int l = 0; l < nlev; ++l) ptr[(size_t)l + 1] += ptr[(size_t)l];
  ord.resize((size_t)ptr[(size_t)nlev]);
  std::vector<int> at(ptr.begin(), ptr.end() - 1);
  for (int64_t i = 0; i < n; ++i)
    if (lv[i] >= 0) ord[(size_t)at[(size_t)lv[i]]++] = (int)i;
}

// The sliced copy of a level's rows: position q (row erow[q], or q) of slice
// q / 64 takes its row's entries in order at (ecb[s] + k) * 64 + q % 64, and
// column -1 with value 0 past its row's end up to the slice's width (the
// positions past the last row included).
__global__ void k_amg_fill_ell(int64_t n, int64_t npos, const int *__restrict__ rp,
                               const int *__restrict__ col, const double *__restrict__ val,
                               const int *__restrict__ ecb, const int *__restrict__ erow,
                               int *__restrict__ ecol, double *__restrict__ eval) {
  const int64_t q = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (q >= npos) return;
  const int64_t sl = q >> 6, lane = q & 63;
  int b = 0, len = 0;
  if (q < n) {
    const int64_t i = erow ? (int64_t)erow[q] : q;
    b = rp[i];
    len = rp[i + 1] - b;
  }
  const int c0 = ecb[sl], m = ecb[sl + 1] - c0;
  for (int k = 0; k < m; ++k) {
    const size_t at = (size_t)(c0 + k) * 64 + lane;
    ecol[at] = k < len ? col[b + k] : -1;
    eval[at] = k < len ? val[b + k] : 0.0;
  }
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
  // sliced copy: each slice as long as its longest row; rows sorted by
  // length (descending, then by index) when the natural order pads more
  // than a quarter
  const int64_t n = A.rows, ns = (n + 63) / 64;
  auto len = [&](int64_t i) { return A.rp[i + 1] - A.rp[i]; };
  auto padded = [&](const std::vector<int64_t> &at) {
    int64_t tot = 0;
    for (int64_t sl = 0; sl < ns; ++sl) {
      int64_t m = 0;
      for (int64_t q = 64 * sl; q < std::min<int64_t>(n, 64 * sl + 64); ++q)
        m = std::max(m, len(at.empty() ? q : at[q]));
      tot += 64 * m;
    }
    return tot;
  };
  std::vector<int64_t> pos;  // row at each position (empty: the identity)
  if (padded(pos) * 4 > A.nnz() * 5) {
    pos.resize((size_t)n);
    for (int64_t i = 0; i < n; ++i) pos[i] = i;
    std::stable_sort(pos.begin(), pos.end(),
                     [&](int64_t a, int64_t b) { return len(a) > len(b); });
  }
  auto row_at = [&](int64_t q) { return pos.empty() ? q : pos[q]; };
  std::vector<int> cb((size_t)ns + 1, 0);
  for (int64_t sl = 0; sl < ns; ++sl) {
    int64_t m = 0;
    for (int64_t q = 64 * sl; q < std::min<int64_t>(n, 64 * sl + 64); ++q)
      m = std::max(m, len(row_at(q)));
    if ((int64_t)cb[sl] + m >= INT32_MAX / 64)
      throw Error(GS_ERR_TOO_LARGE, "AMG level too large for its sliced copy");
    cb[sl + 1] = cb[sl] + (int)m;
  }
  const size_t ne = (size_t)cb[ns] * 64;
  // rows longer than a slice is wide: when they are few and lead the sorted
  // order, the slices that hold them go to one wavefront per row
  if (!pos.empty()) {
    int64_t nl = 0;
    while (nl < n && len(pos[nl]) > 64) ++nl;
    nl = std::min<int64_t>(n, (nl + 63) / 64 * 64);
    if (nl > 0 && nl * 8 <= n) D.n_long = nl;
  }
  int64_t longest = 0;
  for (int64_t q = D.n_long; q < n; ++q) longest = std::max(longest, len(row_at(q)));
  D.ew = longest > 72 ? 8 : 4;
  D.ecb = to_device(ctx, cb);
  if (!pos.empty()) D.erow = to_device(ctx, std::vector<int>(pos.begin(), pos.end()));
  // the slices' entries, filled on the device from the uploaded rows
  D.ecol = dalloc<int>((int64_t)ne);
  D.eval = dalloc<double>((int64_t)ne);
  if (ns > 0) {
    hipLaunchKernelGGL(k_amg_fill_ell, nblk(ns * 64), BS, 0, ctx->stream, n, ns * 64,
                       D.rp.get(), D.col.get(), D.val.get(), D.ecb.get(), D.erow.get(),
                       D.ecol.get(), D.eval.get());
    GS_LAUNCH_CHECK("k_amg_fill_ell");
  }
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

// Row i of a sliced copy (one lane per row of a 64-row slice, entries
// loaded side by side by the slice's lanes): the same sum as row_sum, the
// padding skipped by a select.
struct Ell {
  const int *__restrict__ cb;
  const int *__restrict__ col;
  const double *__restrict__ val;
  const int *__restrict__ row;  // row at each position (nullptr: the same)
};
inline Ell ell(const DevCsr &M) {
  return Ell{M.ecb.get(), M.ecol.get(), M.eval.get(), M.erow.get()};
}

// the row at position q of the slices
__device__ __forceinline__ int64_t ell_row(const Ell &M, int64_t q) {
  return M.row ? (int64_t)M.row[q] : q;
}

// (M x) of the row at position q: W entries per trip, the next trip's
// columns and values loaded while this trip's gathers are in flight (W = 4
// keeps more rows in flight on long operators; W = 8 shortens the chain of
// trips of levels with long rows, whose longest row sets the kernel's time)
template <int W>
__device__ __forceinline__ double ell_sum(const Ell &M, int64_t q, const double *x) {
  const int64_t s = q >> 6;
  const int lane = (int)(q & 63);
  const int b = M.cb[s], e = M.cb[s + 1];
  double sum = 0.0;
  int c[W];
  double a[W];
  auto load = [&](int k) {
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const bool in = k + u < e;
      const size_t at = (size_t)(in ? k + u : b) * 64 + lane;
      c[u] = in ? M.col[at] : -1;
      a[u] = in ? M.val[at] : 0.0;
    }
  };
  if (b < e) load(b);
  for (int k = b; k < e; k += W) {
    double v[W];
    int cc[W];
    double aa[W];
#pragma unroll
    for (int u = 0; u < W; ++u) {
      cc[u] = c[u];
      aa[u] = a[u];
      v[u] = x[cc[u] >= 0 ? cc[u] : 0];
    }
    if (k + W < e) load(k + W);
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const double t = sum + aa[u] * v[u];
      sum = cc[u] >= 0 ? t : sum;
    }
  }
  return sum;
}

__global__ void k_amg_zero(int64_t n, double *__restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

// The V-cycle's row operations, each on (M x)_i of one operator M:
//   JACOBI    weighted_jacobi_step (:250-263): xo_i = xi_i + omega (r_i - ax) dinv_i
//   RESIDUAL  res_i = r_i - ax (x = xi)
//   SPMV      y_i = ax (matvec, :230-247; restriction by P^T)
//   PROLONG   x_i += (P e)_i, the correction formed first, then added (:168-175)
enum AmgOp { OP_JACOBI, OP_RESIDUAL, OP_SPMV, OP_PROLONG };
struct OpArgs {
  const double *dinv, *r, *xi;  // xi: the vector M multiplies
  double *xo;                   // the output (PROLONG: x, updated in place)
};

template <int OP>
__device__ __forceinline__ void amg_epilogue(int64_t i, double ax, const OpArgs &a) {
  if constexpr (OP == OP_JACOBI) {
    const double t = AMG_OMEGA * (a.r[i] - ax);
    a.xo[i] = a.xi[i] + t * a.dinv[i];
  } else if constexpr (OP == OP_RESIDUAL) {
    a.xo[i] = a.r[i] - ax;
  } else if constexpr (OP == OP_SPMV) {
    a.xo[i] = ax;
  } else {
    a.xo[i] = a.xo[i] + ax;
  }
}

// positions [q0, n) of the slices, a lane per row
template <int OP, int W>
__global__ void k_amg_ell(int64_t q0, int64_t n, const Ell M, const OpArgs a) {
  const int64_t q = q0 + (int64_t)blockIdx.x * BS + threadIdx.x;
  if (q >= n) return;
  amg_epilogue<OP>(ell_row(M, q), ell_sum<W>(M, q, a.xi), a);
}

// The rows at positions [0, gridDim.x), longer than a slice is wide, a
// wavefront each: the lanes form 64 products at a time side by side (the
// row's CSR entries are contiguous) and lane 0 adds them in order, so the sum
// is the same as a lane's, with one round trip per 64 entries instead of one
// per W.
template <int OP>
__global__ __launch_bounds__(64) void k_amg_long(const int *__restrict__ pos,
                                                 const int *__restrict__ rp,
                                                 const int *__restrict__ col,
                                                 const double *__restrict__ val,
                                                 const OpArgs a) {
  __shared__ double prod[64];
  const int lane = threadIdx.x;
  const int64_t i = pos[blockIdx.x];
  const int b = rp[i], e = rp[i + 1];
  double sum = 0.0;
  for (int k0 = b; k0 < e; k0 += 64) {
    const int k = k0 + lane;
    prod[lane] = k < e ? val[k] * a.xi[col[k]] : 0.0;
    __syncthreads();
    if (lane == 0) {
      const int m = min(64, e - k0);
      for (int u = 0; u < m; ++u) sum = sum + prod[u];
    }
    __syncthreads();
  }
  if (lane == 0) amg_epilogue<OP>(i, sum, a);
}

// The first step from x = 0 with every operator entry finite: (A x)_i is
// +0.0, so xo_i = 0.0 + omega (r_i - 0.0) dinv_i, the same operations.
__global__ void k_amg_jacobi_zero(int64_t n, const double *__restrict__ dinv,
                                  const double *__restrict__ r, double *__restrict__ xo) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (i >= n) return;
  const double ax = 0.0;
  const double t = AMG_OMEGA * (r[i] - ax);
  xo[i] = 0.0 + t * dinv[i];
}

// One row operation over every row of M.
template <int OP>
void amg_op(hipStream_t s, const DevCsr &M, const OpArgs &a) {
  const int64_t n = M.rows;
  if (M.n_long > 0)
    hipLaunchKernelGGL(k_amg_long<OP>, (unsigned)M.n_long, 64, 0, s, M.erow.get(), M.rp.get(),
                       M.col.get(), M.val.get(), a);
  if (n > M.n_long) {
    const unsigned g = nblk(n - M.n_long);
    if (M.ew == 8)
      hipLaunchKernelGGL((k_amg_ell<OP, 8>), g, BS, 0, s, M.n_long, n, ell(M), a);
    else
      hipLaunchKernelGGL((k_amg_ell<OP, 4>), g, BS, 0, s, M.n_long, n, ell(M), a);
  }
}

// The last level's sweeps of its isolated rows: each row's ten updates at
// once (they read and write only its own entry).
__global__ void k_amg_sweeps_iso(int64_t n, const int *__restrict__ iso,
                                 const int *__restrict__ rp, const int *__restrict__ col,
                                 const double *__restrict__ val, const double *__restrict__ dinv,
                                 const double *__restrict__ r, double *__restrict__ x) {
  const int64_t t = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (t >= n) return;
  const int i = iso[t];
  for (int s = 0; s < AMG_SWEEPS; ++s) {
    const double ax = row_sum(rp[i], rp[i + 1], col, val, x);
    const double d = (r[i] - ax) * dinv[i];
    x[i] = x[i] + d;
  }
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

// The same sweeps with the coupled rows staged in LDS (when they fit): their
// operator rows in a local numbering that keeps the global order (so each row
// still sums in ascending column), x, r and the inverse diagonal; every read
// of a level step then stays on the CU.  Dynamic LDS, laid out as: x, r, dinv
// (m doubles each), values (nnz doubles), row pointers (m + 1), columns (nnz),
// level order (m) and level offsets (nlev + 1) as ints.
__global__ __launch_bounds__(GS_NT) void k_amg_sweeps_lds(
    int m, int nnz, int nlev, const int *__restrict__ grow, const int *__restrict__ lrp,
    const int *__restrict__ lcol, const double *__restrict__ lval,
    const int *__restrict__ lord, const int *__restrict__ lptr,
    const double *__restrict__ dinv, const double *__restrict__ r, double *__restrict__ x) {
  extern __shared__ double lds[];
  double *sx = lds, *sr = sx + m, *sd = sr + m, *sv = sd + m;
  int *srp = (int *)(sv + nnz);
  int *sc = srp + m + 1, *so = sc + nnz, *sp = so + m;
  for (int k = threadIdx.x; k < m; k += GS_NT) {
    const int g = grow[k];
    sx[k] = x[g];
    sr[k] = r[g];
    sd[k] = dinv[g];
    so[k] = lord[k];
  }
  for (int k = threadIdx.x; k <= m; k += GS_NT) srp[k] = lrp[k];
  for (int k = threadIdx.x; k < nnz; k += GS_NT) {
    sc[k] = lcol[k];
    sv[k] = lval[k];
  }
  for (int k = threadIdx.x; k <= nlev; k += GS_NT) sp[k] = lptr[k];
  __syncthreads();
  for (int s = 0; s < AMG_SWEEPS; ++s) {
    for (int l = 0; l < nlev; ++l) {
      for (int t = sp[l] + threadIdx.x; t < sp[l + 1]; t += GS_NT) {
        const int i = so[t];
        double ax = 0.0;
        for (int p = srp[i]; p < srp[i + 1]; ++p) ax += sv[p] * sx[sc[p]];
        const double d = (sr[i] - ax) * sd[i];
        sx[i] = sx[i] + d;
      }
      __syncthreads();
    }
  }
  for (int k = threadIdx.x; k < m; k += GS_NT) x[grow[k]] = sx[k];
}

constexpr size_t AMG_SWEEP_LDS = 64 << 10;  // bytes of the staged sweeps' LDS

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
      const HCsr S = strength(cur, theta);
      const std::vector<unsigned char> cf = coarsen(S, nc);
      last = nc == 0 || nc >= n;
      if (!last) P = prolongation(cur, S, cf, nc);
    }
    // the V-cycle ends here when this level has no P or the loop ends
    const bool end = last || level + 1 == max_levels;
    L.A = upload(ctx, cur);
    L.finite = std::all_of(cur.val.begin(), cur.val.end(),
                           [](double v) { return std::isfinite(v); });
    if (end) {
      std::vector<int> iso, ptr, ord;
      sweep_levels(cur, iso, ptr, ord);
      L.n_iso = (int64_t)iso.size();
      L.iso = to_device(ctx, iso);
      L.nlev = (int)ptr.size() - 1;
      L.lptr = to_device(ctx, ptr);
      L.ord = to_device(ctx, ord);
      // the coupled rows in a local numbering (ascending global row), for the
      // LDS-staged sweeps when they fit
      const int64_t m = (int64_t)ord.size();
      std::vector<int> grow(ord.begin(), ord.end());
      std::sort(grow.begin(), grow.end());
      std::vector<int> loc((size_t)n, -1);
      for (int64_t k = 0; k < m; ++k) loc[grow[k]] = (int)k;
      std::vector<int> lrp((size_t)m + 1, 0), lcol, lord((size_t)m);
      std::vector<double> lval;
      for (int64_t k = 0; k < m; ++k) {
        const int64_t i = grow[k];
        for (int64_t p = cur.rp[i]; p < cur.rp[i + 1]; ++p) {
          lcol.push_back(loc[cur.col[p]]);
          lval.push_back(cur.val[p]);
        }
        lrp[k + 1] = (int)lcol.size();
      }
      for (int64_t t = 0; t < m; ++t) lord[t] = loc[ord[t]];
      const int64_t nz = (int64_t)lcol.size();
      const size_t bytes = (size_t)(3 * m + nz) * sizeof(double) +
                           (size_t)((m + 1) + nz + m + (L.nlev + 1)) * sizeof(int);
      // (a coupled row reads only coupled rows: a column j != i stored in row
      // i makes j read by another row; checked here all the same)
      const bool closed = std::all_of(lcol.begin(), lcol.end(), [](int c) { return c >= 0; });
      if (m > 0 && closed && bytes <= AMG_SWEEP_LDS) {
        L.sweep_lds = bytes;
        L.lnnz = nz;
        L.grow = to_device(ctx, grow);
        L.lrp = to_device(ctx, lrp);
        L.lcol = to_device(ctx, lcol);
        L.lval = to_device(ctx, lval);
        L.lord = to_device(ctx, lord);
      }
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
  const int *rp = L.A.rp.get(), *col = L.A.col.get();
  const double *val = L.A.val.get(), *dinv = L.dinv.get();
  if (l + 1 == M->lv.size()) {
    hipLaunchKernelGGL(k_amg_zero, g, BS, 0, s, n, x);
    if (L.n_iso > 0)
      hipLaunchKernelGGL(k_amg_sweeps_iso, nblk(L.n_iso), BS, 0, s, L.n_iso, L.iso.get(), rp,
                         col, val, dinv, r, x);
    if (L.nlev > 0 && L.sweep_lds > 0)
      hipLaunchKernelGGL(k_amg_sweeps_lds, 1, GS_NT, L.sweep_lds, s, (int)L.ord.n, (int)L.lnnz,
                         L.nlev, L.grow.get(), L.lrp.get(), L.lcol.get(), L.lval.get(),
                         L.lord.get(), L.lptr.get(), dinv, r, x);
    else if (L.nlev > 0)
      hipLaunchKernelGGL(k_amg_sweeps, 1, GS_NT, 0, s, L.nlev, L.lptr.get(), L.ord.get(), rp,
                         col, val, dinv, r, x);
    GS_LAUNCH_CHECK("k_amg_sweeps");
    return;
  }
  double *t = L.t.get();
  // x starts at 0: the first step's product A x is +0.0 in every row (each
  // term a_ij 0.0 is a signed zero and the sum starts at +0.0) when every
  // entry is finite, so that step reads no operator
  auto smooth = [&](bool from_zero) {
    for (int k = 0; k < AMG_SMOOTH; ++k) {
      const double *xi = (k & 1) ? t : x;
      double *xo = (k & 1) ? x : t;
      if (k == 0 && from_zero && L.finite)
        hipLaunchKernelGGL(k_amg_jacobi_zero, g, BS, 0, s, n, dinv, r, xo);
      else
        amg_op<OP_JACOBI>(s, L.A, OpArgs{dinv, r, xi, xo});
    }
  };
  static_assert(AMG_SMOOTH % 2 == 0, "the smoothing steps end in x");
  if (!L.finite) hipLaunchKernelGGL(k_amg_zero, g, BS, 0, s, n, x);
  smooth(true);
  amg_op<OP_RESIDUAL>(s, L.A, OpArgs{nullptr, r, x, t});
  const AmgLevel &C = M->lv[l + 1];
  amg_op<OP_SPMV>(s, L.PT, OpArgs{nullptr, nullptr, t, C.r.get()});
  GS_LAUNCH_CHECK("k_amg_spmv");
  vcycle(M, l + 1, C.r.get(), C.x.get());
  amg_op<OP_PROLONG>(s, L.P, OpArgs{nullptr, nullptr, C.x.get(), x});
  smooth(false);
  GS_LAUNCH_CHECK("k_amg_jacobi");
}

void amg_apply(const gs_amg *M, const double *r, double *z) { vcycle(M, 0, r, z); }

}  // namespace gs
