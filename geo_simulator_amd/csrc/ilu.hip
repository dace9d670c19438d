// ILU(0) preconditioner on the device (ILUPreconditioner,
// src/linalg/preconditioner.rs:78-204).
//
// The factorization is defined by the row order, so it runs in the caller's
// numbering: a matrix held in the library's locality order is first
// renumbered back (rows and columns through dperm, columns sorted).  Row i is
// eliminated by the rows k < i of its lower part, in ascending k, each
// followed by a merge over row k's upper part -- the reference's loop -- and
// needs only the finished rows k.  Rows are grouped in dependency levels
// (level(i) = 1 + the largest level of its lower columns; the backward
// substitution uses the upper columns), computed once on the host from the
// pattern.  The factorization launches once per level with a wavefront per
// row, the substitutions with a thread per row.  Each row's
// arithmetic is the reference's, in its order (compiled without
// contraction), so factors and substitutions are bit-identical to the
// reference's sequential loops.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "internal.hpp"

namespace gs {
namespace {

constexpr int BS = 256;

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + BS - 1) / BS); }

// key of nonzero k of internal row r: (caller row << 32) | caller column
__global__ void k_ilu_keys(int64_t n, const int *__restrict__ rp, const int *__restrict__ col,
                           const int *__restrict__ dperm, uint64_t *__restrict__ keys,
                           int *__restrict__ idx, int *__restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const uint64_t R = (uint64_t)dperm[r];
  cnt[R] = rp[r + 1] - rp[r];
  for (int k = rp[r]; k < rp[r + 1]; ++k) {
    keys[k] = (R << 32) | (uint64_t)(uint32_t)dperm[col[k]];
    idx[k] = k;
  }
}

__global__ void k_ilu_cols(int64_t nnz, const uint64_t *__restrict__ keys,
                           const int *__restrict__ idx, const double *__restrict__ val,
                           int *__restrict__ col, double *__restrict__ f) {
  const int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (k >= nnz) return;
  col[k] = (int)(uint32_t)keys[k];
  f[k] = val[idx[k]];
}

__global__ void k_ilu_copy(int64_t nnz, const int *__restrict__ idx,
                           const double *__restrict__ val, double *__restrict__ f) {
  const int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (k < nnz) f[k] = idx ? val[idx[k]] : val[k];
}

// one level of factorize_in_place (src/linalg/preconditioner.rs:131-160), a
// wavefront per row i.  The eliminating rows k of i's lower part go in
// ascending order, as in the reference; the merge over row k's upper part is
// spread over the lanes: each lane takes columns c of row k past its
// diagonal, finds c among row i's columns after k by binary search, and
// updates that entry.  The reference's merge updates exactly the entries of
// row i whose column appears in both (it stops early only past row i's last
// column), each once per k, so every entry receives the same operations in
// the same order.  A row is a one-wavefront workgroup, and a barrier (with
// its workgroup-scope fence) after each k makes the updated entries visible
// to every lane before the next k reads them.  The first
// failure in the reference's order is kept as the smallest key
// (2 i + phase) << 31 | row: phase 0 an eliminating row's pivot (the row
// reported is that row k), phase 1 row i's own pivot after its elimination.
__global__ __launch_bounds__(64) void k_ilu_factor(const int *__restrict__ ord,
                                                   const int *__restrict__ rp,
                                                   const int *__restrict__ col,
                                                   const int *__restrict__ diag, double *f,
                                                   unsigned long long *__restrict__ err) {
  const int lane = threadIdx.x;
  const int i = ord[blockIdx.x];
  const int row_start = rp[i], diag_i = diag[i], row_end = rp[i + 1];
  for (int kk = row_start; kk < diag_i; ++kk) {
    const int k = col[kk];
    const int k_diag = diag[k], k_end = rp[k + 1];
    const double dvk = f[k_diag];
    if (fabs(dvk) < 1e-15) {
      if (lane == 0) atomicMin(err, ((unsigned long long)(2ll * i) << 31) | (unsigned long long)k);
      return;
    }
    const double val_ik = f[kk] / dvk;
    if (lane == 0) f[kk] = val_ik;
    for (int kj = k_diag + 1 + lane; kj < k_end; kj += 64) {
      const int cj = col[kj];
      int lo = kk + 1, hi = row_end;  // first position of row i with column >= cj
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (col[mid] < cj) lo = mid + 1;
        else hi = mid;
      }
      if (lo < row_end && col[lo] == cj) f[lo] -= val_ik * f[kj];
    }
    __syncthreads();
  }
  if (lane == 0 && fabs(f[diag_i]) < 1e-15)
    atomicMin(err, ((unsigned long long)(2ll * i + 1) << 31) | (unsigned long long)i);
}

// One level of the substitutions, a thread per row: forward
// z_i = r_i - sum_{k<i} L_ik z_k, backward z_i = (z_i - sum_{j>i} U_ij z_j) / U_ii,
// each sum in ascending column.  (One workgroup walking all levels with a
// barrier between them measured slower: a level's row latency then runs on
// one CU.)

// sum_{idx in [b, e)} f[idx] z[col[idx]], added in ascending idx; the loads
// of eight entries are issued together (a row's entries are independent
// loads, only the additions are ordered)
__device__ __forceinline__ double row_dot(int b, int e, const int *__restrict__ col,
                                          const double *__restrict__ f, const double *z) {
  double sum = 0.0;
  int idx = b;
  for (; idx + 8 <= e; idx += 8) {
    int c[8];
    double a[8], x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      c[u] = col[idx + u];
      a[u] = f[idx + u];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = z[c[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u) sum += a[u] * x[u];
  }
  for (; idx < e; ++idx) sum += f[idx] * z[col[idx]];
  return sum;
}

__global__ void k_ilu_lower(int cnt, const int *__restrict__ ord, const int *__restrict__ rp,
                            const int *__restrict__ col, const int *__restrict__ diag,
                            const double *__restrict__ f, const double *__restrict__ r,
                            double *z) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int i = ord[t];
  z[i] = r[i] - row_dot(rp[i], diag[i], col, f, z);
}

__global__ void k_ilu_upper(int cnt, const int *__restrict__ ord, const int *__restrict__ rp,
                            const int *__restrict__ col, const int *__restrict__ diag,
                            const double *__restrict__ f, double *z) {
  const int t = blockIdx.x * BS + threadIdx.x;
  if (t >= cnt) return;
  const int i = ord[t];
  const int d = diag[i];
  z[i] = (z[i] - row_dot(d + 1, rp[i + 1], col, f, z)) / f[d];
}

__global__ void k_ilu_same(int64_t m, const int *__restrict__ a, const int *__restrict__ b,
                           int *__restrict__ flag) {
  const int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (k < m && a[k] != b[k]) *flag = 1;
}

__global__ void k_ilu_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }
__global__ void k_ilu_set_int(int *p, int v) { *p = v; }

// rows grouped by level: ord (rows, level by level, ascending within one)
// and the level offsets
void levels(const std::vector<int> &lv, int nlev, std::vector<int> &ptr, std::vector<int> &ord) {
  ptr.assign((size_t)nlev + 1, 0);
  for (int v : lv) ++ptr[(size_t)v + 1];
  for (int l = 0; l < nlev; ++l) ptr[(size_t)l + 1] += ptr[(size_t)l];
  ord.resize(lv.size());
  std::vector<int> at(ptr.begin(), ptr.end() - 1);
  for (size_t i = 0; i < lv.size(); ++i) ord[(size_t)at[(size_t)lv[i]]++] = (int)i;
}

}  // namespace

// A in caller numbering with sorted columns (the matrix the reference holds).
CallerCsr caller_csr(gs_ctx *ctx, const gs_csr *A) {
  hipStream_t s = ctx->stream;
  const int64_t n = A->rows, nnz = A->nnz;
  CallerCsr C;
  C.val = dalloc<double>(nnz);
  if (!A->dperm) {
    C.rp = dalloc<int>(n + 1);
    C.col = dalloc<int>(nnz);
    GS_HIP(hipMemcpyAsync(C.rp.get(), A->row_ptr.get(), (n + 1) * sizeof(int),
                          hipMemcpyDeviceToDevice, s));
    if (nnz > 0) {
      GS_HIP(hipMemcpyAsync(C.col.get(), A->col.get(), nnz * sizeof(int),
                            hipMemcpyDeviceToDevice, s));
      hipLaunchKernelGGL(k_ilu_copy, nblk(nnz), BS, 0, s, nnz, (const int *)nullptr,
                         A->val.get(), C.val.get());
      GS_LAUNCH_CHECK("k_ilu_copy");
    }
    return C;
  }
  auto keys = dalloc<uint64_t>(nnz), keys2 = dalloc<uint64_t>(nnz);
  auto idx = dalloc<int>(nnz);
  C.src = dalloc<int>(nnz);
  auto cnt = dalloc<int>(n + 1);
  C.rp = dalloc<int>(n + 1);
  GS_HIP(hipMemsetAsync(cnt.get(), 0, (n + 1) * sizeof(int), s));
  if (n > 0) {
    hipLaunchKernelGGL(k_ilu_keys, nblk(n), BS, 0, s, n, A->row_ptr.get(), A->col.get(),
                       A->dperm.get(), keys.get(), idx.get(), cnt.get());
    GS_LAUNCH_CHECK("k_ilu_keys");
  }
  // row pointers: exclusive scan of the caller rows' lengths
  {
    size_t bytes = 0;
    GS_HIP(rocprim::exclusive_scan(nullptr, bytes, cnt.get(), C.rp.get(), 0, (size_t)(n + 1),
                                   rocprim::plus<int>(), s));
    auto tmp = dalloc<char>((int64_t)bytes);
    GS_HIP(rocprim::exclusive_scan(tmp.get(), bytes, cnt.get(), C.rp.get(), 0, (size_t)(n + 1),
                                   rocprim::plus<int>(), s));
    GS_HIP(hipStreamSynchronize(s));
  }
  C.col = dalloc<int>(nnz);
  if (nnz > 0) {
    size_t bytes = 0;
    GS_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys.get(), keys2.get(), idx.get(),
                                     C.src.get(), (size_t)nnz, 0u, 64u, s));
    auto tmp = dalloc<char>((int64_t)bytes);
    GS_HIP(rocprim::radix_sort_pairs(tmp.get(), bytes, keys.get(), keys2.get(), idx.get(),
                                     C.src.get(), (size_t)nnz, 0u, 64u, s));
    hipLaunchKernelGGL(k_ilu_cols, nblk(nnz), BS, 0, s, nnz, keys2.get(), C.src.get(),
                       A->val.get(), C.col.get(), C.val.get());
    GS_LAUNCH_CHECK("k_ilu_cols");
    GS_HIP(hipStreamSynchronize(s));
  }
  return C;
}


void ilu_factor(gs_ilu *M) {
  gs_ctx *ctx = M->ctx;
  hipStream_t s = ctx->stream;
  auto err = dalloc<unsigned long long>(1);
  hipLaunchKernelGGL(k_ilu_set_u64, 1, 1, 0, s, err.get(), ULLONG_MAX);
  for (size_t l = 0; l + 1 < M->lvl_lower.size(); ++l) {
    const int b = M->lvl_lower[l], cnt = M->lvl_lower[l + 1] - b;
    hipLaunchKernelGGL(k_ilu_factor, (unsigned)cnt, 64, 0, s, M->ord_lower.get() + b,
                       M->rp.get(), M->col.get(), M->diag.get(), M->f.get(), err.get());
  }
  GS_LAUNCH_CHECK("k_ilu_factor");
  const unsigned long long e = to_host(ctx, err.get(), 1)[0];
  if (e != ULLONG_MAX) {
    const long long phase = (long long)(e >> 31) & 1, row = (long long)(e & 0x7fffffffull);
    const std::string msg = (phase ? "Zero diag after update at " : "Zero diag at ") +
                            std::to_string(row);
    throw Error(GS_ERR_ZERO_PIVOT, msg);
  }
}

gs_ilu *ilu_create(gs_ctx *ctx, const gs_csr *A) {
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "ILU(0) needs a square matrix");
  ensure_csr_values(ctx, A);
  std::unique_ptr<gs_ilu> M(new gs_ilu());
  M->ctx = ctx;
  M->n = A->rows;
  M->nnz = A->nnz;
  CallerCsr C = caller_csr(ctx, A);
  M->rp = C.rp;
  M->col = C.col;
  M->f = C.val;
  // diagonal positions and dependency levels, on the host from the pattern
  const int64_t n = M->n;
  const std::vector<int> rp = to_host(ctx, M->rp.get(), n + 1);
  const std::vector<int> col = to_host(ctx, M->col.get(), M->nnz);
  std::vector<int> diag((size_t)n), lo((size_t)n, 0), up((size_t)n, 0);
  for (int64_t i = 0; i < n; ++i) {
    int d = -1;
    for (int idx = rp[i]; idx < rp[i + 1]; ++idx)
      if (col[idx] == i) {
        d = idx;
        break;
      }
    if (d < 0) throw Error(GS_ERR_ZERO_PIVOT, "Row " + std::to_string(i) + " has no diag");
    diag[(size_t)i] = d;
  }
  int nlo = n > 0 ? 1 : 0, nup = n > 0 ? 1 : 0;
  for (int64_t i = 0; i < n; ++i) {
    int v = 0;
    for (int idx = rp[i]; idx < diag[(size_t)i]; ++idx) v = std::max(v, lo[(size_t)col[idx]] + 1);
    lo[(size_t)i] = v;
    nlo = std::max(nlo, v + 1);
  }
  for (int64_t i = n - 1; i >= 0; --i) {
    int v = 0;
    for (int idx = diag[(size_t)i] + 1; idx < rp[i + 1]; ++idx)
      v = std::max(v, up[(size_t)col[idx]] + 1);
    up[(size_t)i] = v;
    nup = std::max(nup, v + 1);
  }
  std::vector<int> ord;
  levels(lo, nlo, M->lvl_lower, ord);
  M->ord_lower = to_device(ctx, ord);
  levels(up, nup, M->lvl_upper, ord);
  M->ord_upper = to_device(ctx, ord);
  M->diag = to_device(ctx, diag);
  ilu_factor(M.get());
  return M.release();
}

void ilu_update(gs_ilu *M, const gs_csr *A) {
  gs_ctx *ctx = M->ctx;
  if (A->rows != M->n || A->cols != M->n) throw Error(GS_ERR_INVALID_ARG, "Size mismatch");
  if (A->nnz != M->nnz)
    throw Error(GS_ERR_INVALID_ARG, "the matrix's pattern differs from the factored one");
  ensure_csr_values(ctx, A);
  hipStream_t s = ctx->stream;
  CallerCsr C = caller_csr(ctx, A);
  auto flag = dalloc<int>(1);
  hipLaunchKernelGGL(k_ilu_set_int, 1, 1, 0, s, flag.get(), 0);
  hipLaunchKernelGGL(k_ilu_same, nblk(M->n + 1), BS, 0, s, M->n + 1, C.rp.get(), M->rp.get(),
                     flag.get());
  if (M->nnz > 0)
    hipLaunchKernelGGL(k_ilu_same, nblk(M->nnz), BS, 0, s, M->nnz, C.col.get(), M->col.get(),
                       flag.get());
  GS_LAUNCH_CHECK("k_ilu_same");
  if (to_host(ctx, flag.get(), 1)[0])
    throw Error(GS_ERR_INVALID_ARG, "the matrix's pattern differs from the factored one");
  M->f = C.val;
  ilu_factor(M);
}

void ilu_apply(const gs_ilu *M, const double *r, double *z) {
  hipStream_t s = M->ctx->stream;
  for (size_t l = 0; l + 1 < M->lvl_lower.size(); ++l) {
    const int b = M->lvl_lower[l], cnt = M->lvl_lower[l + 1] - b;
    hipLaunchKernelGGL(k_ilu_lower, nblk(cnt), BS, 0, s, cnt, M->ord_lower.get() + b,
                       M->rp.get(), M->col.get(), M->diag.get(), M->f.get(), r, z);
  }
  for (size_t l = 0; l + 1 < M->lvl_upper.size(); ++l) {
    const int b = M->lvl_upper[l], cnt = M->lvl_upper[l + 1] - b;
    hipLaunchKernelGGL(k_ilu_upper, nblk(cnt), BS, 0, s, cnt, M->ord_upper.get() + b,
                       M->rp.get(), M->col.get(), M->diag.get(), M->f.get(), z);
  }
  GS_LAUNCH_CHECK("k_ilu_apply");
}

}  // namespace gs
