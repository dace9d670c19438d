// SELL-64 view of a square CSR matrix, built on the GPU (layout in sell.hpp).
//
//   order   rows sorted by descending length inside windows of 256 consecutive
//           rows (stable, so equal lengths keep the CSR order); SELL row s is
//           CSR row sperm[s], and the renumbering is applied to rows and
//           columns alike
//   chunks  slice s (SELL rows 64s..64s+63) gets ceil(len(first row) / 4)
//           chunks -- its first row is its longest; chunk offsets by a scan
//   index   16-bit offsets from 64s when every column of the matrix fits,
//           otherwise int32 columns
//   values  entry k of CSR row r goes to chunk k/4, lane of r, slot k%4
//
// The value scatter keeps the CSR nonzero -> SELL entry map (pos) when the
// caller refreshes values later (re-assembly on a fixed pattern).
#include <cstring>
#include <functional>

#include <rocprim/rocprim.hpp>

#include "internal.hpp"
#include "sell.hpp"

namespace gs {
namespace {

constexpr int BS = 256;

inline unsigned nblk(int64_t n, int bs = BS) {
  return (unsigned)std::max<int64_t>(1, (n + bs - 1) / bs);
}

// One workgroup per window of SELL_SIGMA rows: stable rank by descending
// length (rows past n rank last and are not written).
__global__ __launch_bounds__(SELL_SIGMA) void k_sell_order(
    int n, const int *__restrict__ rp, int *__restrict__ sperm,
    int *__restrict__ siperm) {
  __shared__ int len[SELL_SIGMA];
  const int t = threadIdx.x;
  const int r = blockIdx.x * SELL_SIGMA + t;
  const int mine = r < n ? rp[r + 1] - rp[r] : -1;
  len[t] = mine;
  __syncthreads();
  if (r >= n) return;
  int rank = 0;
  for (int u = 0; u < SELL_SIGMA; ++u) {
    const int l = len[u];
    rank += (l > mine) || (l == mine && u < t);
  }
  const int s = blockIdx.x * SELL_SIGMA + rank;
  sperm[s] = r;
  siperm[r] = s;
}

// chunks of each slice (its first row is the longest), CW entries per lane
// per chunk
template <int CW>
__global__ void k_sell_nchunk(int n, int nslices, const int *__restrict__ rp,
                              const int *__restrict__ sperm,
                              int *__restrict__ nchunk) {
  const int s = blockIdx.x * BS + threadIdx.x;
  if (s >= nslices) return;
  const int r = sperm[s * SELL_C];
  nchunk[s] = (rp[r + 1] - rp[r] + CW - 1) / CW;
}

// ---- block (3x3) view ----
// Block row I merges DOF rows 3I..3I+2 (columns ascending in each): the
// distinct column blocks c/3, counted (FILL = false) or written in ascending
// order (FILL = true).
template <bool FILL>
__global__ void k_b3_rows(int nb, const int *__restrict__ rp, const int *__restrict__ col,
                          const int *__restrict__ brp, int *__restrict__ out) {
  const int I = blockIdx.x * BS + threadIdx.x;
  if (I >= nb) return;
  int k[3], e[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    k[m] = rp[3 * I + m];
    e[m] = rp[3 * I + m + 1];
  }
  int cnt = 0;
  int *dst = FILL ? out + brp[I] : nullptr;
  for (;;) {
    int best = INT32_MAX;
#pragma unroll
    for (int m = 0; m < 3; ++m)
      if (k[m] < e[m]) best = min(best, col[k[m]] / 3);
    if (best == INT32_MAX) break;
    if (FILL) dst[cnt] = best;
    ++cnt;
#pragma unroll
    for (int m = 0; m < 3; ++m)
      while (k[m] < e[m] && col[k[m]] / 3 == best) ++k[m];
  }
  if (!FILL) out[I] = cnt;
}

// DOF-level orders from the block-row order: DOF 3I+c <-> 3 bsiperm[I] + c
__global__ void k_b3_perms(int nb, const int *__restrict__ bsperm,
                           const int *__restrict__ bsiperm, int *__restrict__ sperm,
                           int *__restrict__ siperm) {
  const int I = blockIdx.x * BS + threadIdx.x;
  if (I >= nb) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    sperm[3 * I + c] = 3 * bsperm[I] + c;
    siperm[3 * I + c] = 3 * bsiperm[I] + c;
  }
}

// flag = 1 if some block column's offset from its slice base does not fit
__global__ void k_b3_check16(int nb, const int *__restrict__ brp,
                             const int *__restrict__ bcol,
                             const int *__restrict__ bsiperm, int *__restrict__ flag) {
  const int I = blockIdx.x * BS + threadIdx.x;
  if (I >= nb) return;
  const int base = (bsiperm[I] / SELL_C) * SELL_C;
  bool bad = false;
  for (int k = brp[I]; k < brp[I + 1]; ++k) {
    const int d = bsiperm[bcol[k]] - base;
    bad |= d <= SELL_PAD16 || d > 32767;
  }
  if (bad) *flag = 1;
}

// Block index layout, one thread per block row.
__global__ void k_b3_index(int nb, const int *__restrict__ brp,
                           const int *__restrict__ bcol, const int *__restrict__ bsiperm,
                           const int *__restrict__ cb, short *__restrict__ off,
                           int *__restrict__ col32) {
  const int I = blockIdx.x * BS + threadIdx.x;
  if (I >= nb) return;
  const int s = bsiperm[I];
  const int sl = s / SELL_C, lane = s - sl * SELL_C;
  for (int k = brp[I]; k < brp[I + 1]; ++k) {
    const int64_t e = (int64_t)(cb[sl] + (k - brp[I])) * SELL_C + lane;
    const int c = bsiperm[bcol[k]];
    if (off) off[e] = (short)(c - sl * SELL_C);
    else col32[e] = c;
  }
}

// CSR nonzero -> block value slot, one thread per DOF row: entry (3I+a,
// 3J+b) is value 3a+b of the row's block J.
__global__ void k_b3_pos(int n, const int *__restrict__ rp, const int *__restrict__ col,
                         const int *__restrict__ brp, const int *__restrict__ bcol,
                         const int *__restrict__ bsiperm, const int *__restrict__ cb,
                         int *__restrict__ pos) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const int I = r / 3, a = r - 3 * (r / 3);
  const int s = bsiperm[I];
  const int sl = s / SELL_C, lane = s - sl * SELL_C;
  int q = brp[I];
  for (int k = rp[r]; k < rp[r + 1]; ++k) {
    const int J = col[k] / 3, b = col[k] - 3 * (col[k] / 3);
    while (bcol[q] < J) ++q;  // columns ascend, so the block pointer only advances
    const int64_t chunk = cb[sl] + (q - brp[I]);
    pos[k] = (int)((chunk * 9 + 3 * a + b) * SELL_C + lane);
  }
}

// flag = 1 if some column's offset from its slice base does not fit 16 bits
__global__ void k_sell_check16(int n, const int *__restrict__ rp,
                               const int *__restrict__ col,
                               const int *__restrict__ siperm,
                               int *__restrict__ flag) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const int base = (siperm[r] / SELL_C) * SELL_C;
  bool bad = false;
  for (int k = rp[r]; k < rp[r + 1]; ++k) {
    const int d = siperm[col[k]] - base;
    bad |= d <= SELL_PAD16 || d > 32767;
  }
  if (bad) *flag = 1;
}

__global__ void k_fill_short(int64_t n, short *__restrict__ p, short v) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BS)
    p[i] = v;
}
__global__ void k_fill_i32(int64_t n, int *__restrict__ p, int v) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BS)
    p[i] = v;
}
template <typename T>
__global__ void k_fill_zero(int64_t n, T *__restrict__ p) {
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * BS)
    p[i] = (T)0;
}

// Index layout (one thread per CSR row) and the CSR -> SELL entry map.
__global__ void k_sell_index(int n, const int *__restrict__ rp,
                             const int *__restrict__ col,
                             const int *__restrict__ siperm,
                             const int *__restrict__ cb, short *__restrict__ off,
                             int *__restrict__ col32, int *__restrict__ pos) {
  const int r = blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const int s = siperm[r];
  const int sl = s / SELL_C, lane = s - sl * SELL_C;
  const int64_t e0 = (int64_t)cb[sl] * SELL_CHUNK + lane * SELL_CW;
  const int k0 = rp[r];
  for (int k = k0; k < rp[r + 1]; ++k) {
    const int kk = k - k0;
    const int64_t e = e0 + (int64_t)(kk / SELL_CW) * SELL_CHUNK + (kk % SELL_CW);
    const int c = siperm[col[k]];
    if (off) off[e] = (short)(c - sl * SELL_C);
    else col32[e] = c;
    pos[k] = (int)e;
  }
}

template <typename T>
__global__ void k_sell_values(int64_t nnz, const int *__restrict__ pos,
                              const double *__restrict__ val,
                              T *__restrict__ sval) {
  for (int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x; k < nnz;
       k += (int64_t)gridDim.x * BS)
    sval[pos[k]] = (T)val[k];
}

// JacobiPreconditioner::new (src/linalg/preconditioner.rs:37-51) in SELL
// order: d = 1/A_ii when |A_ii| > 1e-14, else 1.
__global__ void k_sell_jacobi(int n, const int *__restrict__ rp,
                              const int *__restrict__ col,
                              const double *__restrict__ val,
                              const int *__restrict__ sperm,
                              double *__restrict__ d64, float *__restrict__ d32) {
  const int s = blockIdx.x * BS + threadIdx.x;
  if (s >= n) return;
  const int r = sperm[s];
  double d = 1.0;
  for (int k = rp[r]; k < rp[r + 1]; ++k)
    if (col[k] == r) {
      if (fabs(val[k]) > 1e-14) d = 1.0 / val[k];
      break;
    }
  d64[s] = d;
  if (d32) d32[s] = (float)d;
}

__global__ void k_compose(int n, const int *__restrict__ sperm,
                          const int *__restrict__ dperm, int *__restrict__ vperm) {
  const int s = blockIdx.x * BS + threadIdx.x;
  if (s < n) vperm[s] = dperm ? dperm[sperm[s]] : sperm[s];
}

constexpr unsigned GFILL = 2048;

}  // namespace

namespace {

// out[0..count] = exclusive scan of in[0..count) (in[count] must be 0);
// returns out[count]
int64_t scan_count(gs_ctx *ctx, int *in, int *out, int64_t count) {
  hipStream_t s = ctx->stream;
  size_t bytes = 0;
  GS_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)count + 1,
                                 rocprim::plus<int>(), s));
  auto tmp = dalloc<char>((int64_t)bytes);
  GS_HIP(rocprim::exclusive_scan(tmp.get(), bytes, in, out, 0, (size_t)count + 1,
                                 rocprim::plus<int>(), s));
  int h = 0;
  GS_HIP(hipMemcpyAsync(ctx->pinned, out + count, sizeof(int), hipMemcpyDeviceToHost, s));
  GS_HIP(hipStreamSynchronize(s));
  memcpy(&h, ctx->pinned, sizeof(int));
  return h;
}

bool offsets_fit16(gs_ctx *ctx, const std::function<void(int *)> &launch) {
  hipStream_t s = ctx->stream;
  auto flag = dalloc<int>(1);
  GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
  launch(flag.get());
  GS_LAUNCH_CHECK("check16");
  int h = 0;
  GS_HIP(hipMemcpyAsync(ctx->pinned, flag.get(), sizeof(int), hipMemcpyDeviceToHost, s));
  GS_HIP(hipStreamSynchronize(s));
  memcpy(&h, ctx->pinned, sizeof(int));
  return h == 0;
}

// The 3x3-block view (layout in sell.hpp).
std::shared_ptr<SellPattern> build_b3_pattern(gs_ctx *ctx, const gs_csr *A, bool force32) {
  hipStream_t s = ctx->stream;
  auto P = std::make_shared<SellPattern>();
  const int64_t n = A->rows, nb = n / 3;
  P->bs = 3;
  P->n = n;
  P->nnz = A->nnz;
  P->nslices = (nb + SELL_C - 1) / SELL_C;
  P->sperm = dalloc<int>(n);
  P->siperm = dalloc<int>(n);
  P->vperm = dalloc<int>(n);
  P->cb = dalloc<int>(P->nslices + 1);
  P->pos = dalloc<int>(A->nnz);
  const int *rp = A->row_ptr.get(), *col = A->col.get();
  // block rows: distinct column blocks of the three DOF rows
  auto blen = dalloc<int>(nb + 1);
  auto brp = dalloc<int>(nb + 1);
  GS_HIP(hipMemsetAsync(blen.get(), 0, (nb + 1) * sizeof(int), s));
  hipLaunchKernelGGL((k_b3_rows<false>), nblk(nb), BS, 0, s, (int)nb, rp, col,
                     (const int *)nullptr, blen.get());
  const int64_t nbnz = scan_count(ctx, blen.get(), brp.get(), nb);
  auto bcol = dalloc<int>(std::max<int64_t>(1, nbnz));
  hipLaunchKernelGGL((k_b3_rows<true>), nblk(nb), BS, 0, s, (int)nb, rp, col,
                     (const int *)brp.get(), bcol.get());
  // order the block rows, then derive the DOF-level orders
  auto bsperm = dalloc<int>(nb), bsiperm = dalloc<int>(nb);
  hipLaunchKernelGGL(k_sell_order, (unsigned)((nb + SELL_SIGMA - 1) / SELL_SIGMA),
                     SELL_SIGMA, 0, s, (int)nb, (const int *)brp.get(), bsperm.get(),
                     bsiperm.get());
  hipLaunchKernelGGL(k_b3_perms, nblk(nb), BS, 0, s, (int)nb, bsperm.get(), bsiperm.get(),
                     P->sperm.get(), P->siperm.get());
  hipLaunchKernelGGL(k_compose, nblk(n), BS, 0, s, (int)n, P->sperm.get(),
                     A->dperm.get(), P->vperm.get());
  auto nchunk = dalloc<int>(P->nslices + 1);
  GS_HIP(hipMemsetAsync(nchunk.get(), 0, (P->nslices + 1) * sizeof(int), s));
  hipLaunchKernelGGL((k_sell_nchunk<1>), nblk(P->nslices), BS, 0, s, (int)nb,
                     (int)P->nslices, (const int *)brp.get(), bsperm.get(), nchunk.get());
  GS_LAUNCH_CHECK("k_b3_rows");
  P->nchunks = scan_count(ctx, nchunk.get(), P->cb.get(), P->nslices);
  P->entries = P->nchunks * SELL_C;
  P->nvals = P->entries * 9;
  if (P->nvals >= (int64_t)INT32_MAX)
    throw Error(GS_ERR_TOO_LARGE, "block SELL layout exceeds int32 indexing");
  bool wide = force32;
  if (!wide && nbnz > 0)
    wide = !offsets_fit16(ctx, [&](int *flag) {
      hipLaunchKernelGGL(k_b3_check16, nblk(nb), BS, 0, s, (int)nb, brp.get(), bcol.get(),
                         bsiperm.get(), flag);
    });
  P->idx16 = !wide;
  if (wide) {
    P->col = dalloc<int>(P->entries);
    P->off = dalloc<short>(0);
    hipLaunchKernelGGL(k_fill_i32, GFILL, BS, 0, s, P->entries, P->col.get(), -1);
  } else {
    P->off = dalloc<short>(P->entries);
    P->col = dalloc<int>(0);
    hipLaunchKernelGGL(k_fill_short, GFILL, BS, 0, s, P->entries, P->off.get(),
                       (short)SELL_PAD16);
  }
  hipLaunchKernelGGL(k_b3_index, nblk(nb), BS, 0, s, (int)nb, brp.get(), bcol.get(),
                     bsiperm.get(), P->cb.get(), wide ? nullptr : P->off.get(),
                     wide ? P->col.get() : nullptr);
  hipLaunchKernelGGL(k_b3_pos, nblk(n), BS, 0, s, (int)n, rp, col, brp.get(), bcol.get(),
                     bsiperm.get(), P->cb.get(), P->pos.get());
  GS_LAUNCH_CHECK("k_b3_index");
  GS_HIP(hipStreamSynchronize(s));
  return P;
}

}  // namespace

std::shared_ptr<SellPattern> build_sell_pattern(gs_ctx *ctx, const gs_csr *A,
                                                bool force32, int block) {
  if (A->rows != A->cols) throw Error(GS_ERR_INVALID_ARG, "SELL needs a square matrix");
  if (block == 3 && A->rows > 0 && A->rows % 3 == 0) return build_b3_pattern(ctx, A, force32);
  hipStream_t s = ctx->stream;
  auto P = std::make_shared<SellPattern>();
  const int64_t n = A->rows;
  P->n = n;
  P->nnz = A->nnz;
  P->nslices = (n + SELL_C - 1) / SELL_C;
  P->sperm = dalloc<int>(n);
  P->siperm = dalloc<int>(n);
  P->vperm = dalloc<int>(n);
  P->cb = dalloc<int>(P->nslices + 1);
  P->pos = dalloc<int>(A->nnz);
  if (n == 0) {
    GS_HIP(hipMemsetAsync(P->cb.get(), 0, sizeof(int), s));
    GS_HIP(hipStreamSynchronize(s));
    P->off = dalloc<short>(0);
    return P;
  }
  const int *rp = A->row_ptr.get(), *col = A->col.get();
  hipLaunchKernelGGL(k_sell_order, (unsigned)((n + SELL_SIGMA - 1) / SELL_SIGMA),
                     SELL_SIGMA, 0, s, (int)n, rp, P->sperm.get(), P->siperm.get());
  hipLaunchKernelGGL(k_compose, nblk(n), BS, 0, s, (int)n, P->sperm.get(),
                     A->dperm.get(), P->vperm.get());
  auto nchunk = dalloc<int>(P->nslices + 1);
  GS_HIP(hipMemsetAsync(nchunk.get(), 0, (P->nslices + 1) * sizeof(int), s));
  hipLaunchKernelGGL((k_sell_nchunk<SELL_CW>), nblk(P->nslices), BS, 0, s, (int)n,
                     (int)P->nslices, rp, P->sperm.get(), nchunk.get());
  GS_LAUNCH_CHECK("k_sell_order");
  {
    size_t bytes = 0;
    GS_HIP(rocprim::exclusive_scan(nullptr, bytes, nchunk.get(), P->cb.get(), 0,
                                   (size_t)P->nslices + 1, rocprim::plus<int>(), s));
    auto tmp = dalloc<char>((int64_t)bytes);
    GS_HIP(rocprim::exclusive_scan(tmp.get(), bytes, nchunk.get(), P->cb.get(), 0,
                                   (size_t)P->nslices + 1, rocprim::plus<int>(), s));
    int h = 0;
    GS_HIP(hipMemcpyAsync(ctx->pinned, P->cb.get() + P->nslices, sizeof(int),
                          hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    memcpy(&h, ctx->pinned, sizeof(int));
    P->nchunks = h;
  }
  P->entries = P->nchunks * SELL_CHUNK;
  P->nvals = P->entries;
  if (P->entries >= (int64_t)INT32_MAX)
    throw Error(GS_ERR_TOO_LARGE, "SELL layout exceeds int32 indexing");
  bool wide = force32;
  if (!wide && A->nnz > 0) {
    auto flag = dalloc<int>(1);
    GS_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), s));
    hipLaunchKernelGGL(k_sell_check16, nblk(n), BS, 0, s, (int)n, rp, col,
                       P->siperm.get(), flag.get());
    GS_LAUNCH_CHECK("k_sell_check16");
    int h = 0;
    GS_HIP(hipMemcpyAsync(ctx->pinned, flag.get(), sizeof(int),
                          hipMemcpyDeviceToHost, s));
    GS_HIP(hipStreamSynchronize(s));
    memcpy(&h, ctx->pinned, sizeof(int));
    wide = h != 0;
  }
  P->idx16 = !wide;
  if (wide) {
    P->col = dalloc<int>(P->entries);
    P->off = dalloc<short>(0);
    hipLaunchKernelGGL(k_fill_i32, GFILL, BS, 0, s, P->entries, P->col.get(), -1);
  } else {
    P->off = dalloc<short>(P->entries);
    P->col = dalloc<int>(0);
    hipLaunchKernelGGL(k_fill_short, GFILL, BS, 0, s, P->entries, P->off.get(),
                       (short)SELL_PAD16);
  }
  hipLaunchKernelGGL(k_sell_index, nblk(n), BS, 0, s, (int)n, rp, col,
                     P->siperm.get(), P->cb.get(), wide ? nullptr : P->off.get(),
                     wide ? P->col.get() : nullptr, P->pos.get());
  GS_LAUNCH_CHECK("k_sell_index");
  GS_HIP(hipStreamSynchronize(s));
  return P;
}

template <typename T>
void sell_values(hipStream_t s, const SellPattern &P, const double *val, T *sval) {
  hipLaunchKernelGGL((k_fill_zero<T>), GFILL, BS, 0, s, P.nvals, sval);
  if (P.nnz > 0)
    hipLaunchKernelGGL((k_sell_values<T>), GFILL, BS, 0, s, P.nnz, P.pos.get(),
                       val, sval);
  GS_LAUNCH_CHECK("k_sell_values");
}
template void sell_values<double>(hipStream_t, const SellPattern &, const double *,
                                  double *);
template void sell_values<float>(hipStream_t, const SellPattern &, const double *,
                                 float *);

__global__ void k_csr_from_sell(int64_t nnz, const int *__restrict__ pos,
                                const double *__restrict__ sval,
                                double *__restrict__ val) {
  const int64_t d = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (d < nnz) val[d] = sval[pos[d]];
}

void ensure_csr_values(gs_ctx *ctx, const gs_csr *Ac) {
  gs_csr *A = const_cast<gs_csr *>(Ac);  // a derived cache
  if (!A->val_lazy || A->val) return;
  if (!A->sell || !A->sval)
    throw Error(GS_ERR_INVALID_ARG, "CSR values of a fused system without its SELL values");
  A->val = dalloc<double>(A->nnz);
  if (A->nnz > 0) {
    hipLaunchKernelGGL(k_csr_from_sell, nblk(A->nnz), BS, 0, ctx->stream, A->nnz,
                       A->sell->pos.get(), A->sval.get(), A->val.get());
    GS_LAUNCH_CHECK("k_csr_from_sell");
  }
}

void sell_zero_holes(hipStream_t s, const SellPattern &P, double *sval, float *sval32) {
  if (P.bs == 1) return;
  hipLaunchKernelGGL((k_fill_zero<double>), GFILL, BS, 0, s, P.nvals, sval);
  hipLaunchKernelGGL((k_fill_zero<float>), GFILL, BS, 0, s, P.nvals, sval32);
  GS_LAUNCH_CHECK("sell_zero_holes");
}

void sell_jacobi(hipStream_t s, const SellPattern &P, const gs_csr *A, double *d64,
                 float *d32) {
  if (P.n == 0) return;
  hipLaunchKernelGGL(k_sell_jacobi, nblk(P.n), BS, 0, s, (int)P.n,
                     A->row_ptr.get(), A->col.get(), A->val.get(), P.sperm.get(),
                     d64, d32);
  GS_LAUNCH_CHECK("k_sell_jacobi");
}

}  // namespace gs
