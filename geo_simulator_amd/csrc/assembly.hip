// Deterministic CSR assembly and Dirichlet elimination on gfx950.
//
// Symbolic phase (once per mesh, cached): every element contributes the 100
// node pairs (conn[e][a], conn[e][b]); a stable LSD radix sort of the packed
// (row, col) keys carrying the triplet id t = e*100 + a*10 + b yields, in one
// pass, the sorted/unique CSR pattern (sprs TriMat::to_csr, src/fem/
// assembly.rs:57,116) and, for every CSR slot, its contributors in ascending
// element order.
//
// Numeric phase: a segmented sum per slot over that contributor list (no
// float atomics), in ascending element order -- the order in which the
// reference's element-major triplet stream (src/fem/assembly.rs:77-113,
// 413-434) presents duplicates.  Loads are gathered the same way through each
// node's diagonal slot (src/fem/assembly.rs:130-154,906-937).
//
// Compiled without FP contraction: the Dirichlet RHS lift
// f_j -= K_ji g_i (src/fem/assembly.rs:198-208) stays a separate multiply and
// subtract, as in the reference.
#pragma clang fp contract(off)

#include <climits>
#include <cstring>
#include <type_traits>

#include <rocprim/rocprim.hpp>

#include "internal.hpp"
#include "sell.hpp"

namespace gs {
namespace {

constexpr int BS = 256;

inline unsigned nblk(int64_t n, int bs = BS) {
  return (unsigned)((n + bs - 1) / bs);
}

// Block b of a one-item-per-thread kernel, mapped so that each XCD works on
// a contiguous range: workgroup i runs on XCD i % 8, and XCD x takes the
// x-th eighth of the blocks in order.  Nearby slots and rows read nearby
// element matrices, so an XCD's L2 holds its own share of them instead of
// every XCD fetching the whole stream.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t G = gridDim.x, i = blockIdx.x;
  if (G < 8) return i;
  const int64_t x = i & 7, q = G >> 3, rem = G & 7;
  return q * x + (x < rem ? x : rem) + (i >> 3);
}
__device__ __forceinline__ int64_t xcd_item() { return xcd_block() * BS + threadIdx.x; }

// Contributions whose loads a gather thread issues together (the sums still
// add them one by one in ascending order).
constexpr int GATHER_GROUP = 8;

// Accumulators per pass of the thermal row gather (a multiple of the SELL
// chunk width, so a pass writes whole chunks): 64 x LW doubles of LDS per
// workgroup.
constexpr int ROW_LW = 68;
static_assert(ROW_LW % SELL_CW == 0, "a pass covers whole SELL chunks");
// The slices whose rows (with their lifts) need at most this many: the
// midpoint rows of a Tet10 mesh, about three slices in four.  Half the LDS
// per wavefront lets twice as many run per CU.
constexpr int ROW_LW_SHORT = 32;
static_assert(ROW_LW_SHORT % SELL_CW == 0, "a pass covers whole SELL chunks");
// incidence rows whose loads a lane of the row gather issues together
constexpr int ROW_UNROLL = 4;

typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));

// The ten accumulator codes of one incidence: 1-byte codes at a 2-byte
// aligned address, 2-byte codes at a 4-byte aligned one.
__device__ __forceinline__ void load_codes(const uint8_t *p, int (&code)[10]) {
  const unsigned short *w = (const unsigned short *)p;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const unsigned v = w[j];
    code[2 * j] = (int)(v & 0xffu);
    code[2 * j + 1] = (int)(v >> 8);
  }
}
__device__ __forceinline__ void load_codes(const unsigned short *p, int (&code)[10]) {
  const unsigned *w = (const unsigned *)p;
#pragma unroll
  for (int j = 0; j < 5; ++j) {
    const unsigned v = w[j];
    code[2 * j] = (int)(v & 0xffffu);
    code[2 * j + 1] = (int)(v >> 16);
  }
}

// |v| as the ordered bit pattern of a non-negative double (0 for NaN)
__device__ __forceinline__ unsigned long long abs_bits(double v) {
  const double a = fabs(v);
  return a == a ? (unsigned long long)__double_as_longlong(a) : 0ull;
}

// Grid of a max-reduction kernel: a grid-stride loop over at most this many
// workgroups, each folding its maximum into one device-wide atomic (atomics
// on one address serialise; one per wavefront over a million rows cost
// about 0.3 ms).
constexpr int64_t MAX_GRID = 1024;
inline unsigned nblk_max(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk(n), MAX_GRID));
}

// Largest of the workgroup's (BS threads) values -> part[blockIdx.x] (0 when
// none); every thread of the workgroup must call it.
__device__ __forceinline__ void wg_max_store(unsigned long long mine,
                                             unsigned long long *part) {
  __shared__ unsigned long long sm[BS / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mine, off, 64);
    mine = o > mine ? o : mine;
  }
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sm[0];
#pragma unroll
    for (int k = 1; k < BS / 64; ++k) m = sm[k] > m ? sm[k] : m;
    part[blockIdx.x] = m;
  }
}


// Largest of the workgroup's (BS threads) values, folded into *out with one
// atomic; bit patterns of non-negative doubles, so the integer maximum is
// the floating-point one and the result is order independent.
__device__ __forceinline__ void wg_max_atomic(unsigned long long mine,
                                              unsigned long long *out) {
  __shared__ unsigned long long sm[BS / 64];
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mine, off, 64);
    mine = o > mine ? o : mine;
  }
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = sm[0];
#pragma unroll
    for (int k = 1; k < BS / 64; ++k) m = sm[k] > m ? sm[k] : m;
    if (m != 0) atomicMax(out, m);
  }
}

// max of np partial maxima folded into *out (a grid-stride pass, one
// atomic per workgroup)
__global__ __launch_bounds__(BS) void k_max_partials(int64_t np,
                                                    const unsigned long long *__restrict__ part,
                                                    unsigned long long *__restrict__ out) {
  unsigned long long mine = 0;
  for (int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x; i < np; i += (int64_t)gridDim.x * BS)
    mine = part[i] > mine ? part[i] : mine;
  wg_max_atomic(mine, out);
}

// ---------------------------------------------------------------------------
// symbolic phase
// ---------------------------------------------------------------------------
// Keys are (row, col) in the internal numbering iperm[caller node].
__global__ void k_make_keys(int64_t E, const int *__restrict__ conn,
                            const int *__restrict__ iperm, int bits,
                            uint64_t *__restrict__ keys,
                            int *__restrict__ vals) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= E * 100) return;
  const int64_t e = t / 100;
  const int ab = (int)(t - e * 100);
  const int a = ab / 10, b = ab - 10 * (ab / 10);
  const uint64_t r = (uint32_t)iperm[conn[e * 10 + a]];
  const uint64_t c = (uint32_t)iperm[conn[e * 10 + b]];
  keys[t] = (r << bits) | c;
  vals[t] = (int)t;
}

__global__ void k_mark_heads(int64_t N, const uint64_t *__restrict__ keys,
                             int *__restrict__ head) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  head[k] = (k == 0 || keys[k] != keys[k - 1]) ? 1 : 0;
}

// For every run head k (slot s = incl[k]-1): record column, row, contributor
// start, diagonal slot, and count the row's slots.
__global__ void k_slots(int64_t N, const uint64_t *__restrict__ keys,
                        const int *__restrict__ incl, int bits,
                        int *__restrict__ col, int *__restrict__ slot_row,
                        int *__restrict__ contrib_ptr,
                        int *__restrict__ diag_slot, int *__restrict__ row_cnt) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  if (k != 0 && keys[k] == keys[k - 1]) return;
  const int s = incl[k] - 1;
  const uint64_t key = keys[k];
  const int r = (int)(key >> bits);
  const int c = (int)(key & ((1ull << bits) - 1));
  col[s] = c;
  slot_row[s] = r;
  contrib_ptr[s] = (int)k;
  if (r == c) diag_slot[r] = s;
  atomicAdd(&row_cnt[r], 1);  // integer count: order-independent
}

__global__ void k_fill_int(int64_t n, int *__restrict__ p, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_set_int(int *__restrict__ p, int v) { *p = v; }

__global__ void k_set_u64(unsigned long long *__restrict__ p,
                          unsigned long long v) {
  *p = v;
}

// ---------------------------------------------------------------------------
// elastic DOF pattern: node pattern (x) 3x3, DOF = 3*node + c
// (src/fem/dof.rs:87-106)
// ---------------------------------------------------------------------------
__global__ void k_dof3_rowptr(int64_t n_nodes, const int *__restrict__ rp,
                              int *__restrict__ drp) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > 3 * n_nodes) return;
  if (i == 3 * n_nodes) { drp[i] = 9 * rp[n_nodes]; return; }
  const int64_t r = i / 3;
  const int c = (int)(i - 3 * r);
  const int deg = rp[r + 1] - rp[r];
  drp[i] = 9 * rp[r] + 3 * c * deg;
}

__global__ void k_dof3_cols(int64_t nnz, const int *__restrict__ rp,
                            const int *__restrict__ col,
                            const int *__restrict__ slot_row,
                            int *__restrict__ dcol) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nnz) return;
  const int r = slot_row[s];
  const int base = rp[r], deg = rp[r + 1] - base;
  const int j = (int)(s - base);
  const int cn = col[s];
#pragma unroll
  for (int ci = 0; ci < 3; ++ci) {
    const int64_t d = 9ll * base + 3ll * ci * deg + 3ll * j;
#pragma unroll
    for (int cj = 0; cj < 3; ++cj) dcol[d + cj] = 3 * cn + cj;
  }
}

// ---------------------------------------------------------------------------
// numeric gathers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tri_index(int a, int b) {
  // packed upper triangle of a 10x10 symmetric matrix, a <= b
  return 10 * a - (a * (a - 1)) / 2 + (b - a);
}

// FULL: element matrices with all 100 entries (Ke[e*100 + a*10 + b], the
// thermal mass, symmetric only to rounding as in the reference); otherwise
// the packed upper triangle of a bitwise-symmetric K_e.
template <bool FULL>
__global__ void k_gather_thermal(int64_t nnz, const int *__restrict__ cptr,
                                 const int *__restrict__ contrib,
                                 const double *__restrict__ Ke,
                                 double *__restrict__ val) {
  const int64_t s = xcd_item();
  if (s >= nnz) return;
  const int b0 = cptr[s], b1 = cptr[s + 1];
  double v = 0.0;
  for (int k0 = b0; k0 < b1; k0 += GATHER_GROUP) {
    int t[GATHER_GROUP];
    double x[GATHER_GROUP];
#pragma unroll
    for (int u = 0; u < GATHER_GROUP; ++u)
      if (k0 + u < b1) t[u] = contrib[k0 + u];
#pragma unroll
    for (int u = 0; u < GATHER_GROUP; ++u)
      if (k0 + u < b1) {
        const int e = t[u] / 100;
        const int ab = t[u] - 100 * e;
        const int a = ab / 10, b = ab - 10 * (ab / 10);
        x[u] = FULL ? Ke[(int64_t)e * 100 + ab]
                    : Ke[(int64_t)e * 55 + (a <= b ? tri_index(a, b) : tri_index(b, a))];
      }
#pragma unroll
    for (int u = 0; u < GATHER_GROUP; ++u)
      if (k0 + u < b1) v = (k0 + u == b0) ? x[u] : v + x[u];
  }
  val[s] = v;
}

__global__ void k_gather_elastic(int64_t nnz, const int *__restrict__ rp,
                                 const int *__restrict__ slot_row,
                                 const int *__restrict__ cptr,
                                 const int *__restrict__ contrib,
                                 const double *__restrict__ Ke,
                                 double *__restrict__ val) {
  const int64_t s = xcd_item();
  if (s >= nnz) return;
  const int r = slot_row[s];
  const int base = rp[r], deg = rp[r + 1] - base;
  const int j = (int)(s - base);
  const int b0 = cptr[s], b1 = cptr[s + 1];
  double v[9];
#pragma unroll
  for (int m = 0; m < 9; ++m) v[m] = 0.0;
  for (int k = b0; k < b1; ++k) {
    const int t = contrib[k];
    const int e = t / 100;
    const int ab = t - 100 * e;
    const int a = ab / 10, b = ab - 10 * (ab / 10);
    const double *Kr = Ke + (int64_t)e * 900 + (a * 10 + b) * 9;
#pragma unroll
    for (int ci = 0; ci < 3; ++ci)
#pragma unroll
      for (int cj = 0; cj < 3; ++cj) {
        const double x = Kr[ci * 3 + cj];
        v[3 * ci + cj] = (k == b0) ? x : v[3 * ci + cj] + x;
      }
  }
#pragma unroll
  for (int ci = 0; ci < 3; ++ci) {
    const int64_t d = 9ll * base + 3ll * ci * deg + 3ll * j;
#pragma unroll
    for (int cj = 0; cj < 3; ++cj) val[d + cj] = v[3 * ci + cj];
  }
}

// f[r] = sum over (e,a) with conn[e][a] == r, ascending e, of Fe[e][a]:
// exactly the contributors of the diagonal slot (r,r).
// Written at node out_map[r] (caller numbering) when out_map is given.
template <int DPN>
__global__ void k_gather_load(int64_t n_nodes, const int *__restrict__ diag,
                              const int *__restrict__ cptr,
                              const int *__restrict__ contrib,
                              const double *__restrict__ Fe,
                              const int *__restrict__ out_map,
                              double *__restrict__ f) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_nodes) return;
  double v[DPN];
#pragma unroll
  for (int c = 0; c < DPN; ++c) v[c] = 0.0;
  const int s = diag[r];
  if (s >= 0) {
    for (int k = cptr[s]; k < cptr[s + 1]; ++k) {
      const int t = contrib[k];
      const int e = t / 100;
      const int a = (t - 100 * e) / 10;
#pragma unroll
      for (int c = 0; c < DPN; ++c) v[c] += Fe[((int64_t)e * 10 + a) * DPN + c];
    }
  }
  const int64_t o = out_map ? (int64_t)out_map[r] : r;
#pragma unroll
  for (int c = 0; c < DPN; ++c) f[o * DPN + c] = v[c];
}

// ---------------------------------------------------------------------------
// Dirichlet elimination (src/fem/assembly.rs:168-230), one thread per row.
// ---------------------------------------------------------------------------
// Pass 1: kept-entry count per row and max |K_ii| (as the ordered bit pattern
// of a non-negative double; max is order independent, so this is exact).
__global__ void k_bc_count(int64_t n, const int *__restrict__ rp,
                           const int *__restrict__ col,
                           const double *__restrict__ val,
                           const uint8_t *__restrict__ dir,
                           int *__restrict__ cnt,
                           unsigned long long *__restrict__ maxbits) {
  unsigned long long mine = 0;
  for (int64_t r = (int64_t)blockIdx.x * BS + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * BS) {
    int c = 0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int j = col[k];
      if (!dir[j]) ++c;
      if (j == r) {
        const double a = fabs(val[k]);
        if (a == a) {
          const unsigned long long bits = (unsigned long long)__double_as_longlong(a);
          mine = bits > mine ? bits : mine;
        }
      }
    }
    cnt[r] = dir[r] ? 1 : c;
  }
  wg_max_atomic(mine, maxbits);
}

// Pass 2: copy kept entries and lift the RHS.  The reference subtracts
// K_ji g_i in ascending caller column order; with an internal numbering
// (dperm) the constrained entries are visited in ascending dperm order by
// repeated minimum search, so the lift is the same sequence of operations.
__global__ void k_bc_fill(int64_t n, const int *__restrict__ rp,
                          const int *__restrict__ col,
                          const double *__restrict__ val,
                          const uint8_t *__restrict__ dir,
                          const double *__restrict__ g,
                          const double *__restrict__ f,
                          const int *__restrict__ dperm,
                          const unsigned long long *__restrict__ maxbits,
                          const int *__restrict__ nrp, int *__restrict__ ncol,
                          double *__restrict__ nval, double *__restrict__ fbc) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int p = nrp[r];
  if (dir[r]) {
    const double c = __longlong_as_double((long long)*maxbits);
    ncol[p] = (int)r;
    nval[p] = c;
    fbc[r] = g[r] * c;
    return;
  }
  const int k0 = rp[r], k1 = rp[r + 1];
  double fr = f[r];
  for (int k = k0; k < k1; ++k) {
    const int j = col[k];
    if (dir[j]) {
      if (!dperm) fr -= val[k] * g[j];
    } else {
      ncol[p] = j;
      nval[p] = val[k];
      ++p;
    }
  }
  if (dperm) {
    int last = -1;
    for (;;) {
      int best = INT_MAX, bk = -1;
      for (int k = k0; k < k1; ++k) {
        const int j = col[k];
        if (dir[j]) {
          const int oc = dperm[j];
          if (oc > last && oc < best) {
            best = oc;
            bk = k;
          }
        }
      }
      if (bk < 0) break;
      fr -= val[bk] * g[col[bk]];
      last = best;
    }
  }
  fbc[r] = fr;
}

__global__ void k_perm_gather(int64_t n, const double *__restrict__ src,
                              const int *__restrict__ map,
                              double *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[map[i]];
}

__global__ void k_perm_gather_u8(int64_t n, const uint8_t *__restrict__ src,
                                 const int *__restrict__ map,
                                 uint8_t *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[map[i]];
}

__global__ void k_perm_scatter(int64_t n, const double *__restrict__ src,
                               const int *__restrict__ map,
                               double *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[map[i]] = src[i];
}

// DOF permutation for 3 DOFs per node: perm3[3 r + c] = 3 perm[r] + c
__global__ void k_perm3(int64_t n_nodes, const int *__restrict__ perm,
                        int *__restrict__ perm3) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * n_nodes) perm3[i] = 3 * perm[i / 3] + (int)(i % 3);
}

// ---------------------------------------------------------------------------
// System plans: the Dirichlet elimination (src/fem/assembly.rs:168-230)
// decided once on the pattern, so the numeric phase writes K_bc directly.
// ---------------------------------------------------------------------------
// Per row: kept entries (constrained row: its diagonal only) and entries of
// the RHS lift (free row, constrained column).
__global__ void k_plan_count(int64_t n, const int *__restrict__ rp,
                             const int *__restrict__ col,
                             const uint8_t *__restrict__ dir,
                             int *__restrict__ kept, int *__restrict__ lift) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  if (dir[r]) {
    kept[r] = 1;
    lift[r] = 0;
    return;
  }
  int a = 0, b = 0;
  for (int k = rp[r]; k < rp[r + 1]; ++k) {
    if (dir[col[k]]) ++b;
    else ++a;
  }
  kept[r] = a;
  lift[r] = b;
}

// K_bc columns, the slot destinations and the lift list.  Lift entries of a
// row are numbered in ascending caller column order (the order in which the
// reference subtracts them), found by repeated minimum search.
__global__ void k_plan_fill(int64_t n, const int *__restrict__ rp,
                            const int *__restrict__ col,
                            const uint8_t *__restrict__ dir,
                            const int *__restrict__ dperm,
                            const int *__restrict__ brp, int *__restrict__ bcol,
                            int *__restrict__ dst, const int *__restrict__ lptr,
                            int *__restrict__ lcol, int *__restrict__ diag) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int k0 = rp[r], k1 = rp[r + 1];
  int p = brp[r];
  if (dir[r]) {
    bcol[p] = (int)r;
    diag[r] = p;
    for (int k = k0; k < k1; ++k) dst[k] = -1;
    return;
  }
  int dg = -1;
  for (int k = k0; k < k1; ++k) {
    const int j = col[k];
    if (!dir[j]) {
      bcol[p] = j;
      dst[k] = p;
      if (j == r) dg = p;
      ++p;
    }
  }
  diag[r] = dg;
  int q = lptr[r], last = -1;
  for (;;) {
    int best = INT_MAX, bk = -1;
    for (int k = k0; k < k1; ++k) {
      const int j = col[k];
      if (dir[j]) {
        const int oc = dperm ? dperm[j] : j;
        if (oc > last && oc < best) {
          best = oc;
          bk = k;
        }
      }
    }
    if (bk < 0) break;
    dst[bk] = -2 - q;
    lcol[q] = col[bk];
    ++q;
    last = best;
  }
}

// Numeric phase, thermal, in the order of the SELL view: a workgroup per
// slice (slices[blockIdx], the plan's order: the short ones launched with
// ROW_LW_SHORT accumulators, the rest with ROW_LW), a lane per SELL row
// (node r).  The element kernel left row a of
// every K_e (and f_e[a]) in the incidence slot of (e, a): step j of the
// slice holds the j-th incidence (ascending element) of each of its 64
// nodes side by side, so the lanes' loads of a step are contiguous.  The
// lane walks its node's incidence rows
// and adds each into accumulators in LDS, one per entry of r's K_bc row
// followed by its lift entries, as icode (from the plan) directs.  Each slot
// so receives its contributions in ascending element order, as in
// k_gather_thermal: the sums start from -0.0, which leaves the first term
// unchanged.  The additions are LDS atomics without return (ds_add_f64): a
// lane only ever adds to its own accumulators, and one wavefront's LDS
// operations complete in program order, so the order of every sum stays
// fixed while the lane issues its additions without waiting on each.  A pass covers LW accumulators; longer rows take further passes
// over their incidence rows.  The lane then writes whole SELL chunks (entries
// and padding zeros, 32 contiguous bytes per lane), its lift values, the
// node's load (k_gather_load's sum, same order) and diagonal, and the
// workgroup leaves the largest |K_ii| in dmax[blockIdx.x].
template <typename C, int LW>
__global__ __launch_bounds__(64) void k_sys_rows_thermal(
    const int *__restrict__ slices, int64_t n, const int *__restrict__ sperm,
    const int *__restrict__ cb,
    const int *__restrict__ brp, const int *__restrict__ lptr,
    const int *__restrict__ diag, const int *__restrict__ ib, const int *__restrict__ icnt,
    const double *__restrict__ rows, const C *__restrict__ icode, int with_load,
    double *__restrict__ sval, float *__restrict__ sval32, double *__restrict__ liftv,
    double *__restrict__ fload, double *__restrict__ dval,
    unsigned long long *__restrict__ dmax) {
  __shared__ double acc[LW * 64];
  const int lane = threadIdx.x;
  const int64_t sl = slices[xcd_block()];
  const int64_t srow = sl * 64 + lane;
  int r = -1, nk = 0, nl = 0, lb = 0, ni = 0, kd = -1;
  if (srow < n) {
    r = sperm[srow];
    const int b0 = brp[r];
    nk = brp[r + 1] - b0;
    lb = lptr[r];
    nl = lptr[r + 1] - lb;
    ni = icnt[srow];
    const int d = diag[r];
    if (d >= 0 && ni > 0) kd = d - b0;
  }
  const int c0 = cb[sl], nch = cb[sl + 1] - c0;
  // incidence slot of step j: (ib[sl] + j) * 64 + lane
  const int64_t q0 = (int64_t)ib[sl] * 64 + lane;
  int span = max(nk + nl, SELL_CW * nch);
  for (int off = 32; off > 0; off >>= 1) span = max(span, __shfl_xor(span, off, 64));
  double *o64 = sval + (int64_t)c0 * SELL_CHUNK + lane * SELL_CW;
  float *o32 = sval32 ? sval32 + (int64_t)c0 * SELL_CHUNK + lane * SELL_CW : nullptr;
  double fl = 0.0;
  unsigned long long mine = 0;
  for (int p0 = 0; p0 < span; p0 += LW) {
#pragma unroll
    for (int k = 0; k < LW; ++k) acc[k * 64 + lane] = -0.0;
    // ROW_UNROLL incidence rows per trip: their loads are issued together,
    // the additions stay in ascending order
    for (int i = 0; i < ni; i += ROW_UNROLL) {
      d2 x[ROW_UNROLL][6];
      int code[ROW_UNROLL][10];
#pragma unroll
      for (int u = 0; u < ROW_UNROLL; ++u)
        if (i + u < ni) {
          const int64_t q = q0 + 64ll * (i + u);
          const d2 *row = (const d2 *)(rows + q * 12);
#pragma unroll
          for (int j = 0; j < 6; ++j) x[u][j] = row[j];
          load_codes(icode + q * 10, code[u]);
        }
#pragma unroll
      for (int u = 0; u < ROW_UNROLL; ++u)
        if (i + u < ni) {
          if (p0 == 0 && with_load) fl += x[u][5].x;
#pragma unroll
          for (int b = 0; b < 10; ++b) {
            const unsigned kk = (unsigned)(code[u][b] - p0);
            if (kk < (unsigned)LW)
              unsafeAtomicAdd(&acc[kk * 64 + lane], (b & 1) ? x[u][b / 2].y : x[u][b / 2].x);
          }
        }
    }
    // this pass's SELL chunks, lift values and diagonal
    const int cend = min(nch, (p0 + LW) / SELL_CW);
    for (int c = p0 / SELL_CW; c < cend; ++c) {
      double v[SELL_CW];
#pragma unroll
      for (int q = 0; q < SELL_CW; ++q) {
        const int k = SELL_CW * c + q;
        v[q] = k < nk ? acc[(k - p0) * 64 + lane] : 0.0;
      }
      d2 *w64 = (d2 *)(o64 + (int64_t)c * SELL_CHUNK);
      w64[0] = (d2){v[0], v[1]};
      w64[1] = (d2){v[2], v[3]};
      if (sval32)
        *(f4 *)(o32 + (int64_t)c * SELL_CHUNK) =
            (f4){(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
    }
    const int kl1 = min(p0 + LW, nk + nl);
    for (int k = max(p0, nk); k < kl1; ++k) liftv[lb + k - nk] = acc[(k - p0) * 64 + lane];
    if (kd >= p0 && kd < p0 + LW) {
      const double dv = acc[(kd - p0) * 64 + lane];
      dval[r] = dv;
      mine = abs_bits(dv);
    }
  }
  if (with_load && ni > 0) fload[r] = fl;
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(mine, off, 64);
    mine = o > mine ? o : mine;
  }
  if (lane == 0) dmax[blockIdx.x] = mine;
}

// The 3x3 block sums of node slot s (contributions in ascending element
// order, as k_gather_elastic); diag: a diagonal slot, whose contributions
// also give the node's three load components (k_gather_load's sums, same
// order) when Fe is given.
__device__ __forceinline__ void elastic_slot_sums(
    int64_t s, const int *__restrict__ cptr, const int *__restrict__ contrib,
    const double *__restrict__ Ke, const double *__restrict__ Fe, double (&v)[9],
    double (&fl)[3], bool &diag) {
  const int b0 = cptr[s], b1 = cptr[s + 1];
  const int ab0 = b1 > b0 ? contrib[b0] % 100 : 1;
  diag = ab0 / 10 == ab0 % 10;
  const bool load = diag && Fe;
#pragma unroll
  for (int m = 0; m < 9; ++m) v[m] = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) fl[c] = 0.0;
  // two contributions' loads issued together; additions in ascending order
  for (int k0 = b0; k0 < b1; k0 += 2) {
    double x[2][9], y[2][3];
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (k0 + u < b1) {
        const int t = contrib[k0 + u];
        const int e = t / 100;
        const int ab = t - 100 * e;
        const int a = ab / 10, b = ab - 10 * (ab / 10);
        const double *Kr = Ke + (int64_t)e * 900 + (a * 10 + b) * 9;
#pragma unroll
        for (int m = 0; m < 9; ++m) x[u][m] = Kr[m];
        if (load) {
#pragma unroll
          for (int c = 0; c < 3; ++c) y[u][c] = Fe[((int64_t)e * 10 + a) * 3 + c];
        }
      }
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (k0 + u < b1) {
#pragma unroll
        for (int m = 0; m < 9; ++m) v[m] = (k0 + u == b0) ? x[u][m] : v[m] + x[u][m];
        if (load) {
#pragma unroll
          for (int c = 0; c < 3; ++c) fl[c] += y[u][c];
        }
      }
  }
}

// Node slot s's nine destinations in the plan (dst of the full K's 3x3 block)
__device__ __forceinline__ void elastic_slot_dst(int64_t s, const int *__restrict__ rp,
                                                 const int *__restrict__ slot_row,
                                                 const int *__restrict__ dst, int (&dd)[9]) {
  const int r = slot_row[s];
  const int base = rp[r], deg = rp[r + 1] - base;
  const int j = (int)(s - base);
#pragma unroll
  for (int ci = 0; ci < 3; ++ci)
#pragma unroll
    for (int cj = 0; cj < 3; ++cj)
      dd[3 * ci + cj] = dst[9ll * base + 3ll * ci * deg + 3ll * j + cj];
}

// Numeric phase, elastic, in the order of the 3x3-block SELL view: one
// thread per stored block position (chunk c, lane l), so a wavefront writes
// each of the nine value planes as one contiguous stretch, and every
// position -- padding and the missing entries of a stored block -- is
// written once (as zero), with no separate clearing pass.  bmask (from the
// plan) marks the block's kept entries; the lift values, and the diagonal
// blocks with no kept entry, are left to k_sys_lift_elastic.  A diagonal
// block leaves the node's load and diagonal values for k_sys_rhs.
// (Scattering node slots into the block view instead evicts partly written
// lines: measured 4.5x the useful write traffic.)
__global__ void k_sys_gather_elastic_b(int64_t nent, const int *__restrict__ bslot,
                                       const unsigned short *__restrict__ bmask,
                                       const int *__restrict__ slot_row,
                                       const int *__restrict__ cptr,
                                       const int *__restrict__ contrib,
                                       const double *__restrict__ Ke,
                                       const double *__restrict__ Fe,
                                       double *__restrict__ sval,
                                       float *__restrict__ sval32,
                                       double *__restrict__ fload,
                                       double *__restrict__ dval,
                                       unsigned long long *__restrict__ dmax) {
  const int64_t eb = xcd_item();
  unsigned long long mine = 0;
  if (eb < nent) {
    const int sl = bslot[eb];
    double out[9];
#pragma unroll
    for (int m = 0; m < 9; ++m) out[m] = 0.0;
    if (sl >= 0) {
      const unsigned mask = bmask[eb];
      double v[9], fl[3];
      bool diag;
      elastic_slot_sums(sl, cptr, contrib, Ke, Fe, v, fl, diag);
      if (diag) {
        const int r = slot_row[sl];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const unsigned long long b = abs_bits(v[4 * c]);
          mine = b > mine ? b : mine;
          dval[3ll * r + c] = v[4 * c];
          if (Fe) fload[3ll * r + c] = fl[c];
        }
      }
#pragma unroll
      for (int m = 0; m < 9; ++m)
        if ((mask >> m) & 1u) out[m] = v[m];
    }
    const int64_t c = eb >> 6, l = eb & 63;
#pragma unroll
    for (int m = 0; m < 9; ++m) {
      const int64_t vi = (c * 9 + m) * 64 + l;
      sval[vi] = out[m];
      sval32[vi] = (float)out[m];
    }
  }
  wg_max_store(mine, dmax);
}

// The node slots the block view does not hold in full (the plan's list):
// lift values f_j -= K_ji g_i, and the diagonal blocks of nodes whose three
// DOFs are all constrained (their diagonal counts in c = max |K_ii|).  One
// thread per slot; one maximum per workgroup into dmax.
__global__ void k_sys_lift_elastic(int64_t nls, const int *__restrict__ lslots,
                                   const int *__restrict__ rp,
                                   const int *__restrict__ slot_row,
                                   const int *__restrict__ cptr,
                                   const int *__restrict__ contrib,
                                   const double *__restrict__ Ke,
                                   const int *__restrict__ dst,
                                   double *__restrict__ liftv,
                                   unsigned long long *__restrict__ dmax) {
  const int64_t i = (int64_t)blockIdx.x * BS + threadIdx.x;
  unsigned long long mine = 0;
  if (i < nls) {
    const int s = lslots[i];
    double v[9], fl[3];
    bool diag;
    elastic_slot_sums(s, cptr, contrib, Ke, nullptr, v, fl, diag);
    int dd[9];
    elastic_slot_dst(s, rp, slot_row, dst, dd);
#pragma unroll
    for (int m = 0; m < 9; ++m)
      if (dd[m] <= -2) liftv[-2 - dd[m]] = v[m];
    if (diag) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const unsigned long long b = abs_bits(v[4 * c]);
        mine = b > mine ? b : mine;
      }
    }
  }
  wg_max_store(mine, dmax);
}

// Plan, elastic: the node slot of each block position of the K_bc block view
// (bslot, -1 for padding) with its kept-entry mask (bmask, bit m: entry m of
// the 3x3 block), and a flag per slot that k_sys_lift_elastic must visit:
// any lift destination, or a diagonal slot with no kept entry.
__global__ void k_plan_bslot(int64_t nnz, const int *__restrict__ rp,
                             const int *__restrict__ slot_row,
                             const int *__restrict__ col,
                             const int *__restrict__ dst, const int *__restrict__ pos,
                             int *__restrict__ bslot, unsigned short *__restrict__ bmask,
                             int *__restrict__ lflag) {
  const int64_t s = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (s >= nnz) return;
  const int r = slot_row[s];
  const int base = rp[r], deg = rp[r + 1] - base;
  const int j = (int)(s - base);
  int keep = -1, lift = 0;
  unsigned mask = 0;
  for (int ci = 0; ci < 3; ++ci)
    for (int cj = 0; cj < 3; ++cj) {
      const int d = dst[9ll * base + 3ll * ci * deg + 3ll * j + cj];
      if (d >= 0) {
        if (keep < 0) keep = d;
        mask |= 1u << (3 * ci + cj);
      }
      if (d <= -2) lift = 1;
    }
  const bool diag = col[s] == r;
  lflag[s] = lift || (diag && keep < 0);
  if (keep >= 0) {
    const int64_t vi = pos[keep];
    const int64_t eb = (vi / 64 / 9) * 64 + vi % 64;
    bslot[eb] = (int)s;
    bmask[eb] = (unsigned short)mask;
  }
}

// Incidences of node r: the contributions (e, a, a) of its diagonal slot,
// counted (cnt) and numbered from base (incpos[e*10 + a]), in slot order.
__global__ void k_inc_count(int64_t n, const int *__restrict__ diag_slot,
                            const int *__restrict__ cptr, const int *__restrict__ contrib,
                            int *__restrict__ cnt) {
  const int64_t r = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (r >= n) return;
  const int ds = diag_slot[r];
  int c = 0;
  if (ds >= 0)
    for (int i = cptr[ds]; i < cptr[ds + 1]; ++i) {
      const int ab = contrib[i] % 100;
      c += ab / 10 == ab % 10;
    }
  cnt[r] = c;
}
// Per slice: the incidence count of each SELL row (icnt) and the slice's
// largest (nib[s], its steps of 64 incidence slots).
// Accumulators slice sl's row gather needs: its longest row with its lift
// entries, or its SELL chunks' entries.
__global__ void k_slice_span(int64_t n, int64_t nslices, const int *__restrict__ sperm,
                             const int *__restrict__ brp, const int *__restrict__ lptr,
                             const int *__restrict__ cb, int *__restrict__ span) {
  const int64_t sl = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (sl >= nslices) return;
  int m = SELL_CW * (cb[sl + 1] - cb[sl]);
  for (int l = 0; l < 64; ++l) {
    const int64_t srow = sl * 64 + l;
    if (srow < n) {
      const int r = sperm[srow];
      m = max(m, brp[r + 1] - brp[r] + lptr[r + 1] - lptr[r]);
    }
  }
  span[sl] = m;
}

__global__ void k_slice_inc(int64_t n, int64_t nslices, const int *__restrict__ sperm,
                            const int *__restrict__ cnt, int *__restrict__ icnt,
                            int *__restrict__ nib) {
  const int64_t sl = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (sl >= nslices) return;
  int m = 0;
  for (int l = 0; l < 64; ++l) {
    const int64_t srow = sl * 64 + l;
    const int c = srow < n ? cnt[sperm[srow]] : 0;
    icnt[srow] = c;
    m = max(m, c);
  }
  nib[sl] = m;
}
// incpos[e*10 + a] = the slot of incidence (e, a): one thread per SELL row
__global__ void k_inc_pos(int64_t n, const int *__restrict__ sperm,
                          const int *__restrict__ diag_slot, const int *__restrict__ cptr,
                          const int *__restrict__ contrib, const int *__restrict__ ib,
                          int *__restrict__ incpos) {
  const int64_t srow = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (srow >= n) return;
  const int ds = diag_slot[sperm[srow]];
  if (ds < 0) return;
  int64_t q = (int64_t)ib[srow / 64] * 64 + (srow & 63);
  for (int i = cptr[ds]; i < cptr[ds + 1]; ++i) {
    const int t = contrib[i];
    const int ab = t % 100;
    if (ab / 10 == ab % 10) {
      incpos[(t / 100) * 10 + ab / 10] = (int)q;
      q += 64;
    }
  }
}

// Plan, thermal: the accumulator of every triplet (e, a, b) in the row
// gather of its row r = conn[e][a], stored at icode[incpos[e*10 + a]*10 + b]:
// k for the row's K_bc entry k, kept + q for its lift entry q; in a
// constrained row the diagonal is entry 0 and the dropped entries go to
// accumulator kept + lift, which is never written out.
template <typename C>
__global__ void k_plan_icode(int64_t nnz, const int *__restrict__ slot_row,
                             const int *__restrict__ col, const int *__restrict__ dst,
                             const int *__restrict__ brp, const int *__restrict__ lptr,
                             const int *__restrict__ cptr, const int *__restrict__ contrib,
                             const int *__restrict__ incpos, C *__restrict__ icode) {
  const int64_t s = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (s >= nnz) return;
  const int r = slot_row[s];
  const int d = dst[s];
  const int nk = brp[r + 1] - brp[r];
  int code;
  if (d >= 0) code = d - brp[r];
  else if (d <= -2) code = nk + (-2 - d) - lptr[r];
  else code = col[s] == r ? 0 : nk + lptr[r + 1] - lptr[r];
  for (int i = cptr[s]; i < cptr[s + 1]; ++i) {
    const int t = contrib[i];
    const int e = t / 100, ab = t - 100 * e;
    icode[(int64_t)incpos[e * 10 + ab / 10] * 10 + ab % 10] = (C)code;
  }
}

// *out = max(*out, longest row of rp)
__global__ void k_max_rowlen(int64_t n, const int *__restrict__ rp, int *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * BS + threadIdx.x;
  int m = r < n ? rp[r + 1] - rp[r] : 0;
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

__global__ void k_fill_int(int64_t n, int v, int *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (k < n) out[k] = v;
}

// out[scan[s] - 1] = s for every flagged s (scan: inclusive sum of the flags)
__global__ void k_compact_flags(int64_t n, const int *__restrict__ flag,
                                const int *__restrict__ scan, int *__restrict__ out) {
  const int64_t s = (int64_t)blockIdx.x * BS + threadIdx.x;
  if (s < n && flag[s]) out[scan[s] - 1] = (int)s;
}



// Per DOF row (internal numbering): the load the gather summed for the
// node's diagonal slot (none when fload is null) plus the caller's vector
// fext (caller numbering, nullable: the f_ext + f_history of
// src/fem/time_integration.rs:101-110), then either the constrained row
// (c e_i, f = g c) or the lift f -= K_ji g_i in the plan's order; the RHS
// lands at caller index dperm[r]; Jacobi inverse diagonal
// (src/linalg/preconditioner.rs:37-51) in SELL order from the gathered
// diagonal value.
template <int DPN>
__global__ void k_sys_rhs(int64_t n, const int *__restrict__ diag_slot,
                          const double *__restrict__ fload,
                          const double *__restrict__ dval,
                          const double *__restrict__ fext,
                          const uint8_t *__restrict__ dir,
                          const double *__restrict__ g,
                          const int *__restrict__ lptr,
                          const int *__restrict__ lcol,
                          const double *__restrict__ liftv,
                          const int *__restrict__ diag,
                          const int *__restrict__ pos,
                          const int *__restrict__ siperm,
                          const unsigned long long *__restrict__ maxbits,
                          double *__restrict__ sval,
                          float *__restrict__ sval32,
                          const int *__restrict__ dperm,
                          double *__restrict__ fbc, double *__restrict__ sdinv,
                          float *__restrict__ sdinv32) {
  const int64_t r = xcd_item();
  if (r >= n) return;
  const int64_t out = dperm ? dperm[r] : r;
  const int d = diag[r];
  double f, dv;
  if (dir[r]) {
    const double cv = __longlong_as_double((long long)*maxbits);
    const int e = pos[d];
    sval[e] = cv;
    if (sval32) sval32[e] = (float)cv;
    f = g[r] * cv;
    dv = cv;
  } else {
    const bool has = diag_slot[r / DPN] >= 0;
    f = (has && fload) ? fload[r] : 0.0;
    if (fext) f = fext[out] + f;
    for (int q = lptr[r]; q < lptr[r + 1]; ++q) f -= liftv[q] * g[lcol[q]];
    dv = d >= 0 ? dval[r] : 0.0;
  }
  fbc[out] = f;
  const double di = fabs(dv) > 1e-14 ? 1.0 / dv : 1.0;
  const int si = siperm[r];
  sdinv[si] = di;
  sdinv32[si] = (float)di;
}

__global__ void k_iota(int64_t n, int *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = (int)k;
}

// f[node] += contributions, each node's in face order (the contributions
// are sorted stably by node; a run head adds its run in order).
__global__ void k_neumann_add(int64_t N, const int *__restrict__ node,
                              const int *__restrict__ idx,
                              const double *__restrict__ contrib, double *__restrict__ f) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int nd = node[k];
  if (k > 0 && node[k - 1] == nd) return;
  double acc = f[nd];
  for (int64_t j = k; j < N && node[j] == nd; ++j) acc += contrib[idx[j]];
  f[nd] = acc;
}

template <class F> void scan_inclusive(gs_ctx *ctx, const int *in, int *out,
                                       int64_t n, F op) {
  size_t bytes = 0;
  GS_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, (size_t)n, op,
                                 ctx->stream));
  auto tmp = dalloc<char>((int64_t)bytes);
  GS_HIP(rocprim::inclusive_scan(tmp.get(), bytes, in, out, (size_t)n, op,
                                 ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
}

int read_int(gs_ctx *ctx, const int *p) {
  int v = 0;
  GS_HIP(hipMemcpyAsync(ctx->pinned, p, sizeof(int), hipMemcpyDeviceToHost,
                        ctx->stream));
  GS_HIP(hipStreamSynchronize(ctx->stream));
  memcpy(&v, ctx->pinned, sizeof(int));
  return v;
}

}  // namespace

int read_flag_and_clear(gs_ctx *ctx, int *flag_dev) {
  const int v = read_int(ctx, flag_dev);
  hipLaunchKernelGGL(k_set_int, 1, 1, 0, ctx->stream, flag_dev, 0);
  GS_LAUNCH_CHECK("k_set_int");
  return v;
}

void build_node_pattern(gs_ctx *ctx, gs_mesh *m) {
  if (m->has_pattern) return;
  const int64_t E = m->n_elems, n = m->n_nodes;
  const int64_t N = E * 100;
  if (N >= (int64_t)INT32_MAX)
    throw Error(GS_ERR_TOO_LARGE, "mesh too large: 100*n_elems must be < 2^31");
  int bits = 1;
  while ((1ll << bits) < n) ++bits;
  hipStream_t s = ctx->stream;
  NodePattern p;
  p.row_ptr = dalloc<int>(n + 1);
  p.diag_slot = dalloc<int>(n);
  hipLaunchKernelGGL(k_fill_int, nblk(n), BS, 0, s, n, p.diag_slot.get(), -1);
  GS_LAUNCH_CHECK("k_fill_int");
  if (N == 0) {
    GS_HIP(hipMemsetAsync(p.row_ptr.get(), 0, (n + 1) * sizeof(int), s));
    p.col = dalloc<int>(0);
    p.slot_row = dalloc<int>(0);
    p.contrib_ptr = dalloc<int>(1);
    GS_HIP(hipMemsetAsync(p.contrib_ptr.get(), 0, sizeof(int), s));
    p.contrib = dalloc<int>(0);
    GS_HIP(hipStreamSynchronize(s));
    m->pat = p;
    m->has_pattern = true;
    return;
  }
  DArr<int> head, incl;
  {
    auto keys = dalloc<uint64_t>(N);
    auto keys2 = dalloc<uint64_t>(N);
    auto vals = dalloc<int>(N);
    p.contrib = dalloc<int>(N);
    hipLaunchKernelGGL(k_make_keys, nblk(N), BS, 0, s, E, m->conn.get(),
                       m->iperm.get(), bits,
                       keys.get(), vals.get());
    GS_LAUNCH_CHECK("k_make_keys");
    size_t bytes = 0;
    GS_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys.get(), keys2.get(),
                                     vals.get(), p.contrib.get(), (size_t)N, 0u,
                                     (unsigned)(2 * bits), s));
    {
      auto tmp = dalloc<char>((int64_t)bytes);
      GS_HIP(rocprim::radix_sort_pairs(tmp.get(), bytes, keys.get(),
                                       keys2.get(), vals.get(), p.contrib.get(),
                                       (size_t)N, 0u, (unsigned)(2 * bits), s));
      GS_HIP(hipStreamSynchronize(s));
    }
    keys = DArr<uint64_t>();
    vals = DArr<int>();
    head = dalloc<int>(N);
    incl = dalloc<int>(N);
    hipLaunchKernelGGL(k_mark_heads, nblk(N), BS, 0, s, N, keys2.get(),
                       head.get());
    GS_LAUNCH_CHECK("k_mark_heads");
    scan_inclusive(ctx, head.get(), incl.get(), N, rocprim::plus<int>());
    const int nnz = read_int(ctx, incl.get() + (N - 1));
    p.nnz = nnz;
    p.col = dalloc<int>(nnz);
    p.slot_row = dalloc<int>(nnz);
    p.contrib_ptr = dalloc<int>(nnz + 1);
    auto row_cnt = dalloc<int>(n + 1);
    GS_HIP(hipMemsetAsync(row_cnt.get(), 0, (n + 1) * sizeof(int), s));
    hipLaunchKernelGGL(k_slots, nblk(N), BS, 0, s, N, keys2.get(), incl.get(),
                       bits, p.col.get(), p.slot_row.get(),
                       p.contrib_ptr.get(), p.diag_slot.get(), row_cnt.get());
    GS_LAUNCH_CHECK("k_slots");
    hipLaunchKernelGGL(k_set_int, 1, 1, 0, s, p.contrib_ptr.get() + nnz, (int)N);
    // row_ptr[0] = 0, row_ptr[1..n] = inclusive scan of counts
    GS_HIP(hipMemsetAsync(p.row_ptr.get(), 0, sizeof(int), s));
    scan_inclusive(ctx, row_cnt.get(), p.row_ptr.get() + 1, n,
                   rocprim::plus<int>());
  }
  GS_HIP(hipStreamSynchronize(s));
  m->pat = p;
  m->has_pattern = true;
}

void build_dof3_pattern(gs_ctx *ctx, gs_mesh *m) {
  if (m->has_dof3) return;
  build_node_pattern(ctx, m);
  const NodePattern &p = m->pat;
  const int64_t n = m->n_nodes;
  if (9 * p.nnz >= (int64_t)INT32_MAX || 3 * n >= (int64_t)INT32_MAX)
    throw Error(GS_ERR_TOO_LARGE, "elastic operator too large for int32 indices");
  DofPattern3 d;
  d.nnz = 9 * p.nnz;
  d.row_ptr = dalloc<int>(3 * n + 1);
  d.col = dalloc<int>(d.nnz);
  hipLaunchKernelGGL(k_dof3_rowptr, nblk(3 * n + 1), BS, 0, ctx->stream, n,
                     p.row_ptr.get(), d.row_ptr.get());
  GS_LAUNCH_CHECK("k_dof3_rowptr");
  if (p.nnz > 0) {
    hipLaunchKernelGGL(k_dof3_cols, nblk(p.nnz), BS, 0, ctx->stream, p.nnz,
                       p.row_ptr.get(), p.col.get(), p.slot_row.get(),
                       d.col.get());
    GS_LAUNCH_CHECK("k_dof3_cols");
  }
  GS_HIP(hipStreamSynchronize(ctx->stream));
  m->dof3 = d;
  m->has_dof3 = true;
}

void gather_thermal(hipStream_t s, const NodePattern &p, int64_t E,
                    const double *Ke, double *val) {
  (void)E;
  if (p.nnz == 0) return;
  hipLaunchKernelGGL((k_gather_thermal<false>), nblk(p.nnz), BS, 0, s, p.nnz,
                     p.contrib_ptr.get(), p.contrib.get(), Ke, val);
  GS_LAUNCH_CHECK("k_gather_thermal");
}

void apply_neumann(gs_ctx *ctx, const gs_mesh *m, int64_t F, const int64_t *elem,
                   const int *lface, const double *flux, double *f) {
  if (F == 0) return;
  hipStream_t s = ctx->stream;
  const int64_t N = 6 * F;
  auto contrib = dalloc<double>(N);
  auto node = dalloc<int>(N), node2 = dalloc<int>(N);
  auto idx = dalloc<int>(N), idx2 = dalloc<int>(N);
  launch_face_flux(s, F, m->xyz.get(), m->conn.get(), elem, lface, flux, contrib.get(),
                   node.get());
  hipLaunchKernelGGL(k_iota, nblk(N), BS, 0, s, N, idx.get());
  int bits = 1;
  while (bits < 31 && (int64_t(1) << bits) < m->n_nodes) ++bits;
  size_t bytes = 0;
  GS_HIP(rocprim::radix_sort_pairs(nullptr, bytes, node.get(), node2.get(), idx.get(),
                                   idx2.get(), (size_t)N, 0u, (unsigned)bits, s));
  auto tmp = dalloc<char>((int64_t)bytes);
  GS_HIP(rocprim::radix_sort_pairs(tmp.get(), bytes, node.get(), node2.get(), idx.get(),
                                   idx2.get(), (size_t)N, 0u, (unsigned)bits, s));
  hipLaunchKernelGGL(k_neumann_add, nblk(N), BS, 0, s, N, node2.get(), idx2.get(),
                     contrib.get(), f);
  GS_LAUNCH_CHECK("k_neumann_add");
  GS_HIP(hipStreamSynchronize(s));
}

void gather_thermal_full(hipStream_t s, const NodePattern &p, const double *Me,
                         double *val) {
  if (p.nnz == 0) return;
  hipLaunchKernelGGL((k_gather_thermal<true>), nblk(p.nnz), BS, 0, s, p.nnz,
                     p.contrib_ptr.get(), p.contrib.get(), Me, val);
  GS_LAUNCH_CHECK("k_gather_thermal_full");
}

void gather_thermal_load(hipStream_t s, const NodePattern &p, int64_t n_nodes,
                         const double *Fe, const int *out_map, double *f) {
  if (n_nodes == 0) return;
  hipLaunchKernelGGL((k_gather_load<1>), nblk(n_nodes), BS, 0, s, n_nodes,
                     p.diag_slot.get(), p.contrib_ptr.get(), p.contrib.get(),
                     Fe, out_map, f);
  GS_LAUNCH_CHECK("k_gather_load<1>");
}

void gather_elastic(hipStream_t s, const NodePattern &p, int64_t n_nodes,
                    const double *Ke, double *val) {
  (void)n_nodes;
  if (p.nnz == 0) return;
  hipLaunchKernelGGL(k_gather_elastic, nblk(p.nnz), BS, 0, s, p.nnz,
                     p.row_ptr.get(), p.slot_row.get(), p.contrib_ptr.get(),
                     p.contrib.get(), Ke, val);
  GS_LAUNCH_CHECK("k_gather_elastic");
}

void gather_gravity_load(hipStream_t s, const NodePattern &p, int64_t n_nodes,
                         const double *Fe, const int *out_map, double *f) {
  if (n_nodes == 0) return;
  hipLaunchKernelGGL((k_gather_load<3>), nblk(n_nodes), BS, 0, s, n_nodes,
                     p.diag_slot.get(), p.contrib_ptr.get(), p.contrib.get(),
                     Fe, out_map, f);
  GS_LAUNCH_CHECK("k_gather_load<3>");
}

gs_csr *dirichlet_eliminate(gs_ctx *ctx, const gs_csr *K, const double *f,
                            const uint8_t *is_dir, const double *g,
                            double *fbc) {
  ensure_csr_values(ctx, K);
  const int64_t n = K->rows;
  hipStream_t s = ctx->stream;
  auto cnt = dalloc<int>(n > 0 ? n : 1);
  auto maxbits = dalloc<unsigned long long>(1);
  const double one = 1.0;
  unsigned long long one_bits;
  memcpy(&one_bits, &one, sizeof(one));
  hipLaunchKernelGGL(k_set_u64, 1, 1, 0, s, maxbits.get(), one_bits);
  GS_LAUNCH_CHECK("k_set_u64");
  std::unique_ptr<gs_csr> out(new gs_csr());
  out->ctx = ctx;
  out->rows = n;
  out->cols = K->cols;
  out->dperm = K->dperm;
  out->row_ptr = dalloc<int>(n + 1);
  GS_HIP(hipMemsetAsync(out->row_ptr.get(), 0, sizeof(int), s));
  if (n > 0) {
    hipLaunchKernelGGL(k_bc_count, nblk_max(n), BS, 0, s, n, K->row_ptr.get(),
                       K->col.get(), K->val.get(), is_dir, cnt.get(),
                       maxbits.get());
    GS_LAUNCH_CHECK("k_bc_count");
    scan_inclusive(ctx, cnt.get(), out->row_ptr.get() + 1, n,
                   rocprim::plus<int>());
    out->nnz = read_int(ctx, out->row_ptr.get() + n);
  }
  out->col = dalloc<int>(out->nnz);
  out->val = dalloc<double>(out->nnz);
  if (n > 0) {
    hipLaunchKernelGGL(k_bc_fill, nblk(n), BS, 0, s, n, K->row_ptr.get(),
                       K->col.get(), K->val.get(), is_dir, g, f,
                       K->dperm.get(), maxbits.get(), out->row_ptr.get(),
                       out->col.get(), out->val.get(), fbc);
    GS_LAUNCH_CHECK("k_bc_fill");
  }
  GS_HIP(hipStreamSynchronize(s));
  return out.release();
}

void build_perm3(gs_ctx *ctx, gs_mesh *m) {
  if (m->perm3) return;
  m->perm3 = dalloc<int>(3 * m->n_nodes);
  if (m->n_nodes > 0) {
    hipLaunchKernelGGL(k_perm3, nblk(3 * m->n_nodes), BS, 0, ctx->stream,
                       m->n_nodes, m->perm.get(), m->perm3.get());
    GS_LAUNCH_CHECK("k_perm3");
  }
}

void permute_gather(hipStream_t s, int64_t n, const double *src, const int *map,
                    double *dst) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_perm_gather, nblk(n), BS, 0, s, n, src, map, dst);
  GS_LAUNCH_CHECK("k_perm_gather");
}

void permute_gather_u8(hipStream_t s, int64_t n, const uint8_t *src,
                       const int *map, uint8_t *dst) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_perm_gather_u8, nblk(n), BS, 0, s, n, src, map, dst);
  GS_LAUNCH_CHECK("k_perm_gather_u8");
}

void permute_scatter(hipStream_t s, int64_t n, const double *src, const int *map,
                     double *dst) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_perm_scatter, nblk(n), BS, 0, s, n, src, map, dst);
  GS_LAUNCH_CHECK("k_perm_scatter");
}

std::shared_ptr<SystemPlan> system_plan(gs_ctx *ctx, gs_mesh *m,
                                        const gs_dofmap *d) {
  for (auto &p : m->plans)
    if (p->dof_uid == d->uid) return p;
  hipStream_t s = ctx->stream;
  const int dpn = d->dpn;
  if (dpn == 1) {
    build_node_pattern(ctx, m);
  } else {
    build_dof3_pattern(ctx, m);
    build_perm3(ctx, m);
  }
  const int *rp = dpn == 1 ? m->pat.row_ptr.get() : m->dof3.row_ptr.get();
  const int *col = dpn == 1 ? m->pat.col.get() : m->dof3.col.get();
  const int64_t nnzK = dpn == 1 ? m->pat.nnz : m->dof3.nnz;
  const DArr<int> &dperm = dpn == 1 ? m->perm : m->perm3;
  const int64_t n = d->n_dofs;
  auto P = std::make_shared<SystemPlan>();
  P->dof_uid = d->uid;
  P->dpn = dpn;
  P->n = n;
  P->dir = dalloc<uint8_t>(n);
  P->g = dalloc<double>(n);
  permute_gather_u8(s, n, d->is_dir.get(), dperm.get(), P->dir.get());
  permute_gather(s, n, d->values.get(), dperm.get(), P->g.get());
  P->row_ptr = dalloc<int>(n + 1);
  P->lift_ptr = dalloc<int>(n + 1);
  GS_HIP(hipMemsetAsync(P->row_ptr.get(), 0, sizeof(int), s));
  GS_HIP(hipMemsetAsync(P->lift_ptr.get(), 0, sizeof(int), s));
  if (n > 0) {
    auto kept = dalloc<int>(n), lift = dalloc<int>(n);
    hipLaunchKernelGGL(k_plan_count, nblk(n), BS, 0, s, n, rp, col, P->dir.get(),
                       kept.get(), lift.get());
    GS_LAUNCH_CHECK("k_plan_count");
    scan_inclusive(ctx, kept.get(), P->row_ptr.get() + 1, n, rocprim::plus<int>());
    scan_inclusive(ctx, lift.get(), P->lift_ptr.get() + 1, n, rocprim::plus<int>());
    P->nnz = read_int(ctx, P->row_ptr.get() + n);
    P->n_lift = read_int(ctx, P->lift_ptr.get() + n);
  }
  P->col = dalloc<int>(P->nnz);
  P->dst = dalloc<int>(nnzK);
  P->lift_col = dalloc<int>(P->n_lift);
  P->diag = dalloc<int>(n);
  if (n > 0) {
    hipLaunchKernelGGL(k_plan_fill, nblk(n), BS, 0, s, n, rp, col, P->dir.get(),
                       dperm.get(), P->row_ptr.get(), P->col.get(), P->dst.get(),
                       P->lift_ptr.get(), P->lift_col.get(), P->diag.get());
    GS_LAUNCH_CHECK("k_plan_fill");
  }
  gs_csr pat;
  pat.ctx = ctx;
  pat.rows = pat.cols = n;
  pat.nnz = P->nnz;
  pat.row_ptr = P->row_ptr;
  pat.col = P->col;
  pat.dperm = dperm;
  P->sell = build_sell_pattern(ctx, &pat, false, P->dpn == 3 ? 3 : 1);
  const NodePattern &np = m->pat;
  if (dpn == 1) {
    // incidence slots of the row gather, and their accumulator codes
    const int64_t E = m->n_elems;
    const SellPattern &S = *P->sell;
    P->ib = dalloc<int>(S.nslices + 1);
    P->icnt = dalloc<int>(S.nslices * 64);
    P->incpos = dalloc<int>(10 * E);
    GS_HIP(hipMemsetAsync(P->ib.get(), 0, sizeof(int), s));
    if (S.nslices > 0) {
      auto cnt = dalloc<int>(n), nib = dalloc<int>(S.nslices);
      hipLaunchKernelGGL(k_inc_count, nblk(n), BS, 0, s, n, np.diag_slot.get(),
                         np.contrib_ptr.get(), np.contrib.get(), cnt.get());
      GS_LAUNCH_CHECK("k_inc_count");
      hipLaunchKernelGGL(k_slice_inc, nblk(S.nslices), BS, 0, s, n, S.nslices,
                         S.sperm.get(), cnt.get(), P->icnt.get(), nib.get());
      GS_LAUNCH_CHECK("k_slice_inc");
      scan_inclusive(ctx, nib.get(), P->ib.get() + 1, S.nslices, rocprim::plus<int>());
      const int steps = read_int(ctx, P->ib.get() + S.nslices);
      if ((int64_t)steps * 64 >= (int64_t)INT32_MAX)
        throw Error(GS_ERR_TOO_LARGE, "too many incidence slots for int32 indices");
      P->n_islots = (int64_t)steps * 64;
      hipLaunchKernelGGL(k_inc_pos, nblk(n), BS, 0, s, n, S.sperm.get(),
                         np.diag_slot.get(), np.contrib_ptr.get(), np.contrib.get(),
                         P->ib.get(), P->incpos.get());
      GS_LAUNCH_CHECK("k_inc_pos");
      // the gather's slice order: short ones first (each part in slice order)
      auto span = dalloc<int>(S.nslices);
      hipLaunchKernelGGL(k_slice_span, nblk(S.nslices), BS, 0, s, n, S.nslices,
                         S.sperm.get(), P->row_ptr.get(), P->lift_ptr.get(), S.cb.get(),
                         span.get());
      GS_LAUNCH_CHECK("k_slice_span");
      const std::vector<int> hs = to_host(ctx, span.get(), S.nslices);
      std::vector<int> order;
      order.reserve((size_t)S.nslices);
      for (int64_t sl = 0; sl < S.nslices; ++sl)
        if (hs[sl] <= ROW_LW_SHORT) order.push_back((int)sl);
      P->n_tshort = (int64_t)order.size();
      for (int64_t sl = 0; sl < S.nslices; ++sl)
        if (hs[sl] > ROW_LW_SHORT) order.push_back((int)sl);
      P->tslices = to_device(ctx, order);
    }
    int maxlen = 0;
    if (n > 0) {
      auto ml = dalloc<int>(1);
      hipLaunchKernelGGL(k_set_int, 1, 1, 0, s, ml.get(), 0);
      hipLaunchKernelGGL(k_max_rowlen, nblk(n), BS, 0, s, n, rp, ml.get());
      GS_LAUNCH_CHECK("k_max_rowlen");
      maxlen = read_int(ctx, ml.get());
    }
    // codes run to the row length (the dropped entries of a constrained row
    // go past its single kept entry)
    if (maxlen > 65535)
      throw Error(GS_ERR_TOO_LARGE, "a node couples to more than 65535 nodes");
    auto go = [&](auto *out) {
      if (nnzK > 0)
        hipLaunchKernelGGL(k_plan_icode, nblk(nnzK), BS, 0, s, nnzK, np.slot_row.get(),
                           np.col.get(), P->dst.get(), P->row_ptr.get(), P->lift_ptr.get(),
                           np.contrib_ptr.get(), np.contrib.get(), P->incpos.get(), out);
      GS_LAUNCH_CHECK("k_plan_icode");
    };
    if (maxlen <= 255) {
      P->icode8 = dalloc<uint8_t>(10 * P->n_islots);
      go(P->icode8.get());
    } else {
      P->icode16 = dalloc<unsigned short>(10 * P->n_islots);
      go(P->icode16.get());
    }
  } else {
    // the block-view order of the elastic numeric gather
    const SellPattern &S = *P->sell;
    P->bslot = dalloc<int>(S.entries);
    P->bmask = dalloc<unsigned short>(S.entries);
    if (S.entries > 0)
      hipLaunchKernelGGL(k_fill_int, nblk(S.entries), BS, 0, s, S.entries, -1,
                         P->bslot.get());
    if (np.nnz > 0) {
      auto lflag = dalloc<int>(np.nnz), lscan = dalloc<int>(np.nnz);
      hipLaunchKernelGGL(k_plan_bslot, nblk(np.nnz), BS, 0, s, np.nnz, np.row_ptr.get(),
                         np.slot_row.get(), np.col.get(), P->dst.get(), S.pos.get(),
                         P->bslot.get(), P->bmask.get(), lflag.get());
      GS_LAUNCH_CHECK("k_plan_bslot");
      scan_inclusive(ctx, lflag.get(), lscan.get(), np.nnz, rocprim::plus<int>());
      P->n_lift_slots = read_int(ctx, lscan.get() + np.nnz - 1);
      P->lift_slots = dalloc<int>(P->n_lift_slots);
      hipLaunchKernelGGL(k_compact_flags, nblk(np.nnz), BS, 0, s, np.nnz, lflag.get(),
                         lscan.get(), P->lift_slots.get());
      GS_LAUNCH_CHECK("k_compact_flags");
    }
  }
  GS_HIP(hipStreamSynchronize(s));
  m->plans.push_back(P);
  if (m->plans.size() > 4) m->plans.erase(m->plans.begin());
  return P;
}

int64_t system_gather_blocks(const gs_mesh *m, const SystemPlan &P) {
  if (P.dpn == 1) return P.sell->nslices;
  return nblk(P.sell->entries) + (int64_t)nblk(P.n_lift_slots);
}

void system_gather(hipStream_t s, const gs_mesh *m, const SystemPlan &P, const double *Ke,
                   const double *Fe, bool with_load, double *sval, float *sval32,
                   double *liftv, double *fload, double *dval, unsigned long long *dmax) {
  const NodePattern &p = m->pat;
  if (p.nnz == 0) return;
  if (P.dpn == 1) {
    const SellPattern &S = *P.sell;
    auto go = [&](const auto *code) {
      using C = std::remove_cv_t<std::remove_pointer_t<decltype(code)>>;
      const int64_t ns = P.n_tshort, nl = S.nslices - P.n_tshort;
      if (ns > 0)
        hipLaunchKernelGGL((k_sys_rows_thermal<C, ROW_LW_SHORT>), (unsigned)ns, 64, 0, s,
                           P.tslices.get(), P.n, S.sperm.get(), S.cb.get(), P.row_ptr.get(),
                           P.lift_ptr.get(), P.diag.get(), P.ib.get(), P.icnt.get(), Ke,
                           code, (int)with_load, sval, sval32, liftv, fload, dval, dmax);
      if (nl > 0)
        hipLaunchKernelGGL((k_sys_rows_thermal<C, ROW_LW>), (unsigned)nl, 64, 0, s,
                           P.tslices.get() + ns, P.n, S.sperm.get(), S.cb.get(),
                           P.row_ptr.get(), P.lift_ptr.get(), P.diag.get(), P.ib.get(),
                           P.icnt.get(), Ke, code, (int)with_load, sval, sval32, liftv,
                           fload, dval, dmax + ns);
    };
    if (S.nslices > 0) {
      if (P.icode8) go((const uint8_t *)P.icode8.get());
      else go((const unsigned short *)P.icode16.get());
    }
  } else {
    const int64_t nent = P.sell->entries;
    if (nent > 0)
      hipLaunchKernelGGL(k_sys_gather_elastic_b, nblk(nent), BS, 0, s, nent,
                         P.bslot.get(), P.bmask.get(), p.slot_row.get(),
                         p.contrib_ptr.get(), p.contrib.get(), Ke, with_load ? Fe : nullptr,
                         sval, sval32, fload, dval, dmax);
    if (P.n_lift_slots > 0)
      hipLaunchKernelGGL(k_sys_lift_elastic, nblk(P.n_lift_slots), BS, 0, s,
                         P.n_lift_slots, P.lift_slots.get(), p.row_ptr.get(),
                         p.slot_row.get(), p.contrib_ptr.get(), p.contrib.get(), Ke,
                         P.dst.get(), liftv, dmax + nblk(nent));
  }
  GS_LAUNCH_CHECK("k_sys_gather");
}

void system_finish(gs_ctx *ctx, const gs_mesh *m, const SystemPlan &P, const double *fload,
                   const double *dval, const double *fext, double *sval, float *sval32,
                   const double *liftv, const int *dperm, double *fbc, double *sdinv,
                   float *sdinv32, const unsigned long long *dmax) {
  if (P.n == 0) return;
  hipStream_t s = ctx->stream;
  const NodePattern &p = m->pat;
  auto maxbits = palloc<unsigned long long>(ctx->pool, 1);
  const double one = 1.0;
  unsigned long long one_bits;
  memcpy(&one_bits, &one, sizeof(one));
  // c = max(1, max |K_ii|): the gathers left one maximum per workgroup
  hipLaunchKernelGGL(k_set_u64, 1, 1, 0, s, maxbits.get(), one_bits);
  const int64_t np = p.nnz > 0 ? system_gather_blocks(m, P) : 0;
  if (np > 0)
    hipLaunchKernelGGL(k_max_partials, (unsigned)std::min<int64_t>(256, nblk(np)), BS, 0, s,
                       np, dmax, maxbits.get());
  const int *sip = P.sell->siperm.get(), *pos = P.sell->pos.get();
  auto go = [&](auto dpn) {
    constexpr int D = decltype(dpn)::value;
    hipLaunchKernelGGL((k_sys_rhs<D>), nblk(P.n), BS, 0, s, P.n, p.diag_slot.get(), fload,
                       dval, fext, P.dir.get(), P.g.get(), P.lift_ptr.get(), P.lift_col.get(),
                       liftv, P.diag.get(), pos, sip, maxbits.get(), sval, sval32, dperm, fbc,
                       sdinv, sdinv32);
  };
  if (P.dpn == 1) go(std::integral_constant<int, 1>{});
  else go(std::integral_constant<int, 3>{});
  GS_LAUNCH_CHECK("k_sys_rhs");
  GS_HIP(hipStreamSynchronize(s));
}

}  // namespace gs
