// Sliced-ELL (SELL-64) SpMV engine for gfx950 (included by solver.hip).
//
// Layout.  Rows are cut into slices of 64, one wavefront per slice and one
// lane per row.  A slice stores its rows' entries in chunks of 256: chunk c
// holds entries 4c..4c+3 of every row, lane l's four entries contiguous at
// 4l, so each lane fetches 16 (FP32) or 32 (FP64) bytes of values and 8 bytes
// of indices per load and a wavefront reads one contiguous 1-2 KiB stretch.
// A slice has as many chunks as its longest row needs; shorter rows are
// padded with entries whose index is the pad marker (they add nothing, and
// their values are never used, so they need no clearing).  To
// keep padding low the rows are sorted by descending length inside windows
// of 256 consecutive rows; the sort is a symmetric renumbering, so vectors
// live in the same order and every lane writes its own row of y
// contiguously.
//
// Indices: a 16-bit offset from the slice's first row (column =
// 64*slice + off, pad = -32768) when every column fits, else int32 columns
// (pad = -1).
//
// Block variant (3-DOF-per-node operators).  Slices of 64 block rows (3 DOF
// rows each), one lane per block row.  Chunk c holds one 3x3 block per lane:
// the block's column index at c*64 + lane and its 9 values component-major,
// value m of lane l at (c*9 + m)*64 + l, so every load of a wavefront is one
// contiguous stretch.  A block is stored when any of its entries is in the
// matrix; the missing ones are zeros.  One index per 9 values instead of one
// per value.  DOF row 3s+c of the SELL order is block row s, component c.
//
// Scheduling.  Workgroups take blocks of 4 slices.  The blocks are split in
// 8 contiguous ranges, one per XCD (workgroup i runs on XCD i mod 8), so the
// rows an XCD works on at any moment are close together and the gathered
// part of x stays in that XCD's L2.
#pragma once

namespace gs {
namespace {

constexpr int SELL_C = 64;        // rows per slice
constexpr int SELL_CW = 4;        // entries per lane per chunk
constexpr int SELL_CHUNK = 256;   // entries per chunk
constexpr int SELL_WPB = 4;       // slices per workgroup (256 threads)
constexpr int SELL_SIGMA = 256;   // length-sorting window (rows)
constexpr int SELL_PAD16 = -32768;

// STREAM: the operator's values bypass the caches (see load_vals)
template <bool STREAM> struct SellIdx16 {
  const short *__restrict__ off;
  static constexpr bool wide = false;
  static constexpr int bs = 1;
  static constexpr bool stream = STREAM;
};
template <bool STREAM> struct SellIdx32 {
  const int *__restrict__ col;
  static constexpr bool wide = true;
  static constexpr int bs = 1;
  static constexpr bool stream = STREAM;
};
// block variant: the same index streams, one entry per 3x3 block
template <bool STREAM> struct BSellIdx16 {
  const short *__restrict__ off;
  static constexpr bool wide = false;
  static constexpr int bs = 3;
  static constexpr bool stream = STREAM;
};
template <bool STREAM> struct BSellIdx32 {
  const int *__restrict__ col;
  static constexpr bool wide = true;
  static constexpr int bs = 3;
  static constexpr bool stream = STREAM;
};

// Operators larger than this (values and indices, in the precision a
// kernel reads them) stream their values: 7/8 of the 256 MB Infinity Cache.
constexpr int64_t SELL_RESIDENT_BYTES = 224ll << 20;

template <typename T> struct SellVec;
template <> struct SellVec<float> {
  typedef float t __attribute__((ext_vector_type(4)));
};
template <> struct SellVec<double> {
  typedef double t __attribute__((ext_vector_type(4)));
};
template <> struct SellVec<int> {
  typedef int t __attribute__((ext_vector_type(4)));
};

// Fixed-point operator values (int32, value * 2^30): the row sums are taken
// on the integers in FP64 and scaled once.
constexpr double SELL_Q30 = 0x1p-30;

// Value loads.  An operator larger than the Infinity Cache streams its values
// past the caches (non-temporal loads), which leaves the cached lines to the
// pass's vectors and to the index stream: the SpMV itself runs at the same
// speed, the solves gain (FP64 at 1.08M thermal DOFs 26.3 -> 25.4 ms, at
// 1.38M elastic DOFs 181 -> 166 ms).  An operator that fits (the thermal FP32
// one, 194 MB) is kept cached across passes with ordinary loads.
template <bool STREAM, typename V, typename T>
__device__ __forceinline__ V load_vals(const T *p) {
  if constexpr (STREAM) return __builtin_nontemporal_load((const V *)p);
  else return *(const V *)p;
}

// One lane's dot product over chunks [c0, c1) of its slice.  The sum is
// accumulated in FP64 for either value type: the product of two FP32 values
// is exact in FP64, so an FP32 row is rounded once, when it is stored (this
// keeps the FP32 recursion's rounding lower; the memory-bound kernel's time
// does not change).
// TV: the value type -- the vector type T by default, or int32 fixed point
// (SELL_Q30) for the mixed solve's scaled operator.
template <typename T, class IDX, typename TV>
__device__ __forceinline__ T sell_lane(const IDX idx, const TV *__restrict__ val,
                                       const T *__restrict__ x, int c0, int c1,
                                       int lane, int base_row) {
  using V = typename SellVec<TV>::t;
  double acc = 0.0;
  int c = c0;
  auto step = [&](const V &v, int j0, int j1, int j2, int j3) {
    const int cc[4] = {j0, j1, j2, j3};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      // Branch-free: a pad entry loads the slice's first row (a valid,
      // cached address) and its product is dropped by a select, so neither
      // its (unwritten) value nor a non-finite x can reach the sum.
      const bool ok = IDX::wide ? cc[j] >= 0 : cc[j] != SELL_PAD16;
      const int col = ok ? (IDX::wide ? cc[j] : base_row + cc[j]) : base_row;
      const double prod = (double)v[j] * (double)x[col];
      acc += ok ? prod : 0.0;
    }
  };
  for (; c + 2 <= c1; c += 2) {
    const size_t e0 = (size_t)c * SELL_CHUNK + lane * SELL_CW;
    const size_t e1 = e0 + SELL_CHUNK;
    V v0 = load_vals<IDX::stream, V>(val + e0);
    V v1 = load_vals<IDX::stream, V>(val + e1);
    if constexpr (IDX::wide) {
      typedef int I4 __attribute__((ext_vector_type(4)));
      const I4 i0 = *(const I4 *)(idx.col + e0);
      const I4 i1 = *(const I4 *)(idx.col + e1);
      step(v0, i0[0], i0[1], i0[2], i0[3]);
      step(v1, i1[0], i1[1], i1[2], i1[3]);
    } else {
      typedef short S4 __attribute__((ext_vector_type(4)));
      const S4 i0 = *(const S4 *)(idx.off + e0);
      const S4 i1 = *(const S4 *)(idx.off + e1);
      step(v0, i0[0], i0[1], i0[2], i0[3]);
      step(v1, i1[0], i1[1], i1[2], i1[3]);
    }
  }
  if (c < c1) {
    const size_t e0 = (size_t)c * SELL_CHUNK + lane * SELL_CW;
    V v0 = load_vals<IDX::stream, V>(val + e0);
    if constexpr (IDX::wide) {
      typedef int I4 __attribute__((ext_vector_type(4)));
      const I4 i0 = *(const I4 *)(idx.col + e0);
      step(v0, i0[0], i0[1], i0[2], i0[3]);
    } else {
      typedef short S4 __attribute__((ext_vector_type(4)));
      const S4 i0 = *(const S4 *)(idx.off + e0);
      step(v0, i0[0], i0[1], i0[2], i0[3]);
    }
  }
  if constexpr (std::is_same<TV, int>::value) acc *= SELL_Q30;
  return (T)acc;
}

// One lane's three dot products (block row base_row + lane) over chunks
// [c0, c1), NB chunks in flight.  Each row sums its entries in ascending
// column order, as the scalar engine does, and in FP64 as it does.
template <typename T, class IDX, typename TV>
__device__ __forceinline__ void b3_lane(const IDX idx, const TV *__restrict__ val,
                                        const T *__restrict__ x, int c0, int c1,
                                        int lane, int base_row, T (&out)[3]) {
  static_assert(!std::is_same<TV, int>::value, "fixed point: scalar view only");
  double acc[3] = {0.0, 0.0, 0.0};
  auto load = [&](int c, int &col, bool &ok, TV (&v)[9]) {
    const size_t e = (size_t)c * SELL_C + lane;
    if constexpr (IDX::wide) {
      const int j = idx.col[e];
      ok = j >= 0;
      col = ok ? j : base_row;
    } else {
      const int j = idx.off[e];
      ok = j != SELL_PAD16;
      col = ok ? base_row + j : base_row;
    }
#pragma unroll
    for (int m = 0; m < 9; ++m) v[m] = load_vals<IDX::stream, TV>(val + ((size_t)c * 9 + m) * SELL_C + lane);
  };
  auto fma3 = [&](int col, bool ok, const TV (&v)[9]) {
    const double x0 = x[3 * col], x1 = x[3 * col + 1], x2 = x[3 * col + 2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const double p0 = (double)v[3 * r] * x0, p1 = (double)v[3 * r + 1] * x1,
                   p2 = (double)v[3 * r + 2] * x2;
      acc[r] += ok ? p0 : 0.0;
      acc[r] += ok ? p1 : 0.0;
      acc[r] += ok ? p2 : 0.0;
    }
  };
  // NB chunks in flight: their index loads, the x gathers that depend on
  // them and the 9 NB value loads overlap (measured on the 1.38M-DOF elastic
  // column: FP64 best at 4, 217 against 230 us with 2; FP32 at 3)
  constexpr int NB = sizeof(TV) == 8 ? 4 : 3;
  int c = c0;
  for (; c + NB <= c1; c += NB) {
    int cs[NB];
    bool os[NB];
    TV vs[NB][9];
#pragma unroll
    for (int k = 0; k < NB; ++k) load(c + k, cs[k], os[k], vs[k]);
#pragma unroll
    for (int k = 0; k < NB; ++k) fma3(cs[k], os[k], vs[k]);
  }
  for (; c < c1; ++c) {
    int ca;
    bool oa;
    TV va[9];
    load(c, ca, oa, va);
    fma3(ca, oa, va);
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) out[r] = (T)acc[r];
}

// Calls epi(row, (A x)[row]) for every row, from the row's lane.  Every
// thread of the workgroup returns (no early exit), so callers may follow
// with workgroup-wide reductions.
template <typename T, class IDX, typename TV, class Epi>
__device__ __forceinline__ void sell_rows(int n, int nslices,
                                          const int *__restrict__ cb,
                                          const IDX idx,
                                          const TV *__restrict__ val,
                                          const T *__restrict__ x, Epi &epi) {
  const int nb = (nslices + SELL_WPB - 1) / SELL_WPB;
  const int G = gridDim.x, i = blockIdx.x;
  int first, stride, last;
  if (G >= 8) {
    const int xcd = i & 7, j = i >> 3;
    const int gx = (G >> 3) + (xcd < (G & 7) ? 1 : 0);
    first = (int)((int64_t)nb * xcd / 8) + j;
    last = (int)((int64_t)nb * (xcd + 1) / 8);
    stride = gx;
  } else {
    first = i;
    last = nb;
    stride = G;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int b = first; b < last; b += stride) {
    const int s = b * SELL_WPB + w;
    if (s < nslices) {
      if constexpr (IDX::bs == 3) {
        if constexpr (!std::is_same<TV, int>::value) {  // no fixed-point block view
          const int brow = s * SELL_C + lane;
          T acc[3];
          b3_lane<T, IDX, TV>(idx, val, x, cb[s], cb[s + 1], lane, s * SELL_C, acc);
#pragma unroll
          for (int r = 0; r < 3; ++r)
            if (3 * brow + r < n) epi(3 * brow + r, acc[r]);
        }
      } else {
        const int row = s * SELL_C + lane;
        const T acc = sell_lane<T, IDX, TV>(idx, val, x, cb[s], cb[s + 1], lane, s * SELL_C);
        if (row < n) epi(row, acc);
      }
    }
  }
}

}  // namespace
}  // namespace gs
