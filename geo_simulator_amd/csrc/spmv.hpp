// CSR SpMV engines for gfx950 (included by solver.hip).
//
// Column indices come through a reader: Cols32 (plain int32) or Cols16, where
// column = anchor[k >> 10] + off[k] with a 16-bit offset and one int32 anchor
// (the row of nonzero k & ~1023) per 1024 nonzeros.  Under the mesh's
// locality numbering every offset of the Tet10 operators fits in 16 bits, so
// the index stream shrinks from 4 to 2 bytes per nonzero; a matrix whose
// offsets do not fit keeps int32 indices.
#pragma once

#include <type_traits>

namespace gs {
namespace {

constexpr int SPMV_BS = 256;
constexpr int ANCHOR_SHIFT = 10;  // one anchor per 1024 nonzeros

struct Cols32 {
  const int *__restrict__ ci;
  __device__ __forceinline__ int operator()(int k) const { return ci[k]; }
};

struct Cols16 {
  const short *__restrict__ off;
  const int *__restrict__ anchor;
  __device__ __forceinline__ int operator()(int k) const {
    return anchor[k >> ANCHOR_SHIFT] + (int)off[k];
  }
};

template <typename T> struct StreamChunk { static constexpr int value = 2048; };

// Both engines call epi(row, (A x)[row]) once per row from one thread.
//
// Stream engine: a workgroup takes a block of R = 256/S consecutive rows; all
// 256 threads load the block's (value, column) pairs in contiguous chunks
// (fully coalesced, CH/256 independent gathers of x in flight per thread) and
// park the products in LDS; then S threads per row sum the row's products.
// Row blocks are visited grid-stride, so the workgroups resident at any
// moment work on nearby rows and share the gathered part of x in L2.
template <int S, typename T, class C, class Epi>
__device__ __forceinline__ void spmv_stream(int n, const int *__restrict__ rp,
                                            const C cols,
                                            const T *__restrict__ val,
                                            const T *__restrict__ x, Epi &epi) {
  constexpr int BS = SPMV_BS;
  constexpr int R = BS / S;
  constexpr int CH = StreamChunk<T>::value;
  __shared__ T sprod[CH];
  __shared__ int srp[R + 1];
  const int t = threadIdx.x;
  const int lr = t / S, sub = t - (t / S) * S;
  const int nblk = (n + R - 1) / R;
  for (int blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int r0 = blk * R;
    const int nr = min(R, n - r0);
    for (int i = t; i <= nr; i += BS) srp[i] = rp[r0 + i];
    __syncthreads();
    const int k0 = srp[0], k1 = srp[nr];
    const bool own = lr < nr;
    const int a = own ? srp[lr] : 0, b = own ? srp[lr + 1] : 0;
    T acc = (T)0;
    for (int base = k0; base < k1; base += CH) {
      const int m = min(CH, k1 - base);
      int cj[CH / BS];
      T vj[CH / BS];
#pragma unroll
      for (int i = 0; i < CH / BS; ++i) {
        const int j = t + i * BS;
        cj[i] = j < m ? cols(base + j) : 0;
        vj[i] = j < m ? val[base + j] : (T)0;
      }
#pragma unroll
      for (int i = 0; i < CH / BS; ++i) {
        const int j = t + i * BS;
        if (j < m) sprod[j] = vj[i] * x[cj[i]];
      }
      __syncthreads();
      const int lo = max(a, base) - base, hi = min(b, base + m) - base;
      for (int j = lo + sub; j < hi; j += S) acc += sprod[j];
      __syncthreads();
    }
#pragma unroll
    for (int off = S / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, S);
    if (own && sub == 0) epi(r0 + lr, acc);
    __syncthreads();
  }
}

// Vector engine: L lanes per row, rows visited grid-stride; each lane issues
// up to four (column, value) loads before consuming them.
template <int L, typename T, class C, class Epi>
__device__ __forceinline__ void spmv_vector(int n, const int *__restrict__ rp,
                                            const C cols,
                                            const T *__restrict__ val,
                                            const T *__restrict__ x, Epi &epi) {
  constexpr int BS = SPMV_BS;
  const int lane = threadIdx.x & (L - 1);
  const int ng = (gridDim.x * BS) / L;
  for (int row = (blockIdx.x * BS + threadIdx.x) / L; row < n; row += ng) {
    const int b = rp[row], e = rp[row + 1];
    T s = (T)0;
    int k = b + lane;
    for (; k + 3 * L < e; k += 4 * L) {
      const int c0 = cols(k), c1 = cols(k + L), c2 = cols(k + 2 * L),
                c3 = cols(k + 3 * L);
      const T v0 = val[k], v1 = val[k + L], v2 = val[k + 2 * L],
              v3 = val[k + 3 * L];
      s += v0 * x[c0];
      s += v1 * x[c1];
      s += v2 * x[c2];
      s += v3 * x[c3];
    }
    for (; k < e; k += L) s += val[k] * x[cols(k)];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, L);
    if (lane == 0) epi(row, s);
  }
}

// ENG = 0: stream with S = W threads per row; ENG = 1: vector with L = W lanes.
template <int ENG, int W, typename T, class C, class Epi>
__device__ __forceinline__ void spmv_rows(int n, const int *__restrict__ rp,
                                          const C cols,
                                          const T *__restrict__ val,
                                          const T *__restrict__ x, Epi &epi) {
  if constexpr (ENG == 0) spmv_stream<W, T, C>(n, rp, cols, val, x, epi);
  else spmv_vector<W, T, C>(n, rp, cols, val, x, epi);
}

// Launch shape: engine, threads (stream) or lanes (vector) per row,
// workgroups, index width.
struct Shape {
  int eng = 0, W = 2, grid = 1;
  bool idx16 = false;
};

template <int V> using IC = std::integral_constant<int, V>;

// Calls f(IC<ENG>, IC<W>) for the shape's engine and width.
template <class F> void dispatch_shape(const Shape &sh, F &&f) {
  if (sh.eng == 0) {
    switch (sh.W) {
    case 1: f(IC<0>{}, IC<1>{}); break;
    case 2: f(IC<0>{}, IC<2>{}); break;
    case 4: f(IC<0>{}, IC<4>{}); break;
    default: f(IC<0>{}, IC<8>{}); break;
    }
  } else {
    switch (sh.W) {
    case 4: f(IC<1>{}, IC<4>{}); break;
    case 8: f(IC<1>{}, IC<8>{}); break;
    case 16: f(IC<1>{}, IC<16>{}); break;
    default: f(IC<1>{}, IC<32>{}); break;
    }
  }
}

}  // namespace
}  // namespace gs
