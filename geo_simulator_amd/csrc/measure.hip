// Achievable HBM rates of this device, for the roofline figures: a
// read-dominant streaming sum (16-byte loads, one FP64 partial per
// workgroup) and a streaming copy over a buffer of the caller's size, timed
// with HIP events on the context's stream.  Buffers larger than the 256 MB
// Infinity Cache measure HBM itself.
#include <algorithm>

#include "internal.hpp"

namespace gs {
namespace {

constexpr int SBS = 512;

typedef double d2 __attribute__((ext_vector_type(2)));

// sum of x over [0, n2) pairs; grid-stride, four 16-byte loads in flight per
// thread; one partial per workgroup
__global__ __launch_bounds__(SBS) void k_stream_sum(int64_t n2, const d2 *__restrict__ x,
                                                    double *__restrict__ part) {
  double acc = 0.0;
  const int64_t stride = (int64_t)gridDim.x * SBS;
  int64_t i = (int64_t)blockIdx.x * SBS + threadIdx.x;
  for (; i + 3 * stride < n2; i += 4 * stride) {
    const d2 a = x[i], b = x[i + stride], c = x[i + 2 * stride], d = x[i + 3 * stride];
    acc += (a.x + a.y) + (b.x + b.y) + (c.x + c.y) + (d.x + d.y);
  }
  for (; i < n2; i += stride) {
    const d2 a = x[i];
    acc += a.x + a.y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ double sh[SBS / 64];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < SBS / 64; ++w) s += sh[w];
    part[blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(SBS) void k_stream_copy(int64_t n2, const d2 *__restrict__ x,
                                                     d2 *__restrict__ y) {
  const int64_t stride = (int64_t)gridDim.x * SBS;
  for (int64_t i = (int64_t)blockIdx.x * SBS + threadIdx.x; i < n2; i += stride) y[i] = x[i];
}

__global__ void k_stream_fill(int64_t n2, d2 *__restrict__ x) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride)
    x[i] = (d2){1.0, 0.5};
}

}  // namespace

double time_stream(gs_ctx *ctx, size_t bytes, int kind, int reps) {
  hipStream_t s = ctx->stream;
  const int64_t n2 = (int64_t)(bytes / 16);
  if (n2 <= 0) throw Error(GS_ERR_INVALID_ARG, "stream size below 16 bytes");
  int cus = 0, dev = 0;
  GS_HIP(hipGetDevice(&dev));
  GS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)std::max(1, cus * 8);  // persistent: 8 workgroups per CU
  auto x = palloc<d2>(ctx->pool, n2);
  auto y = palloc<d2>(ctx->pool, kind == 1 ? n2 : 1);
  auto part = palloc<double>(ctx->pool, grid);
  hipLaunchKernelGGL(k_stream_fill, 4096, 256, 0, s, n2, x.get());
  auto once = [&] {
    if (kind == 1)
      hipLaunchKernelGGL(k_stream_copy, grid, SBS, 0, s, n2, x.get(), y.get());
    else
      hipLaunchKernelGGL(k_stream_sum, grid, SBS, 0, s, n2, x.get(), part.get());
  };
  once();
  hipEvent_t e0, e1;
  GS_HIP(hipEventCreate(&e0));
  GS_HIP(hipEventCreate(&e1));
  GS_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < reps; ++i) once();
  GS_HIP(hipEventRecord(e1, s));
  GS_LAUNCH_CHECK("k_stream");
  GS_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  GS_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

}  // namespace gs
