// Deterministic device reductions (FP64 accumulation): a fixed grid of
// kRedBlocks blocks writes per-block partials, one block folds them in a fixed
// order. Used for means, variances (std), minima and argmax.
#include "emc_internal.hpp"

namespace emc {

__global__ void __launch_bounds__(kRedThreads) sum_sumsq_partial_kernel(const float* __restrict__ d,
                                                                        int64_t n,
                                                                        double* __restrict__ part) {
  double v[2] = {0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double x = (double)d[i];
    v[0] += x;
    v[1] += x * x;
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = v[0];
    part[2 * blockIdx.x + 1] = v[1];
  }
}

// blockIdx.y: volume of a batch, bstride floats apart; partials of volume b
// at part[b * gridDim.x ..]
__global__ void __launch_bounds__(kRedThreads) min_partial_kernel(const float* __restrict__ d,
                                                                  int64_t n, int stride, int64_t bstride,
                                                                  float* __restrict__ part) {
  d += (int64_t)blockIdx.y * bstride;
  part += (int64_t)blockIdx.y * gridDim.x;
  float m = __builtin_inff();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    m = fminf(m, d[i * stride]);
  m = block_min(m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// block b folds the nb partials of volume b into out[b]
__global__ void __launch_bounds__(kRedThreads) fold_min_kernel(const float* __restrict__ part, int nb,
                                                               double* __restrict__ out) {
  part += (int64_t)blockIdx.x * nb;
  out += blockIdx.x;
  float m = __builtin_inff();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fminf(m, part[b]);
  m = block_min(m);
  if (threadIdx.x == 0) out[0] = (double)m;
}

void reduce_sum_sumsq(emc_ctx* ctx, const float* d, int64_t n, double* d_out2) {
  double* part = (double*)scratch(ctx, "red_part", sizeof(double) * 2 * kRedBlocks);
  hipLaunchKernelGGL(sum_sumsq_partial_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream,
                     d, n, part);
  hipLaunchKernelGGL(fold_partials_kernel<2>, dim3(1), dim3(kRedThreads), 0, ctx->stream,
                     (const double*)part, kRedBlocks, d_out2);
  EMC_HIP(hipGetLastError());
}

void reduce_min(emc_ctx* ctx, const float* d, int64_t n, int stride, double* d_out1) {
  reduce_min_batched(ctx, d, n, stride, 0, 1, d_out1);
}

void reduce_min_batched(emc_ctx* ctx, const float* d, int64_t n, int stride, int64_t bstride, int batch,
                        double* d_out) {
  float* part = (float*)scratch(ctx, "red_min", sizeof(float) * kRedBlocks * (size_t)batch);
  hipLaunchKernelGGL(min_partial_kernel, dim3(kRedBlocks, batch), dim3(kRedThreads), 0, ctx->stream, d, n,
                     stride, bstride, part);
  hipLaunchKernelGGL(fold_min_kernel, dim3(batch), dim3(kRedThreads), 0, ctx->stream,
                     (const float*)part, kRedBlocks, d_out);
  EMC_HIP(hipGetLastError());
}

}  // namespace emc
