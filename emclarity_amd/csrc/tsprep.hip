// Template-search chunk pre-processing (alignment/BH_templateSearch3d_2.m:532-589):
//   chunk = (chunk - mean) / rms;                          (:535-536)
//   chunk = C2R(R2C(chunk) .* bandpass_half) / N;          (:540, fourierTransformer.m:92-158)
//   avg  = BH_movingAverage_2(chunk,   statsRadius);       (:558)
//   avg2 = BH_movingAverage_2(chunk.^2, statsRadius);      (:559)
//   chunk = (chunk - avg) ./ sqrt(avg2 - avg.^2);          (:560-561)
//   chunk = (-shouldBeCTF) .* chunk .* validAreaMask;      (:586)
//   tmp_sum = sum(chunk(validAreaMask > 0.1));             (:589)
//
// BH_movingAverage_2 (statistics/BH_movingAverage_2.m:1-25) pads the volume
// symmetrically by ceil(F/2), convolves with the separable normalised
// Gaussian EMC_gaussianKernel([1,F], 3) (masking/EMC_gaussianKernel.m;
// coordinates -F/2+0.5 .. F/2-0.5) through EMC_convn (testScripts/EMC_convn.m:
// convn(.., 'same') along x, y, z) and crops the padding. Here that is three
// 1-D passes with mirrored indices: MATLAB's 'same' takes the full
// convolution from index floor(F/2)+1, so output i reads input
// i + floor(F/2) - m with tap m (0-based), and the symmetric pad of
// ceil(F/2) covers every such read for the kept outputs (no zero enters).
#include "ops.hpp"

namespace emc {

// Taps of EMC_gaussianKernel([1, F], 3), normalised to sum 1 (FP64, one block).
__global__ void __launch_bounds__(kRedThreads) ma_taps_kernel(int F, float* __restrict__ w) {
  double s[1] = {0.0};
  for (int m = threadIdx.x; m < F; m += blockDim.x) {
    const double c = ((double)m - 0.5 * (F - 1)) / 3.0;
    s[0] += exp(-0.5 * c * c);
  }
  block_sum<1>(s);
  __shared__ double tot;
  if (threadIdx.x == 0) tot = s[0];
  __syncthreads();
  for (int m = threadIdx.x; m < F; m += blockDim.x) {
    const double c = ((double)m - 0.5 * (F - 1)) / 3.0;
    w[m] = (float)(exp(-0.5 * c * c) / tot);
  }
}

// symmetric extension incl. the edge sample (padarray 'symmetric'); valid
// for -N <= i < 2N
__device__ inline int mirror_sym(int i, int N) { return i < 0 ? -i - 1 : (i >= N ? 2 * N - 1 - i : i); }

// One 1-D pass along axis AX. SQ: the input is one channel and the output
// carries (x, x^2) filtered together; otherwise In and Out are the same
// channel type. grid (ceil(nx/256), ny, nz).
template <int AX, bool SQ, class In, class Out>
__global__ void __launch_bounds__(256) ma_pass_kernel(const In* __restrict__ in, Out* __restrict__ out, int nx,
                                                      int ny, int nz, const float* __restrict__ w, int F) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const int N = AX == 0 ? nx : (AX == 1 ? ny : nz);
  const int i = AX == 0 ? x : (AX == 1 ? y : z);
  const int64_t stride = AX == 0 ? 1 : (AX == 1 ? (int64_t)nx : (int64_t)nx * ny);
  const int64_t idx = ((int64_t)z * ny + y) * nx + x;
  const int64_t base = idx - (int64_t)i * stride;
  const int h = F / 2;
  Out acc;
  if constexpr (sizeof(Out) == sizeof(float2)) acc = make_float2(0.0f, 0.0f);
  else acc = 0.0f;
  for (int m = 0; m < F; ++m) {
    const int j = mirror_sym(i + h - m, N);
    const float wm = w[m];
    if constexpr (SQ) {
      const float v = in[base + j * stride];
      acc.x = fmaf(wm, v, acc.x);
      acc.y = fmaf(wm, v * v, acc.y);
    } else if constexpr (sizeof(Out) == sizeof(float2)) {
      const float2 v = in[base + j * stride];
      acc.x = fmaf(wm, v.x, acc.x);
      acc.y = fmaf(wm, v.y, acc.y);
    } else {
      acc = fmaf(wm, in[base + j * stride], acc);
    }
  }
  out[idx] = acc;
}

void check_ma(const int dims[3], int F) {
  EMC_REQUIRE(dims && dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
  EMC_REQUIRE(F >= 2, "feature size must be >= 2 (EMC_gaussianKernel rejects a scalar kernel)");
  for (int a = 0; a < 3; ++a)
    EMC_REQUIRE((F + 1) / 2 <= dims[a], "feature size %d: the symmetric pad ceil(F/2) exceeds dimension %d",
                F, dims[a]);
}

const float* ma_taps(emc_ctx* ctx, int F) {
  float* w = (float*)scratch(ctx, "ma_taps", sizeof(float) * F);
  hipLaunchKernelGGL(ma_taps_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream, F, w);
  EMC_HIP(hipGetLastError());
  return w;
}

// Moving averages of x and x^2 into d_out2 (float2 per voxel).
void moving_average_sq(emc_ctx* ctx, const float* d_in, const int dims[3], int F, float2* d_out2) {
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int64_t nv = (int64_t)nx * ny * nz;
  const float* w = ma_taps(ctx, F);
  float2* t = (float2*)scratch(ctx, "ma_tmp2", sizeof(float2) * nv);
  const dim3 grid(ceil_div(nx, 256), ny, nz), block(256);
  hipLaunchKernelGGL((ma_pass_kernel<0, true, float, float2>), grid, block, 0, ctx->stream, d_in, d_out2, nx, ny,
                     nz, w, F);
  hipLaunchKernelGGL((ma_pass_kernel<1, false, float2, float2>), grid, block, 0, ctx->stream,
                     (const float2*)d_out2, t, nx, ny, nz, w, F);
  hipLaunchKernelGGL((ma_pass_kernel<2, false, float2, float2>), grid, block, 0, ctx->stream,
                     (const float2*)t, d_out2, nx, ny, nz, w, F);
  EMC_HIP(hipGetLastError());
}

// (x - mean) / rms with mean and the centred sum of squares on the device.
__global__ void __launch_bounds__(kRedThreads) centred_sumsq_partial_kernel(const float* __restrict__ d, int64_t n,
                                                                           const double* __restrict__ sums,
                                                                           double* __restrict__ part) {
  const double m = sums[0] / (double)n;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double t = (double)d[i] - m;
    v[0] += t * t;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void standardise_kernel(const float* __restrict__ in, int64_t n, const double* __restrict__ sums,
                                   const double* __restrict__ css, float* __restrict__ out) {
  const double m = sums[0] / (double)n;
  const float inv = (float)(1.0 / sqrt(css[0] / (double)n));
  const float mf = (float)m;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (in[i] - mf) * inv;
}

// Half spectrum times the half-grid bandpass and 1/N (the invFFT(.., 2) scale).
__global__ void mul_half_real_kernel(float2* __restrict__ F, const float* __restrict__ bp, int64_t count,
                                     float scale) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float s = bp[i] * scale;
    F[i] = make_float2(F[i].x * s, F[i].y * s);
  }
}

// out = sign * mask * (x - a) / sqrt(a2 - a^2); per-block partial sums of out
// over mask > 0.1 (a null mask is 1 everywhere).
__global__ void __launch_bounds__(kRedThreads) local_norm_kernel(const float* __restrict__ x,
                                                                const float2* __restrict__ a, int64_t n,
                                                                float sign, const float* __restrict__ mask,
                                                                float* __restrict__ out, double* __restrict__ part) {
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 s = a[i];
    const float mk = mask ? mask[i] : 1.0f;
    const float r = sign * ((x[i] - s.x) / sqrtf(s.y - s.x * s.x)) * mk;
    out[i] = r;
    if (mk > 0.1f) v[0] += (double)r;
  }
  block_sum<1>(v);
  if (part && threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void scale_kernel(float* __restrict__ d, int64_t n, float s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] *= s;
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

// medfilt3(x, [K K K]) with its default 'symmetric' padding: index -1 reads
// 0, index n reads n - 1 (every dimension >= K/2 + 1, so one reflection
// stays inside). The K^3 window goes to a private array and the median (the
// element of rank K^3/2) is selected by Hoare-partition quickselect; every
// partition step moves both ends by at least one, and both scans are
// bounded by the current range.
__device__ inline int reflect_index(int i, int n) { return i < 0 ? -i - 1 : (i >= n ? 2 * n - 1 - i : i); }

template <int K>
__global__ void __launch_bounds__(256) median3_kernel(const float* __restrict__ in, float* __restrict__ out, int nx,
                                                      int ny, int nz) {
  constexpr int M = K * K * K, H = K / 2;
  const int64_t nv = (int64_t)nx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    split3(i, nx, ny, x, y, z);
    float v[M];
    int c = 0;
    for (int dz = -H; dz <= H; ++dz) {
      const int zz = reflect_index(z + dz, nz);
      for (int dy = -H; dy <= H; ++dy) {
        const int64_t row = ((int64_t)zz * ny + reflect_index(y + dy, ny)) * nx;
        for (int dx = -H; dx <= H; ++dx) v[c++] = in[row + reflect_index(x + dx, nx)];
      }
    }
    constexpr int k = M / 2;
    int lo = 0, hi = M - 1;
    for (int guard = 0; lo < hi && guard < M; ++guard) {
      const float pivot = v[(lo + hi) / 2];
      int a = lo, b = hi;
      while (a <= b) {
        while (a <= hi && v[a] < pivot) ++a;
        while (b >= lo && v[b] > pivot) --b;
        if (a <= b) {
          const float t = v[a];
          v[a] = v[b];
          v[b] = t;
          ++a;
          --b;
        }
      }
      if (k <= b)
        hi = b;
      else if (k >= a)
        lo = a;
      else
        break;
    }
    out[i] = v[k];
  }
}

// K = 3 by forgetful selection, in registers: the median of 27 is the
// rank-13 element. Start with a working set of the first 15 (27/2 + 2)
// window values; 13 times, move the set's minimum to its low end and its
// maximum to its high end, drop the minimum, and put the next window value
// in the maximum's place (no more values after 12 steps: the last step just
// leaves the middle one): the forgetful selection of M. McGuire, "A fast,
// small-radius GPU median filter" (ShaderX6, 2008). It returns one of the
// inputs, the same value as a full sort's middle element. All indices are
// compile-time, so the 27 values stay in registers.
__device__ inline void cswap(float& a, float& b) {
  const float lo = fminf(a, b), hi = fmaxf(a, b);
  a = lo;
  b = hi;
}

__global__ void __launch_bounds__(256) median27_kernel(const float* __restrict__ in, float* __restrict__ out, int nx,
                                                       int ny, int nz) {
  const int64_t nv = (int64_t)nx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    split3(i, nx, ny, x, y, z);
    float v[27];
    const int xs[3] = {reflect_index(x - 1, nx), x, reflect_index(x + 1, nx)};
#pragma unroll
    for (int dz = 0; dz < 3; ++dz) {
      const int zz = reflect_index(z + dz - 1, nz);
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const int64_t row = ((int64_t)zz * ny + reflect_index(y + dy - 1, ny)) * nx;
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) v[(dz * 3 + dy) * 3 + dx] = in[row + xs[dx]];
      }
    }
#pragma unroll
    for (int it = 0; it < 13; ++it) {  // working set v[it .. 14]
      cswap(v[it], v[14]);
#pragma unroll
      for (int j = it + 1; j < 14; ++j) {
        cswap(v[it], v[j]);
        cswap(v[j], v[14]);
      }
      if (it < 12) v[14] = v[15 + it];
    }
    out[i] = v[13];
  }
}

// K = 5: the same selection over a working set of 125/2 + 2 = 64 registers;
// the other 61 window values are read one per step, as they enter the set.
__global__ void __launch_bounds__(256) median125_kernel(const float* __restrict__ in, float* __restrict__ out, int nx,
                                                        int ny, int nz) {
  constexpr int K = 5, M = K * K * K, R = M / 2 + 2;
  const int64_t nv = (int64_t)nx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    split3(i, nx, ny, x, y, z);
    int xs[K], ys[K];
    int64_t zs[K];
#pragma unroll
    for (int d = 0; d < K; ++d) {
      xs[d] = reflect_index(x + d - K / 2, nx);
      ys[d] = reflect_index(y + d - K / 2, ny);
      zs[d] = (int64_t)reflect_index(z + d - K / 2, nz) * ny;
    }
    auto at = [&](int e) { return in[(zs[e / (K * K)] + ys[(e / K) % K]) * nx + xs[e % K]]; };
    float v[R];
#pragma unroll
    for (int e = 0; e < R; ++e) v[e] = at(e);
#pragma unroll
    for (int it = 0; it < M / 2; ++it) {  // working set v[it .. R - 1]
      cswap(v[it], v[R - 1]);
#pragma unroll
      for (int j = it + 1; j < R - 1; ++j) {
        cswap(v[it], v[j]);
        cswap(v[j], v[R - 1]);
      }
      if (R + it < M) v[R - 1] = at(R + it);
    }
    out[i] = v[M / 2];
  }
}

void median3(emc_ctx* ctx, const float* in, float* out, const int dims[3], int k) {
  const int64_t nv = (int64_t)dims[0] * dims[1] * dims[2];
  const int grid = grid_for(nv);
  if (k == 3)
    hipLaunchKernelGGL(median27_kernel, dim3(grid), dim3(256), 0, ctx->stream, in, out, dims[0], dims[1], dims[2]);
  else if (k == 5)
    hipLaunchKernelGGL(median125_kernel, dim3(grid), dim3(256), 0, ctx->stream, in, out, dims[0], dims[1], dims[2]);
  else
    hipLaunchKernelGGL(median3_kernel<7>, dim3(grid), dim3(256), 0, ctx->stream, in, out, dims[0], dims[1], dims[2]);
  EMC_HIP(hipGetLastError());
}

void check_median(const int dims[3], int k) {
  EMC_REQUIRE(k == 3 || k == 5 || k == 7, "median filter size can only be 3, 5 or 7, not %d", k);
  for (int a = 0; a < 3; ++a)
    EMC_REQUIRE(dims[a] > k / 2, "dimension %d (%d) too small for a %d^3 median", a, dims[a], k);
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_moving_average(emc_ctx* ctx, const float* d_in, const int dims[3], int feature_size,
                              float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_in && d_out, "null pointer");
    check_ma(dims, feature_size);
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t nv = (int64_t)nx * ny * nz;
    const float* w = ma_taps(ctx, feature_size);
    float* t1 = (float*)scratch(ctx, "ma_tmp1", sizeof(float) * nv);
    float* t2 = (float*)scratch(ctx, "ma_tmp1b", sizeof(float) * nv);
    const dim3 grid(ceil_div(nx, 256), ny, nz), block(256);
    hipLaunchKernelGGL((ma_pass_kernel<0, false, float, float>), grid, block, 0, ctx->stream, d_in, t1, nx, ny, nz,
                       w, feature_size);
    hipLaunchKernelGGL((ma_pass_kernel<1, false, float, float>), grid, block, 0, ctx->stream, (const float*)t1, t2,
                       nx, ny, nz, w, feature_size);
    hipLaunchKernelGGL((ma_pass_kernel<2, false, float, float>), grid, block, 0, ctx->stream, (const float*)t2,
                       d_out, nx, ny, nz, w, feature_size);
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_local_normalize(emc_ctx* ctx, const float* d_in, const int dims[3], int feature_size,
                               float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_in && d_out, "null pointer");
    check_ma(dims, feature_size);
    const int64_t nv = (int64_t)dims[0] * dims[1] * dims[2];
    float2* a = (float2*)scratch(ctx, "ma_out2", sizeof(float2) * nv);
    moving_average_sq(ctx, d_in, dims, feature_size, a);
    hipLaunchKernelGGL(local_norm_kernel, dim3(grid_for(nv)), dim3(kRedThreads), 0, ctx->stream, d_in,
                       (const float2*)a, nv, 1.0f, (const float*)nullptr, d_out, (double*)nullptr);
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_median_filter3d(emc_ctx* ctx, const float* d_in, const int dims[3], int k, float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_in && d_out && dims, "null pointer");
    EMC_REQUIRE(d_in != d_out, "d_out must not alias d_in");
    EMC_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    check_median(dims, k);
    median3(ctx, d_in, d_out, dims, k);
  });
}

emc_status emc_tsearch_preprocess_chunk(emc_ctx* ctx, const float* d_in, const int dims[3], const double bp[4],
                                        int med_filt, int stats_radius, float sign, const float* d_valid_mask,
                                        float* d_out, double* h_masked_sum) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_in && d_out && bp, "null pointer");
    check_ma(dims, stats_radius);
    if (med_filt) check_median(dims, med_filt);
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t nv = (int64_t)nx * ny * nz, nh = (int64_t)(nx / 2 + 1) * ny * nz;
    // (x - mean) / rms, two FP64 passes
    double* st = (double*)scratch(ctx, "tsp_stats", sizeof(double) * 4);
    reduce_sum_sumsq(ctx, d_in, nv, st);
    double* part = (double*)scratch(ctx, "tsp_part", sizeof(double) * kRedBlocks);
    hipLaunchKernelGGL(centred_sumsq_partial_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream, d_in, nv,
                       (const double*)st, part);
    hipLaunchKernelGGL(fold_partials_kernel<1>, dim3(1), dim3(kRedThreads), 0, ctx->stream, (const double*)part,
                       kRedBlocks, st + 2);
    float* y = (float*)scratch(ctx, "tsp_y", sizeof(float) * nv);
    hipLaunchKernelGGL(standardise_kernel, dim3(grid_for(nv)), dim3(256), 0, ctx->stream, d_in, nv,
                       (const double*)st, (const double*)(st + 2), y);
    // bandpass: R2C, x bandpass half grid / N, C2R
    float2* F = (float2*)scratch(ctx, "tsp_F", sizeof(float2) * nh);
    float* bph = (float*)scratch(ctx, "tsp_bp", sizeof(float) * nh);
    bandpass_impl(ctx, dims, 1, bp[0], bp[1], bp[2], bp[3], bph);
    emc_fft_plan* plan = cached_plan(ctx, nx, ny, nz, 1);
    fft_r2c(plan, y, F);
    hipLaunchKernelGGL(mul_half_real_kernel, dim3(grid_for(nh)), dim3(256), 0, ctx->stream, F,
                       (const float*)bph, nh, (float)(1.0 / (double)nv));
    fft_c2r(plan, F, y);
    if (med_filt) {  // medfilt3 after the bandpass (:543-548)
      float* m = (float*)scratch(ctx, "tsp_med", sizeof(float) * nv);
      median3(ctx, y, m, dims, med_filt);
      y = m;
    }
    // local normalisation, sign, mask, masked sum
    float2* a = (float2*)scratch(ctx, "ma_out2", sizeof(float2) * nv);
    moving_average_sq(ctx, y, dims, stats_radius, a);
    const int nb = kRedBlocks;
    hipLaunchKernelGGL(local_norm_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, (const float*)y,
                       (const float2*)a, nv, sign, d_valid_mask, d_out, part);
    hipLaunchKernelGGL(fold_partials_kernel<1>, dim3(1), dim3(kRedThreads), 0, ctx->stream, (const double*)part,
                       nb, st + 3);
    EMC_HIP(hipGetLastError());
    if (h_masked_sum) {
      double* h = (double*)pinned(ctx, sizeof(double));
      EMC_HIP(hipMemcpyAsync(h, st + 3, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      EMC_HIP(hipStreamSynchronize(ctx->stream));
      *h_masked_sum = *h;
    }
  });
}

emc_status emc_scale(emc_ctx* ctx, float* d, int64_t count, double s) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d || count == 0, "null pointer");
    EMC_REQUIRE(count >= 0, "negative count");
    if (count == 0) return;
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(count)), dim3(256), 0, ctx->stream, d, count, (float)s);
    EMC_HIP(hipGetLastError());
  });
}

}  // extern "C"
