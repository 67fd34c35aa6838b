// Cross-correlation peak search and CC normalisation: the
// BH_multi_xcf_Translational / BH_multi_xcf_Rotational replacements
// (reference: statistics/BH_multi_xcf_Translational.m:31-103,
// statistics/BH_multi_xcf_Rotational.m:29-80).
//
// The reference fftshifts the map, overwrites masked voxels with the map
// minimum, takes MATLAB max (first column-major index on ties), then two
// centre-of-mass passes over a (2c+1)^3 box. Here the map stays in FFT order;
// the kernels compute each voxel's fftshifted linear index on the fly and
// reduce a 64-bit key (orderable value, inverted index) so the tie-break is
// exactly "first index of the shifted map". No shifted copy is written.
#include <cmath>

#include "ops.hpp"

namespace emc {

// argmax over the masked, fftshifted map -> per-block keys. blockIdx.y: map
// of a batch (map_bs / mask_bs floats apart, its own minimum d_min[b] and
// partials part[b * gridDim.x ..])
__global__ void __launch_bounds__(kRedThreads) peak_argmax_partial_kernel(
    const float* __restrict__ map, int stride, int nx, int ny, int nz, const float* __restrict__ mask,
    const double* __restrict__ d_min, unsigned long long* __restrict__ part, int64_t map_bs, int64_t mask_bs) {
  const int bb = blockIdx.y;
  map += bb * map_bs;
  if (mask) mask += bb * mask_bs;
  d_min += bb;
  part += (int64_t)bb * gridDim.x;
  const int64_t n = (int64_t)nx * ny * nz;
  const float mn = (float)d_min[0];
  unsigned long long best = 0ull;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int x = (int)(j % nx);
    const int64_t yz = j / nx;
    const int y = (int)(yz % ny), z = (int)(yz / ny);
    const int sx = shifted_index(x, nx), sy = shifted_index(y, ny), sz = shifted_index(z, nz);
    const uint32_t L = (uint32_t)(((int64_t)sz * ny + sy) * nx + sx);
    float v = map[j * stride];
    if (mask && mask[L] < 0.95f) v = mn;
    const unsigned long long k = argmax_key(v, L);
    best = k > best ? k : best;
  }
  best = block_max_u64(best);
  if (threadIdx.x == 0) part[blockIdx.x] = best;
}

__device__ inline double matlab_round(double x) { return x >= 0 ? floor(x + 0.5) : ceil(x - 0.5); }

// Fold the keys, then the two COM passes (:47-91) and the origin shift
// (:97), all in one block; block b takes map b of a batch (offsets as in
// peak_argmax_partial_kernel, result res[b]).
__global__ void __launch_bounds__(kRedThreads) peak_com_kernel(
    const float* __restrict__ map, int stride, int nx, int ny, int nz, const float* __restrict__ mask,
    const double* __restrict__ d_min, const unsigned long long* __restrict__ part, int nparts, int cx,
    int cy, int cz, PeakResult* __restrict__ res, const double* __restrict__ d_add, int64_t map_bs,
    int64_t mask_bs) {
  const int bb = blockIdx.x;
  map += bb * map_bs;
  if (mask) mask += bb * mask_bs;
  d_min += bb;
  part += (int64_t)bb * nparts;
  res += bb;
  unsigned long long best = 0ull;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) best = part[b] > best ? part[b] : best;
  best = block_max_u64(best);
  __shared__ unsigned long long s_best;
  if (threadIdx.x == 0) s_best = best;
  __syncthreads();
  best = s_best;
  const float mn = (float)d_min[0];
  const float maxv = key_value(best);
  const uint32_t L = key_index(best);
  double peak[3] = {(double)(L % nx) + 1.0, (double)((L / nx) % ny) + 1.0,
                    (double)(L / ((uint32_t)nx * ny)) + 1.0};  // 1-based (:41-42)
  int status = 0;
  const bool ok = !(maxv != maxv) && maxv != 0.0f;  // :45
  const int nn[3] = {nx, ny, nz};
  const int com[3] = {cx, cy, cz};
  if (ok && cx != 0 && cy != 0 && cz != 0) {
    for (int icom = 1; icom <= 2; ++icom) {
      double pc[3];
      for (int k = 0; k < 3; ++k) pc[k] = icom == 1 ? peak[k] : matlab_round(peak[k]);
      int lo[3], hi[3];
      bool inside = true;
      for (int k = 0; k < 3; ++k) {
        lo[k] = (int)(pc[k] - com[k]);
        hi[k] = (int)(pc[k] + com[k]);
        if (lo[k] < 1 || hi[k] > nn[k]) inside = false;
      }
      if (!inside) {
        status = 1;
        break;
      }
      const int bx = 2 * cx + 1, by = 2 * cy + 1, bz = 2 * cz + 1;
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      for (int i = threadIdx.x; i < bx * by * bz; i += blockDim.x) {
        const int ox = i % bx, oy = (i / bx) % by, oz = i / (bx * by);
        const int s[3] = {lo[0] - 1 + ox, lo[1] - 1 + oy, lo[2] - 1 + oz};  // 0-based shifted
        const int64_t Ls = ((int64_t)s[2] * ny + s[1]) * nx + s[0];
        // shifted -> FFT order: j = (s - n/2) mod n
        const int jx = (s[0] + nx - nx / 2) % nx, jy = (s[1] + ny - ny / 2) % ny,
                  jz = (s[2] + nz - nz / 2) % nz;
        float m = map[(((int64_t)jz * ny + jy) * nx + jx) * stride];
        if (mask && mask[Ls] < 0.95f) m = mn;
        const double b = (double)(m - mn);  // :70 box - minVal
        v[0] += b;
        v[1] += b * (double)(ox - cx);
        v[2] += b * (double)(oy - cy);
        v[3] += b * (double)(oz - cz);
      }
      block_sum<4>(v);
      __shared__ double s_cm[3];
      if (threadIdx.x == 0) {
        s_cm[0] = v[1] / v[0];
        s_cm[1] = v[2] / v[0];
        s_cm[2] = v[3] / v[0];
      }
      __syncthreads();
      for (int k = 0; k < 3; ++k) peak[k] = pc[k] + s_cm[k];  // :88
      __syncthreads();
    }
  } else if (!ok) {
    peak[0] = peak[1] = peak[2] = 0.0;  // :95
  }
  if (threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) {
      double p = peak[k] - ceil(((double)nn[k] + 1.0) / 2.0);  // :97
      if (p != p) p = 0.0;                                       // :99-103
      res->peak[k] = p + (d_add ? d_add[k] : 0.0);
    }
    res->max_val = (double)maxv;
    res->status = status;
  }
}

void peak_search(emc_ctx* ctx, const float* map, int stride, const int dims[3], const float* mask,
                 const int com[3], PeakResult* d_res, const double* d_add) {
  peak_search_batched(ctx, map, stride, dims, mask, com, d_res, d_add, 1, 0, 0);
}

void peak_search_batched(emc_ctx* ctx, const float* map, int stride, const int dims[3], const float* mask,
                         const int com[3], PeakResult* d_res, const double* d_add, int batch, int64_t map_bs,
                         int64_t mask_bs) {
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  EMC_REQUIRE(n < ((int64_t)1 << 32), "map too large for 32-bit peak index");
  EMC_REQUIRE(batch >= 1 && batch <= 65535, "peak search batch %d out of range", batch);
  double* d_min = (double*)scratch(ctx, "peak_min", sizeof(double) * (size_t)batch);
  unsigned long long* part =
      (unsigned long long*)scratch(ctx, "peak_keys", sizeof(unsigned long long) * kRedBlocks * (size_t)batch);
  reduce_min_batched(ctx, map, n, stride, map_bs, batch, d_min);  // :33 min over the whole map
  hipLaunchKernelGGL(peak_argmax_partial_kernel, dim3(kRedBlocks, batch), dim3(kRedThreads), 0, ctx->stream,
                     map, stride, dims[0], dims[1], dims[2], mask, (const double*)d_min, part, map_bs, mask_bs);
  hipLaunchKernelGGL(peak_com_kernel, dim3(batch), dim3(kRedThreads), 0, ctx->stream, map, stride, dims[0],
                     dims[1], dims[2], mask, (const double*)d_min, (const unsigned long long*)part,
                     kRedBlocks, com ? com[0] : 0, com ? com[1] : 0, com ? com[2] : 0, d_res, d_add, map_bs,
                     mask_bs);
  EMC_HIP(hipGetLastError());
}

// X = A .* B / n  (full complex grids)
__global__ void cmul_scale_kernel(const float2* __restrict__ a, const float2* __restrict__ b,
                                  float2* __restrict__ x, int64_t n, float s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 u = a[i], v = b[i];
    x[i] = make_float2((u.x * v.x - u.y * v.y) * s, (u.x * v.y + u.y * v.x) * s);
  }
}

// Full-grid sums of BH_multi_xcf_Rotational (:73-78)
__global__ void __launch_bounds__(kRedThreads) xcf_rot_full_kernel(
    const float2* __restrict__ P, const float2* __restrict__ R, const float* __restrict__ wp,
    const float* __restrict__ wr, int64_t n, double* __restrict__ part) {
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 p = P[i], r = R[i];
    const double a = wp ? (double)wp[i] : 1.0;
    const double c = wr ? (double)wr[i] : 1.0;
    const double re = (double)p.x * r.x - (double)p.y * r.y;  // Re(P R)
    v[0] += c * a * re;
    v[1] += ((double)r.x * r.x + (double)r.y * r.y) * a * a;
    v[2] += ((double)p.x * p.x + (double)p.y * p.y) * c * c;
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = v[k];
}

// Shell-weighted form: block (b, s) reduces shell s's voxel list with
// stride gridDim.x * blockDim.x (fixed partition: deterministic sums)

__global__ void __launch_bounds__(kRedThreads) xcf_rot_shell_kernel(
    const float2* __restrict__ P, const float2* __restrict__ R, const float* __restrict__ wp,
    const float* __restrict__ wr, const int64_t* __restrict__ offsets, const uint32_t* __restrict__ idx,
    double* __restrict__ part) {
  const int s = blockIdx.y;
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t j = offsets[s] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < offsets[s + 1];
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = (int64_t)idx[j] - 1;  // MATLAB linear index
    const float2 p = P[i], r = R[i];
    const double a = wp ? (double)wp[i] : 1.0;
    const double c = wr ? (double)wr[i] : 1.0;
    const double re = (double)p.x * r.x - (double)p.y * r.y;
    v[0] += c * a * re;
    v[1] += ((double)r.x * r.x + (double)r.y * r.y) * a * a;
    v[2] += ((double)p.x * p.x + (double)p.y * p.y) * c * c;
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[((int64_t)s * gridDim.x + blockIdx.x) * 3 + k] = v[k];
}

// Per shell: CCC_s = sum num / sqrt(sum d1 sum d2), then
// sum over shells (in order) of (CCC_s^2 / (CCC_s + w^-1.5))^3 (abs if asked)
// (statistics/BH_multi_xcf_Rotational.m:53-70)
__global__ void __launch_bounds__(kWave) xcf_rot_shell_fold_kernel(const double* __restrict__ part, int n_shells,
                                                                  int nb, int abs_flag, double w,
                                                                  double* __restrict__ out) {
  __shared__ double cc[64];
  for (int s0 = 0; s0 < n_shells; s0 += 64) {
    const int ns = min(64, n_shells - s0);
    if ((int)threadIdx.x < ns) {
      double v[3] = {0.0, 0.0, 0.0};
      const double* p = part + (int64_t)(s0 + threadIdx.x) * nb * 3;
      for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) v[k] += p[3 * b + k];
      cc[threadIdx.x] = v[0] / sqrt(v[1] * v[2]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double acc = s0 == 0 ? 0.0 : out[0];
      const double k = pow(w, -1.5);
      for (int s = 0; s < ns; ++s) {
        const double c = cc[s];
        double t = c * c / (c + k);
        t = t * t * t;
        acc += abs_flag ? fabs(t) : t;
      }
      out[0] = acc;
    }
    __syncthreads();
  }
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_shell_plan_create(emc_ctx* ctx, int abs_flag, double w, int n_shells, const int64_t* h_offsets,
                                 const uint32_t* h_indices, emc_shell_plan** out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(out && h_offsets && n_shells > 0, "null pointer or no shells");
    EMC_REQUIRE(h_offsets[0] == 0, "offsets must start at 0");
    for (int s = 0; s < n_shells; ++s) EMC_REQUIRE(h_offsets[s + 1] >= h_offsets[s], "offsets must not decrease");
    const int64_t total = h_offsets[n_shells];
    EMC_REQUIRE(total == 0 || h_indices, "null index list");
    int64_t mx = 0;
    for (int64_t j = 0; j < total; ++j) {
      EMC_REQUIRE(h_indices[j] >= 1, "shell indices are 1-based");
      mx = std::max<int64_t>(mx, h_indices[j]);
    }
    auto* p = new emc_shell_plan;
    p->abs_flag = abs_flag ? 1 : 0;
    p->w = w;
    p->n_shells = n_shells;
    p->n_idx = total;
    p->max_index = mx;
    if (hipMalloc(&p->d_offsets, sizeof(int64_t) * (n_shells + 1)) != hipSuccess ||
        hipMalloc(&p->d_idx, sizeof(uint32_t) * std::max<int64_t>(total, 1)) != hipSuccess) {
      (void)hipGetLastError();
      if (p->d_offsets) (void)hipFree(p->d_offsets);
      delete p;
      fail(EMC_ERR_ALLOC, "shell plan: cannot allocate");
    }
    EMC_HIP(hipMemcpy(p->d_offsets, h_offsets, sizeof(int64_t) * (n_shells + 1), hipMemcpyHostToDevice));
    if (total) EMC_HIP(hipMemcpy(p->d_idx, h_indices, sizeof(uint32_t) * total, hipMemcpyHostToDevice));
    *out = p;
  });
}

emc_status emc_shell_plan_destroy(emc_ctx* ctx, emc_shell_plan* plan) {
  return guard([&] {
    if (!plan) return;
    use_device(ctx);
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    (void)hipFree(plan->d_offsets);
    (void)hipFree(plan->d_idx);
    delete plan;
  });
}

emc_status emc_xcf_rotational_shells(emc_ctx* ctx, const emc_shell_plan* plan, const void* d_p, const void* d_r,
                                     const float* d_wp, const float* d_wr, const int dims[3], double* cc_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(plan && d_p && d_r && dims && cc_out, "null pointer");
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    EMC_REQUIRE(plan->max_index <= n, "shell index %lld beyond the %lld-voxel volume", (long long)plan->max_index,
                (long long)n);
    const float* wp = d_wp;
    const float* wr = d_wr;
    float h0[2] = {0.0f, 0.0f};
    if (wp) EMC_HIP(hipMemcpyAsync(&h0[0], wp, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (wr) EMC_HIP(hipMemcpyAsync(&h0[1], wr, 4, hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    if (wp && h0[0] != h0[0]) wp = nullptr;  // NaN first element: weight 1 (:34-40)
    if (wr && h0[1] != h0[1]) wr = nullptr;
    double* part =
        (double*)scratch(ctx, "xrot_shell_part", sizeof(double) * 3 * kShellBlocks * (size_t)plan->n_shells);
    double* out = (double*)scratch(ctx, "xrot_shell_out", sizeof(double));
    hipLaunchKernelGGL(xcf_rot_shell_kernel, dim3(kShellBlocks, plan->n_shells), dim3(kRedThreads), 0, ctx->stream,
                       (const float2*)d_p, (const float2*)d_r, wp, wr, (const int64_t*)plan->d_offsets,
                       (const uint32_t*)plan->d_idx, part);
    hipLaunchKernelGGL(xcf_rot_shell_fold_kernel, dim3(1), dim3(kWave), 0, ctx->stream, (const double*)part,
                       plan->n_shells, kShellBlocks, plan->abs_flag, plan->w, out);
    EMC_HIP(hipGetLastError());
    EMC_HIP(hipMemcpyAsync(cc_out, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

emc_status emc_xcf_translational(emc_ctx* ctx, emc_fft_plan* plan, const void* d_a, const void* d_b,
                                 const int dims[3], const float* d_peak_mask, const int com[3],
                                 void* d_work, double peak_out[3], double* max_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(plan && d_a && d_b && dims && com && d_work && peak_out, "null pointer");
    int pd[3];
    fft_plan_dims(plan, pd);
    EMC_REQUIRE(pd[0] == dims[0] && pd[1] == dims[1] && pd[2] == dims[2], "plan dims mismatch");
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    float2* X = (float2*)d_work;
    float2* Y = X + n;
    hipLaunchKernelGGL(cmul_scale_kernel, dim3(1024), dim3(256), 0, ctx->stream, (const float2*)d_a,
                       (const float2*)d_b, X, n, (float)(1.0 / (double)n));
    EMC_HIP(hipGetLastError());
    fft_c2c(plan, X, Y, +1);  // real(ifftn(A.*B)) (:31)
    PeakResult* d_res = (PeakResult*)scratch(ctx, "xcf_res", sizeof(PeakResult));
    peak_search(ctx, (const float*)Y, 2, dims, d_peak_mask, com, d_res, nullptr);
    PeakResult h;
    EMC_HIP(hipMemcpyAsync(&h, d_res, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    if (h.status != 0) fail(EMC_ERR_RANGE, "failed to box out COM calc (peak too close to the edge)");
    for (int k = 0; k < 3; ++k) peak_out[k] = h.peak[k];
    if (max_out) *max_out = h.max_val;
  });
}

emc_status emc_xcf_rotational(emc_ctx* ctx, const void* d_p, const void* d_r, const float* d_wp,
                              const float* d_wr, const int dims[3], double* cc_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_p && d_r && dims && cc_out, "null pointer");
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    // NaN first weight element means weight 1 (:34-40)
    const float* wp = d_wp;
    const float* wr = d_wr;
    float h0[2] = {0.0f, 0.0f};
    if (wp) EMC_HIP(hipMemcpyAsync(&h0[0], wp, 4, hipMemcpyDeviceToHost, ctx->stream));
    if (wr) EMC_HIP(hipMemcpyAsync(&h0[1], wr, 4, hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    if (wp && h0[0] != h0[0]) wp = nullptr;
    if (wr && h0[1] != h0[1]) wr = nullptr;
    double* part = (double*)scratch(ctx, "xrot_part", sizeof(double) * 3 * kRedBlocks);
    double* out = (double*)scratch(ctx, "xrot_out", sizeof(double) * 3);
    hipLaunchKernelGGL(xcf_rot_full_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream,
                       (const float2*)d_p, (const float2*)d_r, wp, wr, n, part);
    hipLaunchKernelGGL(fold_partials_kernel<3>, dim3(1), dim3(kRedThreads), 0, ctx->stream,
                       (const double*)part, kRedBlocks, out);
    EMC_HIP(hipGetLastError());
    double h[3];
    EMC_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    *cc_out = h[0] / std::sqrt(h[1] * h[2]);
  });
}

}  // extern "C"
