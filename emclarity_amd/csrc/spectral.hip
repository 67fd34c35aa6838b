// Bandpass, padding and band-limit/centre/normalise on half spectra
// (reference: masking/BH_bandpass3d.m:54-170, masking/BH_padZeros3d.m:37-229,
// masking/BH_bandLimitCenterNormalize.m:27-61).
//
// BH_bandLimitCenterNormalize works on the full complex grid (C2C fftn). The
// input is real, so the full spectrum is Hermitian: here the R2C half
// spectrum is used and every full-grid sum is taken as sum over the half
// grid with multiplicity 2 for bins whose mirror -k is not stored (x planes
// 1..ceil(nx/2)-1) and 1 for the self-mirrored planes x = 0 and, for even
// nx, x = nx/2. This is exact, and halves FFT work and HBM traffic.
#include <cmath>

#include "ops.hpp"

namespace emc {

__global__ void __launch_bounds__(256) bandpass_kernel(float* __restrict__ out, int nx, int ny, int nz,
                                                       int cxw, double lc, double lroll, double hc,
                                                       double hroll) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= cxw || y >= ny) return;
  // FFT-order coordinates [0..n/2, -(n-1)/2..-1]/n (BH_multi_gridCoordinates.m:209-215)
  const double fx = (double)(x <= nx / 2 ? x : x - nx) / nx;
  const double fy = (double)(y <= ny / 2 ? y : y - ny) / ny;
  const double fz = (double)(z <= nz / 2 ? z : z - nz) / nz;
  const double r = sqrt(fx * fx + fy * fy + fz * fz);
  double w = r < lc ? 1.0 : exp(-(r - lc) * (r - lc) / (2.0 * lroll * lroll));
  if (hc != 0.0 && r <= hc) w = exp(-(r - hc) * (r - hc) / (2.0 * hroll * hroll));
  out[((int64_t)z * ny + y) * cxw + x] = (float)w;
}

void bandpass_impl(emc_ctx* ctx, const int dims[3], int half, double thr, double high_cut,
                   double low_cut, double pixel, float* d_out) {
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  EMC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad dims");
  EMC_REQUIRE(thr >= 0.0 && thr < 1.0, "HIGH_THRESH must be between 0 and 1, not %f", thr);
  EMC_REQUIRE(low_cut != 0.0, "LOW_CUT must be nonzero");
  const double hc = high_cut != 0.0 ? pixel / high_cut : 0.0;  // :138-142
  double lc = pixel / low_cut;                                   // :143
  double lroll;
  if (lc > 0) {
    const int mn = nz == 1 ? std::min(nx, ny) : std::min(nx, std::min(ny, nz));
    lroll = 2.0 / mn;                                            // :149-154
  } else {
    lc = std::fabs(lc);
    lroll = std::sqrt(-((0.5 - lc) * (0.5 - lc)) / (2.0 * std::log(1e-3)));  // :155-159
  }
  const double hroll = high_cut != 0.0 ? std::sqrt(-(hc * hc) / (2.0 * std::log(thr))) : 0.0;
  const int cxw = half ? nx / 2 + 1 : nx;
  hipLaunchKernelGGL(bandpass_kernel, dim3(ceil_div(cxw, 64), ceil_div(ny, 4), nz), dim3(64, 4, 1), 0,
                     ctx->stream, d_out, nx, ny, nz, cxw, lc, lroll, hc, hroll);
  EMC_HIP(hipGetLastError());
}

static __global__ void div_scalar_kernel(const double* in, double n, double* out) {
  out[0] = in[0] / n;
}

// Pad/trim with the 'singleTaper' edge (BH_padZeros3d.m:79-93, :110-117,
// :189-229). Output voxel (X,Y,Z); image voxel (x,y,z) = (X,Y,Z) - low;
// the trimmed image starts at trim_low inside the input. fill value from
// a device scalar (the mean for band-limit normalise) or a constant.
__global__ void __launch_bounds__(256) pad_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                  PadArgs a, const double* __restrict__ d_fill) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int Z = blockIdx.z;
  if (X >= a.ox || Y >= a.oy) return;
  const float ev = d_fill ? (float)d_fill[0] : a.fill;
  const int x = X - a.lx, y = Y - a.ly, z = Z - a.lz;  // in the trimmed image
  float v = ev;
  if ((unsigned)x < (unsigned)a.wx && (unsigned)y < (unsigned)a.wy && (unsigned)z < (unsigned)a.wz) {
    v = in[((int64_t)(z + a.tz) * a.ny + (y + a.ty)) * a.nx + (x + a.tx)];
    if (a.taper) {
      // taper(i) = 0.5 + 0.5 cos(i pi / 7), i = 0..6; low edges use flip()
      auto tp = [](int i) { return 0.5f + 0.5f * cosf((float)i * 3.14159265358979f / 7.0f); };
      if (y < 7) v = v * tp(6 - y) + ev * (1.0f - tp(6 - y));
      if (x < 7) v = v * tp(6 - x) + ev * (1.0f - tp(6 - x));
      if (z < 7) v = v * tp(6 - z) + ev * (1.0f - tp(6 - z));
      if (y >= a.wy - 7) v = v * tp(y - (a.wy - 7)) + ev * (1.0f - tp(y - (a.wy - 7)));
      if (x >= a.wx - 7) v = v * tp(x - (a.wx - 7)) + ev * (1.0f - tp(x - (a.wx - 7)));
      if (z >= a.wz - 7) v = v * tp(z - (a.wz - 7)) + ev * (1.0f - tp(z - (a.wz - 7)));
    }
  }
  out[((int64_t)Z * a.oy + Y) * a.ox + X] = v;
}

PadArgs pad_args(const int dims[3], const int low[3], const int top[3], float fill, int taper) {
  PadArgs a;
  a.nx = dims[0];
  a.ny = dims[1];
  a.nz = dims[2];
  int tl[3], tt[3], pl[3], pt[3];
  for (int k = 0; k < 3; ++k) {
    tl[k] = low[k] < 0 ? -low[k] : 0;
    tt[k] = top[k] < 0 ? -top[k] : 0;
    pl[k] = low[k] > 0 ? low[k] : 0;
    pt[k] = top[k] > 0 ? top[k] : 0;
  }
  a.tx = tl[0];
  a.ty = tl[1];
  a.tz = tl[2];
  a.wx = dims[0] - tl[0] - tt[0];
  a.wy = dims[1] - tl[1] - tt[1];
  a.wz = dims[2] - tl[2] - tt[2];
  EMC_REQUIRE(a.wx > 0 && a.wy > 0 && a.wz > 0, "trim larger than the image");
  a.lx = pl[0];
  a.ly = pl[1];
  a.lz = pl[2];
  a.ox = a.wx + pl[0] + pt[0];
  a.oy = a.wy + pl[1] + pt[1];
  a.oz = a.wz + pl[2] + pt[2];
  a.fill = fill;
  a.taper = taper;
  return a;
}

void launch_pad(emc_ctx* ctx, const float* in, float* out, const PadArgs& a, const double* d_fill) {
  hipLaunchKernelGGL(pad_kernel, dim3(ceil_div(a.ox, 64), ceil_div(a.oy, 4), a.oz), dim3(64, 4, 1), 0,
                     ctx->stream, in, out, a, d_fill);
  EMC_HIP(hipGetLastError());
}

// Half-spectrum pass of band-limit normalise: F(0)=0, F *= bandpass (full
// grid FFT order, read at the same (x,y,z)), and per-block partial sums of
// multiplicity-weighted |F|^2. Batched over `batch` spectra (blockIdx.y).
__global__ void __launch_bounds__(kRedThreads) blcn_pass_kernel(float2* __restrict__ F, int nx, int ny,
                                                                int nz, const float* __restrict__ bp,
                                                                double* __restrict__ part) {
  const int cx = nx / 2 + 1;
  const int64_t nh = (int64_t)cx * ny * nz;
  float2* Fb = F + (int64_t)blockIdx.y * nh;
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    split3(i, cx, ny, x, y, z);
    const int64_t yz = (int64_t)z * ny + y;
    float2 f = Fb[i];
    if (i == 0) f = make_float2(0.0f, 0.0f);
    if (bp) {
      const float w = bp[yz * nx + x];
      f.x *= w;
      f.y *= w;
    }
    Fb[i] = f;
    const double m = (x == 0 || (nx % 2 == 0 && x == nx / 2)) ? 1.0 : 2.0;
    v[0] += m * ((double)f.x * f.x + (double)f.y * f.y);
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = v[0];
}

// fold per-spectrum partials into scale = numel / sqrt(sum) (:61)
__global__ void __launch_bounds__(kRedThreads) blcn_scale_kernel(const double* __restrict__ part, int nb,
                                                                 double numel, double* __restrict__ scale) {
  double v[1] = {0.0};
  const double* p = part + (int64_t)blockIdx.x * nb;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += p[b];
  block_sum<1>(v);
  if (threadIdx.x == 0) scale[blockIdx.x] = numel / sqrt(v[0]);
}

__global__ void scale_c64_kernel(float2* __restrict__ F, int64_t nh, const double* __restrict__ scale) {
  const float s = (float)scale[blockIdx.y];
  float2* Fb = F + (int64_t)blockIdx.y * nh;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (int64_t)gridDim.x * blockDim.x) {
    float2 f = Fb[i];
    Fb[i] = make_float2(f.x * s, f.y * s);
  }
}

void blcn_half(emc_ctx* ctx, float2* F, const int dims[3], int batch, const float* bp,
               double* d_scale, bool apply_scale) {
  const int nb = 128;
  double* part = (double*)scratch(ctx, "blcn_part", sizeof(double) * nb * batch);
  hipLaunchKernelGGL(blcn_pass_kernel, dim3(nb, batch), dim3(kRedThreads), 0, ctx->stream, F, dims[0],
                     dims[1], dims[2], bp, part);
  const double numel = (double)dims[0] * dims[1] * dims[2];
  hipLaunchKernelGGL(blcn_scale_kernel, dim3(batch), dim3(kRedThreads), 0, ctx->stream,
                     (const double*)part, nb, numel, d_scale);
  if (apply_scale) {
    const int64_t nh = (int64_t)(dims[0] / 2 + 1) * dims[1] * dims[2];
    hipLaunchKernelGGL(scale_c64_kernel, dim3(256, batch), dim3(256), 0, ctx->stream, F, nh,
                       (const double*)d_scale);
  }
  EMC_HIP(hipGetLastError());
}

// fourierTransformer.fwdFFT's options on an R2C half spectrum (cx = nx/2+1
// x-planes): f' = f * ctf * norm * bp, then (centre) f'(0) = 0 and
// f' / sqrt(2 * sum |f'|^2 over x < cx - invTrim). Pass 1 (centre only)
// forms per-block partial sums; pass 2 writes f' scaled.
__device__ inline float2 ftc_value(const float2* F, int64_t i, const float* ctf, const float* bp, float norm) {
  float2 f = F[i];
  if (ctf) {
    f.x *= ctf[i];
    f.y *= ctf[i];
  }
  f.x *= norm;
  f.y *= norm;
  if (bp) {
    f.x *= bp[i];
    f.y *= bp[i];
  }
  return f;
}

__global__ void __launch_bounds__(kRedThreads) ftc_sum_kernel(const float2* __restrict__ F, int64_t nh, int cx,
                                                              int x_end, const float* __restrict__ ctf,
                                                              const float* __restrict__ bp, float norm,
                                                              double* __restrict__ part) {
  double v[1] = {0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += (int64_t)gridDim.x * blockDim.x) {
    if (i == 0 || (int)(i % cx) >= x_end) continue;
    const float2 f = ftc_value(F, i, ctf, bp, norm);
    v[0] += (double)f.x * f.x + (double)f.y * f.y;
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void __launch_bounds__(kRedThreads) ftc_fold_kernel(const double* __restrict__ part, int nb,
                                                               double* __restrict__ scale) {
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) v[0] += part[b];
  block_sum<1>(v);
  if (threadIdx.x == 0) scale[0] = 1.0 / sqrt(2.0 * v[0]);
}

__global__ void ftc_apply_kernel(float2* __restrict__ F, int64_t nh, const float* __restrict__ ctf,
                                 const float* __restrict__ bp, float norm, const double* __restrict__ scale) {
  const float s = scale ? (float)scale[0] : 1.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh; i += (int64_t)gridDim.x * blockDim.x) {
    float2 f = ftc_value(F, i, ctf, bp, norm);
    if (scale) {
      f = i == 0 ? make_float2(0.0f, 0.0f) : make_float2(f.x * s, f.y * s);
    }
    F[i] = f;
  }
}

static __global__ void mul_f32_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                      float* __restrict__ o, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    o[i] = b ? a[i] * b[i] : a[i];
}

static __global__ void conj_c64_kernel(float2* __restrict__ F, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    F[i].y = -F[i].y;
}

// pad with the mean -> R2C -> DC 0, bandpass, scale to full-grid
// sum |F|^2 = numel^2 (the whole of BH_bandLimitCenterNormalize)
static void blcn_full(emc_ctx* ctx, emc_fft_plan* plan, const float* d_img, const int dims[3],
                      const int pad_low[3], const int pad_top[3], const float* d_bandpass,
                      void* d_out_c64);

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_ref_ft1(emc_ctx* ctx, emc_fft_plan* plan, const float* d_ref, const float* d_vol_mask,
                       const int dims[3], const int pad_low[3], const int pad_top[3],
                       void* d_out_c64) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(plan && d_ref && dims && pad_low && pad_top && d_out_c64, "null pointer");
    const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
    float* masked = (float*)scratch(ctx, "ref_ft1_masked", sizeof(float) * n);
    hipLaunchKernelGGL(mul_f32_kernel, dim3(1024), dim3(256), 0, ctx->stream, d_ref, d_vol_mask,
                       masked, n);
    EMC_HIP(hipGetLastError());
    blcn_full(ctx, plan, masked, dims, pad_low, pad_top, nullptr, d_out_c64);
    int pd[3];
    fft_plan_dims(plan, pd);
    const int64_t nh = (int64_t)(pd[0] / 2 + 1) * pd[1] * pd[2];
    hipLaunchKernelGGL(conj_c64_kernel, dim3(1024), dim3(256), 0, ctx->stream, (float2*)d_out_c64, nh);
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_bandpass3d(emc_ctx* ctx, const int dims[3], int half, double high_thresh,
                          double high_cut, double low_cut, double pixel, float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(dims && d_out, "null pointer");
    bandpass_impl(ctx, dims, half, high_thresh, high_cut, low_cut, pixel, d_out);
  });
}

emc_status emc_pad_zeros3d(emc_ctx* ctx, const float* d_in, const int dims[3], const int pad_low[3],
                           const int pad_top[3], float extrap_val, int taper, float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_in && dims && pad_low && pad_top && d_out, "null pointer");
    EMC_REQUIRE(d_in != d_out, "d_out must not alias d_in");
    PadArgs a = pad_args(dims, pad_low, pad_top, extrap_val, taper);
    launch_pad(ctx, d_in, d_out, a, nullptr);
  });
}

emc_status emc_ft_condition(emc_ctx* ctx, void* d_ft_c64, const int dims[3], int do_norm, int do_center,
                            const float* d_bp_half, const float* d_ctf_half) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_ft_c64 && dims, "null pointer");
    EMC_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    EMC_REQUIRE(do_norm >= 0 && do_norm <= 2, "do_norm must be 0, 1 or 2");
    const int cx = dims[0] / 2 + 1;
    const int64_t nh = (int64_t)cx * dims[1] * dims[2];
    const int inv_trim = dims[0] % 2 ? 1 : 2;
    // (1/sqrt(numel))^doNorm, only without centring (fourierTransformer.m:130-132)
    const float norm =
        (do_norm && !do_center) ? (float)pow(1.0 / sqrt((double)dims[0] * dims[1] * dims[2]), do_norm) : 1.0f;
    float2* F = (float2*)d_ft_c64;
    const int grid = (int)std::min<int64_t>(1024, ceil_div(nh, (int64_t)256));
    double* scale = nullptr;
    if (do_center) {
      const int nb = 256;
      double* part = (double*)scratch(ctx, "ftc_part", sizeof(double) * (nb + 1));
      scale = part + nb;
      hipLaunchKernelGGL(ftc_sum_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, (const float2*)F, nh, cx,
                         cx - inv_trim, d_ctf_half, d_bp_half, norm, part);
      hipLaunchKernelGGL(ftc_fold_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream, (const double*)part, nb,
                         scale);
    }
    hipLaunchKernelGGL(ftc_apply_kernel, dim3(grid), dim3(256), 0, ctx->stream, F, nh, d_ctf_half, d_bp_half, norm,
                       (const double*)scale);
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_band_limit_center_normalize(emc_ctx* ctx, emc_fft_plan* plan, const float* d_img,
                                           const int dims[3], const int pad_low[3],
                                           const int pad_top[3], const float* d_bandpass,
                                           void* d_out_c64) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(plan && d_img && dims && pad_low && pad_top && d_out_c64, "null pointer");
    blcn_full(ctx, plan, d_img, dims, pad_low, pad_top, d_bandpass, d_out_c64);
  });
}

}  // extern "C"

namespace emc {

static void blcn_full(emc_ctx* ctx, emc_fft_plan* plan, const float* d_img, const int dims[3],
                      const int pad_low[3], const int pad_top[3], const float* d_bandpass,
                      void* d_out_c64) {
  PadArgs a = pad_args(dims, pad_low, pad_top, 0.0f, 0);
  int pdims[3] = {a.ox, a.oy, a.oz};
  int pd[3];
  fft_plan_dims(plan, pd);
  EMC_REQUIRE(pd[0] == a.ox && pd[1] == a.oy && pd[2] == a.oz, "plan dims do not match padded dims");
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  double* stats = (double*)scratch(ctx, "blcn_stats", sizeof(double) * 4);
  reduce_sum_sumsq(ctx, d_img, n, stats);
  // pad value = mean(img) (:48)
  hipLaunchKernelGGL(div_scalar_kernel, dim3(1), dim3(1), 0, ctx->stream, stats, (double)n, stats + 2);
  float* padded = (float*)scratch(ctx, "blcn_padded", sizeof(float) * (size_t)a.ox * a.oy * a.oz);
  launch_pad(ctx, d_img, padded, a, stats + 2);
  fft_r2c(plan, padded, d_out_c64);
  blcn_half(ctx, (float2*)d_out_c64, pdims, 1, d_bandpass, stats + 3, true);
}

}  // namespace emc
