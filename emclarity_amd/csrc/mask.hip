// EMC_maskShape (reference: masking/EMC_maskShape.m:93-215, with
// coordinates/EMC_coordVectors.m origin 1, masking/EMC_gaussianKernel.m,
// masking/EMC_resize.m 'force_taper', testScripts/EMC_convn.m): the peak,
// volume and centre-of-mass masks of the alignment
// (alignment/BH_alignRaw3d_v2.m:415, :482, :493, :504).
//
// One launch evaluates the binary shape (single precision, as the
// reference's default 'precision') with the outer layers of every face
// zeroed when a blurring kernel follows; three launches convolve it with the
// separable Gaussian ('same', zero outside); a reduction finds the maximum
// and one launch normalises and applies the 0.001 cut-off.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ops.hpp"

namespace emc {

constexpr int kMaskTapsMax = 255;

struct MaskShapeArgs {
  int nx, ny, nz;
  int shape;        // 0 sphere, 1 cylinder, 2 rectangle
  float r[3];       // sphere / cylinder x, y radii (already enlarged)
  int ir[3];        // rectangle radii; cylinder z half-extent (r[2] unused)
  float s[3];       // shifts (sphere, cylinder x, y)
  int is[3];        // rounded shifts (rectangle; cylinder z)
  int edge;         // outer layers zeroed per face (0: none)
};

__global__ void mask_shape_kernel(float* __restrict__ out, MaskShapeArgs a) {
  // each square and sum rounded on its own: a fused multiply-add moves
  // voxels whose sum is exactly 1 (3-4-5 triangles on the radius) across the
  // boundary
#pragma clang fp contract(off)
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= a.nx || y >= a.ny) return;
  bool in = false;
  if (a.shape == 2) {
    in = abs(x - a.nx / 2 - a.is[0]) <= a.ir[0] && abs(y - a.ny / 2 - a.is[1]) <= a.ir[1] &&
         abs(z - a.nz / 2 - a.is[2]) <= a.ir[2];
  } else {
    // EMC_coordVectors: v = (i - floor(n/2)) - shift, in single precision
    const float vx = (float)(x - a.nx / 2) - a.s[0], vy = (float)(y - a.ny / 2) - a.s[1];
    const float tx = vx / a.r[0], ty = vy / a.r[1];
    if (a.shape == 0) {
      const float vz = (float)(z - a.nz / 2) - a.s[2];
      const float tz = vz / a.r[2];
      in = (tx * tx + ty * ty) + tz * tz <= 1.0f;
    } else {
      in = tx * tx + ty * ty <= 1.0f && abs(z - a.nz / 2 - a.is[2]) <= a.ir[2];
    }
  }
  if (a.edge > 0 && (x < a.edge || x >= a.nx - a.edge || y < a.edge || y >= a.ny - a.edge || z < a.edge ||
                     z >= a.nz - a.edge))
    in = false;
  out[((int64_t)z * a.ny + y) * a.nx + x] = in ? 1.0f : 0.0f;
}

// out[i] = sum_m k[m] in[i + K/2 - m] along one axis ('same', zero outside);
// axis 0 = x (stride 1), 1 = y (stride nx), 2 = z (stride nx ny)
struct MaskTaps {
  float k[kMaskTapsMax];
  int K;
};

__global__ void mask_conv_kernel(const float* __restrict__ in, float* __restrict__ out, int nx, int ny, int nz,
                                 int axis, MaskTaps t) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const int n = axis == 0 ? nx : axis == 1 ? ny : nz;
  const int i = axis == 0 ? x : axis == 1 ? y : z;
  const int64_t stride = axis == 0 ? 1 : axis == 1 ? nx : (int64_t)nx * ny;
  const int64_t base = ((int64_t)z * ny + y) * nx + x - (int64_t)i * stride;
  const int h = t.K / 2;
  float acc = 0.0f;
  for (int m = 0; m < t.K; ++m) {
    const int j = i + h - m;
    if ((unsigned)j < (unsigned)n) acc = fmaf(t.k[m], in[base + (int64_t)j * stride], acc);
  }
  out[((int64_t)z * ny + y) * nx + x] = acc;
}

__global__ void __launch_bounds__(kRedThreads) mask_max_partial_kernel(const float* __restrict__ d, int64_t n,
                                                                       float* __restrict__ part) {
  float m = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, d[i]);
  m = -block_min(-m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void mask_normalise_kernel(float* __restrict__ d, int64_t n, const float* __restrict__ part, int nb,
                                      float cutoff) {
  float mx = 0.0f;
  for (int b = 0; b < nb; ++b) mx = fmaxf(mx, part[b]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float v = mx > 0.0f ? d[i] / mx : 0.0f;
    d[i] = v <= cutoff ? 0.0f : v;
  }
}

static double matlab_round(double x) { return x >= 0 ? std::floor(x + 0.5) : std::ceil(x - 0.5); }

// The Gaussian taps of EMC_maskShape.m:96-104 (EMC_gaussianKernel, single).
static std::vector<float> mask_taps(const int dims[3], double frac) {
  const int mn = std::min(dims[0], std::min(dims[1], dims[2]));
  int K = (int)matlab_round((mn * frac - 1.0) / 2.0) * 2 + 1;
  if (K < 9) K = 9;
  const double middle = std::ceil(K / 2.0) / 2.0;
  const float sigma = (float)std::sqrt(-1.0 * middle * middle / (2.0 * std::log(0.5)));
  std::vector<float> k(K);
  float sum = 0.0f;
  for (int i = 0; i < K; ++i) {
    const float v = (float)i - (float)((K - 1) / 2);
    const float q = v / sigma;
    k[i] = std::exp(-0.5f * (q * q));
    sum += k[i];
  }
  for (float& v : k) v /= sum;
  return k;
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_mask_shape(emc_ctx* ctx, int shape, const int dims[3], const double radius[3],
                          const double shift[3], double kernel, float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(dims && radius && shift && d_out, "null pointer");
    EMC_REQUIRE(shape >= 0 && shape <= 2, "shape must be 0 (sphere), 1 (cylinder) or 2 (rectangle)");
    EMC_REQUIRE(dims[0] > 1 && dims[1] > 1 && dims[2] > 0, "SIZE should be a 3d size");
    for (int k = 0; k < 3; ++k) {
      EMC_REQUIRE(radius[k] > 1 && std::isfinite(radius[k]), "RADIUS should be greater than 1");
      EMC_REQUIRE(std::isfinite(shift[k]), "shifts should not contain NaNs or Inf");
    }
    EMC_REQUIRE(kernel >= 0 && kernel <= 1, "kernel should be between 0 and 1");
    std::vector<float> taps;
    if (kernel > 0) taps = mask_taps(dims, kernel);
    const int K = (int)taps.size();
    EMC_REQUIRE(K <= kMaskTapsMax, "kernel of %d taps is longer than %d", K, kMaskTapsMax);
    for (int k = 0; k < 3; ++k)
      EMC_REQUIRE(K == 0 || dims[k] >= (K + 1) / 2 * 2 + 1, "with a kernel size of %d, the minimum SIZE is %d", K,
                  (K + 1) / 2 * 2 + 1);
    MaskShapeArgs a{};
    a.nx = dims[0];
    a.ny = dims[1];
    a.nz = dims[2];
    a.shape = shape;
    const int ce = (K + 1) / 2, fl = K / 2;  // ceil(K/2), floor(K/2)
    for (int k = 0; k < 3; ++k) {
      a.r[k] = (float)(radius[k] + ce);
      a.s[k] = (float)shift[k];
      a.is[k] = (int)matlab_round(shift[k]);
      a.ir[k] = (int)matlab_round(radius[k]) + fl;
    }
    a.edge = K > 0 ? ce : 0;
    const int64_t nv = (int64_t)dims[0] * dims[1] * dims[2];
    const dim3 blk(64, 4, 1), grd(ceil_div(dims[0], 64), ceil_div(dims[1], 4), dims[2]);
    float* bin = K > 0 ? (float*)scratch(ctx, "mask_bin", sizeof(float) * nv) : d_out;
    hipLaunchKernelGGL(mask_shape_kernel, grd, blk, 0, ctx->stream, bin, a);
    EMC_HIP(hipGetLastError());
    if (K == 0) return;
    MaskTaps t{};
    t.K = K;
    for (int i = 0; i < K; ++i) t.k[i] = taps[i];
    float* tmp = (float*)scratch(ctx, "mask_tmp", sizeof(float) * nv);
    hipLaunchKernelGGL(mask_conv_kernel, grd, blk, 0, ctx->stream, (const float*)bin, tmp, a.nx, a.ny, a.nz, 0, t);
    hipLaunchKernelGGL(mask_conv_kernel, grd, blk, 0, ctx->stream, (const float*)tmp, bin, a.nx, a.ny, a.nz, 1, t);
    hipLaunchKernelGGL(mask_conv_kernel, grd, blk, 0, ctx->stream, (const float*)bin, d_out, a.nx, a.ny, a.nz, 2, t);
    const int nb = 256;
    float* part = (float*)scratch(ctx, "mask_max", sizeof(float) * nb);
    hipLaunchKernelGGL(mask_max_partial_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream, (const float*)d_out, nv,
                       part);
    hipLaunchKernelGGL(mask_normalise_kernel, dim3(256), dim3(256), 0, ctx->stream, d_out, nv, (const float*)part, nb,
                       0.001f);
    EMC_HIP(hipGetLastError());
  });
}

}  // extern "C"
