// CTF and missing-wedge x CTF weights: the mexCTF / mexSF3D replacements
// (reference: mexFiles/utils/ctf.cu:12-128 kernels, mexFiles/mexCTF.cu,
// mexFiles/mexSF3D.cu:25-303, include/core_headers.cuh:40-75 ctfParams,
// masking/BH_weightMaskMex.m marshalling).
//
// The reference builds SF3D with 2 launches + a texture copy/bind per tilt
// (mexSF3D.cu:238-277) and a per-thread serial z loop accumulating into
// global memory. Here: one launch writes the 2D CTF image of every tilt
// (phase evaluated in FP64, stored FP32), one voxel-parallel launch
// accumulates all tilts in registers and writes each output voxel once.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ops.hpp"

namespace emc {

// Per-tilt constants, derived as the ctfParams constructor does
// (core_headers.cuh:63-75).
struct CtfTilt {
  double cs_term, df_term, d1, d2, astig, amp_phase;
  double fvx, fvy;
  double occupancy, exposure;
  float sin_t, cos_t;  // of -tilt (mexSF3D.cu:131-137)
};

static CtfTilt make_tilt(double pixel, double wl, double cs_mm, double df1, double df2,
                         double astig_deg, double amp, double fvx, double fvy) {
  CtfTilt t;
  std::memset(&t, 0, sizeof(t));
  t.cs_term = M_PI * 0.5 * cs_mm * 1e7 * wl * wl * wl;
  t.df_term = M_PI * wl;
  t.d1 = 0.5 * (df1 + df2);
  t.d2 = 0.5 * (df1 - df2);
  t.astig = astig_deg * M_PI / 180.0 - M_PI * std::nearbyint(astig_deg / 180.0);
  t.amp_phase = std::atan(amp / std::sqrt(1.0 - amp * amp));
  t.fvx = fvx;
  t.fvy = fvy;
  (void)pixel;
  return t;
}

// sin(chi) with optional square and optional radial x exposure weight
// (ctf.cu:45-64, :103-127). (x, y) are centred / wrapped frequency indices.
__device__ inline double ctf_value(const CtfTilt& p, int x, int y, int sq, int weighted) {
  const double ky = (double)y * p.fvy;
  const double kx = (double)x * p.fvx;
  const double phi = atan2(ky, kx);
  const double r2 = kx * kx + ky * ky;
  double v = sin(p.cs_term * r2 * r2 - p.df_term * r2 * (p.d1 + p.d2 * cos(2.0 * (phi - p.astig))) -
                 p.amp_phase);
  if (sq) v *= v;
  if (weighted) {
    const double rw = p.occupancy * (x == 0 ? 0.2 : fabs((double)x));
    const double pw = r2 > 0.0 ? pow(r2, -0.8325) : __builtin_inf();
    v *= rw * exp((-0.5 * p.exposure) / (0.24499 * pw + 2.81410));
  }
  return v;
}

// 2D CTF images for n tilts: out[t][y][x], x in [0,cx). centred: x -= o_x
// (full grid only), y -= o_y; else FFT-order wrap (ctf.cu:31-40, :90-99).
__global__ void __launch_bounds__(256) ctf2d_kernel(float* __restrict__ out,
                                                    const CtfTilt* __restrict__ tilts, int cx,
                                                    int ny, int o_x, int o_y, int half, int sq,
                                                    int centered, int weighted) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int t = blockIdx.z;
  if (x >= cx || y >= ny) return;
  int xx = x, yy = y;
  if (centered) {
    yy -= o_y;
    if (!half) xx -= o_x;
  } else {
    if (yy > o_y) yy -= ny;
    if (!half && xx > o_x) xx -= cx;
  }
  out[((int64_t)t * ny + y) * cx + x] = (float)ctf_value(tilts[t], xx, yy, sq, weighted);
}

// SF3D accumulation (mexSF3D.cu:37-77): for each voxel and tilt, the slab
// coordinate tw = -u sin + size_shift_z + w cos; inside |tw| <= 4 add
// 0.25 (0.5 + 0.5 cos(pi tw / 4)) x CTF_t(row y, linear in x).
__global__ void __launch_bounds__(256) sf3d_kernel(float* __restrict__ out,
                                                   const float* __restrict__ ctfs,
                                                   const CtfTilt* __restrict__ tilts, int n_tilts,
                                                   int nx, int ny, int nz, int cx) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= nx || y >= ny) return;
  const float ssx = (nx % 2 == 0) ? 0.5f : 0.0f;
  const float ssz = (nz % 2 == 0) ? 0.5f : 0.0f;
  const float u = (float)x - (float)nx / 2.0f;   // float division, as :47
  const float w = (float)z - (float)(nz / 2);
  const float xscale = (float)cx / (float)nx;
  const float xoff = 0.5f * (float)cx - 0.5f;
  const float kPi = 3.14159265358979323f;
  float acc = 0.0f;
  for (int t = 0; t < n_tilts; ++t) {
    const float s = tilts[t].sin_t, c = tilts[t].cos_t;
    const float tw = (-u * s + ssz) + w * c;
    if (tw < -4.0f || tw > 4.0f) continue;
    const float zw = (0.5f + 0.5f * cosf(tw * (kPi / 4.0f))) * 0.25f;
    const float xt = (u * c + w * s + ssx) * xscale + xoff;  // texel coordinate
    const float xf = floorf(xt);
    const float f = xt - xf;
    const int x0 = (int)xf;
    const float* row = ctfs + ((int64_t)t * ny + y) * cx;
    const float v0 = (unsigned)x0 < (unsigned)cx ? row[x0] : 0.0f;
    const float v1 = (unsigned)(x0 + 1) < (unsigned)cx ? row[x0 + 1] : 0.0f;
    acc += zw * (v0 + f * (v1 - v0));
  }
  out[((int64_t)z * ny + y) * nx + x] = acc;
}

__global__ void shift_min_kernel(float* __restrict__ d, int64_t n, const double* __restrict__ dmin,
                                 float floor_value) {
  const float sub = (float)dmin[0] - floor_value;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = d[i] - sub;
}

void sf3d_impl(emc_ctx* ctx, int half, int sq, const int dims[3], int n,
               const std::vector<double>& pixel, const std::vector<double>& wl,
               const std::vector<double>& cs, const std::vector<double>& df1,
               const std::vector<double>& df2, const std::vector<double>& astig,
               const std::vector<double>& amp, const std::vector<double>& tilt,
               const std::vector<double>& exposure, const std::vector<double>& occ, float* d_out) {
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  EMC_REQUIRE(nx > 0 && ny > 0 && nz > 0, "bad dims");
  EMC_REQUIRE(n > 0, "n_tilts must be > 0");
  const int cx = half ? nx / 2 + 1 : nx;
  std::vector<CtfTilt> h(n);
  for (int i = 0; i < n; ++i) {
    // Fourier voxel size from ctf_dims after the half-grid reduction; tilt
    // 0 and the others use their own pixel size (mexSF3D.cu:234-245)
    const double fvx = 1.0 / (pixel[i] * (double)cx), fvy = 1.0 / (pixel[i] * (double)ny);
    h[i] = make_tilt(pixel[i], wl[i], cs[i], df1[i], df2[i], astig[i], amp[i], fvx, fvy);
    h[i].occupancy = occ[i];
    h[i].exposure = exposure[i];
    const double a = -tilt[i] * M_PI / 180.0;
    h[i].sin_t = (float)std::sin(a);
    h[i].cos_t = (float)std::cos(a);
  }
  CtfTilt* d_t = (CtfTilt*)scratch(ctx, "ctf_tilts", sizeof(CtfTilt) * n);
  float* d_ctf = (float*)scratch(ctx, "ctf_stack", sizeof(float) * (size_t)n * cx * ny);
  EMC_HIP(hipMemcpyAsync(d_t, h.data(), sizeof(CtfTilt) * n, hipMemcpyHostToDevice, ctx->stream));
  dim3 b2(64, 4, 1);
  hipLaunchKernelGGL(ctf2d_kernel, dim3(ceil_div(cx, 64), ceil_div(ny, 4), n), b2, 0, ctx->stream,
                     d_ctf, (const CtfTilt*)d_t, cx, ny, half ? 0 : cx / 2, ny / 2, half, sq, 1, 1);
  hipLaunchKernelGGL(sf3d_kernel, dim3(ceil_div(nx, 64), ceil_div(ny, 4), nz), b2, 0, ctx->stream,
                     d_out, (const float*)d_ctf, (const CtfTilt*)d_t, n, nx, ny, nz, cx);
  EMC_HIP(hipGetLastError());
  // the tilt table lives in pageable memory; finish before it goes away
  EMC_HIP(hipStreamSynchronize(ctx->stream));
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_sf3d(emc_ctx* ctx, int half, int sq_ctf, const int dims[3], int n_tilts,
                    const float* pixel, const float* wavelength, const float* cs_mm,
                    const float* df1, const float* df2, const float* astig_deg,
                    const float* amp_contrast, const float* tilt_deg, const float* exposure,
                    const float* occupancy, float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(dims && d_out && pixel && wavelength && cs_mm && df1 && df2 && astig_deg &&
                    amp_contrast && tilt_deg && exposure && occupancy,
                "null pointer");
    auto v = [&](const float* p) { return std::vector<double>(p, p + n_tilts); };
    sf3d_impl(ctx, half, sq_ctf, dims, n_tilts, v(pixel), v(wavelength), v(cs_mm), v(df1), v(df2),
              v(astig_deg), v(amp_contrast), v(tilt_deg), v(exposure), v(occupancy), d_out);
  });
}

emc_status emc_weight_mask(emc_ctx* ctx, const int dims[3], double sampling,
                           const double* tlt, int n, const double xyz[3],
                           const double recon_size[3], const double recon_shift[3],
                           float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(dims && tlt && xyz && recon_size && recon_shift && d_out, "null pointer");
    EMC_REQUIRE(n > 0, "empty TLT");
    // sortrows(TLT,1) (BH_weightMaskMex.m:12), stable
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return tlt[22 * a] < tlt[22 * b]; });
    auto col = [&](int row, int c) { return tlt[22 * order[row] + c]; };
    // everything passes through single() in the reference (:12)
    auto sgl = [](double x) { return (double)(float)x; };
    double origin[3], prj[3];
    for (int k = 0; k < 3; ++k) {
      origin[k] = std::ceil((recon_size[k] + 1.0) / 2.0);
      prj[k] = xyz[k] - origin[k] + recon_shift[k];  // :17-20
    }
    std::vector<double> pixel(n), wl(n), cs(n), dfa(n), dfb(n), ast(n), amp(n), tilt(n), ex(n),
        occ(n);
    double dose_mean = 0.0;
    for (int i = 0; i < n; ++i) dose_mean += sgl(col(i, 13));
    dose_mean /= n;
    std::vector<double> elastic(n);
    double emax = 0.0;
    for (int i = 0; i < n; ++i) {
      elastic[i] = std::exp(-75.0 / (std::cos(sgl(col(i, 3)) * M_PI / 180.0) * 400.0));  // :43
      emax = std::max(emax, elastic[i]);
    }
    for (int i = 0; i < n; ++i) {
      tilt[i] = sgl(col(i, 3));
      pixel[i] = sgl(sgl(col(i, 15)) * sampling * 1e10);  // :16
      cs[i] = sgl(sgl(col(i, 16)) * 1e3);                 // :22
      wl[i] = sgl(sgl(col(i, 17)) * 1e10);                // :23
      amp[i] = sgl(col(i, 18));                           // :24
      double defocus = sgl(-sgl(col(i, 14)) * 1e10);      // :25
      const double ddf = sgl(sgl(col(i, 11)) * 1e10);     // :26
      ast[i] = sgl(sgl(col(i, 12)) * 180.0 / M_PI);       // :27
      ex[i] = sgl(col(i, 10));                            // :30
      // particle height along the beam after the tilt (:33-37):
      // BH_defineMatrix(tilt,'TILT','forwardVector') row 3 . prjVector
      const double a = tilt[i] * M_PI / 180.0;
      const double z = -std::sin(a) * prj[0] + std::cos(a) * prj[2];
      defocus = defocus - z * pixel[i];
      dfa[i] = sgl(defocus + ddf);  // :48-49
      dfb[i] = sgl(defocus - ddf);
      occ[i] = sgl(elastic[i] / emax * (sgl(col(i, 13)) / dose_mean));  // :42-44
    }
    sf3d_impl(ctx, 0, 1, dims, n, pixel, wl, cs, dfa, dfb, ast, amp, tilt, ex, occ, d_out);
    const int64_t nv = (int64_t)dims[0] * dims[1] * dims[2];
    double* d_min = (double*)scratch(ctx, "wm_min", sizeof(double));
    reduce_min(ctx, d_out, nv, 1, d_min);
    hipLaunchKernelGGL(shift_min_kernel, dim3(256), dim3(256), 0, ctx->stream, d_out, nv,
                       (const double*)d_min, 1e-4f);  // :55
    EMC_HIP(hipGetLastError());
  });
}

emc_status emc_ctf2d(emc_ctx* ctx, int half, int sq_ctf, int nx, int ny, float pixel,
                     float wavelength, float cs_mm, float df1, float df2, float astig_deg,
                     float amp_contrast, int centered, float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_out && nx > 0 && ny > 0, "bad arguments");
    const int cx = half ? nx / 2 + 1 : nx;
    // mexCTF's Fourier voxel size uses the full nX (mexCTF.cu:44)
    CtfTilt t = make_tilt(pixel, wavelength, cs_mm, df1, df2, astig_deg, amp_contrast,
                          1.0 / ((double)pixel * nx), 1.0 / ((double)pixel * ny));
    CtfTilt* d_t = (CtfTilt*)scratch(ctx, "ctf2d_tilt", sizeof(CtfTilt));
    EMC_HIP(hipMemcpyAsync(d_t, &t, sizeof(t), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(ctf2d_kernel, dim3(ceil_div(cx, 64), ceil_div(ny, 4), 1), dim3(64, 4, 1), 0,
                       ctx->stream, d_out, (const CtfTilt*)d_t, cx, ny, half ? 0 : nx / 2, ny / 2,
                       half, sq_ctf, centered, 0);
    EMC_HIP(hipGetLastError());
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
