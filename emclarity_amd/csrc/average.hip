// Subtomogram averaging: the per-particle body of
// transformations/BH_average3d.m:1417-1441 and its half-set accumulation
// (:1465-1490), for a batch of particles cut from a resident tomogram.
//
// Per particle p (matrix M_p = geometry columns 17:25, shift t_p):
//   P = (1/nsym) sum_s trilinear(window_p, M_p S_s, t_p, 'inv')
//   W = (1/nsym) sum_s trilinear(wedge_p,  M_p S_s, 0,   'inv')
//   P = (P - mean(P[mask > 0.01])) / rms(.. [mask > 0.01]) .* mask
//   sum[half_p] += P * weight_p,  wedge_sum[half_p] += W * weight_p
// The reference runs this particle by particle (two texture rebuilds,
// ~6 elementwise kernels and two host reductions each). Here one launch
// rotates a whole chunk of particles (2-D grid: voxel blocks x particles)
// with fixed-partition FP64 partial sums for the masked statistics, one
// launch folds the statistics, and one launch walks the chunk's particles in
// order for each output voxel, so the FP32 running sums add the particles in
// the reference's order. HBM traffic per particle: window + rotated copy
// written and read once (12 B per voxel) + the wedge likewise.
#include <cmath>
#include <cstring>
#include <vector>

#include "emc_internal.hpp"
#include "ops.hpp"

namespace emc {
namespace {

constexpr int kAvgThreads = 256;
constexpr int kAvgParts = 64;          // partial sums per particle (fixed: deterministic folds)
constexpr float kBinaryCut = 0.01f;    // interpMask > 0.01 (BH_average3d.m:1087-1092)

// Rotate every particle of the chunk; with `mask`, per-block FP64 partials
// of sum r and sum r^2 over mask > 0.01, the voxel count there, and the sum
// of r over all voxels (part[(p * kAvgParts + block) * 4 + k]).
__global__ void __launch_bounds__(kAvgThreads) avg_rotate_kernel(
    const float* __restrict__ in, int nx, int ny, int nz, const float* __restrict__ rows, int nsym,
    float inv_nsym, const float* __restrict__ shifts, const float* __restrict__ mask, float* __restrict__ out,
    double* __restrict__ part) {
  const int p = blockIdx.y;
  const int64_t nv = (int64_t)nx * ny * nz;
  const float* vin = in + (int64_t)p * nv;
  float* vout = out + (int64_t)p * nv;
  const float* m0 = rows + (int64_t)p * nsym * 9;
  const float tx = shifts ? shifts[3 * p] : 0.0f, ty = shifts ? shifts[3 * p + 1] : 0.0f,
              tz = shifts ? shifts[3 * p + 2] : 0.0f;
  const float cx = (float)(nx / 2), cy = (float)(ny / 2), cz = (float)(nz / 2);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  // the threads cover 8 x 8 x 4 output bricks, one wavefront a 4 x 4 x 4
  // cube of it, so each wavefront's samples fall in a compact region of
  // the input
  static_assert(kAvgThreads == 256, "8 x 8 x 4 bricks");
  const int bx = (nx + 7) / 8, by = (ny + 7) / 8, bz = (nz + 3) / 4;
  const int nbrick = bx * by * bz;
  const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
  const int ox = (wv % 2) * 4 + lane % 4, oy = (wv / 2) * 4 + (lane / 4) % 4, oz = lane / 16;
  for (int br = blockIdx.x; br < nbrick; br += gridDim.x) {
    const int x = (br % bx) * 8 + ox, y = ((br / bx) % by) * 8 + oy, z = (br / (bx * by)) * 4 + oz;
    if (x >= nx || y >= ny || z >= nz) continue;
    const int64_t v = ((int64_t)z * ny + y) * nx + x;
    const float u = (float)x - cx, w = (float)y - cy, q = (float)z - cz;
    float r = 0.0f;
    for (int s = 0; s < nsym; ++s) {
      const float* m = m0 + 9 * s;
      const float px = u * m[0] + w * m[1] + q * m[2] + tx;
      const float py = u * m[3] + w * m[4] + q * m[5] + ty;
      const float pz = u * m[6] + w * m[7] + q * m[8] + tz;
      r += sample_trilinear_zero(vin, nx, ny, nz, px + cx, py + cy, pz + cz);
    }
    if (nsym > 1) r *= inv_nsym;
    vout[v] = r;
    if (mask) {
      if (mask[v] > kBinaryCut) {
        acc[0] += r;
        acc[1] += (double)r * r;
        acc[2] += 1.0;
      }
      acc[3] += r;
    }
  }
  if (!mask) return;
  block_sum<4>(acc);
  if (threadIdx.x == 0) {
    double* o = part + ((int64_t)p * gridDim.x + blockIdx.x) * 4;
    for (int k = 0; k < 4; ++k) o[k] = acc[k];
  }
}

// One wavefront per particle folds its partials in a fixed order:
// stats[p] = {mean, rms of (r - mean) over the mask}; used[p] = 0 when the
// normalised particle would not be finite (BH_average3d.m:1445-1462).
__global__ void __launch_bounds__(kWave) avg_stats_kernel(const double* __restrict__ part, int nparts,
                                                         double* __restrict__ stats, int* __restrict__ used) {
  const int p = blockIdx.x;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nparts; b += kWave)
    for (int k = 0; k < 4; ++k) a[k] += part[((int64_t)p * nparts + b) * 4 + k];
  for (int k = 0; k < 4; ++k) a[k] = wave_sum(a[k]);
  if (threadIdx.x == 0) {
    const double n = a[2];
    const double mean = n > 0 ? a[0] / n : 0.0;
    const double var = n > 0 ? a[1] / n - mean * mean : 0.0;
    const double rms = sqrt(var > 0 ? var : 0.0);
    const bool ok = n > 0 && rms > 0 && isfinite(mean) && isfinite(rms) && isfinite(a[3]);
    stats[2 * p] = mean;
    stats[2 * p + 1] = rms;
    used[p] = ok ? 1 : 0;
  }
}

// For each voxel, the chunk's particles in order: the reference's FP32
// operations (- mean, ./ rms, .* mask, .* weight, +) into the half set's sum.
__global__ void __launch_bounds__(kAvgThreads) avg_accumulate_kernel(
    const float* __restrict__ rot, int64_t nv, int np, const double* __restrict__ stats,
    const int* __restrict__ used, const int* __restrict__ half, const float* __restrict__ weight,
    const float* __restrict__ mask, float* __restrict__ odd, float* __restrict__ eve) {
  const int64_t v = (int64_t)blockIdx.x * kAvgThreads + threadIdx.x;
  if (v >= nv) return;
  float so = odd[v], se = eve[v];
  const float mk = mask[v];
  for (int p = 0; p < np; ++p) {
    if (!used[p]) continue;
    float t = rot[(int64_t)p * nv + v] - (float)stats[2 * p];
    t = t / (float)stats[2 * p + 1];
    t = t * mk;
    t = t * weight[p];
    if (half[p] == 1)
      so += t;
    else
      se += t;
  }
  odd[v] = so;
  eve[v] = se;
}

// Wedge sums: wsum[half_p] += W_p * weight_p for the used particles.
__global__ void __launch_bounds__(kAvgThreads) avg_wedge_accumulate_kernel(
    const float* __restrict__ wrot, int64_t nv, int np, const int* __restrict__ used, const int* __restrict__ half,
    const float* __restrict__ weight, float* __restrict__ odd, float* __restrict__ eve) {
  const int64_t v = (int64_t)blockIdx.x * kAvgThreads + threadIdx.x;
  if (v >= nv) return;
  float so = odd[v], se = eve[v];
  for (int p = 0; p < np; ++p) {
    if (!used[p]) continue;
    const float t = wrot[(int64_t)p * nv + v] * weight[p];
    if (half[p] == 1)
      so += t;
    else
      se += t;
  }
  odd[v] = so;
  eve[v] = se;
}

void check_dims3(const int d[3], const char* what) {
  EMC_REQUIRE(d && d[0] > 0 && d[1] > 0 && d[2] > 0, "bad %s dims", what);
  EMC_REQUIRE((int64_t)d[0] * d[1] * d[2] < ((int64_t)1 << 31), "%s too large", what);
}

}  // namespace
}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_average_accumulate(emc_ctx* ctx, int n, const int window[3], const float* d_windows,
                                  const double* h_mats, const double* h_shifts, const int32_t* h_half,
                                  const double* h_weight, int nsym, const double* S_colmajor,
                                  const float* d_interp_mask, const int wedge_dims[3], const float* d_wedges,
                                  float* d_sum_odd, float* d_sum_eve, float* d_wedge_odd, float* d_wedge_eve,
                                  int32_t* h_used) {
  return guard([&] {
    use_device(ctx);
    check_dims3(window, "window");
    EMC_REQUIRE(n >= 0, "negative particle count");
    EMC_REQUIRE(nsym >= 1 && nsym <= kMaxSym, "nsym must be 1..%d", kMaxSym);
    EMC_REQUIRE(nsym == 1 || S_colmajor, "symmetry matrices required for nsym > 1");
    if (n == 0) return;
    EMC_REQUIRE(d_windows && h_mats && h_shifts && h_half && h_weight && d_interp_mask && d_sum_odd &&
                    d_sum_eve && h_used,
                "null pointer");
    if (d_wedges) {
      check_dims3(wedge_dims, "wedge");
      EMC_REQUIRE(d_wedge_odd && d_wedge_eve, "wedge sums required with wedges");
    }
    for (int p = 0; p < n; ++p) {
      EMC_REQUIRE(h_half[p] == 1 || h_half[p] == 2, "particle %d: half set must be 1 or 2", p);
      for (int k = 0; k < 3; ++k) EMC_REQUIRE(std::isfinite(h_shifts[3 * p + k]), "particle %d: shift", p);
      for (int k = 0; k < 9; ++k) EMC_REQUIRE(std::isfinite(h_mats[9 * p + k]), "particle %d: matrix", p);
    }
    const int64_t nv = (int64_t)window[0] * window[1] * window[2];
    const int64_t nw = d_wedges ? (int64_t)wedge_dims[0] * wedge_dims[1] * wedge_dims[2] : 0;
    // host tables: rows of M_p S_s (products in FP64, as emc_interp3d),
    // shifts, halves, weights
    std::vector<float> rows((size_t)n * nsym * 9), shifts((size_t)n * 3), weight((size_t)n);
    std::vector<int> half((size_t)n);
    for (int p = 0; p < n; ++p) {
      const double* M = h_mats + 9 * (int64_t)p;
      for (int s = 0; s < nsym; ++s) {
        double P[9];
        if (S_colmajor) {
          const double* S = S_colmajor + 9 * s;
          for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
              double acc = 0.0;
              for (int k = 0; k < 3; ++k) acc += M[k * 3 + r] * S[c * 3 + k];
              P[c * 3 + r] = acc;
            }
        } else {
          std::memcpy(P, M, sizeof(P));
        }
        colmajor_to_rows(P, &rows[((size_t)p * nsym + s) * 9]);
      }
      for (int k = 0; k < 3; ++k) shifts[3 * p + k] = (float)h_shifts[3 * p + k];
      weight[(size_t)p] = (float)h_weight[p];
      half[(size_t)p] = h_half[p];
    }
    // chunk the particles so that the rotated copies stay within ~1 GiB
    const int64_t per = 4 * std::max(nv, nw);
    // (and grid.y, which carries the chunk's particles, within 65535)
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(std::min(n, 65535), ((int64_t)1 << 30) / per));
    const size_t tab_bytes = rows.size() * 4 + shifts.size() * 4 + weight.size() * 4 + half.size() * 4;
    unsigned char* d_tab = static_cast<unsigned char*>(scratch(ctx, "avg_tables", tab_bytes));
    float* d_rows = reinterpret_cast<float*>(d_tab);
    float* d_shifts = d_rows + rows.size();
    float* d_weight = d_shifts + shifts.size();
    int* d_half = reinterpret_cast<int*>(d_weight + weight.size());
    // the pinned staging is shared: wait for earlier users, then copy
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    unsigned char* stage = static_cast<unsigned char*>(pinned(ctx, tab_bytes));
    std::memcpy(stage, rows.data(), rows.size() * 4);
    std::memcpy(stage + rows.size() * 4, shifts.data(), shifts.size() * 4);
    std::memcpy(stage + (rows.size() + shifts.size()) * 4, weight.data(), weight.size() * 4);
    std::memcpy(stage + (rows.size() + shifts.size() + weight.size()) * 4, half.data(), half.size() * 4);
    EMC_HIP(hipMemcpyAsync(d_tab, stage, tab_bytes, hipMemcpyHostToDevice, ctx->stream));
    float* d_rot = static_cast<float*>(scratch(ctx, "avg_rotated", (size_t)chunk * per));
    double* d_part = static_cast<double*>(scratch(ctx, "avg_partials", (size_t)chunk * kAvgParts * 4 * 8));
    double* d_stats = static_cast<double*>(scratch(ctx, "avg_stats", (size_t)n * 2 * 8));
    int* d_used = static_cast<int*>(scratch(ctx, "avg_used", (size_t)n * 4));
    const float inv_nsym = 1.0f / (float)nsym;
    for (int p0 = 0; p0 < n; p0 += chunk) {
      const int np = std::min(chunk, n - p0);
      {
        KTimer t(ctx, "avg_rotate");
        hipLaunchKernelGGL(avg_rotate_kernel, dim3(kAvgParts, np), dim3(kAvgThreads), 0, ctx->stream,
                           d_windows + (int64_t)p0 * nv, window[0], window[1], window[2],
                           d_rows + (size_t)p0 * nsym * 9, nsym, inv_nsym, d_shifts + 3 * p0, d_interp_mask, d_rot,
                           d_part);
        EMC_HIP(hipGetLastError());
      }
      hipLaunchKernelGGL(avg_stats_kernel, dim3(np), dim3(kWave), 0, ctx->stream, d_part, kAvgParts,
                         d_stats + 2 * p0, d_used + p0);
      EMC_HIP(hipGetLastError());
      {
        KTimer t(ctx, "avg_accumulate");
        hipLaunchKernelGGL(avg_accumulate_kernel, dim3(ceil_div(nv, kAvgThreads)), dim3(kAvgThreads), 0,
                           ctx->stream, d_rot, nv, np, d_stats + 2 * p0, d_used + p0, d_half + p0,
                           d_weight + p0, d_interp_mask, d_sum_odd, d_sum_eve);
        EMC_HIP(hipGetLastError());
      }
      if (d_wedges) {
        hipLaunchKernelGGL(avg_rotate_kernel, dim3(kAvgParts, np), dim3(kAvgThreads), 0, ctx->stream,
                           d_wedges + (int64_t)p0 * nw, wedge_dims[0], wedge_dims[1], wedge_dims[2],
                           d_rows + (size_t)p0 * nsym * 9, nsym, inv_nsym, nullptr, nullptr, d_rot, nullptr);
        EMC_HIP(hipGetLastError());
        hipLaunchKernelGGL(avg_wedge_accumulate_kernel, dim3(ceil_div(nw, kAvgThreads)), dim3(kAvgThreads), 0,
                           ctx->stream, d_rot, nw, np, d_used + p0, d_half + p0, d_weight + p0, d_wedge_odd,
                           d_wedge_eve);
        EMC_HIP(hipGetLastError());
      }
    }
    std::vector<int> used((size_t)n);
    EMC_HIP(hipMemcpyAsync(used.data(), d_used, (size_t)n * 4, hipMemcpyDeviceToHost, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    for (int p = 0; p < n; ++p) h_used[p] = used[(size_t)p];
  });
}

}  // extern "C"
