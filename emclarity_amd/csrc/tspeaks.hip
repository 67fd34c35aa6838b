// Template-search peak extraction (alignment/BH_templateSearch3d_2.m:1100-1275)
// and the soft sphere mask it erases with (masking/BH_mask3d.m, 'sphere').
//
// The reference repeats, while fewer than 2*peakThreshold peaks are taken,
// fewer than max_tries maxima were visited and the maximum exceeds highThr:
//   [MAX, coord] = max(mag(:))            (first maximum, x fastest)
//   if Ang(coord) > 0:
//     c = ind2sub(coord); centre of mass of mag over the 3^3 box around c;
//     peak = (c + com - rmDim, ANGLE_LIST(Ang(coord)), MAX);
//     rmMask = interp3d(removalMask, angles, [0,0,0], 'forward');
//     mag(c +- rmDim) .*= (rmMask < maskCutOff)
//   else: mag(coord) = 0
// Here the map's maxima are kept per 8^3 tile on the device, so a step is an
// argmax over the tile keys, one small kernel for the centre of mass, the
// rotation of the removal mask and the erase, and a refresh of the tiles the
// erase touched; one 32-byte result copy per step.
#include <algorithm>
#include <cmath>
#include <vector>

#include "ops.hpp"

namespace emc {

void define_matrix(const double ang_deg[3], int convention, int direction, double out_colmajor[9]);

constexpr int kPeakTile = 8;

// key of each tile in the tile box [t0, t0 + tn): max argmax_key over its voxels
__global__ void __launch_bounds__(256) tile_keys_kernel(const float* __restrict__ mag, int nx, int ny, int nz,
                                                       int tx, int ty, int t0x, int t0y, int t0z, int tnx, int tny,
                                                       unsigned long long* __restrict__ keys) {
  const int b = blockIdx.x;
  const int ix = t0x + b % tnx, iy = t0y + (b / tnx) % tny, iz = t0z + b / (tnx * tny);
  unsigned long long k = 0ull;
  for (int v = threadIdx.x; v < kPeakTile * kPeakTile * kPeakTile; v += blockDim.x) {
    const int x = ix * kPeakTile + v % kPeakTile, y = iy * kPeakTile + (v / kPeakTile) % kPeakTile,
              z = iz * kPeakTile + v / (kPeakTile * kPeakTile);
    if (x >= nx || y >= ny || z >= nz) continue;
    const int64_t i = ((int64_t)z * ny + y) * nx + x;
    const unsigned long long kk = argmax_key(mag[i], (uint32_t)i);
    k = kk > k ? kk : k;
  }
  k = block_max_u64(k);
  if (threadIdx.x == 0) keys[((int64_t)iz * ty + iy) * tx + ix] = k;
}

__global__ void __launch_bounds__(256) keys_max_kernel(const unsigned long long* __restrict__ keys, int64_t n,
                                                      unsigned long long* __restrict__ out) {
  unsigned long long k = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    k = keys[i] > k ? keys[i] : k;
  k = block_max_u64(k);
  if (threadIdx.x == 0) out[blockIdx.x] = k;
}

// Result of one step: the maximum's key, its angle id and the centre of mass
// of the 3^3 box around it.
struct PeakStep {
  unsigned long long key;
  int ang;
  int pad;
  double com[3];
};

// bd: half edge of the centre-of-mass box, 1 + round(log(nPeaks)) (:1156);
// box_mag / box_ang (or null): that box's map values and angle ids, x
// fastest, 0 outside the map, for the extra peaks.
__global__ void peak_step_kernel(const unsigned long long* __restrict__ part, int nb, const float* __restrict__ mag,
                                 const int32_t* __restrict__ ang, int nx, int ny, int nz, int bd,
                                 PeakStep* __restrict__ out, float* __restrict__ box_mag,
                                 int32_t* __restrict__ box_ang) {
  __shared__ uint32_t s_i;
  unsigned long long k = 0ull;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) k = part[i] > k ? part[i] : k;
  k = block_max_u64(k);
  if (threadIdx.x == 0) {
    uint32_t i = key_index(k);
    if ((int64_t)i >= (int64_t)nx * ny * nz) i = 0;  // keys always carry a voxel index; stay in bounds
    s_i = i;
    const int x = (int)(i % (uint32_t)nx), y = (int)((i / (uint32_t)nx) % (uint32_t)ny),
              z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
    double s = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int dz = -bd; dz <= bd; ++dz)
      for (int dy = -bd; dy <= bd; ++dy)
        for (int dx = -bd; dx <= bd; ++dx) {
          const int xx = x + dx, yy = y + dy, zz = z + dz;
          const double v = (xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz)
                               ? (double)mag[((int64_t)zz * ny + yy) * nx + xx]
                               : 0.0;
          s += v;
          cx += v * dx;
          cy += v * dy;
          cz += v * dz;
        }
    out->key = k;
    out->ang = ang[i];
    out->com[0] = cx / s;
    out->com[1] = cy / s;
    out->com[2] = cz / s;
  }
  if (box_mag) {  // uniform
    __syncthreads();
    const uint32_t i = s_i;
    const int x = (int)(i % (uint32_t)nx), y = (int)((i / (uint32_t)nx) % (uint32_t)ny),
              z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
    const int e = 2 * bd + 1;
    for (int v = threadIdx.x; v < e * e * e; v += blockDim.x) {
      const int xx = x + v % e - bd, yy = y + (v / e) % e - bd, zz = z + v / (e * e) - bd;
      const bool in = xx >= 0 && xx < nx && yy >= 0 && yy < ny && zz >= 0 && zz < nz;
      const int64_t j = ((int64_t)zz * ny + yy) * nx + xx;
      box_mag[v] = in ? mag[j] : 0.0f;
      box_ang[v] = in ? ang[j] : 0;
    }
  }
}

// mag(box around c) .*= (rot < cutoff); box edge 2 rm + 1, voxels outside
// the map are skipped.
__global__ void erase_kernel(float* __restrict__ mag, int nx, int ny, int nz, int cx, int cy, int cz, int rm,
                             const float* __restrict__ rot, float cutoff) {
  const int e = 2 * rm + 1;
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= e * e * e) return;
  const int bx = v % e, by = (v / e) % e, bz = v / (e * e);
  const int x = cx - rm + bx, y = cy - rm + by, z = cz - rm + bz;
  if (x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz) return;
  if (!(rot[v] < cutoff)) mag[((int64_t)z * ny + y) * nx + x] = 0.0f;
}

__global__ void zero_voxel_kernel(float* __restrict__ mag, int64_t i) { mag[i] = 0.0f; }

// ------------------------------------------------------ BH_mask3d('sphere')

// shells (masking/BH_mask3d.m:126-150): core (ell <= 1) = 1, then for
// s = 1..6 w = w (1 - taper_s) + [inside radius + s] taper_s with
// taper_s = 0.5 + 0.5 cos(s pi / 7); coordinates about floor(N/2)
__global__ void sphere_shells_kernel(int nx, int ny, int nz, double rx, double ry, double rz,
                                     float* __restrict__ w, uint8_t* __restrict__ core) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const double gx = x - nx / 2, gy = y - ny / 2, gz = z - nz / 2;
  auto ell = [&](double s) {
    const double a = gx / (rx + s), b = gy / (ry + s), c = gz / (rz + s);
    return a * a + b * b + c * c;
  };
  const bool in = ell(0.0) <= 1.0;
  float m = in ? 1.0f : 0.0f;
  for (int s = 1; s <= 6; ++s) {
    const float tp = (float)(0.5 + 0.5 * cospi((double)s / 7.0));
    const float inside = ell((double)s) <= 1.0 ? 1.0f : 0.0f;
    m = m + (inside - m) * tp;
  }
  const int64_t i = ((int64_t)z * ny + y) * nx + x;
  w[i] = m;
  core[i] = in ? 1 : 0;
}

// convn(.., 'same') along one axis with the 5-tap normalised Gaussian of
// BH_multi_gaussian3d(5, 0.5) (separable), zero outside
__global__ void gauss5_pass_kernel(const float* __restrict__ in, float* __restrict__ out, int nx, int ny, int nz,
                                   int axis, float w0, float w1, float w2) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
  if (x >= nx) return;
  const int N = axis == 0 ? nx : (axis == 1 ? ny : nz);
  const int i = axis == 0 ? x : (axis == 1 ? y : z);
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)nx : (int64_t)nx * ny);
  const int64_t idx = ((int64_t)z * ny + y) * nx + x;
  const float w[5] = {w2, w1, w0, w1, w2};
  float acc = 0.0f;
  for (int m = 0; m < 5; ++m) {
    const int j = i + 2 - m;
    if (j >= 0 && j < N) acc = fmaf(w[m], in[idx + (int64_t)(j - i) * stride], acc);
  }
  out[idx] = acc;
}

__global__ void __launch_bounds__(256) max_partial_kernel(const float* __restrict__ d, int64_t n,
                                                         float* __restrict__ part) {
  float m = -__builtin_inff();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, d[i]);
  m = -block_min(-m);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

__global__ void sphere_finish_kernel(float* __restrict__ w, const uint8_t* __restrict__ core, int64_t n,
                                     const float* __restrict__ part, int nb) {
  float m = -__builtin_inff();
  for (int b = 0; b < nb; ++b) m = fmaxf(m, part[b]);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    w[i] = core[i] ? 1.0f : w[i] / m;
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_mask3d_sphere(emc_ctx* ctx, const int dims[3], const double radius[3], float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(dims && radius && d_out, "null pointer");
    EMC_REQUIRE(dims[0] >= 14 && dims[1] >= 14 && dims[2] >= 14, "sphere mask edges must be >= 14");
    EMC_REQUIRE(radius[0] > 0 && radius[1] > 0 && radius[2] > 0, "radius must be positive");
    const int nx = dims[0], ny = dims[1], nz = dims[2];
    const int64_t nv = (int64_t)nx * ny * nz;
    float* a = (float*)scratch(ctx, "msk_a", sizeof(float) * nv);
    float* b = (float*)scratch(ctx, "msk_b", sizeof(float) * nv);
    uint8_t* core = (uint8_t*)scratch(ctx, "msk_core", nv);
    const dim3 grid(ceil_div(nx, 64), ny, nz), block(64);
    hipLaunchKernelGGL(sphere_shells_kernel, grid, block, 0, ctx->stream, nx, ny, nz, radius[0], radius[1],
                       radius[2], a, core);
    // BH_multi_gaussian3d(5, 0.5): exp(-0.5 (d / 0.5)^2), d = -2..2, normalised
    double g[3], s = 0.0;
    for (int d = 0; d < 3; ++d) g[d] = std::exp(-0.5 * (d / 0.5) * (d / 0.5));
    s = g[0] + 2 * g[1] + 2 * g[2];
    const float w0 = (float)(g[0] / s), w1 = (float)(g[1] / s), w2 = (float)(g[2] / s);
    hipLaunchKernelGGL(gauss5_pass_kernel, grid, block, 0, ctx->stream, (const float*)a, b, nx, ny, nz, 0, w0, w1, w2);
    hipLaunchKernelGGL(gauss5_pass_kernel, grid, block, 0, ctx->stream, (const float*)b, a, nx, ny, nz, 1, w0, w1, w2);
    hipLaunchKernelGGL(gauss5_pass_kernel, grid, block, 0, ctx->stream, (const float*)a, b, nx, ny, nz, 2, w0, w1, w2);
    const int nb = 256;
    float* part = (float*)scratch(ctx, "msk_part", sizeof(float) * nb);
    hipLaunchKernelGGL(max_partial_kernel, dim3(nb), dim3(256), 0, ctx->stream, (const float*)b, nv, part);
    hipLaunchKernelGGL(sphere_finish_kernel, dim3(256), dim3(256), 0, ctx->stream, b, (const uint8_t*)core, nv,
                       (const float*)part, nb);
    // final BH_padZeros3d(mask, 0, 0, 'singleTaper') (masking/BH_mask3d.m:462-466)
    const int zero[3] = {0, 0, 0};
    PadArgs pa = pad_args(dims, zero, zero, 0.0f, 1);
    launch_pad(ctx, b, d_out, pa, nullptr);
    EMC_HIP(hipGetLastError());
  });
}

// The extra peaks of one taken peak (:1198-1240), on the host from the
// centre-of-mass box: possible = box map values with the peak's angle id
// zeroed; per extra peak, if any value is non-zero: cAng = first maximum,
// iSNR = mean of the values of its id (then zeroed), and unless that id or
// Ang(cAng) is <= 0 its angles are ANGLE_LIST(Ang(cAng)) -- Ang indexed
// with the box's linear index, as the reference does -- else, and when every
// value is zero, random: angles = peak position (times the sampling rate) +
// nRandom^2 x the caller's normals, nRandom = 2, 3, .. 10, with iSNR = 0.75
// x the mean score of the peaks before it (only where a value was non-zero).
// Every extra peak has the position of the taken one.
static void extra_peaks(double* row, int n_peaks, int box, const float* box_mag, const int32_t* box_ang,
                        const int32_t* ang_head, int top_id, const double* angle_list, int n_angles,
                        const double* normals) {
  std::vector<float> pa(box_mag, box_mag + box);
  for (int v = 0; v < box; ++v)
    if (box_ang[v] == top_id) pa[v] = 0.0f;
  double snr = 0.0;
  int n_random = 2;
  for (int ip = 2; ip <= n_peaks; ++ip) {
    bool random = false;
    double a3[3] = {0.0, 0.0, 0.0};
    bool any = false;
    for (int v = 0; v < box && !any; ++v) any = pa[v] != 0.0f;
    if (any) {
      int c = 0;  // first maximum over the whole box, zeros included
      for (int v = 1; v < box; ++v)
        if (pa[v] > pa[c]) c = v;
      const int top = box_ang[c];
      double sum = 0.0;
      int cnt = 0;
      for (int v = 0; v < box; ++v)
        if (box_ang[v] == top) {
          sum += pa[v];
          ++cnt;
          pa[v] = 0.0f;
        }
      snr = (double)(float)(sum / cnt);  // the mean of single values
      const int head = ang_head[c];
      if (top <= 0 || head <= 0) {
        random = true;
        double m = 0.0;
        for (int q = 1; q < ip; ++q) m += row[10 * q - 1];
        snr = m / (ip - 1) * 0.75;
      } else {
        EMC_REQUIRE(head <= n_angles, "angle id %d outside the angle list (%d)", head, n_angles);
        for (int d = 0; d < 3; ++d) a3[d] = angle_list[3 * (head - 1) + d];
      }
    } else {
      random = true;
    }
    if (random) {
      const double* z = normals + 3 * (ip - 2);
      for (int d = 0; d < 3; ++d) a3[d] = z[d] * (double)(n_random * n_random) + row[d];
      if (n_random < 10) ++n_random;
    }
    double* sub = row + 10 * (ip - 1);
    for (int d = 0; d < 3; ++d) {
      sub[d] = row[d];
      sub[3 + d] = a3[d];
    }
    sub[9] = snr;
  }
}

// The loop for both entry points: peakMat rows of 10 n_peaks doubles.
static void tsearch_peaks_impl(emc_ctx* ctx, float* d_mag, const int32_t* d_ang, const int dims[3],
                               const float* d_removal_mask, int rm_dim, const double* angle_list, int n_angles,
                               int convention, double mask_cutoff, double high_thr, int max_peaks, int max_tries,
                               int n_peaks, double sampling_rate, const double* normals, double* h_rows,
                               int* n_found) {
  use_device(ctx);
  EMC_REQUIRE(d_mag && d_ang && dims && d_removal_mask && angle_list && h_rows && n_found, "null pointer");
  EMC_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
  EMC_REQUIRE((int64_t)dims[0] * dims[1] * dims[2] < ((int64_t)1 << 32), "map too large for 32-bit indices");
  EMC_REQUIRE(rm_dim >= 0 && rm_dim < 512, "bad removal mask radius");
  EMC_REQUIRE(n_angles > 0 && max_peaks >= 0 && max_tries >= 0, "bad counts");
  EMC_REQUIRE(n_peaks >= 1 && n_peaks <= 1000, "n_peaks must be 1..1000");
  EMC_REQUIRE(n_peaks == 1 || normals, "null normals");
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  const int64_t nvox = (int64_t)nx * ny * nz;
  const int bd = 1 + (int)std::lround(std::log((double)n_peaks));
  const int be = 2 * bd + 1, box = be * be * be;
  const int tx = ceil_div(nx, kPeakTile), ty = ceil_div(ny, kPeakTile), tz = ceil_div(nz, kPeakTile);
  const int64_t nt = (int64_t)tx * ty * tz;
  auto* keys = (unsigned long long*)scratch(ctx, "pk_keys", sizeof(unsigned long long) * nt);
  const int nb = 256;
  auto* part = (unsigned long long*)scratch(ctx, "pk_part", sizeof(unsigned long long) * nb);
  auto* step = (PeakStep*)scratch(ctx, "pk_step", sizeof(PeakStep));
  const int e = 2 * rm_dim + 1;
  float* rot = (float*)scratch(ctx, "pk_rot", sizeof(float) * (size_t)e * e * e);
  float* d_box = nullptr;
  int32_t* d_box_ang = nullptr;
  std::vector<int32_t> ang_head;
  if (n_peaks > 1) {
    EMC_REQUIRE(box <= nvox, "map smaller than the %d^3 centre-of-mass box", be);
    d_box = (float*)scratch(ctx, "pk_box", sizeof(float) * box);
    d_box_ang = (int32_t*)scratch(ctx, "pk_box_ang", sizeof(int32_t) * box);
    ang_head.resize(box);  // Ang(1 .. box), read by the extra peaks
    EMC_HIP(hipMemcpyAsync(ang_head.data(), d_ang, sizeof(int32_t) * box, hipMemcpyDeviceToHost, ctx->stream));
  }
  const size_t hb = sizeof(PeakStep) + (n_peaks > 1 ? (sizeof(float) + sizeof(int32_t)) * box : 0);
  auto* h = (PeakStep*)pinned(ctx, hb);
  float* h_box = (float*)(h + 1);
  int32_t* h_box_ang = (int32_t*)(h_box + box);
  EMC_HIP(hipStreamSynchronize(ctx->stream));  // the pinned buffer may be in use
  auto refresh = [&](int x0, int y0, int z0, int x1, int y1, int z1) {  // voxel box, inclusive
    const int a0 = std::max(0, x0 / kPeakTile), b0 = std::max(0, y0 / kPeakTile), c0 = std::max(0, z0 / kPeakTile);
    const int a1 = std::min(tx - 1, x1 / kPeakTile), b1 = std::min(ty - 1, y1 / kPeakTile),
              c1 = std::min(tz - 1, z1 / kPeakTile);
    if (a1 < a0 || b1 < b0 || c1 < c0) return;
    const int tnx = a1 - a0 + 1, tny = b1 - b0 + 1, tnz = c1 - c0 + 1;
    hipLaunchKernelGGL(tile_keys_kernel, dim3(tnx * tny * tnz), dim3(256), 0, ctx->stream, (const float*)d_mag,
                       nx, ny, nz, tx, ty, a0, b0, c0, tnx, tny, keys);
  };
  auto next = [&]() {
    hipLaunchKernelGGL(keys_max_kernel, dim3(nb), dim3(256), 0, ctx->stream, (const unsigned long long*)keys, nt,
                       part);
    hipLaunchKernelGGL(peak_step_kernel, dim3(1), dim3(256), 0, ctx->stream, (const unsigned long long*)part, nb,
                       (const float*)d_mag, d_ang, nx, ny, nz, bd, step, d_box, d_box_ang);
    EMC_HIP(hipGetLastError());
    EMC_HIP(hipMemcpyAsync(h, step, sizeof(PeakStep), hipMemcpyDeviceToHost, ctx->stream));
    if (d_box) {
      EMC_HIP(hipMemcpyAsync(h_box, d_box, sizeof(float) * box, hipMemcpyDeviceToHost, ctx->stream));
      EMC_HIP(hipMemcpyAsync(h_box_ang, d_box_ang, sizeof(int32_t) * box, hipMemcpyDeviceToHost, ctx->stream));
    }
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    return *h;
  };
  refresh(0, 0, 0, nx - 1, ny - 1, nz - 1);
  PeakStep s = next();
  int n = 0, tries = 0;
  const int64_t w = 10 * (int64_t)n_peaks;
  while (n < max_peaks && tries < max_tries && (double)key_value(s.key) > high_thr) {
    ++tries;
    uint32_t i = key_index(s.key);
    if ((int64_t)i >= nvox) i = 0;
    const int x = (int)(i % (uint32_t)nx), y = (int)((i / (uint32_t)nx) % (uint32_t)ny),
              z = (int)(i / ((uint32_t)nx * (uint32_t)ny));
    if (s.ang > 0) {
      EMC_REQUIRE(s.ang <= n_angles, "angle id %d outside the angle list (%d)", s.ang, n_angles);
      const double* ang = angle_list + 3 * (s.ang - 1);
      double* o = h_rows + w * n;
      std::fill(o, o + w, 0.0);
      // MATLAB 1-based centre plus centre of mass, minus the padding
      o[0] = sampling_rate * (x + 1 + s.com[0] - rm_dim);
      o[1] = sampling_rate * (y + 1 + s.com[1] - rm_dim);
      o[2] = sampling_rate * (z + 1 + s.com[2] - rm_dim);
      o[3] = ang[0];
      o[4] = ang[1];
      o[5] = ang[2];
      o[9] = (double)key_value(s.key);
      if (n_peaks > 1)
        extra_peaks(o, n_peaks, box, h_box, h_box_ang, ang_head.data(), s.ang, angle_list, n_angles,
                    normals + 3 * (n_peaks - 1) * (int64_t)n);
      // removal mask rotated by the peak's angles ('forward', no shift)
      double M[9];
      define_matrix(ang, convention, 0, M);
      float rows[9];
      colmajor_to_rows(M, rows);
      const int md[3] = {e, e, e};
      const float t0[3] = {0.0f, 0.0f, 0.0f};
      XformArgs xa = xform_args(md, rows, 1, t0, 1);
      launch_xform(ctx, d_removal_mask, rot, xa);
      hipLaunchKernelGGL(erase_kernel, dim3(ceil_div((int64_t)e * e * e, 256)), dim3(256), 0, ctx->stream, d_mag,
                         nx, ny, nz, x, y, z, rm_dim, (const float*)rot, (float)mask_cutoff);
      refresh(x - rm_dim, y - rm_dim, z - rm_dim, x + rm_dim, y + rm_dim, z + rm_dim);
      ++n;
    } else {
      hipLaunchKernelGGL(zero_voxel_kernel, dim3(1), dim3(1), 0, ctx->stream, d_mag, (int64_t)i);
      refresh(x, y, z, x, y, z);
    }
    s = next();
  }
  *n_found = n;
}

emc_status emc_tsearch_peaks(emc_ctx* ctx, float* d_mag, const int32_t* d_ang, const int dims[3],
                             const float* d_removal_mask, int rm_dim, const double* angle_list, int n_angles,
                             int convention, double mask_cutoff, double high_thr, int max_peaks, int max_tries,
                             double* h_peaks, int* n_found) {
  return guard([&] {
    EMC_REQUIRE(h_peaks && max_peaks >= 0, "null pointer");
    std::vector<double> rows(10 * (size_t)std::max(max_peaks, 1));
    tsearch_peaks_impl(ctx, d_mag, d_ang, dims, d_removal_mask, rm_dim, angle_list, n_angles, convention,
                       mask_cutoff, high_thr, max_peaks, max_tries, 1, 1.0, nullptr, rows.data(), n_found);
    for (int k = 0; k < *n_found; ++k) {
      for (int d = 0; d < 6; ++d) h_peaks[7 * (int64_t)k + d] = rows[10 * (size_t)k + d];
      h_peaks[7 * (int64_t)k + 6] = rows[10 * (size_t)k + 9];
    }
  });
}

emc_status emc_tsearch_peaks_multi(emc_ctx* ctx, float* d_mag, const int32_t* d_ang, const int dims[3],
                                   const float* d_removal_mask, int rm_dim, const double* angle_list, int n_angles,
                                   int convention, double mask_cutoff, double high_thr, int max_peaks,
                                   int max_tries, int n_peaks, double sampling_rate, const double* normals,
                                   double* h_peak_mat, int* n_found) {
  return guard([&] {
    tsearch_peaks_impl(ctx, d_mag, d_ang, dims, d_removal_mask, rm_dim, angle_list, n_angles, convention,
                       mask_cutoff, high_thr, max_peaks, max_tries, n_peaks, sampling_rate, normals, h_peak_mat,
                       n_found);
  });
}

}  // extern "C"
