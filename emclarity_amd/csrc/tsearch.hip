// Template search over one tomogram chunk: the loop body of
// BH_templateSearch3d_2 (reference: alignment/BH_templateSearch3d_2.m:632-641
// swapQuadrants, :713-764 rotate/pad/correlate/normalise, :851-887 max/argmax
// update).
//
// Per orientation: rotate the template into a template-sized block, form
// the padded template's half spectrum by direct DFTs over that block fused
// with the product by the pre-centred tomogram spectrum (or, for templates
// over 64 voxels, rotate straight into the zero-padded chunk and R2C it),
// C2R, std over the valid area (FP64 sums), then a fused per-voxel
// max/argmax update of the score and angle maps.
// Orientations are processed in blocks: the rotate, FFTs and multiply are
// batched over the block; the updates are applied in orientation order, so
// the strict '<' tie rule of the reference is preserved.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "ops.hpp"

namespace emc {

void define_matrix(const double ang_deg[3], int convention, int direction, double out_colmajor[9]);

// tomoFou = swapQuadrants .* R2C(chunk) (:632-641, :713-714). swapQuadrants on
// the half grid: exp(-2 pi i sum_k c_k (n_k/2 + 1)) with the centred
// coordinate c_k = (j_k - n_k/2)/n_k of the array position j_k; the phase is
// reduced exactly in integers before the sincos.
__global__ void swapq_kernel(float2* __restrict__ F, int nx, int ny, int nz) {
  const int cx = nx / 2 + 1;
  const int64_t nh = (int64_t)cx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    split3(i, cx, ny, x, y, z);
    auto frac = [](int j, int n) {
      int64_t a = ((int64_t)(j - n / 2) * (n / 2 + 1)) % n;
      if (a < 0) a += n;
      return (double)a / (double)n;
    };
    double f = frac(x, nx) + frac(y, ny) + frac(z, nz);
    f -= floor(f);
    double s, c;
    sincospi(-2.0 * f, &s, &c);
    const float2 v = F[i];
    F[i] = make_float2((float)(v.x * c - v.y * s), (float)(v.x * s + v.y * c));
  }
}

// X_b = T0 * conj(T_b) for a block of spectra (no scaling: mexFFT does not
// scale, :759-761)
__global__ void conj_mul_batch_kernel(const float2* __restrict__ T0, float2* __restrict__ T,
                                      int64_t nh) {
  float2* Tb = T + (int64_t)blockIdx.y * nh;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float2 a = T0[i], b = Tb[i];
    Tb[i] = make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
  }
}

// ------------------------------------------------ pruned template spectrum
// The zero-padded template's R2C has non-zero samples only in the
// t_x x t_y x t_z block at pad_low, so its half spectrum is the separable
// sum over that block,
//   F(k) = sum_z e_z(kz, z) sum_y e_y(ky, y) sum_x e_x(kx, x) t(x, y, z),
//   e_a(k, j) = exp(-2 pi i k (p_a + j) / N_a),
// taken as three direct DFTs over the block: along x for the block's rows,
// along y for the block's planes, and along z fused with the product by the
// tomogram spectrum (tomoFou .* conj(tempFou), :756-759). No padded volume
// is written and no 3D R2C runs per orientation.
constexpr int kTMax = 64;  // template edge limit of the pruned path

// e(k, j) = exp(-2 pi i k (p + j) / n) for k < nk, j < kTMax (zero for j >= jt);
// the phase is reduced exactly in integers before the FP64 sincos
__global__ void dft_twiddle_kernel(float2* __restrict__ e, int nk, int jt, int p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk * kTMax) return;
  const int k = i / kTMax, j = i % kTMax;
  if (j >= jt) {
    e[i] = make_float2(0.0f, 0.0f);
    return;
  }
  const int64_t m = ((int64_t)k * (p + j)) % n;
  double sn, cs;
  sincospi(-2.0 * (double)m / (double)n, &sn, &cs);
  e[i] = make_float2((float)cs, (float)sn);
}

// X1[b][z][kx][y] = sum_x t[b][z][y][x] e_x(kx, x) for z < tz, kx in this
// block's chunk, y < YP (rows y >= ty are zero). One thread per (z, y) row,
// the row in registers; the twiddle e_x(kx, .) is the same for the whole
// wave and consecutive lanes write consecutive y.
template <int XP, int YP>
__global__ void __launch_bounds__(256) tdft_x_kernel(const float* __restrict__ t, int tx, int ty, int tz, int cx,
                                                     const float2* __restrict__ ex, int kx_chunk,
                                                     float2* __restrict__ X1) {
  const int b = blockIdx.y;
  const int rowi = blockIdx.x * blockDim.x + threadIdx.x;  // z * YP + y
  if (rowi >= tz * YP) return;
  const int z = rowi / YP, y = rowi % YP;
  float v[XP];
  const int yy = y < ty ? y : 0;  // padding rows read row 0 and keep zeros
  const float* row = t + (int64_t)b * tx * ty * tz + ((int64_t)z * ty + yy) * tx;
#pragma unroll
  for (int x = 0; x < XP; ++x) {
    const float r = x < tx ? row[x < tx ? x : 0] : 0.0f;
    v[x] = y < ty ? r : 0.0f;
  }
  const int kx0 = blockIdx.z * kx_chunk, kx1 = min(cx, kx0 + kx_chunk);
  float2* ob = X1 + (int64_t)b * tz * cx * YP + (int64_t)z * cx * YP + y;
  for (int kx = kx0; kx < kx1; ++kx) {
    const float2* e = ex + (int64_t)kx * kTMax;
    float sr = 0.0f, si = 0.0f;
#pragma unroll
    for (int x = 0; x < XP; ++x) {
      const float2 w = e[x];
      sr = fmaf(v[x], w.x, sr);
      si = fmaf(v[x], w.y, si);
    }
    ob[(int64_t)kx * YP] = make_float2(sr, si);
  }
}

// X2[b][z][ky][kx] = sum_y X1[b][z][kx][y] e_y(ky, y) for z < ZP (planes
// z >= tz are zero), ky in this block's chunk, kx < cx; one thread per (z, kx)
template <int YP>
__global__ void __launch_bounds__(256) tdft_y_kernel(const float2* __restrict__ X1, int tz, int zp, int ny, int cx,
                                                     const float2* __restrict__ ey, int ky_chunk,
                                                     float2* __restrict__ X2) {
  const int b = blockIdx.y;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= zp * cx) return;
  const int z = col / cx, kx = col % cx;
  const int ky0 = blockIdx.z * ky_chunk, ky1 = min(ny, ky0 + ky_chunk);
  float2* ob = X2 + (int64_t)b * zp * ny * cx + (int64_t)z * ny * cx + kx;
  if (z >= tz) {
    for (int ky = ky0; ky < ky1; ++ky) ob[(int64_t)ky * cx] = make_float2(0.0f, 0.0f);
    return;
  }
  const float2* ib = X1 + (int64_t)b * tz * cx * YP + ((int64_t)z * cx + kx) * YP;
  float2 a[YP];
#pragma unroll
  for (int y = 0; y < YP; ++y) a[y] = ib[y];
  for (int ky = ky0; ky < ky1; ++ky) {
    const float2* e = ey + (int64_t)ky * kTMax;
    float sr = 0.0f, si = 0.0f;
#pragma unroll
    for (int y = 0; y < YP; ++y) {
      const float2 w = e[y];
      sr = fmaf(a[y].x, w.x, fmaf(-a[y].y, w.y, sr));
      si = fmaf(a[y].x, w.y, fmaf(a[y].y, w.x, si));
    }
    ob[(int64_t)ky * cx] = make_float2(sr, si);
  }
}

// out[b][kz][ky][kx] = tomo[kz][ky][kx] * conj(sum_z X2[b][z][ky][kx] e_z(kz, z))
// for kz in this block's chunk; one thread per (ky, kx) column
template <int ZP>
__global__ void __launch_bounds__(256) tdft_z_mul_kernel(const float2* __restrict__ X2,
                                                         const float2* __restrict__ tomo, int nz, int ny, int cx,
                                                         const float2* __restrict__ ez, int kz_chunk,
                                                         float2* __restrict__ out) {
  const int b = blockIdx.y;
  const int64_t ncol = (int64_t)ny * cx;
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncol) return;
  const float2* xb = X2 + (int64_t)b * ZP * ncol + col;
  float2 a[ZP];
#pragma unroll
  for (int z = 0; z < ZP; ++z) a[z] = xb[(int64_t)z * ncol];
  const int kz0 = blockIdx.z * kz_chunk, kz1 = min(nz, kz0 + kz_chunk);
  float2* ob = out + (int64_t)b * nz * ncol + col;
  for (int kz = kz0; kz < kz1; ++kz) {
    const float2* e = ez + (int64_t)kz * kTMax;
    float fr = 0.0f, fi = 0.0f;
#pragma unroll
    for (int z = 0; z < ZP; ++z) {
      const float2 w = e[z];
      fr = fmaf(a[z].x, w.x, fmaf(-a[z].y, w.y, fr));
      fi = fmaf(a[z].x, w.y, fmaf(a[z].y, w.x, fi));
    }
    const float2 tv = tomo[(int64_t)kz * ncol + col];
    ob[(int64_t)kz * ncol] = make_float2(tv.x * fr + tv.y * fi, tv.y * fr - tv.x * fi);
  }
}

// The same z stage on the matrix cores. For one orientation b, a tile of 32
// columns (ky, kx) and a tile of 32 kz, the product of the 32 x ZP twiddle
// block e_z(kz, z) by the ZP x 32 block of partial DFTs X2[b][z][col] is a
// complex matrix product: four FP32 MFMA products per pair of z
// (Re = Er Xr - Ei Xi, Im = Er Xi + Ei Xr) with v_mfma_f32_32x32x2f32, then
// each lane multiplies its 16 outputs by the tomogram spectrum and writes
// them (per output register, lanes 0-31 and 32-63 each write 32 consecutive
// columns of one kz row: 256-byte segments).
// Operand layout of the 32x32x2 FP32 MFMA (one wave): A (32 x 2), lane l
// holds A[l % 32][l / 32]; B (2 x 32), lane l holds B[l / 32][l % 32];
// C (32 x 32), lane l holds C[8 (i / 4) + 4 (l / 32) + i % 4][l % 32] in
// its register i = 0..15.
// Block: 4 waves, one kz tile each (grid z covers kz in steps of 128),
// grid x the column tiles, grid y the orientations of the block.
typedef float floatx16 __attribute__((ext_vector_type(16)));

template <int ZP>
__global__ void __launch_bounds__(256) tdft_z_mul_mfma_kernel(const float2* __restrict__ X2,
                                                              const float2* __restrict__ tomo, int nz, int ny,
                                                              int cx, const float2* __restrict__ ez,
                                                              float2* __restrict__ out) {
  const int b = blockIdx.y;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int kt = blockIdx.z * 4 + wave;  // kz tile of this wave
  if (kt * 32 >= nz) return;             // uniform per wave; no barrier below
  const int64_t ncol = (int64_t)ny * cx;
  const int64_t col = (int64_t)blockIdx.x * 32 + (lane % 32);
  const bool col_ok = col < ncol;
  const int kh = lane / 32;  // which z of each pair this lane supplies
  // B operand: X2[b][2 s + kh][col] for s < ZP / 2
  const float2* xb = X2 + (int64_t)b * ZP * ncol + (col_ok ? col : 0);
  float2 xv[ZP / 2];
#pragma unroll
  for (int s2 = 0; s2 < ZP / 2; ++s2)
    xv[s2] = col_ok ? xb[(int64_t)(2 * s2 + kh) * ncol] : make_float2(0.0f, 0.0f);
  // A operand: e_z(kz, 2 s + kh) for the lane's row kz of the tile
  const int kza = kt * 32 + (lane % 32);
  const bool kz_ok = kza < nz;
  const float2* ea = ez + (int64_t)(kz_ok ? kza : 0) * kTMax + kh;
  floatx16 cr, ci;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    cr[i] = 0.0f;
    ci[i] = 0.0f;
  }
#pragma unroll
  for (int s2 = 0; s2 < ZP / 2; ++s2) {
    const float2 e = kz_ok ? ea[2 * s2] : make_float2(0.0f, 0.0f);
    const float2 x = xv[s2];
    cr = __builtin_amdgcn_mfma_f32_32x32x2f32(e.x, x.x, cr, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f32_32x32x2f32(-e.y, x.y, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_32x32x2f32(e.x, x.y, ci, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_32x32x2f32(e.y, x.x, ci, 0, 0, 0);
  }
  if (!col_ok) return;
  float2* ob = out + (int64_t)b * nz * ncol + col;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int kz = kt * 32 + 8 * (i / 4) + 4 * kh + (i % 4);
    if (kz < nz) {
      const float fr = cr[i], fi = ci[i];
      const float2 tv = tomo[(int64_t)kz * ncol + col];
      ob[(int64_t)kz * ncol] = make_float2(tv.x * fr + tv.y * fi, tv.y * fr - tv.x * fi);
    }
  }
}

int pruned_pad(int t) { return t <= 16 ? 16 : t <= 32 ? 32 : t <= 48 ? 48 : 64; }

// the separable pruned template spectrum times the tomogram spectrum for a
// block of nb rotated templates tb[nb][tz][ty][tx] (pad_low = p)
void pruned_spectrum_mul(emc_ctx* ctx, const float* tb, const int td[3], const int cd[3], const float2* tomo,
                         const float2* ex, const float2* ey, const float2* ez, int nb, float2* X1, float2* X2,
                         float2* out, bool mfma) {
  const int cx = cd[0] / 2 + 1, ny = cd[1], nz = cd[2];
  const int xp = pruned_pad(td[0]), yp = pruned_pad(td[1]), zp = pruned_pad(td[2]);
  const int kx_chunk = 32;
  const dim3 g1(ceil_div((int64_t)td[2] * yp, 256), nb, ceil_div(cx, kx_chunk));
#define EMC_TDFT_X(XP, YP)                                                                                    \
  hipLaunchKernelGGL((tdft_x_kernel<XP, YP>), g1, dim3(256), 0, ctx->stream, tb, td[0], td[1], td[2], cx, ex, \
                     kx_chunk, X1)
#define EMC_TDFT_X_Y(XP)          \
  switch (yp) {                   \
    case 16: EMC_TDFT_X(XP, 16); break; \
    case 32: EMC_TDFT_X(XP, 32); break; \
    case 48: EMC_TDFT_X(XP, 48); break; \
    default: EMC_TDFT_X(XP, 64); break; \
  }
  switch (xp) {
    case 16: EMC_TDFT_X_Y(16); break;
    case 32: EMC_TDFT_X_Y(32); break;
    case 48: EMC_TDFT_X_Y(48); break;
    default: EMC_TDFT_X_Y(64); break;
  }
#undef EMC_TDFT_X_Y
#undef EMC_TDFT_X
  const int ky_chunk = 64;
  const dim3 g2(ceil_div((int64_t)zp * cx, 256), nb, ceil_div(ny, ky_chunk));
  switch (yp) {
    case 16: hipLaunchKernelGGL(tdft_y_kernel<16>, g2, dim3(256), 0, ctx->stream, (const float2*)X1, td[2], zp, ny, cx, ey, ky_chunk, X2); break;
    case 32: hipLaunchKernelGGL(tdft_y_kernel<32>, g2, dim3(256), 0, ctx->stream, (const float2*)X1, td[2], zp, ny, cx, ey, ky_chunk, X2); break;
    case 48: hipLaunchKernelGGL(tdft_y_kernel<48>, g2, dim3(256), 0, ctx->stream, (const float2*)X1, td[2], zp, ny, cx, ey, ky_chunk, X2); break;
    default: hipLaunchKernelGGL(tdft_y_kernel<64>, g2, dim3(256), 0, ctx->stream, (const float2*)X1, td[2], zp, ny, cx, ey, ky_chunk, X2); break;
  }
  if (mfma) {
    const dim3 gm(ceil_div((int64_t)ny * cx, 32), nb, ceil_div(nz, 128));
    switch (zp) {
      case 16: hipLaunchKernelGGL(tdft_z_mul_mfma_kernel<16>, gm, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, out); break;
      case 32: hipLaunchKernelGGL(tdft_z_mul_mfma_kernel<32>, gm, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, out); break;
      case 48: hipLaunchKernelGGL(tdft_z_mul_mfma_kernel<48>, gm, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, out); break;
      default: hipLaunchKernelGGL(tdft_z_mul_mfma_kernel<64>, gm, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, out); break;
    }
    EMC_HIP(hipGetLastError());
    return;
  }
  const int kz_chunk = 128;
  const dim3 g3(ceil_div((int64_t)ny * cx, 256), nb, ceil_div(nz, kz_chunk));
  switch (zp) {
    case 16: hipLaunchKernelGGL(tdft_z_mul_kernel<16>, g3, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, kz_chunk, out); break;
    case 32: hipLaunchKernelGGL(tdft_z_mul_kernel<32>, g3, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, kz_chunk, out); break;
    case 48: hipLaunchKernelGGL(tdft_z_mul_kernel<48>, g3, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, kz_chunk, out); break;
    default: hipLaunchKernelGGL(tdft_z_mul_kernel<64>, g3, dim3(256), 0, ctx->stream, (const float2*)X2, tomo, nz, ny, cx, ez, kz_chunk, out); break;
  }
  EMC_HIP(hipGetLastError());
}

// per-map sums over the valid area (partials[b][block][2]): blocks stride
// over the rows (y, z) of the valid area, lanes run along x (32-bit row
// index, no per-element division); a fixed grid keeps the sums deterministic
constexpr int kValidBlocks = 1024;

__global__ void __launch_bounds__(kRedThreads) valid_sums_kernel(const float* __restrict__ ccf,
                                                                 int nx, int ny, int nz, int vx0,
                                                                 int vy0, int vz0, int vx, int vy,
                                                                 int vz, double* __restrict__ part) {
  const float* m = ccf + (int64_t)blockIdx.y * nx * ny * nz;
  double v[2] = {0.0, 0.0};
  const int rows = vy * vz;
  // four rows per step, their loads issued before any is summed
  constexpr int kRows = 4;
  for (int r0 = blockIdx.x; r0 < rows; r0 += kRows * gridDim.x) {
    const float* row[kRows];
    bool ok[kRows];
#pragma unroll
    for (int q = 0; q < kRows; ++q) {
      const int r = r0 + q * gridDim.x;
      ok[q] = r < rows;
      const int rr = ok[q] ? r : r0;
      const int z = rr / vy, y = rr - z * vy;
      row[q] = m + ((int64_t)(z + vz0) * ny + (y + vy0)) * nx + vx0;
    }
    for (int x = threadIdx.x; x < vx; x += blockDim.x) {
      float a[kRows];
#pragma unroll
      for (int q = 0; q < kRows; ++q) a[q] = row[q][x];
#pragma unroll
      for (int q = 0; q < kRows; ++q)
        if (ok[q]) {
          v[0] += (double)a[q];
          v[1] += (double)a[q] * (double)a[q];
        }
    }
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    double* p = part + 2 * ((int64_t)blockIdx.y * gridDim.x + blockIdx.x);
    p[0] = v[0];
    p[1] = v[1];
  }
}

// std with N-1 normalisation (MATLAB std, :764) per map -> inv_std[b]
__global__ void __launch_bounds__(kRedThreads) std_fold_kernel(const double* __restrict__ part, int nb,
                                                               double nvalid,
                                                               float* __restrict__ stdv) {
  double v[2] = {0.0, 0.0};
  const double* p = part + (int64_t)2 * nb * blockIdx.x;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    v[0] += p[2 * b];
    v[1] += p[2 * b + 1];
  }
  block_sum<2>(v);
  if (threadIdx.x == 0) {
    const double mean = v[0] / nvalid;
    double var = (v[1] - nvalid * mean * mean) / (nvalid - 1.0);
    if (var < 0.0) var = 0.0;
    stdv[blockIdx.x] = (float)sqrt(var);
  }
}

// score/angle update over the valid area for a block of maps, in order;
// blocks stride over rows (y, z), lanes along x
__global__ void __launch_bounds__(256) mip_update_kernel(const float* __restrict__ ccf, int nb_maps,
                                                         int nx, int ny, int nz, int vx0, int vy0,
                                                         int vz0, int vx, int vy, int vz,
                                                         const float* __restrict__ stdv,
                                                         float* __restrict__ peak,
                                                         int32_t* __restrict__ angle, int first_id,
                                                         int init_first) {
  const int64_t nc = (int64_t)nx * ny * nz;
  const int rows = vy * vz;
  for (int r = blockIdx.x; r < rows; r += gridDim.x) {
    const int z = r / vy, y = r - z * vy;
    const int64_t src0 = ((int64_t)(z + vz0) * ny + (y + vy0)) * nx + vx0;
    const int64_t dst0 = (int64_t)r * vx;
    for (int x = threadIdx.x; x < vx; x += blockDim.x) {
      float pk = peak[dst0 + x];
      int32_t an = angle[dst0 + x];
      for (int b = 0; b < nb_maps; ++b) {
        const float v = ccf[b * nc + src0 + x] / stdv[b];
        if (b == 0 && init_first) {  // magTmp = ccfmap; angTmp = 1 (:802-823)
          pk = v;
          an = 1;
        } else if (pk < v) {  // strict < (:851-854)
          pk = v;
          an = first_id + b;
        }
      }
      peak[dst0 + x] = pk;
      angle[dst0 + x] = an;
    }
  }
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_template_search_chunk(emc_ctx* ctx, const float* d_chunk, const int cd[3],
                                     const float* d_template, const int td[3], const int pad_low[3],
                                     const int valid_low[3], const int vd[3],
                                     const double* angles_pts, int n_orient, int first_angle_id,
                                     int init_first, float* d_peak, int32_t* d_angle) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(d_chunk && cd && d_template && td && pad_low && valid_low && vd && angles_pts &&
                    d_peak && d_angle,
                "null pointer");
    EMC_REQUIRE(n_orient > 0, "n_orient must be > 0");
    for (int k = 0; k < 3; ++k) {
      EMC_REQUIRE(pad_low[k] >= 0 && pad_low[k] + td[k] <= cd[k], "template does not fit the chunk");
      EMC_REQUIRE(valid_low[k] >= 0 && vd[k] > 0 && valid_low[k] + vd[k] <= cd[k],
                  "valid area outside the chunk");
    }
    const int64_t nc = (int64_t)cd[0] * cd[1] * cd[2];
    const int64_t nh = (int64_t)(cd[0] / 2 + 1) * cd[1] * cd[2];
    const int64_t nvalid = (int64_t)vd[0] * vd[1] * vd[2];
    EMC_REQUIRE(nvalid > 1, "valid area must hold more than one voxel");
    // orientation block: up to 16 orientations, bounded by 48 GiB of batch
    // buffers (a sixth of MI355X's HBM; a 480^3 chunk takes ~0.93 GB per
    // orientation: the C2R map, the product spectrum and the template's
    // partial DFT; a whole 1152 x 1152 x 576 tomogram chunk ~6.5 GB); one pass
    // of the max/argmax update then serves the whole block
    const int64_t per_orient = nc * 4 + nh * 8 + (int64_t)64 * cd[1] * (cd[0] / 2 + 1) * 8;
    int blk = (int)std::max<int64_t>(1, std::min<int64_t>(16, ((int64_t)48 << 30) / per_orient));
    blk = std::min(blk, n_orient);
    emc_fft_plan* plan1 = cached_plan(ctx, cd[0], cd[1], cd[2], 1);
    emc_fft_plan* planb = cached_plan(ctx, cd[0], cd[1], cd[2], blk);

    float2* tomo = (float2*)scratch(ctx, "ts_tomo", sizeof(float2) * nh);
    float* tpad = (float*)scratch(ctx, "ts_tpad", sizeof(float) * nc * blk);
    float2* T = (float2*)scratch(ctx, "ts_T", sizeof(float2) * nh * blk);
    float* mats = (float*)scratch(ctx, "ts_mats", sizeof(float) * 12 * blk);
    double* part = (double*)scratch(ctx, "ts_part", sizeof(double) * 2 * kValidBlocks * blk);
    float* stdv = (float*)scratch(ctx, "ts_std", sizeof(float) * blk);

    // tomoFou = swapQuadrants .* R2C(chunk); copy first: the FFT input is not
    // guaranteed to be preserved
    float* tmp = tpad;
    EMC_HIP(hipMemcpyAsync(tmp, d_chunk, sizeof(float) * nc, hipMemcpyDeviceToDevice, ctx->stream));
    fft_r2c(plan1, tmp, tomo);
    hipLaunchKernelGGL(swapq_kernel, dim3(2048), dim3(256), 0, ctx->stream, tomo, cd[0], cd[1], cd[2]);
    EMC_HIP(hipGetLastError());

    // pruned template spectrum (every template edge <= kTMax; "tsearch_pruned"
    // 0 selects the padded-volume R2C instead)
    const int pmode = tune_value(ctx, "tsearch_pruned", 0, 2);
    const bool pruned = pmode != 0 && td[0] <= kTMax && td[1] <= kTMax && td[2] <= kTMax;
    const int cx = cd[0] / 2 + 1;
    float* tblk = nullptr;
    float2 *X1 = nullptr, *X2 = nullptr, *ex = nullptr, *ey = nullptr, *ez = nullptr;
    if (pruned) {
      const int yp = pruned_pad(td[1]), zp = pruned_pad(td[2]);
      tblk = (float*)scratch(ctx, "ts_tblk", sizeof(float) * (size_t)td[0] * td[1] * td[2] * blk);
      X1 = (float2*)scratch(ctx, "ts_X1", sizeof(float2) * (size_t)td[2] * cx * yp * blk);
      X2 = (float2*)scratch(ctx, "ts_X2", sizeof(float2) * (size_t)zp * cd[1] * cx * blk);
      float2* tw = (float2*)scratch(ctx, "ts_tw", sizeof(float2) * (size_t)kTMax * (cx + cd[1] + cd[2]));
      ex = tw;
      ey = ex + (size_t)cx * kTMax;
      ez = ey + (size_t)cd[1] * kTMax;
      hipLaunchKernelGGL(dft_twiddle_kernel, dim3(ceil_div((int64_t)cx * kTMax, 256)), dim3(256), 0, ctx->stream,
                         ex, cx, td[0], pad_low[0], cd[0]);
      hipLaunchKernelGGL(dft_twiddle_kernel, dim3(ceil_div((int64_t)cd[1] * kTMax, 256)), dim3(256), 0,
                         ctx->stream, ey, cd[1], td[1], pad_low[1], cd[1]);
      hipLaunchKernelGGL(dft_twiddle_kernel, dim3(ceil_div((int64_t)cd[2] * kTMax, 256)), dim3(256), 0,
                         ctx->stream, ez, cd[2], td[2], pad_low[2], cd[2]);
      EMC_HIP(hipGetLastError());
    }

    float* h_par = (float*)pinned(ctx, sizeof(float) * 12 * blk);
    const float one3[3] = {1.0f, 1.0f, 1.0f};
    for (int o0 = 0; o0 < n_orient; o0 += blk) {
      const int nb = std::min(blk, n_orient - o0);
      // the pinned staging buffer is rewritten below: wait for the previous block
      EMC_HIP(hipStreamSynchronize(ctx->stream));
      for (int b = 0; b < nb; ++b) {
        const double* a = angles_pts + 3 * (o0 + b);
        const double ang[3] = {a[0], a[1], a[2] - a[0]};  // [phi, theta, psi-phi] (:744-747)
        double R[9];
        define_matrix(ang, 0, 0, R);  // 'Bah', 'forward'
        colmajor_to_rows(R, h_par + 9 * b);
        for (int k = 0; k < 3; ++k) h_par[9 * blk + 3 * b + k] = one3[k];  // shift [1,1,1]
      }
      EMC_HIP(hipMemcpyAsync(mats, h_par, sizeof(float) * 12 * blk, hipMemcpyHostToDevice,
                             ctx->stream));
      emc_fft_plan* p = nb == blk ? planb : cached_plan(ctx, cd[0], cd[1], cd[2], nb);
      if (pruned) {
        // rotate into a template-sized block (:744-747); its spectrum at the
        // padded position (:751-756) times the tomogram spectrum
        XformArgs xa = xform_args(td, nullptr, 1, one3, 1);
        launch_xform_batched(ctx, d_template, tblk, mats, mats + 9 * blk, nb, xa);
        pruned_spectrum_mul(ctx, tblk, td, cd, (const float2*)tomo, ex, ey, ez, nb, X1, X2, T, pmode == 2);
      } else {
        // rotate into the zero-padded chunk (:744-754)
        XformArgs xa = xform_args(td, nullptr, 1, one3, 1);
        xa.ox = cd[0];
        xa.oy = cd[1];
        xa.oz = cd[2];
        xa.px = pad_low[0];
        xa.py = pad_low[1];
        xa.pz = pad_low[2];
        xa.pad_value = 0.0f;
        launch_xform_batched(ctx, d_template, tpad, mats, mats + 9 * blk, nb, xa);
        fft_r2c(p, tpad, T);  // tempFou (:756)
        hipLaunchKernelGGL(conj_mul_batch_kernel, dim3(1024, nb), dim3(256), 0, ctx->stream,
                           (const float2*)tomo, T, nh);
        EMC_HIP(hipGetLastError());
      }
      fft_c2r(p, T, tpad);  // ccfmap (:759-761), C2R reuses the rotate buffer
      hipLaunchKernelGGL(valid_sums_kernel, dim3(kValidBlocks, nb), dim3(kRedThreads), 0, ctx->stream,
                         (const float*)tpad, cd[0], cd[1], cd[2], valid_low[0], valid_low[1],
                         valid_low[2], vd[0], vd[1], vd[2], part);
      hipLaunchKernelGGL(std_fold_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                         (const double*)part, kValidBlocks, (double)nvalid, stdv);
      hipLaunchKernelGGL(mip_update_kernel, dim3(4096), dim3(256), 0, ctx->stream, (const float*)tpad,
                         nb, cd[0], cd[1], cd[2], valid_low[0], valid_low[1], valid_low[2], vd[0],
                         vd[1], vd[2], (const float*)stdv, d_peak, d_angle, first_angle_id + o0,
                         (init_first && o0 == 0) ? 1 : 0);
      EMC_HIP(hipGetLastError());
    }
  });
}

}  // extern "C"
