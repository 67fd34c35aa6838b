// Trilinear 3D rotation + shift: the mexXform3d replacement
// (reference: mexFiles/mexXform3d.cu:22-87 kernel, :93-324 gateway;
// testScripts/interpolator.m:99-171 symmetry averaging).
//
// The reference binds the input to a 3D texture on every call (cudaArray
// allocation + copy, mexXform3d.cu:231-279). Here the input stays where it is
// in HBM; each block stages the input brick its 8x8x8 output tile maps to in
// LDS with coalesced row loads and interpolates from there, so HBM sees each
// input byte about once and the 8 neighbour reads of a sample are LDS reads.
//
// Geometry (index units, u = x - w/2 on the output window grid):
//   fwd: p = M (u - t)        inv: p = M u + t
// sampled at p + n/2 of the input grid with zero outside. This equals the
// reference's normalised texture coordinate ((p + half-texel)/n + 0.5) for
// both even and odd n.
#include <cmath>
#include <cstring>

#include "ops.hpp"

namespace emc {

// The sample point of window voxel (x, y, z) under one matrix (row-major
// m) and shift t, in input index coordinates.
__device__ inline void xform_point(const XformArgs& a, const float* __restrict__ m, float tx, float ty, float tz,
                                   int x, int y, int z, float& px, float& py, float& pz) {
  const float u0 = (float)x - (float)(a.wx / 2);
  const float v0 = (float)y - (float)(a.wy / 2);
  const float w0 = (float)z - (float)(a.wz / 2);
  const float cx = (float)(a.nx / 2), cy = (float)(a.ny / 2), cz = (float)(a.nz / 2);
  if (a.fwd) {
    const float u = u0 - tx, v = v0 - ty, w = w0 - tz;
    px = u * m[0] + v * m[1] + w * m[2] + cx;
    py = u * m[3] + v * m[4] + w * m[5] + cy;
    pz = u * m[6] + v * m[7] + w * m[8] + cz;
  } else {
    px = u0 * m[0] + v0 * m[1] + w0 * m[2] + tx + cx;
    py = u0 * m[3] + v0 * m[4] + w0 * m[5] + ty + cy;
    pz = u0 * m[6] + v0 * m[7] + w0 * m[8] + tz + cz;
  }
}

// LDS-staged trilinear rotation. A block owns an 8x8x8 tile of the output
// grid (2 voxels per thread). Per symmetry mate it stages the input box
// that the tile's window part maps to: the map is affine, so the sample
// points of the tile lie inside the box spanned by its 8 corners, at most
// 7 sqrt(3) + 2 < 16 voxels on a side for a rotation. The box is read with
// coalesced row loads (zeros outside the volume: the border addressing of
// the reference's texture) and the 8 neighbours of every sample come from
// LDS, with the same arithmetic as a direct gather. A footprint outside the
// staged box (a matrix that is not a rotation, or rounding at the box edge)
// is gathered from global memory instead, so any matrix is exact.
constexpr int kTile = 8;
constexpr int kBox = 16;

__device__ inline float lerp_corners(float c000, float c001, float c010, float c011, float c100, float c101,
                                     float c110, float c111, float fx, float fy, float fz) {
  const float c00 = c000 + fx * (c001 - c000);
  const float c01 = c010 + fx * (c011 - c010);
  const float c10 = c100 + fx * (c101 - c100);
  const float c11 = c110 + fx * (c111 - c110);
  const float c0 = c00 + fy * (c01 - c00);
  const float c1 = c10 + fy * (c11 - c10);
  return c0 + fz * (c1 - c0);
}

__global__ void __launch_bounds__(256) xform3d_tile_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                           XformArgs a, const float* __restrict__ d_mats,
                                                           const float* __restrict__ d_shifts, int tiles_x,
                                                           int tiles_y, int tiles_z) {
  __shared__ float box[kBox * kBox * kBox];
  const int b = blockIdx.y;  // batch entry (d_mats / d_shifts), 0 otherwise
  // XCD-aware: workgroups are dealt round-robin over the 8 XCDs, so XCD i
  // takes the i-th slab of consecutive tiles (z-major) and its L2 holds only
  // the input region that slab maps to
  const int ntiles = tiles_x * tiles_y * tiles_z;
  const int per_xcd = (ntiles + 7) / 8;
  const int tile = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
  if (tile >= ntiles) return;  // uniform per block
  const int X0 = (tile % tiles_x) * kTile, Y0 = ((tile / tiles_x) % tiles_y) * kTile,
            Z0 = (tile / (tiles_x * tiles_y)) * kTile;
  const int t = threadIdx.x;
  const int lx = t % kTile, ly = (t / kTile) % kTile, lz = t / (kTile * kTile);  // lz in 0..3
  float* ob = out + (int64_t)b * a.ox * a.oy * a.oz;
  // the tile's part of the rotated window, in window coordinates
  const int xlo = max(X0 - a.px, 0), xhi = min(X0 + kTile - 1 - a.px, a.wx - 1);
  const int ylo = max(Y0 - a.py, 0), yhi = min(Y0 + kTile - 1 - a.py, a.wy - 1);
  const int zlo = max(Z0 - a.pz, 0), zhi = min(Z0 + kTile - 1 - a.pz, a.wz - 1);
  const bool any = xlo <= xhi && ylo <= yhi && zlo <= zhi;
  float r[2] = {0.0f, 0.0f};
  bool inwin[2];
  int vx[2], vy[2], vz[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    vx[k] = X0 + lx - a.px;
    vy[k] = Y0 + ly - a.py;
    vz[k] = Z0 + lz + 4 * k - a.pz;
    inwin[k] = (unsigned)vx[k] < (unsigned)a.wx && (unsigned)vy[k] < (unsigned)a.wy && (unsigned)vz[k] < (unsigned)a.wz;
  }
  if (any) {
    float tx = a.t[0], ty = a.t[1], tz = a.t[2];
    if (d_shifts) {
      tx = d_shifts[3 * b];
      ty = d_shifts[3 * b + 1];
      tz = d_shifts[3 * b + 2];
    }
    if (a.d_shift) {
      tx += (float)a.d_shift[0];
      ty += (float)a.d_shift[1];
      tz += (float)a.d_shift[2];
    }
    const int nsym = d_mats ? 1 : a.nsym;
    for (int s = 0; s < nsym; ++s) {
      const float* m = d_mats ? d_mats + 9 * b : &a.m[s][0];
      // box of the 8 corners (uniform over the block)
      float mn[3] = {3.0e38f, 3.0e38f, 3.0e38f}, mx[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        float p[3];
        xform_point(a, m, tx, ty, tz, (c & 1) ? xhi : xlo, (c & 2) ? yhi : ylo, (c & 4) ? zhi : zlo, p[0], p[1],
                    p[2]);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          mn[i] = fminf(mn[i], p[i]);
          mx[i] = fmaxf(mx[i], p[i]);
        }
      }
      int lo[3], dim[3];
      bool staged = true;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        staged = staged && mn[i] > -1.0e6f && mx[i] < 1.0e6f;
        lo[i] = staged ? (int)floorf(mn[i]) : 0;
        dim[i] = staged ? (int)floorf(mx[i]) + 2 - lo[i] : kBox + 1;
        staged = staged && dim[i] <= kBox;
      }
      __syncthreads();  // the previous mate's box is no longer read
      if (staged) {
        // 16 lanes per box row, 16 rows per pass: every load of the thread is
        // issued before the first LDS store, so the L2 round trips overlap
        const int ix = t % kBox, gx = lo[0] + ix;
        const bool xin = ix < dim[0] && (unsigned)gx < (unsigned)a.nx;
        const int rows = dim[1] * dim[2];
        float v[kBox];
#pragma unroll
        for (int j = 0; j < kBox; ++j) {
          const int r = t / kBox + kBox * j;
          const int iz = r / dim[1], iy = r - iz * dim[1];
          const int gy = lo[1] + iy, gz = lo[2] + iz;
          v[j] = 0.0f;
          if (r < rows && xin && (unsigned)gy < (unsigned)a.ny && (unsigned)gz < (unsigned)a.nz)
            v[j] = in[((int64_t)gz * a.ny + gy) * a.nx + gx];
        }
#pragma unroll
        for (int j = 0; j < kBox; ++j) {
          const int r = t / kBox + kBox * j;
          const int iz = r / dim[1], iy = r - iz * dim[1];
          if (r < rows) box[(iz * kBox + iy) * kBox + ix] = v[j];
        }
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!inwin[k]) continue;
        float px, py, pz;
        xform_point(a, m, tx, ty, tz, vx[k], vy[k], vz[k], px, py, pz);
        const float xf = floorf(px), yf = floorf(py), zf = floorf(pz);
        const int ix = (int)xf - lo[0], iy = (int)yf - lo[1], iz = (int)zf - lo[2];
        if (staged && xf > -1.0e6f && xf < 1.0e6f && yf > -1.0e6f && yf < 1.0e6f && zf > -1.0e6f && zf < 1.0e6f &&
            ix >= 0 && iy >= 0 && iz >= 0 && ix + 1 < dim[0] && iy + 1 < dim[1] && iz + 1 < dim[2]) {
          const float* q = box + (iz * kBox + iy) * kBox + ix;
          r[k] += lerp_corners(q[0], q[1], q[kBox], q[kBox + 1], q[kBox * kBox], q[kBox * kBox + 1],
                               q[kBox * kBox + kBox], q[kBox * kBox + kBox + 1], px - xf, py - yf, pz - zf);
        } else {
          r[k] += sample_trilinear_zero(in, a.nx, a.ny, a.nz, px, py, pz);
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int X = X0 + lx, Y = Y0 + ly, Z = Z0 + lz + 4 * k;
    if (X >= a.ox || Y >= a.oy || Z >= a.oz) continue;
    float v = a.pad_value;
    if (inwin[k]) {
      v = (!d_mats && a.nsym > 1) ? r[k] * a.inv_nsym : r[k];
      if (a.d_mul) v *= a.d_mul[((int64_t)vz[k] * a.wy + vy[k]) * a.wx + vx[k]];
    }
    ob[((int64_t)Z * a.oy + Y) * a.ox + X] = v;
  }
}

XformArgs xform_args(const int in_dims[3], const float* Rrow, int nsym, const float t[3], int fwd) {
  XformArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nx = in_dims[0];
  a.ny = in_dims[1];
  a.nz = in_dims[2];
  a.ox = a.wx = a.nx;
  a.oy = a.wy = a.ny;
  a.oz = a.wz = a.nz;
  a.nsym = nsym;
  a.inv_nsym = 1.0f / (float)nsym;
  a.fwd = fwd ? 1 : 0;
  for (int i = 0; i < 3; ++i) a.t[i] = t ? t[i] : 0.0f;
  if (Rrow) std::memcpy(&a.m[0][0], Rrow, sizeof(float) * 9 * nsym);
  return a;
}

static void launch_tiles(emc_ctx* ctx, const float* in, float* out, const XformArgs& a, const float* d_mats,
                         const float* d_shifts, int batch) {
  const int tx = ceil_div(a.ox, kTile), ty = ceil_div(a.oy, kTile), tz = ceil_div(a.oz, kTile);
  EMC_REQUIRE((int64_t)tx * ty * tz < ((int64_t)1 << 31) && batch <= 65535, "rotation grid too large");
  const int64_t ntiles = (int64_t)tx * ty * tz;
  hipLaunchKernelGGL(xform3d_tile_kernel, dim3((unsigned)(8 * ((ntiles + 7) / 8)), batch), dim3(256), 0, ctx->stream,
                     in, out, a, d_mats, d_shifts, tx, ty, tz);
  EMC_HIP(hipGetLastError());
}

void launch_xform(emc_ctx* ctx, const float* in, float* out, const XformArgs& a) {
  KTimer kt(ctx, "xform3d");
  launch_tiles(ctx, in, out, a, nullptr, nullptr, 1);
}

void launch_xform_batched(emc_ctx* ctx, const float* in, float* out, const float* d_mats,
                          const float* d_shifts, int batch, const XformArgs& a) {
  KTimer kt(ctx, "xform3d");
  launch_tiles(ctx, in, out, a, d_mats, d_shifts, batch);
}

// column-major MATLAB 9-vector (double) -> row-major float rows, as the
// reference reads rm_1..rm_3 (mexXform3d.cu:170-172) after single() (:163)
void colmajor_to_rows(const double* cm, float* rows) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) rows[r * 3 + c] = (float)cm[c * 3 + r];
}

static void check_dims(const int d[3]) {
  EMC_REQUIRE(d && d[0] > 0 && d[1] > 0 && d[2] > 0, "bad dims");
  EMC_REQUIRE((int64_t)d[0] * d[1] * d[2] < (int64_t)1 << 31, "volume too large");
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_xform3d(emc_ctx* ctx, const float* d_in, const int dims[3], const float R_colmajor[9],
                       const float t[3], int fwd, float* d_out) {
  return guard([&] {
    use_device(ctx);
    check_dims(dims);
    EMC_REQUIRE(d_in && d_out && R_colmajor && t, "null pointer");
    EMC_REQUIRE(d_in != d_out, "d_out must not alias d_in");
    float rows[9];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) rows[r * 3 + c] = R_colmajor[c * 3 + r];
    // the gateway negates the shift for the forward transform (:160-164);
    // xform_voxel applies u - t for fwd, so t is passed as given
    XformArgs a = xform_args(dims, rows, 1, t, fwd);
    launch_xform(ctx, d_in, d_out, a);
  });
}

emc_status emc_interp3d(emc_ctx* ctx, const float* d_in, const int dims[3],
                        const double M_colmajor[9], int nsym, const double* S_colmajor,
                        const float t[3], int fwd, float* d_out) {
  return guard([&] {
    use_device(ctx);
    check_dims(dims);
    EMC_REQUIRE(d_in && d_out && M_colmajor && t, "null pointer");
    EMC_REQUIRE(d_in != d_out, "d_out must not alias d_in");
    EMC_REQUIRE(nsym >= 1 && nsym <= kMaxSym, "nsym must be 1..%d", kMaxSym);
    EMC_REQUIRE(nsym == 1 || S_colmajor, "symmetry matrices required for nsym > 1");
    std::vector<float> rows(9 * nsym);
    for (int s = 0; s < nsym; ++s) {
      double P[9];
      if (S_colmajor) {
        const double* S = S_colmajor + 9 * s;
        // P = M * S (column-major): P(r,c) = sum_k M(r,k) S(k,c)
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += M_colmajor[k * 3 + r] * S[c * 3 + k];
            P[c * 3 + r] = acc;
          }
      } else {
        std::memcpy(P, M_colmajor, sizeof(P));
      }
      colmajor_to_rows(P, &rows[9 * s]);
    }
    XformArgs a = xform_args(dims, rows.data(), nsym, t, fwd);
    launch_xform(ctx, d_in, d_out, a);
  });
}

}  // extern "C"
