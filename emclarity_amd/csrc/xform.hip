// Trilinear 3D rotation + shift: the mexXform3d replacement
// (reference: mexFiles/mexXform3d.cu:22-87 kernel, :93-324 gateway;
// testScripts/interpolator.m:99-171 symmetry averaging).
//
// The reference binds the input to a 3D texture on every call (cudaArray
// allocation + copy, mexXform3d.cu:231-279). Here the input stays where it is
// in HBM and each output voxel gathers its 8 neighbours directly: volumes of
// the alignment path (<= 128^3 FP32 = 8 MiB) stay resident in the 4 MiB-per-
// XCD L2 / 256 MiB MALL, so the gathers hit cache and HBM sees each input
// byte about once. Threads run along x in 64-lane wavefronts so stores are
// fully coalesced.
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

__device__ inline float xform_voxel(const float* __restrict__ in, const XformArgs& a, int nsym,
                                    const float* __restrict__ mats, float tx, float ty, float tz,
                                    int x, int y, int z) {
  const float u0 = (float)x - (float)(a.wx / 2);
  const float v0 = (float)y - (float)(a.wy / 2);
  const float w0 = (float)z - (float)(a.wz / 2);
  const float cx = (float)(a.nx / 2), cy = (float)(a.ny / 2), cz = (float)(a.nz / 2);
  float acc = 0.0f;
  for (int s = 0; s < nsym; ++s) {
    const float* m = mats + 9 * s;
    float px, py, pz;
    if (a.fwd) {
      const float u = u0 - tx, v = v0 - ty, w = w0 - tz;
      px = u * m[0] + v * m[1] + w * m[2];
      py = u * m[3] + v * m[4] + w * m[5];
      pz = u * m[6] + v * m[7] + w * m[8];
    } else {
      px = u0 * m[0] + v0 * m[1] + w0 * m[2] + tx;
      py = u0 * m[3] + v0 * m[4] + w0 * m[5] + ty;
      pz = u0 * m[6] + v0 * m[7] + w0 * m[8] + tz;
    }
    acc += sample_trilinear_zero(in, a.nx, a.ny, a.nz, px + cx, py + cy, pz + cz);
  }
  return nsym > 1 ? acc * a.inv_nsym : acc;
}

__global__ void __launch_bounds__(256) xform3d_kernel(const float* __restrict__ in,
                                                      float* __restrict__ out, XformArgs a) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int Z = blockIdx.z;
  if (X >= a.ox || Y >= a.oy) return;
  float tx = a.t[0], ty = a.t[1], tz = a.t[2];
  if (a.d_shift) {
    tx += (float)a.d_shift[0];
    ty += (float)a.d_shift[1];
    tz += (float)a.d_shift[2];
  }
  const int x = X - a.px, y = Y - a.py, z = Z - a.pz;
  float r = a.pad_value;
  if ((unsigned)x < (unsigned)a.wx && (unsigned)y < (unsigned)a.wy && (unsigned)z < (unsigned)a.wz) {
    r = xform_voxel(in, a, a.nsym, &a.m[0][0], tx, ty, tz, x, y, z);
    if (a.d_mul) r *= a.d_mul[((int64_t)z * a.wy + y) * a.wx + x];
  }
  out[((int64_t)Z * a.oy + Y) * a.ox + X] = r;
}

// Batched form: orientation b uses the row-major matrix d_mats[9*b] and the
// shift d_shifts[3*b] (nsym == 1); output volume b at out + b*ox*oy*oz.
__global__ void __launch_bounds__(256) xform3d_batched_kernel(const float* __restrict__ in,
                                                              float* __restrict__ out,
                                                              const float* __restrict__ d_mats,
                                                              const float* __restrict__ d_shifts,
                                                              XformArgs a) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y * blockDim.y + threadIdx.y;
  const int Z = blockIdx.z % a.oz;
  const int b = blockIdx.z / a.oz;
  if (X >= a.ox || Y >= a.oy) return;
  const int x = X - a.px, y = Y - a.py, z = Z - a.pz;
  float r = a.pad_value;
  if ((unsigned)x < (unsigned)a.wx && (unsigned)y < (unsigned)a.wy && (unsigned)z < (unsigned)a.wz) {
    r = xform_voxel(in, a, 1, d_mats + 9 * b, d_shifts[3 * b], d_shifts[3 * b + 1],
                    d_shifts[3 * b + 2], x, y, z);
    if (a.d_mul) r *= a.d_mul[((int64_t)z * a.wy + y) * a.wx + x];
  }
  out[(((int64_t)b * a.oz + Z) * a.oy + Y) * a.ox + X] = r;
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

void launch_xform(emc_ctx* ctx, const float* in, float* out, const XformArgs& a) {
  dim3 block(64, 4, 1);
  dim3 grid(ceil_div(a.ox, 64), ceil_div(a.oy, 4), a.oz);
  KTimer kt(ctx, "xform3d");
  hipLaunchKernelGGL(xform3d_kernel, grid, block, 0, ctx->stream, in, out, a);
  EMC_HIP(hipGetLastError());
}

void launch_xform_batched(emc_ctx* ctx, const float* in, float* out, const float* d_mats,
                          const float* d_shifts, int batch, const XformArgs& a) {
  dim3 block(64, 4, 1);
  dim3 grid(ceil_div(a.ox, 64), ceil_div(a.oy, 4), a.oz * batch);
  hipLaunchKernelGGL(xform3d_batched_kernel, grid, block, 0, ctx->stream, in, out, d_mats, d_shifts,
                     a);
  EMC_HIP(hipGetLastError());
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
