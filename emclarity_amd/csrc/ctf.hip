// CTF and missing-wedge x CTF weights: the mexCTF / mexSF3D replacements
// (reference: mexFiles/utils/ctf.cu:12-128 kernels, mexFiles/mexCTF.cu,
// mexFiles/mexSF3D.cu:25-303, include/core_headers.cuh:40-75 ctfParams,
// masking/BH_weightMaskMex.m marshalling).
//
// The reference builds SF3D with 2 launches + a texture copy/bind per tilt
// (mexSF3D.cu:238-277) and a per-thread serial z loop accumulating into
// global memory. Here: one launch writes the 2D CTF image of every tilt
// (FP32 phase, as the reference's), one voxel-parallel launch
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
  double cos2a, sin2a;  // of 2 astig
  float sin_t, cos_t;  // of -tilt (mexSF3D.cu:131-137)
  // FP32 copies the device evaluates the phase with (the reference's own
  // precision, ctf.cu:45-64)
  float cs_f, dfd1_f, dfd2_f, amp_f, fvx_f, fvy_f, cos2a_f, sin2a_f;
};

// the FP32 copies of a tilt's constants (df_term folded into d1, d2)
static void finish_tilt(CtfTilt& t) {
  t.cs_f = (float)t.cs_term;
  t.dfd1_f = (float)(t.df_term * t.d1);
  t.dfd2_f = (float)(t.df_term * t.d2);
  t.amp_f = (float)t.amp_phase;
  t.fvx_f = (float)t.fvx;
  t.fvy_f = (float)t.fvy;
  t.cos2a_f = (float)t.cos2a;
  t.sin2a_f = (float)t.sin2a;
}

static CtfTilt make_tilt(double pixel, double wl, double cs_mm, double df1, double df2,
                         double astig_deg, double amp, double fvx, double fvy) {
  CtfTilt t;
  std::memset(&t, 0, sizeof(t));
  t.cs_term = M_PI * 0.5 * cs_mm * 1e7 * wl * wl * wl;
  t.df_term = M_PI * wl;
  t.d1 = 0.5 * (df1 + df2);
  t.d2 = 0.5 * (df1 - df2);
  t.astig = astig_deg * M_PI / 180.0 - M_PI * std::nearbyint(astig_deg / 180.0);
  t.cos2a = std::cos(2.0 * t.astig);
  t.sin2a = std::sin(2.0 * t.astig);
  t.amp_phase = std::atan(amp / std::sqrt(1.0 - amp * amp));
  t.fvx = fvx;
  t.fvy = fvy;
  (void)pixel;
  finish_tilt(t);
  return t;
}

// sin(chi) with optional square and optional radial x exposure weight
// (ctf.cu:45-64, :103-127). (x, y) are centred / wrapped frequency indices.
// The astigmatic term cos(2(phi - astig)), phi = atan2(ky, kx), is formed
// without the angle: cos 2phi = (kx^2 - ky^2) / r^2, sin 2phi = 2 kx ky / r^2
// (phi = 0 at the origin, as atan2(0, 0)). All FP32, as the reference: the
// phase reaches ~100 rad at the Nyquist frequency of a 2.5 A/px box with
// 4 um defocus, where FP32's relative 6e-8 leaves ~1e-5 rad; sinf reduces
// the argument exactly. The exposure weight is
// exp(-E / 2 / (0.245 k^-1.665 + 2.81)).
// r2 = kx^2 + ky^2 with every product and sum rounded on its own (no FMA
// contraction, whatever the inlining context), so the exposure table and
// the inline evaluation see the same r2
__device__ inline float ctf_r2(float fvx, float fvy, int x, int y) {
#pragma clang fp contract(off)
  const float ky = (float)y * fvy;
  const float kx = (float)x * fvx;
  return kx * kx + ky * ky;
}

// r2^-0.8325 of the exposure weight (+inf at the origin). One out-of-line
// function: the table kernel and the inline evaluation run the same code, so
// inlining cannot fold the power's last steps differently in either.
__device__ __attribute__((noinline)) float ctf_pw(float r2) {
  return r2 > 0.0f ? powf(r2, -0.8325f) : __builtin_inff();
}

// pw_tab: ctf_pw of this tilt's (x, y), or null to evaluate it here
__device__ inline float ctf_value(const CtfTilt& p, int x, int y, int sq, int weighted, const float* pw_tab) {
  const float ky = (float)y * p.fvy_f;
  const float kx = (float)x * p.fvx_f;
  const float r2 = ctf_r2(p.fvx_f, p.fvy_f, x, y);
  const float c2 = r2 > 0.0f ? (kx * kx - ky * ky) / r2 : 1.0f;
  const float s2 = r2 > 0.0f ? 2.0f * kx * ky / r2 : 0.0f;
  const float ca = c2 * p.cos2a_f + s2 * p.sin2a_f;
  float v = sinf(p.cs_f * r2 * r2 - r2 * (p.dfd1_f + p.dfd2_f * ca) - p.amp_f);
  if (sq) v *= v;
  if (weighted) {
    const float rw = (float)p.occupancy * (x == 0 ? 0.2f : fabsf((float)x));
    const float pw = pw_tab ? *pw_tab : ctf_pw(r2);
    v *= rw * expf((-0.5f * (float)p.exposure) / (0.24499f * pw + 2.81410f));
  }
  return v;
}

// 2D CTF images for n tilts: out[t][y][x], x in [0,cx). centred: x -= o_x
// (full grid only), y -= o_y; else FFT-order wrap (ctf.cu:31-40, :90-99).
__device__ inline void ctf_index(int x, int y, int cx, int ny, int o_x, int o_y, int half, int centered, int& xx,
                                 int& yy) {
  xx = x;
  yy = y;
  if (centered) {
    yy -= o_y;
    if (!half) xx -= o_x;
  } else {
    if (yy > o_y) yy -= ny;
    if (!half && xx > o_x) xx -= cx;
  }
}

// pw_tab ([y][x], of the Fourier voxel size (tfvx, tfvy)) serves the tilts
// with that voxel size; others evaluate the power themselves.
__global__ void __launch_bounds__(256) ctf2d_kernel(float* __restrict__ out,
                                                    const CtfTilt* __restrict__ tilts, int cx,
                                                    int ny, int o_x, int o_y, int half, int sq,
                                                    int centered, int weighted, const float* __restrict__ pw_tab,
                                                    float tfvx, float tfvy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int t = blockIdx.z;
  if (x >= cx || y >= ny) return;
  int xx, yy;
  ctf_index(x, y, cx, ny, o_x, o_y, half, centered, xx, yy);
  const CtfTilt& p = tilts[t];
  const bool tab = pw_tab && p.fvx_f == tfvx && p.fvy_f == tfvy;
  out[((int64_t)t * ny + y) * cx + x] =
      ctf_value(p, xx, yy, sq, weighted, tab ? pw_tab + (int64_t)y * cx + x : nullptr);
}

// The exposure weight's r2^-0.8325 over one CTF grid (the same indexing and
// arithmetic as ctf2d_kernel) for the Fourier voxel size (fvx, fvy).
__global__ void __launch_bounds__(256) ctf_pw_kernel(float* __restrict__ out, int cx, int ny, int o_x, int o_y,
                                                     int half, int centered, float fvx, float fvy) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  if (x >= cx || y >= ny) return;
  int xx, yy;
  ctf_index(x, y, cx, ny, o_x, o_y, half, centered, xx, yy);
  out[(int64_t)y * cx + x] = ctf_pw(ctf_r2(fvx, fvy, xx, yy));
}

// SF3D accumulation (mexSF3D.cu:37-77): for each voxel and tilt, the slab
// coordinate tw = -u sin + size_shift_z + w cos; inside |tw| <= 4 add
// 0.25 (0.5 + 0.5 cos(pi tw / 4)) x CTF_t(row y, linear in x).
//
// tw, the slab weight and the CTF x coordinate depend on (x, z, tilt) only.
// A wave covers an 8 x 8 patch of (x, z) (lane = x + 8 z within the patch)
// and each lane kSlabRows consecutive y of its (x, z) column; the block's
// four waves take consecutive row segments. The block first lists, in
// table order, the tilts whose slab reaches its patch (tw is linear in (u,
// w), so the patch corners bound it; widened by 1e-3 against rounding):
// slabs 8 voxels thick through the origin cross a compact patch at a few
// angles only. Every voxel still runs the exact slab test; tilts left out
// add nothing, and per voxel the tilts are summed in table order, as the
// reference's per-voxel loop. A lane reads its CTF row at its own x
// coordinate: the 64 lanes read inside one row of n floats per y.
constexpr int kSlabListMax = 1024;
constexpr int kSlabRows = 16;
constexpr int kPatch = 8;

// The block's tilt list in LDS (built by wave 0, in table order) for tilts
// [0, nt) and the patch x in [x0, x1], z in [z0, z1]; returns its length,
// or -1 when it would not fit (the caller then loops over every tilt).
__device__ inline int slab_list(const CtfTilt* __restrict__ tilts, int nt, int nx, int nz, int x0, int x1, int z0,
                                int z1, int* list) {
  __shared__ int count;
  const int lane = threadIdx.x + blockDim.x * threadIdx.y;
  const float ssz = (nz % 2 == 0) ? 0.5f : 0.0f;
  const float ua = (float)x0 - (float)nx / 2.0f, ub = (float)x1 - (float)nx / 2.0f;
  const float wa = (float)z0 - (float)(nz / 2), wb = (float)z1 - (float)(nz / 2);
  if (lane < kWave) {
    int n = 0;
    for (int base = 0; base < nt; base += kWave) {
      const int t = base + lane;
      bool hit = false;
      if (t < nt) {
        const float s = tilts[t].sin_t, c = tilts[t].cos_t;
        const float t0 = (-ua * s + ssz) + wa * c, t1 = (-ub * s + ssz) + wa * c;
        const float t2 = (-ua * s + ssz) + wb * c, t3 = (-ub * s + ssz) + wb * c;
        const float lo = fminf(fminf(t0, t1), fminf(t2, t3)), hi = fmaxf(fmaxf(t0, t1), fmaxf(t2, t3));
        hit = lo <= 4.001f && hi >= -4.001f;
      }
      const unsigned long long m = __ballot(hit);
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      if (hit && n + before < kSlabListMax) list[n + before] = t;
      n += __popcll(m);
    }
    if (lane == 0) count = n <= kSlabListMax ? n : -1;
  }
  __syncthreads();
  return count;
}

// One lane's column: voxels (x, y0 .. y0 + kSlabRows - 1, z) of the volume
// `out`, each stored plus `add`; ctfs, tilts: its tilts [0, nt). Returns
// the minimum of the lane's sums before `add` (+inf when it has none).
__device__ inline float sf3d_column(float* __restrict__ out, const float* __restrict__ ctfs,
                                    const CtfTilt* __restrict__ tilts, int nt, int nx, int ny, int nz, int cx,
                                    int x, int y0, int z, const int* list, int nlist, float add) {
  const float kPi = 3.14159265358979323f;
  if (x >= nx || z >= nz || y0 >= ny) return __builtin_inff();
  const float ssx = (nx % 2 == 0) ? 0.5f : 0.0f;
  const float ssz = (nz % 2 == 0) ? 0.5f : 0.0f;
  const float u = (float)x - (float)nx / 2.0f;   // float division, as mexSF3D.cu:47
  const float w = (float)z - (float)(nz / 2);
  const float xscale = (float)cx / (float)nx;
  const float xoff = 0.5f * (float)cx - 0.5f;
  float acc[kSlabRows];
#pragma unroll
  for (int r = 0; r < kSlabRows; ++r) acc[r] = 0.0f;
  const int rows = min(kSlabRows, ny - y0);  // uniform over the wave
  const int n = nlist >= 0 ? nlist : nt;
  for (int i = 0; i < n; ++i) {
    const int t = nlist >= 0 ? list[i] : i;
    const float s = tilts[t].sin_t, c = tilts[t].cos_t;
    const float tw = (-u * s + ssz) + w * c;
    if (tw < -4.0f || tw > 4.0f) continue;
    const float zw = (0.5f + 0.5f * cosf(tw * (kPi / 4.0f))) * 0.25f;
    const float xt = (u * c + w * s + ssx) * xscale + xoff;  // texel coordinate
    const float xf = floorf(xt);
    const float f = xt - xf;
    const int x0 = (int)xf;
    const bool in0 = (unsigned)x0 < (unsigned)cx, in1 = (unsigned)(x0 + 1) < (unsigned)cx;
    // Every load is unconditional from an address inside this tilt's CTF
    // (x clamped to the row, rows past the volume re-read its last row), so
    // the 2 x kSlabRows loads issue together; texels outside the row read
    // as 0 through the selects, and sums of rows past the volume are never
    // stored.
    const int xa = min(max(x0, 0), cx - 1), xb = min(max(x0 + 1, 0), cx - 1);
    const float* row = ctfs + ((int64_t)t * ny + y0) * cx;
    float v0[kSlabRows], v1[kSlabRows];
#pragma unroll
    for (int r = 0; r < kSlabRows; ++r) {
      const int rr = min(r, rows - 1);
      v0[r] = row[rr * cx + xa];
      v1[r] = row[rr * cx + xb];
    }
#pragma unroll
    for (int r = 0; r < kSlabRows; ++r) {
      // explicit FMAs: with a, b, zw >= 0 and f < 1 every term is >= 0
      const float a = in0 ? v0[r] : 0.0f, b = in1 ? v1[r] : 0.0f;
      acc[r] = fmaf(zw, fmaf(f, b - a, a), acc[r]);
    }
  }
  float m = __builtin_inff();
  float* o = out + ((int64_t)z * ny + y0) * nx + x;
#pragma unroll
  for (int r = 0; r < kSlabRows; ++r)
    if (r < rows) {
      o[(int64_t)r * nx] = acc[r] + add;
      m = fminf(m, acc[r]);
    }
  return m;
}

// blockDim (64, 4); bx = patch (px + ceil(nx / 8) * pz), by = segment of
// 4 x kSlabRows rows.
struct SlabPatch {
  int x, z, x0, x1, z0, z1, y0;
};

__device__ inline SlabPatch slab_patch(int nx, int ny, int nz, int bx, int by) {
  SlabPatch q;
  const int npx = (nx + kPatch - 1) / kPatch;
  const int px = bx % npx, pz = bx / npx;
  q.x0 = px * kPatch;
  q.z0 = pz * kPatch;
  q.x1 = min(nx, q.x0 + kPatch) - 1;
  q.z1 = min(nz, q.z0 + kPatch) - 1;
  q.x = q.x0 + (int)threadIdx.x % kPatch;
  q.z = q.z0 + (int)threadIdx.x / kPatch;
  q.y0 = (by * blockDim.y + threadIdx.y) * kSlabRows;
  (void)ny;
  return q;
}

__global__ void __launch_bounds__(256) sf3d_kernel(float* __restrict__ out,
                                                   const float* __restrict__ ctfs,
                                                   const CtfTilt* __restrict__ tilts, int n_tilts,
                                                   int nx, int ny, int nz, int cx) {
  __shared__ int list[kSlabListMax];
  const SlabPatch q = slab_patch(nx, ny, nz, blockIdx.x, blockIdx.y);
  const int nl = slab_list(tilts, n_tilts, nx, nz, q.x0, q.x1, q.z0, q.z1, list);
  (void)sf3d_column(out, ctfs, tilts, n_tilts, nx, ny, nz, cx, q.x, q.y0, q.z, list, nl, 0.0f);
}

__global__ void shift_min_kernel(float* __restrict__ d, int64_t n, const double* __restrict__ dmin,
                                 float floor_value) {
  const float sub = (float)dmin[0] - floor_value;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    d[i] = d[i] - sub;
}

// Per-tilt constants of an SF3D call (mexSF3D.cu:125-245): Fourier voxel
// size from ctf_dims after the half-grid reduction, each tilt with its own
// pixel size (mexSF3D.cu:234-245); the slab rotation of -tilt (:131-137).
std::vector<CtfTilt> sf3d_tilts(int half, const int dims[3], int n, const std::vector<double>& pixel,
                                const std::vector<double>& wl, const std::vector<double>& cs,
                                const std::vector<double>& df1, const std::vector<double>& df2,
                                const std::vector<double>& astig, const std::vector<double>& amp,
                                const std::vector<double>& tilt, const std::vector<double>& exposure,
                                const std::vector<double>& occ) {
  const int cx = half ? dims[0] / 2 + 1 : dims[0], ny = dims[1];
  std::vector<CtfTilt> h(n);
  for (int i = 0; i < n; ++i) {
    const double fvx = 1.0 / (pixel[i] * (double)cx), fvy = 1.0 / (pixel[i] * (double)ny);
    h[i] = make_tilt(pixel[i], wl[i], cs[i], df1[i], df2[i], astig[i], amp[i], fvx, fvy);
    h[i].occupancy = occ[i];
    h[i].exposure = exposure[i];
    const double a = -tilt[i] * M_PI / 180.0;
    h[i].sin_t = (float)std::sin(a);
    h[i].cos_t = (float)std::cos(a);
  }
  return h;
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
  const std::vector<CtfTilt> h = sf3d_tilts(half, dims, n, pixel, wl, cs, df1, df2, astig, amp, tilt, exposure, occ);
  CtfTilt* d_t = (CtfTilt*)scratch(ctx, "ctf_tilts", sizeof(CtfTilt) * n);
  float* d_ctf = (float*)scratch(ctx, "ctf_stack", sizeof(float) * (size_t)n * cx * ny);
  EMC_HIP(hipMemcpyAsync(d_t, h.data(), sizeof(CtfTilt) * n, hipMemcpyHostToDevice, ctx->stream));
  dim3 b2(64, 4, 1);
  hipLaunchKernelGGL(ctf2d_kernel, dim3(ceil_div(cx, 64), ceil_div(ny, 4), n), b2, 0, ctx->stream,
                     d_ctf, (const CtfTilt*)d_t, cx, ny, half ? 0 : cx / 2, ny / 2, half, sq, 1, 1, nullptr, 0.0f,
                     0.0f);
  hipLaunchKernelGGL(sf3d_kernel, dim3(ceil_div(nx, kPatch) * ceil_div(nz, kPatch), ceil_div(ny, 4 * kSlabRows), 1),
                     b2, 0, ctx->stream, d_out, (const float*)d_ctf, (const CtfTilt*)d_t, n, nx, ny, nz, cx);
  EMC_HIP(hipGetLastError());
  // the tilt table lives in pageable memory; finish before it goes away
  EMC_HIP(hipStreamSynchronize(ctx->stream));
}

// The SF3D arguments BH_weightMaskMex marshals from one particle's TLT
// (masking/BH_weightMaskMex.m:12-51): rows sorted by column 1 (stable),
// every value through single(), the particle's defocus lowered by its
// height along the beam after the tilt, occupancy = dose fraction x the
// elastic fraction exp(-75 / (400 cos(tilt))) relative to its maximum.
std::vector<CtfTilt> weight_mask_tilts(const int dims[3], double sampling, const double* tlt, int n,
                                       const double xyz[3], const double recon_size[3],
                                       const double recon_shift[3]) {
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return tlt[22 * a] < tlt[22 * b]; });
  auto col = [&](int row, int c) { return tlt[22 * order[row] + c]; };
  auto sgl = [](double x) { return (double)(float)x; };
  double origin[3], prj[3];
  for (int k = 0; k < 3; ++k) {
    origin[k] = std::ceil((recon_size[k] + 1.0) / 2.0);
    prj[k] = xyz[k] - origin[k] + recon_shift[k];  // :17-20
  }
  std::vector<double> pixel(n), wl(n), cs(n), dfa(n), dfb(n), ast(n), amp(n), tilt(n), ex(n), occ(n);
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
  return sf3d_tilts(0, dims, n, pixel, wl, cs, dfa, dfb, ast, amp, tilt, ex, occ);
}

// SF3D of many particles of one box size, the tilts of particle p being
// entries [off[p] - base, off[p+1] - base) of the chunk's tilt table and CTF
// stack. The same slab sum as sf3d_kernel; each block also writes the
// minimum of its voxels (part_min[p][block]).
// XCD-aware: a 1-D grid of 8 x ceil(np / 8) x (gx gy) blocks; block L runs
// on XCD L % 8, which takes particles L % 8, L % 8 + 8, ... one after
// another, so each particle's CTF stack (41 tilts of 128^2: 2.7 MB) is read
// into one XCD's L2 instead of all eight.
__global__ void __launch_bounds__(256) sf3d_many_kernel(float* __restrict__ out, const float* __restrict__ ctfs,
                                                        const CtfTilt* __restrict__ tilts,
                                                        const int* __restrict__ off, int base, int nx, int ny,
                                                        int nz, int cx, int gx, int gy, int np, float add,
                                                        float* __restrict__ part_min) {
  __shared__ int list[kSlabListMax];
  __shared__ float lds[4];
  const int per = gx * gy;
  const int xcd = (int)(blockIdx.x % 8), qb = (int)(blockIdx.x / 8);
  const int p = xcd + 8 * (qb / per), r = qb % per;
  if (p >= np) return;  // uniform per block, before any barrier
  const int bx = r % gx, by = r / gx;
  const int t0 = off[p] - base, t1 = off[p + 1] - base;
  const SlabPatch q = slab_patch(nx, ny, nz, bx, by);
  const int nl = slab_list(tilts + t0, t1 - t0, nx, nz, q.x0, q.x1, q.z0, q.z1, list);
  float m = sf3d_column(out + (int64_t)p * nz * ny * nx, ctfs + (int64_t)t0 * ny * cx, tilts + t0, t1 - t0, nx, ny,
                        nz, cx, q.x, q.y0, q.z, list, nl, add);
  // block minimum: one wave per threadIdx.y (blockDim = 64 x 4)
  m = wave_min(m);
  if (threadIdx.x == 0) lds[threadIdx.y] = m;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const float b = fminf(fminf(lds[0], lds[1]), fminf(lds[2], lds[3]));
    part_min[(int64_t)p * per + (int64_t)by * gx + bx] = b;
  }
}

// Per-particle minimum of the block minima (blockIdx.x = particle).
__global__ void __launch_bounds__(kRedThreads) fold_min_many_kernel(const float* __restrict__ part, int nb,
                                                                    float* __restrict__ out) {
  const float* q = part + (int64_t)blockIdx.x * nb;
  float m = __builtin_inff();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fminf(m, q[b]);
  m = block_min(m);
  if (threadIdx.x == 0) out[blockIdx.x] = m;
}

// SF3D - (min - floor) per particle (BH_weightMaskMex.m:55); blockIdx.y = particle.
__global__ void shift_min_many_kernel(float* __restrict__ d, int64_t nv, const float* __restrict__ mins,
                                      float floor_value) {
  float* v = d + (int64_t)blockIdx.y * nv;
  const float sub = mins[blockIdx.y] - floor_value;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = v[i] - sub;
}

// True when a voxel of an (nx, *, nz) SF3D of these tilts provably gets no
// term (sum exactly 0) and no term is negative: every tilt's occupancy is
// >= 0 and the voxels of column (x = nx/2, z = 0), on the z axis at the
// box edge, lie outside every tilt's slab (|tw| > 4 with a margin against
// the device's FP32 rounding); true for any tilt range short of +-90 deg
// minus the slab's reach at the box edge.
static bool sf3d_has_empty_voxel(const CtfTilt* tilts, int nt, int nx, int nz) {
  const double ssz = (nz % 2 == 0) ? 0.5 : 0.0;
  const double u = (double)(nx / 2) - (double)nx / 2.0;
  const double w = 0.0 - (double)(nz / 2);
  for (int t = 0; t < nt; ++t) {
    if (!(tilts[t].occupancy >= 0.0)) return false;
    const double tw = (-u * tilts[t].sin_t + ssz) + w * tilts[t].cos_t;
    if (std::fabs(tw) <= 4.01) return false;
  }
  return true;
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
    EMC_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    const std::vector<CtfTilt> h = weight_mask_tilts(dims, sampling, tlt, n, xyz, recon_size, recon_shift);
    const int cx = dims[0], ny = dims[1];
    CtfTilt* d_t = (CtfTilt*)scratch(ctx, "ctf_tilts", sizeof(CtfTilt) * n);
    float* d_ctf = (float*)scratch(ctx, "ctf_stack", sizeof(float) * (size_t)n * cx * ny);
    EMC_HIP(hipMemcpyAsync(d_t, h.data(), sizeof(CtfTilt) * n, hipMemcpyHostToDevice, ctx->stream));
    const dim3 b2(64, 4, 1);
    hipLaunchKernelGGL(ctf2d_kernel, dim3(ceil_div(cx, 64), ceil_div(ny, 4), n), b2, 0, ctx->stream, d_ctf,
                       (const CtfTilt*)d_t, cx, ny, cx / 2, ny / 2, 0, 1, 1, 1, nullptr, 0.0f, 0.0f);
    hipLaunchKernelGGL(sf3d_kernel,
                       dim3(ceil_div(dims[0], kPatch) * ceil_div(dims[2], kPatch), ceil_div(ny, 4 * kSlabRows), 1), b2,
                       0, ctx->stream, d_out, (const float*)d_ctf, (const CtfTilt*)d_t, n, dims[0], ny, dims[2], cx);
    EMC_HIP(hipGetLastError());
    const int64_t nv = (int64_t)dims[0] * dims[1] * dims[2];
    double* d_min = (double*)scratch(ctx, "wm_min", sizeof(double));
    reduce_min(ctx, d_out, nv, 1, d_min);
    hipLaunchKernelGGL(shift_min_kernel, dim3(256), dim3(256), 0, ctx->stream, d_out, nv,
                       (const double*)d_min, 1e-4f);  // :55
    EMC_HIP(hipGetLastError());
    // the tilt table lives in pageable memory; finish before it goes away
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

emc_status emc_weight_mask_batched(emc_ctx* ctx, const int dims[3], double sampling, int n_part,
                                   const double* tlt, const int32_t* n_tilts, const double* xyz,
                                   const double recon_size[3], const double recon_shift[3], float* d_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(dims && recon_size && recon_shift, "null pointer");
    EMC_REQUIRE(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, "bad dims");
    EMC_REQUIRE(n_part >= 0, "n_part must be >= 0");
    if (n_part == 0) return;
    EMC_REQUIRE(tlt && n_tilts && xyz && d_out, "null pointer");
    const int nx = dims[0], ny = dims[1], nz = dims[2], cx = nx;
    const int64_t nv = (int64_t)nx * ny * nz;
    // every particle's tilt constants, and its first entry in the table
    std::vector<CtfTilt> all;
    std::vector<int> first(n_part + 1, 0);
    const double* rows = tlt;
    for (int p = 0; p < n_part; ++p) {
      EMC_REQUIRE(n_tilts[p] > 0, "particle %d: empty TLT", p);
      const std::vector<CtfTilt> h =
          weight_mask_tilts(dims, sampling, rows, n_tilts[p], xyz + 3 * (int64_t)p, recon_size, recon_shift);
      all.insert(all.end(), h.begin(), h.end());
      first[p + 1] = first[p] + n_tilts[p];
      rows += 22 * (int64_t)n_tilts[p];
    }
    // chunks of particles whose CTF stack stays within 256 MiB
    const size_t tilt_bytes = sizeof(float) * (size_t)cx * ny;
    const int max_tilts = (int)std::min<size_t>(65535, std::max<size_t>(1, (size_t(256) << 20) / tilt_bytes));
    const int max_parts = 65535;  // grid z = particles of a chunk
    const dim3 b2(64, 4, 1);
    const int gx = ceil_div(nx, kPatch) * ceil_div(nz, kPatch), gy = ceil_div(ny, 4 * kSlabRows);
    const int per = gx * gy;  // blocks (minima) per particle
    std::vector<int> cuts = {0};
    int max_nt = 0;
    while (cuts.back() < n_part) {
      const int p0 = cuts.back();
      int p1 = p0 + 1;
      while (p1 < n_part && p1 - p0 < max_parts && first[p1 + 1] - first[p0] <= max_tilts) ++p1;
      max_nt = std::max(max_nt, first[p1] - first[p0]);
      cuts.push_back(p1);
    }
    EMC_REQUIRE(max_nt <= 65535, "a particle has more than 65535 tilts");
    // tilt table and absolute first-tilt offsets, staged without waiting for
    // the work already queued
    const size_t tab_bytes = sizeof(CtfTilt) * all.size();
    const size_t table_bytes = tab_bytes + sizeof(int) * (n_part + 1);
    char* d_tab = (char*)scratch(ctx, "wmb_tilts", table_bytes);
    upload_async(ctx, d_tab, all.data(), tab_bytes);
    upload_async(ctx, d_tab + tab_bytes, first.data(), sizeof(int) * (n_part + 1));
    const CtfTilt* d_t = (const CtfTilt*)d_tab;
    const int* d_off = (const int*)(d_tab + tab_bytes);
    float* d_ctf = (float*)scratch(ctx, "wmb_stack", tilt_bytes * (size_t)max_nt);
    float* d_part = (float*)scratch(ctx, "wmb_min", sizeof(float) * ((size_t)per + 1) * n_part);
    float* d_mins = d_part + (size_t)per * n_part;
    // the exposure weight's power of r2, once for the grid of the first
    // tilt's Fourier voxel size (every tilt of that size reads it)
    float* d_pw = (float*)scratch(ctx, "wmb_pw", sizeof(float) * (size_t)cx * ny);
    hipLaunchKernelGGL(ctf_pw_kernel, dim3(ceil_div(cx, 64), ceil_div(ny, 4), 1), b2, 0, ctx->stream, d_pw, cx, ny,
                       cx / 2, ny / 2, 0, 1, all[0].fvx_f, all[0].fvy_f);
    for (size_t c = 0; c + 1 < cuts.size(); ++c) {
      const int p0 = cuts[c], p1 = cuts[c + 1];
      const int np = p1 - p0, nt = first[p1] - first[p0];
      const CtfTilt* dt = d_t + first[p0];
      {
        KTimer kt(ctx, "ctf2d_many");
        hipLaunchKernelGGL(ctf2d_kernel, dim3(ceil_div(cx, 64), ceil_div(ny, 4), nt), b2, 0, ctx->stream, d_ctf,
                           dt, cx, ny, cx / 2, ny / 2, 0, 1, 1, 1, (const float*)d_pw, all[0].fvx_f, all[0].fvy_f);
      }
      // SF3D - (min - 1e-4) (BH_weightMaskMex.m:55). Every term is >= 0
      // (squared CTF, slab weight >= 0, occupancy >= 0), so where a voxel
      // provably gets no term its sum is exactly 0, the minimum is 0 and the
      // shift is +1e-4 for the whole chunk, added at the kernel's store
      // (v - (0 - 1e-4f) = v + 1e-4f exactly); otherwise a second pass
      // subtracts each particle's minimum.
      bool zero_min = true;
      for (int p = p0; p < p1 && zero_min; ++p) zero_min = sf3d_has_empty_voxel(&all[first[p]], n_tilts[p], nx, nz);
      {
        KTimer kt(ctx, "sf3d_many");
        const int64_t nblk = (int64_t)8 * ceil_div(np, 8) * gx * gy;
        EMC_REQUIRE(nblk < ((int64_t)1 << 31), "too many SF3D blocks in one launch");
        hipLaunchKernelGGL(sf3d_many_kernel, dim3((unsigned)nblk), b2, 0, ctx->stream, d_out + p0 * nv,
                           (const float*)d_ctf, dt, d_off + p0, first[p0], nx, ny, nz, cx, gx, gy, np,
                           zero_min ? 1e-4f : 0.0f, d_part);
      }
      if (!zero_min) {
        hipLaunchKernelGGL(fold_min_many_kernel, dim3(np), dim3(kRedThreads), 0, ctx->stream, (const float*)d_part,
                           per, d_mins);
        hipLaunchKernelGGL(shift_min_many_kernel, dim3(ceil_div(nv, 256 * 8), np), dim3(256), 0, ctx->stream,
                           d_out + p0 * nv, nv, (const float*)d_mins, 1e-4f);
      }
      EMC_HIP(hipGetLastError());
    }
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
                       half, sq_ctf, centered, 0, nullptr, 0.0f, 0.0f);
    EMC_HIP(hipGetLastError());
    EMC_HIP(hipStreamSynchronize(ctx->stream));
  });
}

}  // extern "C"
