// The plain per-particle grid search of BH_alignRaw3d_v2 on the GPU
// (reference: alignment/BH_alignRaw3d_v2.m:1119-1419 coarse loop, :1501-1639
// refinement body, :1679-1765 final translational pass).
//
// Every orientation runs the reference's two-stage body with the work kept
// on the device and no host synchronisation until the whole list is done:
//   stage 1 (translation): rotate reference, reference weight and peak mask
//     by RotMat' (forward); R2C; X = P1 conj(F_ref) w_sym / N; C2R; masked
//     fftshifted argmax + two COM passes -> estPeak (device);
//   stage 2 (score): rotate the particle (inverse, shift estPeak read from
//     device memory) times volMask, and its SF3D wedge; R2C; bandpass;
//     rotational CC with FP64 sums -> ccc (device).
// The reference works on full complex grids (fftn/ifftn). Both stages only
// need real(ifftn(.)) and real sums, so the half spectrum is used with the
// weight product symmetrised over k and -k, which reproduces the full-grid
// result exactly (the products of the two Hermitian spectra are Hermitian;
// the weights are not, so w(k) and w(-k) both enter).
// Normalisation factors of BH_bandLimitCenterNormalize scale the map and both
// CC factors by positive constants; argmax, COM and the CC ratio do not depend
// on them, so they are not computed in this loop (DC removal and the bandpass
// are applied).
#include <cmath>
#include <cstring>

#include "ops.hpp"

namespace emc {

void define_matrix(const double ang_deg[3], int convention, int direction, double out_colmajor[9]);

// X(k) = P1(k) conj(F(k)) ws(k) / N over the half grid; F(0) treated as 0
// (BH_bandLimitCenterNormalize.m:51). ws = (wa(k)wb(k) + wa(-k)wb(-k))/2
// with wa = ifftshift(rotated ref weight), wb = ifftshift(SF3D wedge).
// blockIdx.y: orientation of a block (F, wa and X sF / sWa / sX elements
// apart; P1 and wb shared).
__global__ void __launch_bounds__(256) xcorr_half_kernel(const float2* __restrict__ P1,
                                                         const float2* __restrict__ F,
                                                         const float* __restrict__ wa_c,
                                                         const float* __restrict__ wb_c,
                                                         float2* __restrict__ X, int nx, int ny, int nz,
                                                         float inv_n, int64_t sF, int64_t sWa, int64_t sX) {
  F += blockIdx.y * sF;
  if (wa_c) wa_c += blockIdx.y * sWa;
  X += blockIdx.y * sX;
  const int cx = nx / 2 + 1;
  const int64_t nh = (int64_t)cx * ny * nz;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (int64_t)gridDim.x * blockDim.x) {
    int x, y, z;
    split3(i, cx, ny, x, y, z);
    const int64_t cp = ((int64_t)shifted_index(z, nz) * ny + shifted_index(y, ny)) * nx +
                       shifted_index(x, nx);
    const int mx = mirror_index(x, nx), my = mirror_index(y, ny), mz = mirror_index(z, nz);
    const int64_t cm = ((int64_t)shifted_index(mz, nz) * ny + shifted_index(my, ny)) * nx +
                       shifted_index(mx, nx);
    const float wa0 = wa_c ? wa_c[cp] : 1.0f, wa1 = wa_c ? wa_c[cm] : 1.0f;
    const float wb0 = wb_c ? wb_c[cp] : 1.0f, wb1 = wb_c ? wb_c[cm] : 1.0f;
    const float ws = 0.5f * (wa0 * wb0 + wa1 * wb1) * inv_n;
    float2 p = P1[i], f = F[i];
    if (i == 0) f = make_float2(0.0f, 0.0f);
    // P1 * conj(F)
    X[i] = make_float2((p.x * f.x + p.y * f.y) * ws, (p.y * f.x - p.x * f.y) * ws);
  }
}

// Rotational CC sums (BH_multi_xcf_Rotational.m:73-78) over the half grid:
//   P = F2 * bp (DC = 0), R = ref_ft1, Wp = ifftshift(rotated wedge),
//   Wr = ref weight (FFT order, null = 1). Writes per-block partials of
//   (num, sum|R Wp|^2, sum|P Wr|^2). blockIdx.y: orientation of a block
//   (F2 and Wp sF2 / sWp elements apart, partials part[3 b gridDim.x ..]).
__global__ void __launch_bounds__(kRedThreads) rot_sums_half_kernel(
    const float2* __restrict__ F2, const float* __restrict__ bp, const float2* __restrict__ R,
    const float* __restrict__ wp_c, const float* __restrict__ wr, int nx, int ny, int nz,
    double* __restrict__ part, int64_t sF2, int64_t sWp) {
  F2 += blockIdx.y * sF2;
  if (wp_c) wp_c += blockIdx.y * sWp;
  part += (int64_t)3 * blockIdx.y * gridDim.x;
  const int cx = nx / 2 + 1;
  const int64_t nh = (int64_t)cx * ny * nz;
  // NaN first element -> weight 1 (:34-40); the first element of
  // ifftshift(wedge) is the centred volume at the shifted origin.
  const int64_t c0 = ((int64_t)(nz / 2) * ny + ny / 2) * nx + nx / 2;
  const bool use_wp = wp_c && !(wp_c[c0] != wp_c[c0]);
  const bool use_wr = wr && !(wr[0] != wr[0]);
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nh;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (i == 0) continue;  // P(0) = 0
    int x, y, z;
    split3(i, cx, ny, x, y, z);
    const bool self = (x == 0) || (nx % 2 == 0 && x == nx / 2);
    const int mx = mirror_index(x, nx), my = mirror_index(y, ny), mz = mirror_index(z, nz);
    const int64_t fp = ((int64_t)z * ny + y) * nx + x;      // full-grid FFT order
    const int64_t fm = ((int64_t)mz * ny + my) * nx + mx;
    const int64_t cp = ((int64_t)shifted_index(z, nz) * ny + shifted_index(y, ny)) * nx +
                       shifted_index(x, nx);
    const int64_t cm = ((int64_t)shifted_index(mz, nz) * ny + shifted_index(my, ny)) * nx +
                       shifted_index(mx, nx);
    const double b = bp ? (double)bp[fp] : 1.0;
    const float2 f = F2[i];
    const double px = f.x * b, py = f.y * b;
    const float2 r = R[i];
    const double re = px * r.x - py * r.y;
    const double r2 = (double)r.x * r.x + (double)r.y * r.y;
    const double p2 = px * px + py * py;
    const double wp0 = use_wp ? (double)wp_c[cp] : 1.0, wp1 = use_wp ? (double)wp_c[cm] : 1.0;
    const double wr0 = use_wr ? (double)wr[fp] : 1.0, wr1 = use_wr ? (double)wr[fm] : 1.0;
    if (self) {
      v[0] += re * wr0 * wp0;
      v[1] += r2 * wp0 * wp0;
      v[2] += p2 * wr0 * wr0;
    } else {
      v[0] += re * (wr0 * wp0 + wr1 * wp1);
      v[1] += r2 * (wp0 * wp0 + wp1 * wp1);
      v[2] += p2 * (wr0 * wr0 + wr1 * wr1);
    }
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = v[k];
}

// Shell sums of the rotational CC (BH_multi_xcf_Rotational.m:50-70) from the
// half spectrum: shell voxel j is the full-grid FFT-order index of k; for
// x > n/2 the half grid holds -k, whose conjugate gives P(k) and R(k). The
// bandpass and the weights are read at k itself. Block (b, s) writes the
// partial (num, sum|R Wp|^2, sum|P Wr|^2) of shell s.
__global__ void __launch_bounds__(kRedThreads) rot_shell_sums_half_kernel(
    const float2* __restrict__ F2, const float* __restrict__ bp, const float2* __restrict__ R,
    const float* __restrict__ wp_c, const float* __restrict__ wr, int n, const int64_t* __restrict__ offsets,
    const uint32_t* __restrict__ idx, double* __restrict__ part) {
  const int s = blockIdx.y;
  const int cx = n / 2 + 1;
  const int64_t c0 = ((int64_t)(n / 2) * n + n / 2) * n + n / 2;
  const bool use_wp = wp_c && !(wp_c[c0] != wp_c[c0]);  // NaN first element -> 1 (:34-40)
  const bool use_wr = wr && !(wr[0] != wr[0]);
  double v[3] = {0.0, 0.0, 0.0};
  for (int64_t j = offsets[s] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < offsets[s + 1];
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = (int64_t)idx[j] - 1;  // MATLAB linear index, x fastest
    int x, y, z;
    split3(f, n, n, x, y, z);
    if (x == 0 && y == 0 && z == 0) continue;  // P(0) = 0
    const bool mirror = x > n / 2;
    const int hx = mirror ? n - x : x, hy = mirror ? (n - y) % n : y, hz = mirror ? (n - z) % n : z;
    const int64_t h = ((int64_t)hz * n + hy) * cx + hx;
    const float2 fp = F2[h], fr = R[h];
    const double b = bp ? (double)bp[f] : 1.0;
    // Re(P R) is the same for k and -k (conj(a) conj(b) = conj(ab))
    const double px = fp.x * b, py = fp.y * b;
    const double re = px * fr.x - py * fr.y;
    const double r2 = (double)fr.x * fr.x + (double)fr.y * fr.y;
    const double p2 = px * px + py * py;
    const int64_t cp = ((int64_t)shifted_index(z, n) * n + shifted_index(y, n)) * n + shifted_index(x, n);
    const double a = use_wp ? (double)wp_c[cp] : 1.0;
    const double c = use_wr ? (double)wr[f] : 1.0;
    v[0] += re * c * a;
    v[1] += r2 * a * a;
    v[2] += p2 * c * c;
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[((int64_t)s * gridDim.x + blockIdx.x) * 3 + k] = v[k];
}

// block b folds the partials of orientation b into cc[b * cc_stride]
__global__ void __launch_bounds__(kRedThreads) rot_cc_fold_kernel(const double* __restrict__ part,
                                                                  int nb, double* __restrict__ cc,
                                                                  int cc_stride) {
  part += (int64_t)3 * blockIdx.x * nb;
  cc += (int64_t)blockIdx.x * cc_stride;
  double v[3] = {0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    for (int k = 0; k < 3; ++k) v[k] += part[3 * b + k];
  block_sum<3>(v);
  if (threadIdx.x == 0) cc[0] = v[0] / sqrt(v[1] * v[2]);
}

// est[3 b + k] = r[b].peak[k] for b < count
__global__ void copy_peak_kernel(const PeakResult* __restrict__ r, double* __restrict__ est, int count) {
  for (int i = threadIdx.x; i < 3 * count; i += blockDim.x) est[i] = r[i / 3].peak[i % 3];
}

// the rotation kernel's per-entry FP32 shifts from FP64 peak estimates
__global__ void shifts_to_float_kernel(const double* __restrict__ est, float* __restrict__ out, int n3) {
  for (int i = threadIdx.x; i < n3; i += blockDim.x) out[i] = (float)est[i];
}

struct AlignBuffers {
  float *rref, *rwdg, *rmask, *map, *rpart, *rwedge, *cpad;
  float2 *fref, *P1, *X, *F2;
  PeakResult* res;
  double *ccc, *est, *est0, *part, *stats;
};

// mean = sum / n for the pad value (stats[0] = sum)
__global__ void mean_of_sum_kernel(const double* __restrict__ stats, double n, double* __restrict__ mean) {
  if (threadIdx.x == 0) mean[0] = stats[0] / n;
}

}  // namespace emc

using namespace emc;

extern "C" {

emc_status emc_align_grid_padded(emc_ctx* ctx, int w, int n, const int pad_low[3], int nsym,
                                 const double* S_colmajor, int n_ref, const float* d_particle,
                                 const float* d_ref, const float* d_ref_wgt_rot, const float* const* d_ref_wgts,
                                 const void* const* d_ref_ft1s, const emc_shell_plan* const* shells,
                                 const float* d_vol_mask, const float* d_peak_mask,
                                 const float* const* d_bandpasses, const float* d_wedge,
                                 const double angles_colmajor[9], const double est0[3], const int peak_com[3],
                                 const double* angles_pts, int n_orient, int mode, double* ccc_out,
                                 double* shift_out) {
  return guard([&] {
    use_device(ctx);
    EMC_REQUIRE(n > 0 && n <= 1024, "bad calc size %d", n);
    EMC_REQUIRE(w > 0 && w <= n, "window size %d must be in 1..calc size %d", w, n);
    EMC_REQUIRE(n_ref >= 1 && n_ref <= 1024, "n_ref must be 1..1024");
    EMC_REQUIRE(d_particle && d_ref && d_ref_ft1s && d_bandpasses && d_wedge && angles_colmajor && est0 &&
                    peak_com && angles_pts && ccc_out && shift_out && pad_low,
                "null pointer");
    int max_shells = 0;
    for (int r = 0; r < n_ref; ++r) {
      EMC_REQUIRE(d_ref_ft1s[r] && d_bandpasses[r], "reference %d: null ref_ft1 or bandpass", r);
      const emc_shell_plan* sp = shells ? shells[r] : nullptr;
      if (sp) {
        EMC_REQUIRE(sp->max_index <= (int64_t)n * n * n, "reference %d: shell index beyond the %d^3 calc grid", r, n);
        max_shells = std::max(max_shells, sp->n_shells);
      }
    }
    const float* d_bandpass = d_bandpasses[0];  // stage 1 (the first reference of the list)
    int pad_top[3];
    for (int k = 0; k < 3; ++k) {
      pad_top[k] = n - w - pad_low[k];
      EMC_REQUIRE(pad_low[k] >= 0 && pad_top[k] >= 0, "padCalc low %d does not fit window %d in calc %d",
                  pad_low[k], w, n);
    }
    EMC_REQUIRE(n_orient > 0, "n_orient must be > 0");
    EMC_REQUIRE(mode >= 1 && mode <= 3, "mode must be 1 (translation), 2 (score) or 3 (both)");
    EMC_REQUIRE(!(mode & 1) || d_peak_mask, "the translation stage (mode 1 or 3) needs a peak mask");
    EMC_REQUIRE(nsym >= 1 && nsym <= kMaxSym, "nsym must be 1..%d", kMaxSym);
    EMC_REQUIRE(nsym == 1 || S_colmajor, "symmetry matrices required for nsym > 1");
    const int dims[3] = {n, n, n};
    const int wdims[3] = {w, w, w};
    const bool padded = w < n;
    const int64_t nv = (int64_t)n * n * n;
    const int64_t wv = (int64_t)w * w * w;
    const int64_t nh = (int64_t)(n / 2 + 1) * n * n;
    emc_fft_plan* plan = cached_plan(ctx, n, n, n, 1);

    AlignBuffers B;
    char* base = (char*)scratch(ctx, "align_buf",
                                sizeof(float) * (nv * 5 + wv * 2) + sizeof(float2) * nh * 4 + 256 * 8);
    B.rwdg = (float*)base;
    B.rmask = B.rwdg + nv;
    B.map = B.rmask + nv;
    B.rwedge = B.map + nv;
    B.cpad = B.rwedge + nv;
    B.rref = B.cpad + nv;
    B.rpart = B.rref + wv;
    B.fref = (float2*)(B.rpart + wv);
    B.P1 = B.fref + nh;
    B.X = B.P1 + nh;
    B.F2 = B.X + nh;
    B.res = (PeakResult*)scratch(ctx, "align_res", sizeof(PeakResult) * n_orient);
    B.ccc = (double*)scratch(ctx, "align_ccc", sizeof(double) * n_orient * (size_t)n_ref);
    B.est = (double*)scratch(ctx, "align_est", sizeof(double) * 3 * n_orient);
    B.est0 = (double*)scratch(ctx, "align_est0", sizeof(double) * 3);
    B.part = (double*)scratch(ctx, "align_part", sizeof(double) * 3 * kRedBlocks);
    B.stats = (double*)scratch(ctx, "align_stats", sizeof(double) * 4);
    double* shell_part =
        max_shells ? (double*)scratch(ctx, "align_shell_part", sizeof(double) * 3 * kShellBlocks * (size_t)max_shells)
                   : nullptr;
    // host staging: est0, the CC results and (batched orientations) the
    // orientation matrices after them
    double* h_stage = (double*)pinned(ctx, sizeof(double) * (3 + (size_t)n_ref) * n_orient + 64 +
                                               sizeof(float) * 18 * (size_t)n_orient);
    std::memcpy(h_stage, est0, sizeof(double) * 3);
    EMC_HIP(hipMemcpyAsync(B.est0, h_stage, sizeof(double) * 3, hipMemcpyHostToDevice, ctx->stream));
    EMC_HIP(hipStreamSynchronize(ctx->stream));

    // A window volume on its way to the calc-size FFT: padded to C with its
    // own mean (BH_bandLimitCenterNormalize.m:48, padCalc), or used as is.
    const PadArgs pa = pad_args(wdims, pad_low, pad_top, 0.0f, 0);
    auto to_calc = [&](const float* vol) -> const float* {
      if (!padded) return vol;
      reduce_sum_sumsq(ctx, vol, wv, B.stats);
      hipLaunchKernelGGL(mean_of_sum_kernel, dim3(1), dim3(1), 0, ctx->stream, (const double*)B.stats,
                         (double)wv, B.stats + 2);
      launch_pad(ctx, vol, B.cpad, pa, B.stats + 2);
      return B.cpad;
    };

    const bool stage1 = mode & 1, stage2 = mode & 2;
    const float zero3[3] = {0.0f, 0.0f, 0.0f};
    const float est0f[3] = {(float)est0[0], (float)est0[1], (float)est0[2]};
    if (stage1) {
      // rotPart_FT of stage 1 is orientation independent (:1316-1318):
      // R2C(particle padded to C) with DC = 0 and the bandpass, computed once.
      fft_r2c(plan, to_calc(d_particle), B.P1);
      double* d_scale = B.part;  // unused scale
      blcn_half(ctx, B.P1, dims, 1, d_bandpass, d_scale, false);
    }
    const int g1 = 1024;
    // RotMat of orientation o: column-major FP64 R and FP32 rows
    auto rot_rows = [&](int o, float rows[9], double R[9]) {
      const double* a = angles_pts + 3 * o;
      const double ang[3] = {a[0], a[1], a[2] - a[0]};  // [phi, theta, psi - phi] (:1175)
      double Rl[9];
      define_matrix(ang, 0, 1, Rl);  // 'Bah', 'inv'
      // RotMat = reshape(angles,3,3) * RotMat (:1176), column-major
      for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) {
          double s = 0.0;
          for (int k = 0; k < 3; ++k) s += angles_colmajor[k * 3 + r] * Rl[c * 3 + k];
          R[c * 3 + r] = s;
        }
      colmajor_to_rows(R, rows);
    };
    // Blocks of orientations in one launch chain per kernel (rotations with
    // per-entry matrices, batched FFTs, per-entry correlation, peak search
    // and CC sums), for one symmetry mate, no padding and unshelled scores;
    // the loop below handles the rest one orientation at a time.
    const int bmax = tune_value(ctx, "align_grid_batch", n, 16);
    const bool batched = bmax > 1 && n_orient > 1 && nsym == 1 && !padded && max_shells == 0;
    if (batched) {
      const int64_t per = (int64_t)sizeof(float) * nv * 6 + (int64_t)sizeof(float2) * nh * 3;
      int nbm = (int)std::max<int64_t>(1, std::min<int64_t>(bmax, ((int64_t)1 << 30) / per));
      nbm = std::min(nbm, n_orient);
      float* b_rref = (float*)scratch(ctx, "alb_rref", sizeof(float) * nv * nbm);
      float* b_rwdg = (float*)scratch(ctx, "alb_rwdg", sizeof(float) * nv * nbm);
      float* b_rmask = (float*)scratch(ctx, "alb_rmask", sizeof(float) * nv * nbm);
      float* b_map = (float*)scratch(ctx, "alb_map", sizeof(float) * nv * nbm);
      float* b_rpart = (float*)scratch(ctx, "alb_rpart", sizeof(float) * nv * nbm);
      float* b_rwedge = (float*)scratch(ctx, "alb_rwedge", sizeof(float) * nv * nbm);
      float2* b_fref = (float2*)scratch(ctx, "alb_fref", sizeof(float2) * nh * nbm);
      float2* b_X = (float2*)scratch(ctx, "alb_X", sizeof(float2) * nh * nbm);
      float2* b_F2 = (float2*)scratch(ctx, "alb_F2", sizeof(float2) * nh * nbm);
      double* b_part = (double*)scratch(ctx, "alb_part", sizeof(double) * 3 * kRedBlocks * nbm);
      float* b_shift = (float*)scratch(ctx, "alb_shift", sizeof(float) * 3 * nbm);
      // stage-1 rows (RotMat') and stage-2 rows (RotMat) of every orientation
      float* d_m = (float*)scratch(ctx, "alb_mats", sizeof(float) * 18 * (size_t)n_orient);
      float* h_m = (float*)((char*)h_stage + sizeof(double) * (3 + (size_t)n_ref) * n_orient + 64);
      for (int o = 0; o < n_orient; ++o) {
        float rows[9];
        double R[9];
        rot_rows(o, rows, R);
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) h_m[9 * o + r * 3 + c] = rows[c * 3 + r];
        float* m2 = h_m + 9 * ((size_t)n_orient + o);
        if (S_colmajor) {  // the one symmetry mate: R * S_1
          double P[9];
          for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
              double acc = 0.0;
              for (int k = 0; k < 3; ++k) acc += R[k * 3 + r] * S_colmajor[c * 3 + k];
              P[c * 3 + r] = acc;
            }
          colmajor_to_rows(P, m2);
        } else {
          std::memcpy(m2, rows, sizeof(rows));
        }
      }
      EMC_HIP(hipMemcpyAsync(d_m, h_m, sizeof(float) * 18 * (size_t)n_orient, hipMemcpyHostToDevice,
                             ctx->stream));
      const float* d_mT = d_m;
      const float* d_mR = d_m + 9 * (size_t)n_orient;
      for (int o0 = 0; o0 < n_orient; o0 += nbm) {
        const int nb = std::min(nbm, n_orient - o0);
        emc_fft_plan* bplan = cached_plan(ctx, n, n, n, nb);
        const double* d_shift2 = B.est0;
        const float* d_shift2f = nullptr;
        if (stage1) {
          launch_xform_batched(ctx, d_ref, b_rref, d_mT + 9 * (size_t)o0, nullptr, nb,
                               xform_args(wdims, nullptr, 1, est0f, 1));  // :1287-1290
          const XformArgs a0 = xform_args(dims, nullptr, 1, zero3, 1);
          const float* wa = nullptr;
          if (d_ref_wgt_rot) {
            launch_xform_batched(ctx, d_ref_wgt_rot, b_rwdg, d_mT + 9 * (size_t)o0, nullptr, nb, a0);  // :1294-1297
            wa = b_rwdg;
          }
          launch_xform_batched(ctx, d_peak_mask, b_rmask, d_mT + 9 * (size_t)o0, nullptr, nb, a0);  // :1300-1303
          fft_r2c(bplan, b_rref, b_fref);  // :1311-1314
          hipLaunchKernelGGL(xcorr_half_kernel, dim3(g1, nb), dim3(256), 0, ctx->stream, (const float2*)B.P1,
                             (const float2*)b_fref, wa, d_wedge, b_X, n, n, n, (float)(1.0 / (double)nv), nh, nv,
                             nh);
          EMC_HIP(hipGetLastError());
          fft_c2r(bplan, b_X, b_map);
          peak_search_batched(ctx, b_map, 1, dims, b_rmask, peak_com, B.res + o0, B.est0, nb, nv, nv);
          hipLaunchKernelGGL(copy_peak_kernel, dim3(1), dim3(256), 0, ctx->stream,
                             (const PeakResult*)(B.res + o0), B.est + 3 * o0, nb);
          hipLaunchKernelGGL(shifts_to_float_kernel, dim3(1), dim3(256), 0, ctx->stream,
                             (const double*)(B.est + 3 * o0), b_shift, 3 * nb);
          EMC_HIP(hipGetLastError());
          d_shift2f = b_shift;
        }
        if (stage2) {
          XformArgs a2 = xform_args(wdims, nullptr, 1, zero3, 0);
          a2.d_mul = d_vol_mask;
          if (!d_shift2f) a2.d_shift = d_shift2;  // stage 2 alone: estPeak = est0 for every entry
          launch_xform_batched(ctx, d_particle, b_rpart, d_mR + 9 * (size_t)o0, d_shift2f, nb, a2);  // :1213-1216
          launch_xform_batched(ctx, d_wedge, b_rwedge, d_mR + 9 * (size_t)o0, nullptr, nb,
                               xform_args(dims, nullptr, 1, zero3, 0));  // :1251-1254
          fft_r2c(bplan, b_rpart, b_F2);  // :1375-1377
          for (int r = 0; r < n_ref; ++r) {
            const float* wr = d_ref_wgts ? d_ref_wgts[r] : nullptr;
            hipLaunchKernelGGL(rot_sums_half_kernel, dim3(kRedBlocks, nb), dim3(kRedThreads), 0, ctx->stream,
                               (const float2*)b_F2, d_bandpasses[r], (const float2*)d_ref_ft1s[r],
                               (const float*)b_rwedge, wr, n, n, n, b_part, nh, nv);
            hipLaunchKernelGGL(rot_cc_fold_kernel, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                               (const double*)b_part, kRedBlocks, B.ccc + (int64_t)o0 * n_ref + r, n_ref);
          }
          EMC_HIP(hipGetLastError());
        }
      }
    }
    for (int o = 0; o < (batched ? 0 : n_orient); ++o) {
      float rows[9], rowsT[9];
      double R[9];
      rot_rows(o, rows, R);
      // stage 2 rotates with the symmetry mates R * S_i and averages them
      // (the particle and wedge interp3d calls pass `symmetry`, :1213-1254)
      float srows[9 * kMaxSym];
      for (int sy = 0; sy < nsym; ++sy) {
        if (!S_colmajor) {
          std::memcpy(srows, rows, sizeof(rows));
          break;
        }
        const double* S = S_colmajor + 9 * sy;
        double P[9];
        for (int c = 0; c < 3; ++c)
          for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += R[k * 3 + r] * S[c * 3 + k];
            P[c * 3 + r] = acc;
          }
        colmajor_to_rows(P, srows + 9 * sy);
      }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) rowsT[r * 3 + c] = rows[c * 3 + r];  // RotMat'
      const double* d_shift2 = B.est0;
      if (stage1) {
        XformArgs a1 = xform_args(wdims, rowsT, 1, est0f, 1);
        launch_xform(ctx, d_ref, B.rref, a1);  // :1287-1290 (window size)
        XformArgs a0 = xform_args(dims, rowsT, 1, zero3, 1);
        const float* wa = nullptr;
        if (d_ref_wgt_rot) {
          launch_xform(ctx, d_ref_wgt_rot, B.rwdg, a0);  // :1294-1297
          wa = B.rwdg;
        }
        launch_xform(ctx, d_peak_mask, B.rmask, a0);  // :1300-1303
        fft_r2c(plan, to_calc(B.rref), B.fref);        // :1311-1314
        hipLaunchKernelGGL(xcorr_half_kernel, dim3(g1), dim3(256), 0, ctx->stream, (const float2*)B.P1,
                           (const float2*)B.fref, wa, d_wedge, B.X, n, n, n,
                           (float)(1.0 / (double)nv), (int64_t)0, (int64_t)0, (int64_t)0);
        EMC_HIP(hipGetLastError());
        fft_c2r(plan, B.X, B.map);
        // estPeak = peakCoord + estPeakCoord (:1341-1344)
        peak_search(ctx, B.map, 1, dims, B.rmask, peak_com, B.res + o, B.est0);
        hipLaunchKernelGGL(copy_peak_kernel, dim3(1), dim3(64), 0, ctx->stream,
                           (const PeakResult*)(B.res + o), B.est + 3 * o, 1);
        d_shift2 = B.est + 3 * o;
      }
      if (stage2) {
        XformArgs a2 = xform_args(wdims, srows, nsym, zero3, 0);
        a2.d_shift = d_shift2;
        a2.d_mul = d_vol_mask;
        launch_xform(ctx, d_particle, B.rpart, a2);  // :1213-1216, x volMask (:1375)
        XformArgs a3 = xform_args(dims, srows, nsym, zero3, 0);
        launch_xform(ctx, d_wedge, B.rwedge, a3);  // :1251-1254
        fft_r2c(plan, to_calc(B.rpart), B.F2);     // :1375-1377
        // one rotated particle, scored against every reference of the list
        // with that reference's bandpass, ref_FT1 and weight (:1278, :1375-1387)
        for (int r = 0; r < n_ref; ++r) {
          const float* wr = d_ref_wgts ? d_ref_wgts[r] : nullptr;
          double* cc = B.ccc + (int64_t)o * n_ref + r;
          const emc_shell_plan* sp = shells ? shells[r] : nullptr;
          if (sp) {  // wCCC a cell: shell-weighted score (:50-70)
            hipLaunchKernelGGL(rot_shell_sums_half_kernel, dim3(kShellBlocks, sp->n_shells), dim3(kRedThreads), 0,
                               ctx->stream, (const float2*)B.F2, d_bandpasses[r], (const float2*)d_ref_ft1s[r],
                               (const float*)B.rwedge, wr, n, (const int64_t*)sp->d_offsets,
                               (const uint32_t*)sp->d_idx, shell_part);
            hipLaunchKernelGGL(xcf_rot_shell_fold_kernel, dim3(1), dim3(kWave), 0, ctx->stream,
                               (const double*)shell_part, sp->n_shells, kShellBlocks, sp->abs_flag, sp->w, cc);
          } else {
            hipLaunchKernelGGL(rot_sums_half_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, ctx->stream,
                               (const float2*)B.F2, d_bandpasses[r], (const float2*)d_ref_ft1s[r],
                               (const float*)B.rwedge, wr, n, n, n, B.part, (int64_t)0, (int64_t)0);
            hipLaunchKernelGGL(rot_cc_fold_kernel, dim3(1), dim3(kRedThreads), 0, ctx->stream,
                               (const double*)B.part, kRedBlocks, cc, 0);
          }
        }
        EMC_HIP(hipGetLastError());
      }
    }
    // one synchronisation for the whole list
    double* h_ccc = h_stage;
    PeakResult* h_res = nullptr;
    std::vector<PeakResult> res;
    if (stage2)
      EMC_HIP(hipMemcpyAsync(h_ccc, B.ccc, sizeof(double) * n_orient * n_ref, hipMemcpyDeviceToHost,
                             ctx->stream));
    if (stage1) {
      res.resize(n_orient);
      EMC_HIP(hipMemcpyAsync(res.data(), B.res, sizeof(PeakResult) * n_orient,
                             hipMemcpyDeviceToHost, ctx->stream));
      h_res = res.data();
    }
    EMC_HIP(hipStreamSynchronize(ctx->stream));
    for (int o = 0; o < n_orient; ++o) {
      if (stage1 && h_res[o].status != 0)
        fail(EMC_ERR_RANGE, "orientation %d: failed to box out COM calc", o);
      for (int r = 0; r < n_ref; ++r) ccc_out[(int64_t)o * n_ref + r] = stage2 ? h_ccc[(int64_t)o * n_ref + r] : 0.0;
      for (int k = 0; k < 3; ++k) shift_out[3 * o + k] = stage1 ? h_res[o].peak[k] : est0[k];
    }
  });
}

emc_status emc_align_grid(emc_ctx* ctx, int n, const float* d_particle, const float* d_ref,
                          const float* d_ref_wgt_rot, const float* d_ref_wgt, const void* d_ref_ft1,
                          const float* d_vol_mask, const float* d_peak_mask, const float* d_bandpass,
                          const float* d_wedge, const double angles_colmajor[9],
                          const double est0[3], const int peak_com[3], const double* angles_pts,
                          int n_orient, int mode, double* ccc_out, double* shift_out) {
  const int zero[3] = {0, 0, 0};
  const float* wgts[1] = {d_ref_wgt};
  const void* ft1s[1] = {d_ref_ft1};
  const float* bps[1] = {d_bandpass};
  return emc_align_grid_padded(ctx, n, n, zero, 1, nullptr, 1, d_particle, d_ref, d_ref_wgt_rot, wgts, ft1s,
                               nullptr, d_vol_mask, d_peak_mask, bps, d_wedge, angles_colmajor, est0, peak_com, angles_pts,
                               n_orient, mode, ccc_out, shift_out);
}

}  // extern "C"
