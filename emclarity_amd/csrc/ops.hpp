// Cross-file launch helpers of libemc_hip.so (host-side declarations and the
// argument structs the kernels take by value).
#pragma once

#include "emc_internal.hpp"

namespace emc {

// ---------------------------------------------------------------- rotate

struct XformArgs {
  int nx, ny, nz;     // input grid
  int ox, oy, oz;     // output grid
  int wx, wy, wz;     // rotated window inside the output
  int px, py, pz;     // window offset in the output
  float pad_value;    // value outside the window
  int nsym;
  float inv_nsym;
  int fwd;
  float t[3];
  const double* d_shift;  // optional device shift added to t
  const float* d_mul;     // optional window-sized multiplier applied to the result
  float m[kMaxSym][9];    // row-major 3x3 per symmetry mate
};

XformArgs xform_args(const int in_dims[3], const float* Rrow, int nsym, const float t[3], int fwd);
void launch_xform(emc_ctx* ctx, const float* in, float* out, const XformArgs& a);
// d_mats: batch row-major 3x3 (FP32); d_shifts: batch x 3 (FP32)
void launch_xform_batched(emc_ctx* ctx, const float* in, float* out, const float* d_mats,
                          const float* d_shifts, int batch, const XformArgs& a);
void colmajor_to_rows(const double* cm, float* rows);

// ---------------------------------------------------------------- padding

struct PadArgs {
  int nx, ny, nz;   // input dims
  int tx, ty, tz;   // trim at the low side
  int wx, wy, wz;   // trimmed image dims
  int lx, ly, lz;   // pad at the low side
  int ox, oy, oz;   // output dims
  float fill;
  int taper;
};

PadArgs pad_args(const int dims[3], const int low[3], const int top[3], float fill, int taper);
void launch_pad(emc_ctx* ctx, const float* in, float* out, const PadArgs& a, const double* d_fill);

// ---------------------------------------------------------------- spectra

void bandpass_impl(emc_ctx* ctx, const int dims[3], int half, double thr, double high_cut,
                   double low_cut, double pixel, float* d_out);
// In place on `batch` half spectra of real dims `dims`: F(0)=0, F *= bp
// (full grid, may be null) and d_scale[b] = numel / sqrt(full-grid sum |F|^2);
// if apply_scale the spectra are also multiplied by d_scale[b].
void blcn_half(emc_ctx* ctx, float2* F, const int dims[3], int batch, const float* bp,
               double* d_scale, bool apply_scale);

// ---------------------------------------------------------------- FFT

void fft_plan_dims(const emc_fft_plan* p, int dims[3]);

// ---------------------------------------------------------------- peaks

// Result of a translational peak search, in device memory.
struct PeakResult {
  double peak[3];   // shift (x,y,z) in pixels (reference return value)
  double max_val;   // map maximum after masking
  int status;       // 0 ok, 1 COM box outside the map
  int pad_;
};

// Peak search of BH_multi_xcf_Translational on a real map stored in FFT order
// (the map the reference fftshifts), read with `stride` floats between
// voxels (2 = real parts of complex data). mask: centred FP32 volume or null.
// d_add: optional device shift (3 doubles) added to the returned peak.
void peak_search(emc_ctx* ctx, const float* map, int stride, const int dims[3], const float* mask,
                 const int com[3], PeakResult* d_res, const double* d_add);
// The same for `batch` maps map_bs floats apart, each with its own mask
// (mask_bs floats apart; 0 = one mask for all), results d_res[0 .. batch).
void peak_search_batched(emc_ctx* ctx, const float* map, int stride, const int dims[3], const float* mask,
                         const int com[3], PeakResult* d_res, const double* d_add, int batch, int64_t map_bs,
                         int64_t mask_bs);

// Shells of the shell-weighted CCC (emc_shell_plan_create): concatenated
// 1-based full-grid voxel lists, shell s at d_idx[d_offsets[s] .. d_offsets[s+1]).
constexpr int kShellBlocks = 64;
__global__ void __launch_bounds__(kWave) xcf_rot_shell_fold_kernel(const double* __restrict__ part, int n_shells,
                                                                  int nb, int abs_flag, double w,
                                                                  double* __restrict__ out);

// ---------------------------------------------------------------- adjoint

// Terms of an adjoint evaluation's partial sums per workgroup: num, d num /
// d s (3), D, d num / d angle (3), d D / d angle (3).
constexpr int kGradTerms = 11;

// Support list of the half grid for (n, radius) (host memory, the order the
// adjoint walks); see batched.hip.
std::vector<int2> support_host(int n, double radius);

// Bricks of side S covering the half ball of table positions the support of
// radius `radius` reaches (host memory; brick-major adjoint, batched.hip).
std::vector<int4> brick_host(int n, double radius, int S);

// Device-side inputs of adjoint launches. kind 1 (brick-major): list = the
// brick list (ns bricks of side S), g = ns; kind 0: list = the support list
// (ns entries), g workgroups per particle. part: n_part x g x kGradTerms
// partial sums. r2: the support radius squared.
struct EvalSetup {
  const void* list;
  int ns, g;
  double* part;
  int kind, S, waves;
  double r2;
};

int eval_grad_blocks(emc_ctx* ctx, int n, int n_part, int ns);
void eval_grad_kind(emc_ctx* ctx, int n, int* kind, int* S, int* waves);
// the adjoint kernel + adjoint_fold_kernel on the context's stream (no
// allocation, copy or host wait); timed adds HIP events around the kernel;
// gen: d_MdM holds the derivatives of right-multiplied rotation increments
// (dM_a = G_a^T M, the refinement's), which the brick kernel forms from M
void eval_grad_launch(emc_ctx* ctx, const EvalSetup& e, int n, int n_part, int table_format, const void* d_F2,
                      const void* d_Q, const float* d_N, const double* d_A, const float* d_MdM, const float* d_shift,
                      double* d_out, bool timed, bool gen);

}  // namespace emc

struct emc_shell_plan {
  int abs_flag = 0;  // wCCC{1}
  double w = 0.0;    // wCCC{2}
  int n_shells = 0;
  int64_t n_idx = 0;
  int64_t max_index = 0;
  int64_t* d_offsets = nullptr;
  uint32_t* d_idx = nullptr;
};
