/*
 * emc_hip.h -- C ABI of the MI355X (gfx950) rotate-weight-correlate library.
 *
 * Every entry point replaces one MATLAB-facing function of emClarity's
 * subtomogram-alignment / template-matching hot path; the reference
 * interface is cited beside each declaration (paths relative to the
 * emClarity source tree).
 *
 * Conventions
 *   - Volumes are FP32, x fastest: linear index x + nx*(y + ny*z), the same
 *     order as MATLAB's column-major arrays (mexFiles/mexXform3d.cu:86).
 *   - dims[3] = {nx, ny, nz}.
 *   - Half spectra (R2C output) are complex64 interleaved, (nx/2+1) x ny x nz,
 *     x fastest (mexFiles/mexFFT.cu:70-86, 136-146).
 *   - Pointers named d_* are device pointers (allocate with emc_malloc or any
 *     allocator of the same HIP runtime); everything else is host memory.
 *   - Calls are stream-ordered on the context's stream and return after the
 *     work is queued unless stated otherwise ("synchronous" entries return
 *     results to host memory and wait). Calls on one context are not
 *     thread-safe; use one context per host thread.
 *   - Errors never exit the process: every call returns an emc_status and
 *     emc_last_error() gives the message of the last failing call on the
 *     calling thread (the reference exits MATLAB on a CUDA error,
 *     mexFiles/include/helper_cuda.h:541-548).
 */
#ifndef EMC_HIP_H
#define EMC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  EMC_OK = 0,
  EMC_ERR_INVALID = 1,   /* bad argument (null pointer, bad size, bad flag) */
  EMC_ERR_HIP = 2,       /* HIP runtime error */
  EMC_ERR_FFT = 3,       /* rocFFT error */
  EMC_ERR_ALLOC = 4,     /* device allocation failed */
  EMC_ERR_RANGE = 5,     /* data-dependent range error (e.g. COM box outside map) */
  EMC_ERR_NODEVICE = 6,  /* no HIP device */
  EMC_ERR_IO = 7         /* a file could not be opened, read or written */
} emc_status;

typedef struct emc_ctx emc_ctx;
typedef struct emc_fft_plan emc_fft_plan;

/* ---------------------------------------------------------------- context */

const char* emc_last_error(void);
const char* emc_version(void);

/* Create a context on HIP device `device` with its own default-priority
 * stream. Replaces the implicit per-MATLAB-process CUDA context
 * (alignment/BH_alignRaw3d_v2.m:774-776 gpuDevice selection). */
emc_status emc_ctx_create(int device, emc_ctx** out);
emc_status emc_ctx_destroy(emc_ctx* ctx);
emc_status emc_ctx_synchronize(emc_ctx* ctx);
/* Make later work on `waiter`'s stream wait for the work queued on
 * `signaller`'s stream so far (an event recorded on one, waited on by the
 * other; no host wait). Both contexts on the same device. */
emc_status emc_ctx_wait(emc_ctx* waiter, emc_ctx* signaller);
/* The hipStream_t of the context, as void*, for callers that interleave their
 * own work on the same stream. */
void* emc_ctx_stream(emc_ctx* ctx);

/* Kernel timing with HIP events on the context stream (for roofline reports):
 * enable/disable, reset, and read the accumulated milliseconds and launch
 * count of one named kernel ("xform3d", "align_fused", ...). Reading
 * synchronises the stream. */
emc_status emc_timing_enable(emc_ctx* ctx, int on);
emc_status emc_timing_reset(emc_ctx* ctx);
emc_status emc_timing_read(emc_ctx* ctx, const char* name, double* ms_total, int64_t* launches);

/* Launch-shape parameters of the batched search, per context (no
 * counterpart in the reference: they are what the autotuner sets).
 *   "search_tile"  kx extent of the (kx, ky) column patches of the search
 *                  schedule, whose ky extent is 256 / search_tile (one
 *                  workgroup's columns), 2..128 (default 16: 16 x 16)
 *   "grad_blocks"  workgroups an adjoint evaluation aims for, 256..65536
 *                  (default 4096)
 *   "search_passes" passes of the search schedule per launch, 1..4
 *                  (default 2): they share the launch's last, partly
 *                  filled round of workgroups
 *   "search_zsplit" kz segments per support column, each in its own pass
 *                  of the search schedule, 1..4 (default 1)
 *   "tsearch_pruned" 2 (default): emc_template_search_chunk forms the
 *                  padded template's spectrum as direct DFTs over the
 *                  template block (every template edge <= 64) fused with
 *                  the product by the tomogram spectrum, the z-direction DFT
 *                  and the product as FP32 matrix-core (MFMA) products; 1:
 *                  the same with the z stage on the vector units; 0: rocFFT
 *                  R2C of the zero-padded volume (set with box 0)
 *   "refine_graph" 1 (default): emc_align_refine replays its iteration
 *                  chain as a captured HIP graph; 0: eager launches
 *   "grad_kernel"  adjoint evaluation order: 1 (default) brick-major (the
 *                  table in bricks, one workgroup per (brick, particle),
 *                  the bricks dealt to the XCDs so each XCD's L2 holds the
 *                  brick all particles read), 0 support-list order
 *                  (workgroups over chunks of each particle's support)
 *   "grad_brick"   brick side of the brick-major adjoint, 16 or 32
 *                  (default 32)
 *   "grad_waves"   waves per SIMD the 32-brick adjoint kernel's registers
 *                  are allocated for, 4 (default) or 5
 *   "align_grid_batch" orientations per launch chain of emc_align_grid /
 *                  emc_align_grid_padded, 1..64 (default 16; capped so the
 *                  block's buffers stay within 1 GiB). Applies to one
 *                  symmetry mate, no window padding and unshelled scores;
 *                  otherwise, or with 1, one orientation at a time
 * box = 0 sets the value for every box size, else for that box size only
 * (which wins). emc_tune_get returns the value in effect for `box`. */
emc_status emc_tune_set(emc_ctx* ctx, const char* name, int box, int value);
emc_status emc_tune_get(emc_ctx* ctx, const char* name, int box, int* value);

/* Device memory helpers so a host without a GPU array type (C harness, MEX
 * gateway on AMD, ctypes) can drive the library. Copies are synchronous.
 * emc_free returns an emc_malloc block to the context's cache without
 * waiting: a later emc_malloc on the same context may hand it out again,
 * ordered after the work already queued on the context's stream (work the
 * caller queued on other streams must be finished first). Blocks not from
 * emc_malloc are freed after the stream drains. emc_ctx_destroy releases
 * the cache. */
emc_status emc_malloc(emc_ctx* ctx, size_t bytes, void** d_ptr);
emc_status emc_free(emc_ctx* ctx, void* d_ptr);
emc_status emc_memcpy_h2d(emc_ctx* ctx, void* d_dst, const void* src, size_t bytes);
/* A host-to-device copy that does not wait for the work already queued on
 * the context's stream: src is staged through a pinned ring buffer of the
 * context and the call returns once src may be reused; the copy is
 * stream-ordered. Copies above 8 MiB wait for the stream as emc_memcpy_h2d
 * does. No reference counterpart (MATLAB's gpuArray uploads). */
emc_status emc_memcpy_h2d_async(emc_ctx* ctx, void* d_dst, const void* src, size_t bytes);
emc_status emc_memcpy_d2h(emc_ctx* ctx, void* dst, const void* d_src, size_t bytes);
emc_status emc_memcpy_d2d(emc_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
emc_status emc_memset_zero(emc_ctx* ctx, void* d_dst, size_t bytes);

/* --------------------------------------------------------- host geometry */

/* BH_defineMatrix (coordinates/BH_defineMatrix.m:51-126), FP64.
 * convention: 0 Bah (ZXZ), 1 TILT (angles[0] about Y), 2 SPIDER, 3 Helical,
 * 4 IMOD. direction: 0 'forward', 1 'inv', 2 'forwardVector', 3 'invVector'.
 * R_colmajor receives the 3x3 matrix in MATLAB column-major order. */
emc_status emc_define_matrix(const double angles_deg[3], int convention, int direction,
                             double R_colmajor[9]);

/* eulerSearch grid (coordinates/eulerSearch.m:38-177), random start off.
 * symmetry: "C<n>", "D<n>", "O", "I". Writes up to `capacity` (phi,theta,psi)
 * triples in the reference loop order theta->phi->psi and the total count
 * in *n_out (call with capacity 0 to size the buffer). */
emc_status emc_euler_search(const char* symmetry, double theta_max, double theta_step,
                            double psi_max, double psi_step, double* angles_out,
                            int64_t capacity, int64_t* n_out);

/* ----------------------------------------------------- mexXform3d (A1/A2) */

/* mexXform3d(size, vol, single R[9], single t[3], logical doFwd)
 * (mexFiles/mexXform3d.cu:93-324, kernel :22-87). Trilinear rotate + shift
 * about index n/2 with zero outside the volume; output has the input's dims.
 * fwd != 0: p = R (u - t); fwd == 0: p = R u + t (u = index - n/2).
 * R is the 9-vector MATLAB passes (column-major), read row-wise. Exact FP32
 * trilinear weights (the reference's texture unit uses 8-bit fixed-point
 * weights). d_out must not alias d_in. */
emc_status emc_xform3d(emc_ctx* ctx, const float* d_in, const int dims[3],
                       const float R_colmajor[9], const float t[3], int fwd, float* d_out);

/* interpolator.interp3d with symmetry (testScripts/interpolator.m:99-154):
 * out = (1/nsym) sum_i xform(vol, M * S_i, t, fwd), products in FP64 then
 * cast to FP32 as the reference does (:110, :163). S holds nsym column-major
 * 3x3 matrices (nsym <= 60). nsym == 1 with S == NULL means C1. */
emc_status emc_interp3d(emc_ctx* ctx, const float* d_in, const int dims[3],
                        const double M_colmajor[9], int nsym, const double* S_colmajor,
                        const float t[3], int fwd, float* d_out);

/* --------------------------------------------------- mexSF3D / mexCTF (A5-A7) */

/* mexSF3D(half, sqCTF, size, pix[], lambda[], Cs[], df1[], df2[], astig[],
 * amp[], nTilts, tilt[], exposure[], occupancy[]) (mexFiles/mexSF3D.cu:85-303;
 * ctf kernel mexFiles/utils/ctf.cu:68-128). Per-tilt arrays are host arrays
 * of length n_tilts; units as the reference: pixel and wavelength in
 * Angstrom, Cs in mm, defocus in Angstrom, astigmatism angle in degrees,
 * amplitude contrast as a fraction, tilt in degrees. d_out: nx*ny*nz FP32,
 * full grid, centred. All tilts are handled in two launches. */
emc_status emc_sf3d(emc_ctx* ctx, int half, int sq_ctf, const int dims[3], int n_tilts,
                    const float* pixel, const float* wavelength, const float* cs_mm,
                    const float* df1, const float* df2, const float* astig_deg,
                    const float* amp_contrast, const float* tilt_deg,
                    const float* exposure, const float* occupancy, float* d_out);

/* BH_weightMaskMex (masking/BH_weightMaskMex.m:1-59): TLT table (n_tilts rows
 * x 22 columns, row-major doubles as loaded from a *_ctf.tlt file), sampling,
 * particle centre xyz and recon geometry rows (origin size [3], shift [3])
 * -> SF3D - (min(SF3D) - 1e-4). */
emc_status emc_weight_mask(emc_ctx* ctx, const int dims[3], double sampling,
                           const double* tlt_rowmajor, int n_tilts,
                           const double xyz_subtomo[3], const double recon_size[3],
                           const double recon_shift[3], float* d_out);

/* BH_weightMaskMex for many particles of one box size (masking/BH_weightMaskMex.m:1-59 per
 * particle, as alignment/BH_alignRaw3d_v2.m:1093-1094 calls it inside its particle loop):
 * particle p has its own TLT, n_tilts[p] rows x 22 columns (row-major doubles, the
 * particles' tables one after another in tlt_rowmajor), and its own centre
 * xyz_subtomo[3p..3p+2]; the recon geometry is shared. d_out: n_part volumes of dims,
 * one after another, each SF3D_p - (min(SF3D_p) - 1e-4). One launch writes the CTF
 * images of every (particle, tilt), one the SF3D voxels of every particle (with each
 * block's minimum), one folds the minima and one applies the shift (particles in chunks
 * whose CTF stack stays within 256 MiB). Stream-ordered; returns once the tilt table is
 * queued for the device. */
emc_status emc_weight_mask_batched(emc_ctx* ctx, const int dims[3], double sampling, int n_part,
                                   const double* tlt_rowmajor, const int32_t* n_tilts,
                                   const double* xyz_subtomo, const double recon_size[3],
                                   const double recon_shift[3], float* d_out);

/* mexCTF(half, sq, nX, nY, pix, lambda, Cs, df1, df2, astig, amp [, centered])
 * (mexFiles/mexCTF.cu:1-121, kernel mexFiles/utils/ctf.cu:12-65). Output
 * (half ? nx/2+1 : nx) x ny FP32. The reference gateway accepts occupancy and
 * exposure but launches the unweighted kernel (mexCTF.cu:101-111); so does
 * this entry. */
emc_status emc_ctf2d(emc_ctx* ctx, int half, int sq_ctf, int nx, int ny, float pixel,
                     float wavelength, float cs_mm, float df1, float df2, float astig_deg,
                     float amp_contrast, int centered, float* d_out);

/* ------------------------------------------------------------ mexFFT (A8/A9) */

/* mexFFT plan handles (mexFiles/mexFFT.cu:178-210): rank-3 FP32 plans for
 * dims {nx,ny,nz} and `batch` contiguous volumes, backed by rocFFT. One
 * handle holds the R2C, C2R and (on demand) C2C plans. No scaling, as the
 * reference (do_scale=false, :37). */
emc_status emc_fft_plan_create(emc_ctx* ctx, const int dims[3], int batch, emc_fft_plan** plan);
emc_status emc_fft_plan_destroy(emc_fft_plan* plan);
/* Real -> half spectrum (mexFFT real input, :98-100). */
emc_status emc_fft_r2c(emc_fft_plan* plan, const float* d_in, void* d_out_c64);
/* Half spectrum -> real (mexFFT complex input, nx_out = 2*nx_half - invTrim,
 * :98-128). d_in may be overwritten, as cuFFT's C2R does. */
emc_status emc_fft_c2r(emc_fft_plan* plan, void* d_in_c64, float* d_out);
/* Full complex 3D FFT (MATLAB gpuArray fftn/ifftn, A9). direction -1 forward
 * (unscaled), +1 inverse (unscaled; fftn/ifftn's 1/N is the caller's). */
emc_status emc_fft_c2c(emc_fft_plan* plan, const void* d_in_c64, void* d_out_c64, int direction);

/* ------------------------------------------ spectral helpers (A10-A12, A20) */

/* BH_bandpass3d(SIZE, HIGH_THRESH, HIGH_CUT, LOW_CUT, 'GPU', PIXEL_SIZE)
 * (masking/BH_bandpass3d.m:54-170), numeric pixel size. Full grid, FFT
 * order; half != 0 writes only the first nx/2+1 x-planes
 * (testScripts/fourierTransformer.m:313-319). */
emc_status emc_bandpass3d(emc_ctx* ctx, const int dims[3], int half, double high_thresh,
                          double high_cut, double low_cut, double pixel, float* d_out);

/* BH_padZeros3d(IMG, padLow, padTop, 'GPU', precision[, extrapVal])
 * (masking/BH_padZeros3d.m:37-229). Negative pads trim. taper != 0 applies
 * the 'singleTaper' 7-voxel cosine edge. d_out has dims in + low + top. */
emc_status emc_pad_zeros3d(emc_ctx* ctx, const float* d_in, const int dims[3],
                           const int pad_low[3], const int pad_top[3], float extrap_val,
                           int taper, float* d_out);

/* fourierTransformer.fwdFFT(vol, doNorm, doCenter, bp, ctf) options
 * (testScripts/fourierTransformer.m:92-145) applied in place to the R2C half
 * spectrum d_ft of a real volume of `dims` (nx/2+1 x-planes): f *= ctf
 * (d_ctf_half, NULL = none); without centring f *= (1/sqrt(numel))^do_norm;
 * f *= bandpass (d_bp_half: BH_bandpass3d's first nx/2+1 x-planes, NULL =
 * none); with do_center f(0) = 0 and f /= sqrt(2 sum|f|^2 over the x-planes
 * x < nx/2+1 - invTrim), invTrim = 2 for even nx, 1 for odd. Sums in FP64. */
emc_status emc_ft_condition(emc_ctx* ctx, void* d_ft_c64, const int dims[3], int do_norm, int do_center,
                            const float* d_bp_half, const float* d_ctf_half);

/* BH_bandLimitCenterNormalize(IMG, BANDPASS, '', PAD, 'single')
 * (masking/BH_bandLimitCenterNormalize.m:27-61) on the half spectrum: pad
 * with the image mean, R2C, F(0)=0, x bandpass (NULL = none; full-grid FFT
 * order FP32 volume of the padded size), scale so that the FULL-grid
 * sum |F|^2 equals numel^2. Means and sums in FP64. `plan` must match the
 * padded dims (batch 1). d_out: half spectrum of the padded size. */
emc_status emc_band_limit_center_normalize(emc_ctx* ctx, emc_fft_plan* plan, const float* d_img,
                                           const int dims[3], const int pad_low[3],
                                           const int pad_top[3], const float* d_bandpass,
                                           void* d_out_c64);

/* ref_FT1 = conj(BH_bandLimitCenterNormalize(ref .* volMask, 1, '', padCalc))
 * (alignment/BH_alignRaw3d_v2.m:607-608), the reference side of the
 * rotational CC. d_vol_mask may be NULL (mask of ones). `plan` must match the
 * padded dims. d_out: half spectrum of the padded size. */
emc_status emc_ref_ft1(emc_ctx* ctx, emc_fft_plan* plan, const float* d_ref, const float* d_vol_mask,
                       const int dims[3], const int pad_low[3], const int pad_top[3],
                       void* d_out_c64);

/* ------------------------------------------------ correlation (A13, A14) */

/* BH_multi_xcf_Translational(rotPART_FT, REF_FT, peakMask, PEAK_COM)
 * (statistics/BH_multi_xcf_Translational.m:1-109). Full-grid complex inputs
 * (nx*ny*nz complex64, FFT order), peak mask FP32 centred (NULL = none),
 * com[3] COM half-widths (0 = no COM). Synchronous: writes the shift
 * (x,y,z) in pixels relative to ceil((N+1)/2) to peak_out[3] and the map
 * maximum to *max_out (may be NULL). `plan` must match dims.
 * d_work: 2*nx*ny*nz complex64 scratch. Returns EMC_ERR_RANGE when the
 * COM box leaves the map, as the reference errors (:64-81). */
emc_status emc_xcf_translational(emc_ctx* ctx, emc_fft_plan* plan, const void* d_a_c64,
                                 const void* d_b_c64, const int dims[3], const float* d_peak_mask,
                                 const int com[3], void* d_work, double peak_out[3],
                                 double* max_out);

/* BH_multi_xcf_Rotational(rotPART_FT, REF_FT, wdgMask, refWdg, wCCC=0)
 * (statistics/BH_multi_xcf_Rotational.m:29-80), un-shelled form:
 * CC = sum Re(Wr P R Wp) / sqrt(sum|R Wp|^2 sum|P Wr|^2) over full-grid
 * complex64 inputs; weights are FP32 full grids (NULL = 1). FP64 sums.
 * Synchronous. */
emc_status emc_xcf_rotational(emc_ctx* ctx, const void* d_p_c64, const void* d_r_c64,
                              const float* d_wp, const float* d_wr, const int dims[3],
                              double* cc_out);

/* Shell-weighted form of BH_multi_xcf_Rotational (wCCC a cell,
 * statistics/BH_multi_xcf_Rotational.m:42-70, the cell built by
 * BH_multi_cRef's calc_shells): a plan holds the whole cell, abs_flag =
 * wCCC{1}, w = wCCC{2} and the shells' voxel lists (wCCC{3..end}: 1-based
 * linear indices into the full grid, concatenated, h_offsets[s] ..
 * h_offsets[s+1] for shell s), made once per cell. Per shell CCC_s =
 * sum Re(Wr P R Wp) / sqrt(sum|R Wp|^2 sum|P Wr|^2) over its voxels;
 * cc = sum_s f(CCC_s), f(c) = (c^2 / (c + w^-1.5))^3, |f| when abs_flag.
 * Synchronous. */
typedef struct emc_shell_plan emc_shell_plan;
emc_status emc_shell_plan_create(emc_ctx* ctx, int abs_flag, double w, int n_shells, const int64_t* h_offsets,
                                 const uint32_t* h_indices, emc_shell_plan** out);
emc_status emc_shell_plan_destroy(emc_ctx* ctx, emc_shell_plan* plan);
emc_status emc_xcf_rotational_shells(emc_ctx* ctx, const emc_shell_plan* plan, const void* d_p_c64,
                                     const void* d_r_c64, const float* d_wp, const float* d_wr, const int dims[3],
                                     double* cc_out);

/* -------------------------------------------- alignment loop (A15, A17) */

/* One particle's plain grid search, the two-stage body of
 * alignment/BH_alignRaw3d_v2.m:1119-1419 run on the GPU for every orientation
 * of `angles_pts` ((phi, theta, psi) triples, degrees, n_orient of them):
 *   stage 1: rotate the reference by RotMat' (shift = shift_val), the
 *            reference weight and peak mask; band-limit-normalise both sides;
 *            translational xcf with the particle wedge -> estPeak (:1282-1344);
 *   stage 2: rotate the particle (inv, shift estPeak) and its SF3D wedge,
 *            mask, band-limit-normalise, rotational CC (:1199-1256, :1371-1400).
 * RotMat = angles_colmajor * BH_defineMatrix([phi,theta,psi-phi],'Bah','inv')
 * (:1175-1176). Window W^3 = calc C^3 (padCalc zero; emc_align_grid_padded
 * takes W < C), C1 symmetry, one reference. Inputs (device, FP32):
 * particle W^3, ref W^3, ref_wgt_rot C^3
 * centred, ref_wgt C^3 FFT order (NULL = 1), ref_ft1 half spectrum of C^3
 * (conj(BH_bandLimitCenterNormalize(ref.*volMask,1,...)), :607), vol_mask W^3,
 * peak_mask C^3, bandpass C^3 FFT order, wedge = SF3D C^3 centred.
 * est0: the stage-1 input shift (shiftVAL in the coarse loop :1186, rXYZest
 * in the refinement :1512, finalrXYZest in the final pass :1706); stage 2
 * alone (mode 2) uses est0 as the particle shift (the initial-CCC call,
 * :1222-1232, with angles_pts = {0,0,0}). mode: 1 stage 1 only, 2 stage 2
 * only, 3 both. The refinement list passes (phi, theta, psi) with the
 * matrix built from [phi, theta, psi - phi] in every mode (:1489-1515).
 * Outputs (host): ccc_out[n_orient] (0 without stage 2) and
 * shift_out[3*n_orient] (estPeak = stage-1 shift + est0, or est0).
 * Synchronous; EMC_ERR_RANGE if a COM box leaves the map. */
emc_status emc_align_grid(emc_ctx* ctx, int n, const float* d_particle, const float* d_ref,
                          const float* d_ref_wgt_rot, const float* d_ref_wgt,
                          const void* d_ref_ft1_c64, const float* d_vol_mask,
                          const float* d_peak_mask, const float* d_bandpass, const float* d_wedge,
                          const double angles_colmajor[9], const double est0[3],
                          const int peak_com[3], const double* angles_pts, int n_orient, int mode,
                          double* ccc_out, double* shift_out);
/* The same with window W^3 < calc C^3 (BH_multi_validArea's sizeWindow and
 * sizeCalc, alignment/BH_alignRaw3d_v2.m:304): particle, ref and vol_mask
 * are W^3, the other volumes C^3. The rotated reference (stage 1), the
 * particle (stage 1, once) and the rotated, masked particle (stage 2) are
 * padded to C^3 with their own mean, pad_low[k] planes below and
 * C - W - pad_low[k] above (padCalc, BH_bandLimitCenterNormalize.m:48)
 * before their FFT; every rotation runs at its volume's own size. Stage 2
 * rotates the particle and its wedge with the nsym symmetry mates R * S_i
 * (S_colmajor: nsym column-major 3x3, NULL for nsym = 1) and averages them,
 * as the reference's interp3d(..., 'inv', symmetry) calls (:1213-1254);
 * stage 1 is C1 as in the reference. Several references (refToAlign,
 * :1267-1278): host arrays of n_ref device pointers d_ref_wgts (the array or
 * its entries may be NULL), d_ref_ft1s and d_bandpasses; stage 1 rotates
 * d_ref (the reference code always rotates ref_FT2_tmp{1}{1}, :1108) with
 * the first reference's bandpass on the particle, whose shift the reference
 * code keeps (its sortrows over all-zero CCC columns keeps the first row,
 * :1198-1201); stage 2 scores the one rotated particle against every
 * reference: ccc_out[o * n_ref + r]; shells (NULL, or n_ref plans with
 * NULL entries) scores reference r with the shell-weighted CCC of its wCCC
 * cell (flgWeightCCC, :1355-1359, shells over the C^3 grid) instead of the
 * plain CCC. emc_align_grid is this call with
 * W = C, nsym = 1 and one reference. */
emc_status emc_align_grid_padded(emc_ctx* ctx, int w, int c, const int pad_low[3], int nsym,
                                 const double* S_colmajor, int n_ref, const float* d_particle,
                                 const float* d_ref, const float* d_ref_wgt_rot, const float* const* d_ref_wgts,
                                 const void* const* d_ref_ft1s_c64, const emc_shell_plan* const* shells,
                                 const float* d_vol_mask,
                                 const float* d_peak_mask, const float* const* d_bandpasses, const float* d_wedge,
                                 const double angles_colmajor[9], const double est0[3], const int peak_com[3],
                                 const double* angles_pts, int n_orient, int mode, double* ccc_out,
                                 double* shift_out);

/* ------------------------------- batched alignment and its adjoint (new) */

/* The batched form of the alignment chain (many particles x many
 * orientations per launch) and its hand-written adjoint. No single reference
 * function corresponds: they replace the per-orientation, per-particle loop
 * of alignment/BH_alignRaw3d_v2.m:1119-1419 for callers that align
 * thousands of particles. The reference is rotated in Fourier space
 * (2x-oversampled spectrum, trilinear), the correlation is evaluated only on
 * the candidate-shift cube by a pruned inverse DFT. Box sizes 8..128.
 *
 * CC_o(s) = sum_k Re[a conj(c_o) e^{2 pi i k.s/n}] / sqrt(sum|a|^2 sum|c_o|^2)
 *   a = P_centred * bandpass, c_o = F_ref(R_o^T k) * wedge_sym * bandpass,
 * summed over the half grid inside `support_radius` (index units). */

/* Reference table formats: EMC_TABLE_C64 = complex64 texels (8 bytes);
 * EMC_TABLE_F16 = complex with IEEE half parts (4 bytes), holding F * 2^s
 * with the power of two chosen so every part is below 2^14 (the normalised
 * correlation and its gradient do not depend on that scale). */
enum { EMC_TABLE_C64 = 0, EMC_TABLE_F16 = 1 };

/* Reference spectrum for one box size: the half spectrum F of the reference
 * zero-padded to 2n with its centre n/2 at the origin, stored centred as
 * texel quadruples: d_F2 holds (2n+1) x (2n+1) x (n+2) entries of four
 * texels of table_format, entry [(kz + n)(2n+1) + (ky + n)](n+2) + kx =
 * (F(kx, ky, kz), F(kx, ky+1, kz), F(kx, ky, kz+1), F(kx, ky+1, kz+1)) for
 * kx in [0, n+1], ky and kz in [-n, n] (indices taken mod 2n; kx = n+1 by
 * Hermitian symmetry), so entries kx and kx+1 hold a trilinear footprint
 * contiguously. */
emc_status emc_batch_reference_spectrum(emc_ctx* ctx, int n, const float* d_ref, int table_format,
                                        void* d_F2);

/* Per-particle preparation: d_particle n^3 (centred), d_bandpass n^3 FFT
 * order, d_wedge n^3 centred SF3D (NULL = 1). Writes Q (half grid complex64,
 * layout [kz][kx][ky]), N (same layout, float) and *d_A (one double). */
emc_status emc_batch_prepare_particle(emc_ctx* ctx, int n, const float* d_particle,
                                      const float* d_bandpass, const float* d_wedge, void* d_Q,
                                      float* d_N, double* d_A);

/* The same preparation for n_part particles of one box size in one call
 * (batched FFTs): h_particles / h_wedges are host arrays of n_part device
 * pointers (h_wedges NULL or an entry NULL = no wedge); particle i writes
 * d_Q + i*nh, d_N + i*nh and d_A[i], nh = (n/2+1) n^2. */
emc_status emc_batch_prepare_particles(emc_ctx* ctx, int n, int n_part, const float* const* h_particles,
                                       const float* d_bandpass, const float* const* h_wedges, void* d_Q,
                                       float* d_N, double* d_A);
/* The same with a centred real-space mask (n^3 FP32, NULL = none) applied to
 * every particle before its transform: the volMask of the reference's score
 * stage (alignment/BH_alignRaw3d_v2.m:1375, the rotated particle times
 * volMask; a mask symmetric about the centre commutes with the rotation), for
 * windows that hold neighbouring density. */
emc_status emc_batch_prepare_particles_masked(emc_ctx* ctx, int n, int n_part, const float* const* h_particles,
                                              const float* d_mask, const float* d_bandpass,
                                              const float* const* h_wedges, void* d_Q, float* d_N, double* d_A);

/* Score every (particle, orientation): particles p = 0..n_part-1 with their
 * Q/N/A stacked contiguously, start matrices d_start_cm (n_part x 9 doubles,
 * column-major), local angles (phi, theta, psi) in angles_pts (host), R =
 * start * BH_defineMatrix([phi,theta,psi-phi],'Bah','inv') composed in FP64
 * on the device. peak_radius 4, 6 or 8 (shift search sphere). Outputs
 * (device): d_score[n_part*n_orient] and d_shift[3*n_part*n_orient]
 * (x,y,z, parabolic sub-pixel), orientation fastest. Stream-ordered. */
emc_status emc_align_batched(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2, const void* d_Q,
                             const float* d_N, const double* d_A, const double* d_start_cm,
                             const double* angles_pts, int n_orient, int peak_radius,
                             double support_radius, float* d_score, float* d_shift);

/* Adjoint: score and gradient at one pose per particle. d_MdM: per particle
 * 36 floats = M = 2 R^T and dM/dphi, dM/dtheta, dM/dpsi (row-major, angles
 * in radians); d_shift: 3 floats per particle. d_out: 7 doubles per particle
 * = [CC, dCC/dphi, dCC/dtheta, dCC/dpsi, dCC/dsx, dCC/dsy, dCC/dsz].
 * Stream-ordered. */
emc_status emc_align_eval_grad(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2,
                               const void* d_Q,
                               const float* d_N, const double* d_A, const float* d_MdM,
                               const float* d_shift, double support_radius, double* d_out);

/* emc_align_eval_grad without allocating: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results (bit for bit: the same kernels, support order and workgroup split);
 * the support list and the partial sums live in a static 16 MiB device array
 * per GPU (held until the process ends) instead of context scratch. Rejects a
 * call whose support list plus n_part x workgroups x 88 bytes of partial sums
 * exceed 16 MiB (EMC_ERR_INVALID). Calls on one GPU share the array: calls that
 * overlap on one GPU (two threads, or two contexts' streams) race. A new
 * (n, support_radius) copies its support list into the array and waits for
 * the copy; the same pair as the previous call on that GPU reuses it. */
emc_status emc_align_eval_grad_noalloc(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2,
                                       const void* d_Q, const float* d_N, const double* d_A, const float* d_MdM,
                                       const float* d_shift, double support_radius, double* d_out);

/* Gradient refinement of one pose per particle, replacing the dense local
 * refinement grid of alignment/BH_alignRaw3d_v2.m:1448-1644: iters steps of
 * batched quasi-Newton ascent of the CC over a right-multiplied rotation
 * increment and the shift, all on the device (iters + 1 adjoint
 * evaluations, no host round trip). Step: (w, ds) = t H g, accepted when the
 * CC rises by >= 1e-4 t (H g).g, then the BFGS update of H and t = min(1, 2t),
 * else t *= 0.3; H starts as diag(step_angle/|g_w| x3, step_shift/|g_s| x3)
 * and is reset to it when H g is not an ascent direction. d_R: 9 doubles per
 * particle (row-major R, R maps particle to reference frame as in
 * emc_align_batched), d_s: 3 doubles (shift, px); both updated in place;
 * d_cc: the final CC per particle. Stream-ordered. */
/* The refinement's start from emc_align_batched's scores without a host
 * round trip: per particle the best orientation (first maximum, as the
 * grid search takes it), d_R0 = A * BH_defineMatrix([phi, theta, psi - phi],
 * 'Bah', 'inv') of that orientation (9 doubles, row-major, A the particle's
 * d_start_cm), d_s0 = its shift (3 doubles); d_best / d_best_score (NULL =
 * not wanted) the index and score. d_angles_pts: the n_orient (phi, theta,
 * psi) triples in device memory. Stream-ordered. */
emc_status emc_align_pick(emc_ctx* ctx, int n_part, int n_orient, const float* d_score, const float* d_shift,
                          const double* d_start_cm, const double* d_angles_pts, double* d_R0, double* d_s0,
                          int32_t* d_best, float* d_best_score);

emc_status emc_align_refine(emc_ctx* ctx, int n, int n_part, int table_format, const void* d_F2, const void* d_Q,
                            const float* d_N, const double* d_A, double* d_R, double* d_s, int iters,
                            double step_angle, double step_shift, double support_radius, double* d_cc);

/* ------------------------------------------------- template search (A16) */

/* One tomogram chunk of alignment/BH_templateSearch3d_2.m:713-887 for the
 * orientations angles_pts[n_orient] ((phi,theta,psi), global ids
 * first_angle_id .. first_angle_id+n_orient-1):
 *   tempRot = interp3d(template, [phi,theta,psi-phi], [1,1,1], 'forward');
 *   zero-pad into the chunk at pad_low; ccf = C2R(swapQuadrants R2C(chunk)
 *   conj(R2C(tempPAD))) cropped to the valid area [valid_low, valid_low +
 *   valid_dims); ccf /= std(ccf) (N-1); per voxel, if peak < ccf then
 *   peak = ccf, angle = id (strict <, :851-854, :883-887).
 * d_chunk: chunk dims (FP32, already preprocessed); d_template: template
 * dims; d_peak / d_angle: valid_dims FP32 / int32, updated in place. If
 * init_first != 0 the first orientation initialises peak = ccf and angle = 1
 * (the first chunk of a run, :802-823). Stream-ordered. */
emc_status emc_template_search_chunk(emc_ctx* ctx, const float* d_chunk, const int chunk_dims[3],
                                     const float* d_template, const int template_dims[3],
                                     const int pad_low[3], const int valid_low[3],
                                     const int valid_dims[3], const double* angles_pts,
                                     int n_orient, int first_angle_id, int init_first,
                                     float* d_peak, int32_t* d_angle);

/* -------------------------------- template-search pre-processing (next) */

/* statistics/BH_movingAverage_2.m:1-25: symmetric pad by ceil(F/2), the
 * separable normalised Gaussian EMC_gaussianKernel([1,F], 3)
 * (masking/EMC_gaussianKernel.m) through convn(.., 'same') along x, y, z
 * (testScripts/EMC_convn.m), crop. F = feature_size >= 2 with ceil(F/2) <=
 * every dimension. d_out may equal d_in. Stream-ordered. */
emc_status emc_moving_average(emc_ctx* ctx, const float* d_in, const int dims[3], int feature_size,
                              float* d_out);

/* alignment/BH_templateSearch3d_2.m:558-561: (x - avg) ./ sqrt(avg2 - avg.^2)
 * with avg, avg2 the moving averages of x and x.^2 (emc_moving_average).
 * d_out may equal d_in. Stream-ordered. */
emc_status emc_local_normalize(emc_ctx* ctx, const float* d_in, const int dims[3], int feature_size,
                               float* d_out);

/* medfilt3(x, [k k k]) (MATLAB's default 'symmetric' padding), k = 3, 5 or
 * 7 (Tmp_medianFilter, alignment/BH_templateSearch3d_2.m:492-499, :543-548).
 * Every dimension > k/2; d_out must not alias d_in. Stream-ordered. */
emc_status emc_median_filter3d(emc_ctx* ctx, const float* d_in, const int dims[3], int k, float* d_out);

/* One chunk of alignment/BH_templateSearch3d_2.m:532-589: x = (x - mean) / rms;
 * bandpass x = C2R(R2C(x) .* BH_bandpass3d half grid) / N with bp = {HIGH_THRESH,
 * HIGH_CUT, LOW_CUT, pixel} (fourierTransformer fwdFFT(x,0,0,bp) then
 * invFFT(.., 2)); med_filt 3, 5 or 7: medfilt3 of the band-passed chunk
 * (0 = none, the default Tmp_medianFilter); local normalisation over
 * stats_radius; d_out = sign * x .*
 * d_valid_mask (sign = -shouldBeCTF; mask NULL = 1). *h_masked_sum (if not
 * NULL) = sum(d_out(mask > 0.1)), the chunk's term of the global variance
 * (:589-592). Synchronous when h_masked_sum is given. */
emc_status emc_tsearch_preprocess_chunk(emc_ctx* ctx, const float* d_in, const int dims[3], const double bp[4],
                                        int med_filt, int stats_radius, float sign, const float* d_valid_mask,
                                        float* d_out, double* h_masked_sum);

/* BH_mask3d('sphere', dims, radius, [0,0,0]) (masking/BH_mask3d.m:126-160,
 * :462-466): ellipsoid of the given radii (x, y, z) about floor(N/2), six
 * cosine-weighted shells, convn with BH_multi_gaussian3d(5, 0.5), divided by
 * its maximum, the core reset to 1, then the 7-voxel edge taper. Every
 * dimension >= 14. Stream-ordered. */
emc_status emc_mask3d_sphere(emc_ctx* ctx, const int dims[3], const double radius[3], float* d_out);

/* EMC_maskShape(SHAPE, SIZE, RADIUS, METHOD, {'shift', shift; 'kernel', kernel}) for a 3-D SIZE,
 * origin 1, single precision (masking/EMC_maskShape.m:93-215): the peak, volume and
 * centre-of-mass masks of alignment/BH_alignRaw3d_v2.m:415, :482, :493, :504. shape 0
 * 'sphere', 1 'cylinder', 2 'rectangle'; dims, radius (each > 1) and shift (x, y, z);
 * kernel 0 = the binary shape, else the fraction of min(SIZE) the Gaussian roll-off spans
 * (the reference's default 0.04; at least 9 taps): the shape enlarged by half the kernel,
 * the outer ceil(K/2) layers zeroed, a separable 'same' convolution, normalised by its
 * maximum, values <= 0.001 set to 0. d_out: nx*ny*nz floats. Stream-ordered. */
emc_status emc_mask_shape(emc_ctx* ctx, int shape, const int dims[3], const double radius[3],
                          const double shift[3], double kernel, float* d_out);

/* Greedy peak extraction of alignment/BH_templateSearch3d_2.m:1100-1275 for
 * nPeaks = 1, on the padded score map d_mag (modified: erased in place) and
 * angle-id map d_ang (1-based ids into angle_list rows (phi, theta, psi-phi),
 * <= 0 = none), both dims. Repeats while fewer than max_peaks peaks
 * (2 * peakThreshold) were taken, fewer than max_tries maxima visited and
 * the maximum > high_thr: take the first maximum (x fastest); if its id > 0,
 * record (x, y, z) = 1-based voxel + centre of mass of the 3^3 box - rm_dim,
 * the angles and the score, rotate d_removal_mask ((2 rm_dim + 1)^3) by the
 * angles ('forward', convention as emc_define_matrix) and zero the map
 * where it is >= mask_cutoff around the voxel; else zero that voxel.
 * h_peaks: max_peaks x 7 doubles (x, y, z, phi, theta, psi-phi, score);
 * *n_found peaks. Synchronous. */
emc_status emc_tsearch_peaks(emc_ctx* ctx, float* d_mag, const int32_t* d_ang, const int dims[3],
                             const float* d_removal_mask, int rm_dim, const double* angle_list, int n_angles,
                             int convention, double mask_cutoff, double high_thr, int max_peaks, int max_tries,
                             double* h_peaks, int* n_found);

/* The same loop with nPeaks >= 1 peaks per taken maximum, as
 * alignment/BH_templateSearch3d_2.m:1156-1240 and :1097 (replaces its
 * peakMat): the centre of mass over the (2 bd + 1)^3 box, bd = 1 +
 * round(log(nPeaks)); then per extra peak 2 .. nPeaks, from that box's map
 * values with the taken id zeroed: if any is non-zero, cAng = the first
 * maximum and iSNR = the mean of the values with its id (then zeroed); the
 * angles are ANGLE_LIST(Ang(cAng)) when that id and Ang(cAng) (the map's
 * id at the box's linear index cAng, as the reference indexes it) are > 0,
 * else random: the taken peak's (x, y, z) times sampling_rate + nRandom^2 x
 * normals (nRandom = 2, 3, .. 10), with iSNR = 0.75 x the mean of the earlier
 * scores where a value was non-zero. h_peak_mat: max_peaks rows of
 * 10 n_peaks doubles, peak k at columns 10 k: x, y, z (times sampling_rate),
 * phi, theta, psi-phi, 0, 0, 0, score (iSNR for k > 0). normals: 3
 * (n_peaks - 1) standard normals per row (the reference draws randn; null
 * when n_peaks = 1). The map must hold at least the box's voxels. */
emc_status emc_tsearch_peaks_multi(emc_ctx* ctx, float* d_mag, const int32_t* d_ang, const int dims[3],
                                   const float* d_removal_mask, int rm_dim, const double* angle_list, int n_angles,
                                   int convention, double mask_cutoff, double high_thr, int max_peaks,
                                   int max_tries, int n_peaks, double sampling_rate, const double* normals,
                                   double* h_peak_mat, int* n_found);

/* d[i] *= s (the global-variance scaling of the chunk stack, :605-611). */
emc_status emc_scale(emc_ctx* ctx, float* d, int64_t count, double s);

/* ------------------------------------------- MRC files and subtomograms */

/* The 1024-byte MRC header in the field order and names of
 * @MRCImage/private/readHeader.m:43-172 (IMOD/MRC2014 layout). `swapped` is
 * 1 when the file's byte order is not the host's: readHeader.m:54-78
 * reopens the file in the other order when nX, nY or nZ read negative, or
 * all three exceed 65535. data_offset = 1024 + nBytesExtended. */
typedef struct emc_mrc_header {
  int32_t nx, ny, nz, mode;
  int32_t nxstart, nystart, nzstart;
  int32_t mx, my, mz;
  float cell[3];        /* cellDimensionX/Y/Z */
  float cell_angle[3];  /* cellAngleX/Y/Z */
  int32_t mapc, mapr, maps;
  float dmin, dmax, dmean;
  int16_t space_group, n_symmetry_bytes;
  int32_t n_bytes_extended;
  int16_t creator_id, n_bytes_per_section, serialem_type;
  int32_t imod_stamp, imod_flags;
  int16_t idtype, lens, nd1, nd2, vd1, vd2;
  float tilt_angles[6];
  float origin[3];      /* xOrigin, yOrigin, zOrigin */
  char map[4], machine_stamp[4];
  float rms;            /* densityRMS */
  int32_t n_labels;
  char labels[10][80];
  int32_t swapped;
  int64_t data_offset;
} emc_mrc_header;

/* MRCImage(path, 0) + getHeader (readHeader.m:43-186). Host only. */
emc_status emc_mrc_read_header(const char* path, emc_mrc_header* hdr);

/* getVolume(MRCImage(path), [i0 i1], [j0 j1], [k0 k1]) (@MRCImage/getVolume.m:
 * 40-298): range = {i0, i1, j0, j1, k0, k1}, 1-based and inclusive (x, y,
 * z), NULL = the whole volume. Modes 0 (read as signed int8, as getVolume's
 * 'int8' precision does), 1 (int16), 2 (float32) and 6 (uint16) convert to
 * float32; other modes are rejected. h_out: (i1-i0+1) x (j1-j0+1) x
 * (k1-k0+1) floats, x fastest. Host only. */
emc_status emc_mrc_read_volume(const char* path, const int range[6], float* h_out);

/* Sections k0..k1 (1-based, inclusive) of the file, as stored (nx*ny*mode
 * bytes per section, modes 0/1/2/6), into device memory d_raw, in host byte
 * order: read into two pinned staging buffers in turn while the previous one
 * is copied, byte order fixed on the device. The tomogram then stays
 * resident in its stored type (half the bytes of float32 for int16 maps).
 * Replaces the whole-tomogram load of BH_multi_loadOrBuild.m:180-183.
 * Synchronous. */
emc_status emc_mrc_upload(emc_ctx* ctx, const char* path, int k0, int k1, void* d_raw);

/* d_out[i] = float(d_raw[i]) for `count` values of MRC mode `mode`
 * (0 int8, 1 int16, 2 float32, 6 uint16). Stream-ordered. */
emc_status emc_mrc_to_float(emc_ctx* ctx, const void* d_raw, int mode, int64_t count, float* d_out);

/* SAVE_IMG(MRCImage(vol), path[, pixel[, origin]]) for a float32 volume
 * (@MRCImage/SAVE_IMG.m, private/setVolumeAndHeaderFromVolume.m,
 * private/writeHeader.m): mode 2, cell = dims (x pixel when pixel is not
 * NULL), angles 90, space group 1, IMOD stamp, 'MAP ', machine stamp
 * 'DA\0\0', one empty label counted; min, max, mean and rms =
 * std(std(std(vol))) (SAVE_IMG.m, the nested std of the reference). origin:
 * NULL = 0; n_origin 1 = -dims/2 (set before the pixel scaling, as
 * SAVE_IMG.m does); n_origin 3 = the given values. Host only. */
emc_status emc_mrc_write(const char* path, const float* h_data, const int dims[3], const double* pixel,
                         const double* origin, int n_origin);

/* Batched subtomogram extraction. For each of n particles:
 * BH_isWindowValid(tomo_dims, window, mask_radius, centre)
 * (logicals/BH_isWindowValid.m:42-65), the window of the resident tomogram
 * (alignment/BH_alignRaw3d_v2.m:1074-1076) and BH_padZeros3d(.., padLow,
 * padTop, 'GPU', 'singleTaper') (:1085-1086), in one launch for all
 * particles; taper 0 pads with plain zeros ('single', the averaging's cut,
 * transformations/BH_average3d.m:1328-1346). d_tomo: the tomogram as stored (mode 0/1/2/6, see
 * emc_mrc_upload), tomo_dims {nx, ny, nz}. h_centres: n x 3 (x, y, z) in the
 * reference's 1-based CENTER convention (positionList(:, 11:13) /
 * samplingRate + binShift, :1034). d_out: n windows of window[0] x
 * window[1] x window[2] floats. h_valid[i] = 1, or 0 when the window keeps
 * less than the mask needs ('noUse': that window is zero-filled);
 * h_shifts: n x 3 (shiftVAL = centre - floor(centre)). The window table is
 * staged through the context's pinned ring without waiting for queued work;
 * h_valid and h_shifts are filled on return, the cut is stream-ordered. */
emc_status emc_extract_subtomos(emc_ctx* ctx, const void* d_tomo, int mode, const int tomo_dims[3], int n,
                                const double* h_centres, const int window[3], const double mask_radius[3],
                                int taper, float* d_out, int32_t* h_valid, double* h_shifts);

/* ------------------------------------------------------------ averaging */

/* The per-particle body of transformations/BH_average3d.m:1417-1441 and its
 * half-set accumulation (:1465-1490) for n particles of one window size, in
 * the given order. Particle p: P = interpolator(window_p, M_p, t_p, 'inv',
 * symmetry) (M_p = h_mats + 9p, the geometry's column-major 17:25; t_p =
 * h_shifts + 3p, shiftVAL), then P = (P - mean(P(mask > 0.01))) ./
 * rms(..(mask > 0.01)) .* mask (d_interp_mask, window-sized), and
 * d_sum_odd (h_half[p] == 1) or d_sum_eve (== 2) += P * h_weight[p]. With
 * d_wedges (n x wedge_dims, e.g. emc_weight_mask per particle) the rotated
 * wedges interpolator(wedge_p, M_p, 0, 'inv', symmetry) * weight go to
 * d_wedge_odd / d_wedge_eve. S_colmajor: nsym C_n mates (NULL with nsym 1).
 * h_used[p] = 0 where the normalised particle would not be finite (the
 * reference zeroes and ignores it, :1445-1462). The sums are read and
 * updated in place: call again with the next batch. Synchronous. */
emc_status emc_average_accumulate(emc_ctx* ctx, int n, const int window[3], const float* d_windows,
                                  const double* h_mats, const double* h_shifts, const int32_t* h_half,
                                  const double* h_weight, int nsym, const double* S_colmajor,
                                  const float* d_interp_mask, const int wedge_dims[3], const float* d_wedges,
                                  float* d_sum_odd, float* d_sum_eve, float* d_wedge_odd, float* d_wedge_eve,
                                  int32_t* h_used);

#ifdef __cplusplus
}
#endif

#endif /* EMC_HIP_H */
