/* BH_multi_xcf_Translational — a MEX of the same name, which MATLAB picks
 * before the .m (statistics/BH_multi_xcf_Translational.m:1-109), as
 * alignment/BH_alignRaw3d_v2.m:1333, :1607, :1749 call it:
 *   PEAK_COORD = BH_multi_xcf_Translational(rotPART_FT, REF_FT, peakMask, PEAK_COM)
 * rotPART_FT, REF_FT: full-grid complex single spectra of one size; peakMask:
 * single of that size, centred, or [] for none; PEAK_COM: the centre-of-mass
 * half-widths (scalar or [x y z]; 0 = no centre of mass). Returns the 1x3
 * double shift of the correlation peak relative to ceil((size+1)/2). An FFT
 * plan is kept per MEX file for the last size used.
 * -> emc_xcf_translational. */
#include "emc_mex.h"

static emc_fft_plan* g_plan = NULL;
static int g_plan_dims[3] = {0, 0, 0};

static void release_plan(void) {
  if (g_plan) emc_fft_plan_destroy(g_plan);
  g_plan = NULL;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static const char* id = "emc:BH_multi_xcf_Translational";
  (void)nlhs;
  if (nrhs != 4) mexErrMsgIdAndTxt(id, "expected rotPART_FT, REF_FT, peakMask, PEAK_COM");
  if (!mxIsSingle(prhs[0]) || !mxIsComplex(prhs[0]) || !mxIsSingle(prhs[1]) || !mxIsComplex(prhs[1]))
    mexErrMsgIdAndTxt(id, "rotPART_FT and REF_FT must be complex single");
  int dims[3], d2[3];
  emc_mex_dims3(prhs[0], dims);
  emc_mex_dims3(prhs[1], d2);
  if (memcmp(dims, d2, sizeof(dims)) != 0) mexErrMsgIdAndTxt(id, "rotPART_FT and REF_FT differ in size");
  const size_t nv = (size_t)dims[0] * dims[1] * dims[2];
  const int have_mask = !mxIsEmpty(prhs[2]);
  if (have_mask && (!mxIsSingle(prhs[2]) || mxGetNumberOfElements(prhs[2]) != nv))
    mexErrMsgIdAndTxt(id, "peakMask must be single of the spectra's size, or []");
  int com[3];
  for (int k = 0; k < 3; ++k)
    com[k] = (int)emc_mex_num(prhs[3], mxGetNumberOfElements(prhs[3]) == 1 ? 0 : k, id, "PEAK_COM");
  emc_ctx* ctx = emc_mex_ctx();
  emc_mex_on_exit = release_plan;
  if (!g_plan || memcmp(g_plan_dims, dims, sizeof(dims)) != 0) {
    release_plan();
    emc_mex_check(emc_fft_plan_create(ctx, dims, 1, &g_plan), id);
    memcpy(g_plan_dims, dims, sizeof(dims));
  }
  const size_t cbytes = nv * 2 * sizeof(float);
  void* d_a = emc_mex_upload(ctx, mxGetComplexSingles(prhs[0]), cbytes, id);
  void* d_b = NULL;
  void* d_mask = NULL;
  void* d_work = NULL;
  emc_status s = emc_malloc(ctx, cbytes, &d_b);
  if (s == EMC_OK) s = emc_memcpy_h2d(ctx, d_b, mxGetComplexSingles(prhs[1]), cbytes);
  if (s == EMC_OK && have_mask) s = emc_malloc(ctx, nv * sizeof(float), &d_mask);
  if (s == EMC_OK && have_mask) s = emc_memcpy_h2d(ctx, d_mask, mxGetSingles(prhs[2]), nv * sizeof(float));
  if (s == EMC_OK) s = emc_malloc(ctx, 2 * cbytes, &d_work);
  double peak[3] = {0.0, 0.0, 0.0};
  if (s == EMC_OK)
    s = emc_xcf_translational(ctx, g_plan, d_a, d_b, dims, (const float*)d_mask, com, d_work, peak, NULL);
  emc_free(ctx, d_a);
  if (d_b) emc_free(ctx, d_b);
  if (d_mask) emc_free(ctx, d_mask);
  if (d_work) emc_free(ctx, d_work);
  emc_mex_check(s, id);
  plhs[0] = mxCreateDoubleMatrix(1, 3, mxREAL);
  memcpy(mxGetDoubles(plhs[0]), peak, sizeof(peak));
}
