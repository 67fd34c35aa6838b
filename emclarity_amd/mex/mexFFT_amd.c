/* mexFFT_amd — drop-in for mexFiles/mexFFT.cu:16-239 as
 * testScripts/fourierTransformer.m:123, :125, :154, :181 calls it:
 *   [ft, plan, planInv] = mexFFT_amd(vol, invTrim)        make the plans
 *   out = mexFFT_amd(vol, invTrim, plan, planInv)          reuse them
 *   mexFFT_amd(1, invTrim, plan, planInv)                  destroy them
 * Real single input: forward R2C, output (nx/2+1) x ny x nz complex single.
 * Complex single input: inverse C2R, output (2*nxh - invTrim) x ny x nz real
 * single. Unscaled both ways, as the reference (do_scale = false, :37).
 * One library plan serves both directions; plan and planInv are the same
 * uint64 handle (a heap record of the plan and its real dimensions), kept
 * by the caller between calls as the reference keeps its cufft handles.
 * 1-D and 2-D inputs are planned as nz = 1 (and ny = 1) volumes.
 * -> emc_fft_plan_create / emc_fft_r2c / emc_fft_c2r / emc_fft_plan_destroy. */
#include "emc_mex.h"

#define EMC_MEX_FFT_MAGIC 0x454d43464654ull /* "EMCFFT" */

typedef struct {
  uint64_t magic;
  emc_fft_plan* plan;
  int dims[3]; /* real-space nx, ny, nz */
} emc_mex_fft;

static emc_mex_fft* handle_of(const mxArray* a, const char* id) {
  if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1) mexErrMsgIdAndTxt(id, "plan must be a uint64 handle");
  emc_mex_fft* h = (emc_mex_fft*)(uintptr_t)((const uint64_T*)mxGetData(a))[0];
  if (!h || h->magic != EMC_MEX_FFT_MAGIC) mexErrMsgIdAndTxt(id, "not a live mexFFT_amd plan handle");
  return h;
}

static mxArray* handle_array(emc_mex_fft* h) {
  mxArray* a = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
  ((uint64_T*)mxGetData(a))[0] = (uint64_T)(uintptr_t)h;
  return a;
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static const char* id = "emc:mexFFT";
  if (nrhs != 2 && nrhs != 4) mexErrMsgIdAndTxt(id, "expected (vol, invTrim) or (vol, invTrim, plan, planInv)");
  const int inv_trim = (int)emc_mex_num(prhs[1], 0, id, "invTrim");
  emc_mex_fft* h = nrhs == 4 ? handle_of(prhs[2], id) : NULL;
  if (h && mxGetNumberOfElements(prhs[0]) == 1) {  /* destroy (:74-84) */
    emc_fft_plan_destroy(h->plan);
    h->magic = 0;
    free(h);
    return;
  }
  if (!mxIsSingle(prhs[0])) mexErrMsgIdAndTxt(id, "vol must be single");
  const int fwd = !mxIsComplex(prhs[0]);
  int in_dims[3], real_dims[3];
  emc_mex_dims3(prhs[0], in_dims);
  real_dims[0] = fwd ? in_dims[0] : 2 * in_dims[0] - inv_trim;
  real_dims[1] = in_dims[1];
  real_dims[2] = in_dims[2];
  if (real_dims[0] < 1) mexErrMsgIdAndTxt(id, "invTrim leaves no output columns");
  const int nxh = real_dims[0] / 2 + 1;
  if (!fwd && nxh != in_dims[0]) mexErrMsgIdAndTxt(id, "complex input of %d columns and invTrim %d", in_dims[0], inv_trim);
  emc_ctx* ctx = emc_mex_ctx();
  if (!h) {
    if (nlhs < 3) mexErrMsgIdAndTxt(id, "the first call returns [out, plan, planInv]");
    h = (emc_mex_fft*)calloc(1, sizeof(emc_mex_fft));
    if (!h) mexErrMsgIdAndTxt(id, "out of host memory");
    memcpy(h->dims, real_dims, sizeof(real_dims));
    if (emc_fft_plan_create(ctx, real_dims, 1, &h->plan) != EMC_OK) {
      free(h);
      emc_mex_check(EMC_ERR_FFT, id);
    }
    h->magic = EMC_MEX_FFT_MAGIC;
    plhs[1] = handle_array(h);
    plhs[2] = handle_array(h);
  } else if (memcmp(h->dims, real_dims, sizeof(real_dims)) != 0) {
    mexErrMsgIdAndTxt(id, "plan made for %d x %d x %d, called with %d x %d x %d", h->dims[0], h->dims[1], h->dims[2],
                      real_dims[0], real_dims[1], real_dims[2]);
  }
  const size_t nreal = (size_t)real_dims[0] * real_dims[1] * real_dims[2];
  const size_t ncplx = (size_t)nxh * real_dims[1] * real_dims[2];
  const size_t in_bytes = fwd ? nreal * sizeof(float) : ncplx * 2 * sizeof(float);
  const size_t out_bytes = fwd ? ncplx * 2 * sizeof(float) : nreal * sizeof(float);
  const void* host_in = fwd ? (const void*)mxGetSingles(prhs[0]) : (const void*)mxGetComplexSingles(prhs[0]);
  void* d_in = emc_mex_upload(ctx, host_in, in_bytes, id);
  void* d_out = NULL;
  emc_status s = emc_malloc(ctx, out_bytes, &d_out);
  if (s == EMC_OK) s = fwd ? emc_fft_r2c(h->plan, (const float*)d_in, d_out) : emc_fft_c2r(h->plan, d_in, (float*)d_out);
  if (s == EMC_OK) {
    const mwSize nd = mxGetNumberOfDimensions(prhs[0]);
    mwSize od[3] = {(mwSize)(fwd ? nxh : real_dims[0]), (mwSize)real_dims[1], (mwSize)real_dims[2]};
    plhs[0] = mxCreateNumericArray(nd < 3 ? nd : 3, od, mxSINGLE_CLASS, fwd ? mxCOMPLEX : mxREAL);
    void* host_out = fwd ? (void*)mxGetComplexSingles(plhs[0]) : (void*)mxGetSingles(plhs[0]);
    s = emc_memcpy_d2h(ctx, host_out, d_out, out_bytes);
  }
  emc_free(ctx, d_in);
  if (d_out) emc_free(ctx, d_out);
  emc_mex_check(s, id);
}
