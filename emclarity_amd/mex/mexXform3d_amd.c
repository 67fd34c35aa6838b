/* mexXform3d_amd — drop-in for mexFiles/mexXform3d.cu:93-324 as
 * interpolator.m calls it (:48, :56, :114, :124, :135):
 *   out = mexXform3d_amd(size[3], vol, R[9], t[3], doFwd)
 * size: the volume's [nx ny nz] (any numeric class), vol: real single
 * nx x ny x nz, R: 3x3 rotation (column-major, as MATLAB holds it; the
 * reference reads it through a float pointer), t: shift [x y z] in pixels,
 * doFwd: logical. Returns the rotated, shifted volume (single, same size):
 * trilinear, zero outside, about the centre floor(n/2). -> emc_xform3d. */
#include "emc_mex.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static const char* id = "emc:mexXform3d";
  (void)nlhs;
  if (nrhs != 5) mexErrMsgIdAndTxt(id, "expected 5 inputs: size, vol, R, t, doFwd");
  if (!mxIsSingle(prhs[1]) || mxIsComplex(prhs[1])) mexErrMsgIdAndTxt(id, "vol must be real single");
  int dims[3];
  for (int k = 0; k < 3; ++k) dims[k] = (int)emc_mex_num(prhs[0], k, id, "size");
  if ((size_t)dims[0] * dims[1] * dims[2] != mxGetNumberOfElements(prhs[1]))
    mexErrMsgIdAndTxt(id, "size does not match vol");
  float R[9], t[3];
  for (int k = 0; k < 9; ++k) R[k] = (float)emc_mex_num(prhs[2], k, id, "R");
  for (int k = 0; k < 3; ++k) t[k] = (float)emc_mex_num(prhs[3], k, id, "t");
  const int fwd = emc_mex_num(prhs[4], 0, id, "doFwd") != 0.0;
  emc_ctx* ctx = emc_mex_ctx();
  const size_t bytes = (size_t)dims[0] * dims[1] * dims[2] * sizeof(float);
  void* d_in = emc_mex_upload(ctx, mxGetData(prhs[1]), bytes, id);
  void* d_out = NULL;
  emc_status s = emc_malloc(ctx, bytes, &d_out);
  if (s == EMC_OK) s = emc_xform3d(ctx, (const float*)d_in, dims, R, t, fwd, (float*)d_out);
  if (s == EMC_OK) {
    const mwSize od[3] = {(mwSize)dims[0], (mwSize)dims[1], (mwSize)dims[2]};
    plhs[0] = mxCreateNumericArray(3, od, mxSINGLE_CLASS, mxREAL);
    s = emc_memcpy_d2h(ctx, mxGetData(plhs[0]), d_out, bytes);
  }
  emc_free(ctx, d_in);
  if (d_out) emc_free(ctx, d_out);
  emc_mex_check(s, id);
}
