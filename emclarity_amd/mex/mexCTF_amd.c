/* mexCTF_amd — drop-in for mexFiles/mexCTF.cu:1-121 as
 * BH_synthetic_mapBack.m:1788, :1854 calls it:
 *   CTF = mexCTF_amd(doHalfGrid, doSqCTF, nX, nY, pixelSize, waveLength, CS,
 *                    defocus1, defocus2, defocusAst, AmpContrast
 *                    [, calcCentered [, occupancy, exposure]])
 * Scalars of any numeric class (the reference reads nX and nY through int
 * pointers although its callers pass int16); units as the reference's.
 * occupancy and exposure are accepted and, as in the reference
 * (mexCTF.cu:101-111), not applied. Returns (doHalfGrid ? nX/2+1 : nX) x nY
 * single. -> emc_ctf2d. */
#include "emc_mex.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static const char* id = "emc:mexCTF";
  (void)nlhs;
  if (nrhs != 11 && nrhs != 12 && nrhs != 14) mexErrMsgIdAndTxt(id, "requires 11, 12 or 14 inputs");
  const int half = emc_mex_num(prhs[0], 0, id, "doHalfGrid") != 0.0;
  const int sq = emc_mex_num(prhs[1], 0, id, "doSqCTF") != 0.0;
  const int nx = (int)emc_mex_num(prhs[2], 0, id, "nX");
  const int ny = (int)emc_mex_num(prhs[3], 0, id, "nY");
  float v[7];
  static const char* names[7] = {"pixelSize", "waveLength", "CS", "defocus1", "defocus2", "defocusAst",
                                 "AmpContrast"};
  for (int k = 0; k < 7; ++k) v[k] = (float)emc_mex_num(prhs[4 + k], 0, id, names[k]);
  const int centered = nrhs >= 12 ? emc_mex_num(prhs[11], 0, id, "calcCentered") != 0.0 : 0;
  if (nx < 1 || ny < 1) mexErrMsgIdAndTxt(id, "nX and nY must be positive");
  const int ox = half ? nx / 2 + 1 : nx;
  emc_ctx* ctx = emc_mex_ctx();
  const size_t bytes = (size_t)ox * ny * sizeof(float);
  void* d_out = NULL;
  emc_status s = emc_malloc(ctx, bytes, &d_out);
  if (s == EMC_OK)
    s = emc_ctf2d(ctx, half, sq, nx, ny, v[0], v[1], v[2], v[3], v[4], v[5], v[6], centered, (float*)d_out);
  if (s == EMC_OK) {
    plhs[0] = mxCreateNumericMatrix((mwSize)ox, (mwSize)ny, mxSINGLE_CLASS, mxREAL);
    s = emc_memcpy_d2h(ctx, mxGetData(plhs[0]), d_out, bytes);
  }
  if (d_out) emc_free(ctx, d_out);
  emc_mex_check(s, id);
}
