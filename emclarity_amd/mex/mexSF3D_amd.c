/* mexSF3D_amd — drop-in for mexFiles/mexSF3D.cu:85-303 as
 * masking/BH_weightMaskMex.m:47-51 calls it:
 *   SF3D = mexSF3D_amd(doHalfGrid, doSqCTF, wantedSize[3], pixelSize, waveLength,
 *                      CS, defocus1, defocus2, defocusAst, AmpContrast, nTilts,
 *                      tiltAngle, exposure, occupancy [, launch])
 * Per-tilt arguments have nTilts elements (or one, used for every tilt);
 * units as the reference's: pixel size and wavelength in Angstrom, CS in mm,
 * defoci in Angstrom, astigmatism angle and tilt angles in degrees. The
 * optional launch argument is accepted and ignored (all tilts go in one
 * launch). Returns the full-grid, centred nx x ny x nz single volume.
 * -> emc_sf3d. */
#include "emc_mex.h"

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  static const char* id = "emc:mexSF3D";
  (void)nlhs;
  if (nrhs != 14 && nrhs != 15) mexErrMsgIdAndTxt(id, "requires 14 inputs");
  const int half = emc_mex_num(prhs[0], 0, id, "doHalfGrid") != 0.0;
  const int sq = emc_mex_num(prhs[1], 0, id, "doSqCTF") != 0.0;
  int dims[3];
  for (int k = 0; k < 3; ++k) dims[k] = (int)emc_mex_num(prhs[2], k, id, "wantedSize");
  const double nt_in = emc_mex_num(prhs[10], 0, id, "nTilts");
  if (nt_in < 1 || nt_in > 1e6) mexErrMsgIdAndTxt(id, "nTilts must be positive");
  const mwSize nt = (mwSize)nt_in;
  /* mxMalloc'd copies: MATLAB frees them if an argument check raises */
  float* pix = emc_mex_floats(prhs[3], nt, id, "pixelSize");
  float* lam = emc_mex_floats(prhs[4], nt, id, "waveLength");
  float* cs = emc_mex_floats(prhs[5], nt, id, "CS");
  float* df1 = emc_mex_floats(prhs[6], nt, id, "defocus1");
  float* df2 = emc_mex_floats(prhs[7], nt, id, "defocus2");
  float* ast = emc_mex_floats(prhs[8], nt, id, "defocusAst");
  float* amp = emc_mex_floats(prhs[9], nt, id, "AmpContrast");
  float* tilt = emc_mex_floats(prhs[11], nt, id, "tiltAngle");
  float* expo = emc_mex_floats(prhs[12], nt, id, "exposure");
  float* occ = emc_mex_floats(prhs[13], nt, id, "occupancy");
  emc_ctx* ctx = emc_mex_ctx();
  const size_t bytes = (size_t)dims[0] * dims[1] * dims[2] * sizeof(float);
  void* d_out = NULL;
  emc_status s = emc_malloc(ctx, bytes, &d_out);
  if (s == EMC_OK)
    s = emc_sf3d(ctx, half, sq, dims, (int)nt, pix, lam, cs, df1, df2, ast, amp, tilt, expo, occ, (float*)d_out);
  if (s == EMC_OK) {
    const mwSize od[3] = {(mwSize)dims[0], (mwSize)dims[1], (mwSize)dims[2]};
    plhs[0] = mxCreateNumericArray(3, od, mxSINGLE_CLASS, mxREAL);
    s = emc_memcpy_d2h(ctx, mxGetData(plhs[0]), d_out, bytes);
  }
  if (d_out) emc_free(ctx, d_out);
  emc_mex_check(s, id);
}
