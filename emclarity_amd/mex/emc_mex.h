/* Shared pieces of the MEX gateways over libemc_hip.so.
 *
 * MATLAB's gpuArray (mxGPUArray) exists only for NVIDIA devices, so these
 * gateways take host mxArrays, copy them into device buffers of the
 * library, call the C ABI of emc_hip.h and return host arrays. Each MEX file
 * holds one context (device 0, or the device named by the EMC_DEVICE
 * environment variable) for the life of the MATLAB session; mexAtExit
 * destroys it. Errors never exit MATLAB: every failing call raises
 * mexErrMsgIdAndTxt with the library's message.
 *
 * Build where MATLAB exists (interleaved complex API):
 *   mex -R2018a -I<repo>/include <gateway>.c -L<repo>/emclarity_amd/lib -lemc_hip
 */
#ifndef EMC_MEX_H
#define EMC_MEX_H

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "emc_hip.h"
#include "mex.h"

static emc_ctx* emc_mex_context_ = NULL;

/* What a gateway holds beyond the context (plans, resident volumes): MATLAB
 * keeps one exit hook per MEX file, so a gateway sets this instead of
 * calling mexAtExit itself; it runs before the context goes. */
static void (*emc_mex_on_exit)(void) = NULL;

static inline void emc_mex_release_(void) {
  if (emc_mex_on_exit) emc_mex_on_exit();
  if (emc_mex_context_) emc_ctx_destroy(emc_mex_context_);
  emc_mex_context_ = NULL;
}

/* The MEX file's context, created on first use. */
static inline emc_ctx* emc_mex_ctx(void) {
  if (!emc_mex_context_) {
    const char* dev = getenv("EMC_DEVICE");
    if (emc_ctx_create(dev ? atoi(dev) : 0, &emc_mex_context_) != EMC_OK)
      mexErrMsgIdAndTxt("emc:ctx", "%s", emc_last_error());
    mexAtExit(emc_mex_release_);
  }
  return emc_mex_context_;
}

/* Raise the library's last error as a MATLAB error when s is not EMC_OK. */
static inline void emc_mex_check(emc_status s, const char* id) {
  if (s != EMC_OK) mexErrMsgIdAndTxt(id, "%s", emc_last_error());
}

/* Element i of a numeric input of any class as a double (the reference
 * gateways read their arguments through typed pointers and break when MATLAB
 * passes another class; these accept any real numeric class). */
static inline double emc_mex_num(const mxArray* a, mwSize i, const char* id, const char* what) {
  if (!mxIsNumeric(a) && !mxIsLogical(a)) mexErrMsgIdAndTxt(id, "%s must be numeric", what);
  if (mxIsComplex(a)) mexErrMsgIdAndTxt(id, "%s must be real", what);
  if (i >= mxGetNumberOfElements(a)) mexErrMsgIdAndTxt(id, "%s has %d elements, needs %d", what,
                                                       (int)mxGetNumberOfElements(a), (int)i + 1);
  switch (mxGetClassID(a)) {
    case mxDOUBLE_CLASS: return ((const double*)mxGetData(a))[i];
    case mxSINGLE_CLASS: return ((const float*)mxGetData(a))[i];
    case mxINT8_CLASS: return ((const int8_T*)mxGetData(a))[i];
    case mxUINT8_CLASS: return ((const uint8_T*)mxGetData(a))[i];
    case mxINT16_CLASS: return ((const int16_T*)mxGetData(a))[i];
    case mxUINT16_CLASS: return ((const uint16_T*)mxGetData(a))[i];
    case mxINT32_CLASS: return ((const int32_T*)mxGetData(a))[i];
    case mxUINT32_CLASS: return ((const uint32_T*)mxGetData(a))[i];
    case mxINT64_CLASS: return (double)((const int64_T*)mxGetData(a))[i];
    case mxUINT64_CLASS: return (double)((const uint64_T*)mxGetData(a))[i];
    case mxLOGICAL_CLASS: return ((const mxLogical*)mxGetData(a))[i] ? 1.0 : 0.0;
    default: mexErrMsgIdAndTxt(id, "%s: unsupported class", what);
  }
  return 0.0;
}

/* n floats from a numeric input: its elements, or its one element repeated
 * (per-tilt arguments given as scalars). The caller frees the result. */
static inline float* emc_mex_floats(const mxArray* a, mwSize n, const char* id, const char* what) {
  const mwSize m = mxGetNumberOfElements(a);
  if (m != n && m != 1) mexErrMsgIdAndTxt(id, "%s has %d elements, expected 1 or %d", what, (int)m, (int)n);
  float* v = (float*)mxMalloc(sizeof(float) * (n ? n : 1));
  for (mwSize i = 0; i < n; ++i) v[i] = (float)emc_mex_num(a, m == 1 ? 0 : i, id, what);
  return v;
}

/* A device copy of a host input (bytes of mxGetData). */
static inline void* emc_mex_upload(emc_ctx* ctx, const void* host, size_t bytes, const char* id) {
  void* d = NULL;
  emc_mex_check(emc_malloc(ctx, bytes ? bytes : 1, &d), id);
  const emc_status s = bytes ? emc_memcpy_h2d(ctx, d, host, bytes) : EMC_OK;
  if (s != EMC_OK) {
    emc_free(ctx, d);
    emc_mex_check(s, id);
  }
  return d;
}

/* The three dimensions of an array (trailing singleton dimensions filled). */
static inline void emc_mex_dims3(const mxArray* a, int dims[3]) {
  const mwSize nd = mxGetNumberOfDimensions(a);
  const mwSize* d = mxGetDimensions(a);
  for (int k = 0; k < 3; ++k) dims[k] = k < (int)nd ? (int)d[k] : 1;
}

#endif
