/* Declarations of the MATLAB C Matrix API and MEX functions that the
 * gateways under emclarity_amd/mex call, with the types and signatures
 * MathWorks documents for the interleaved complex API (mex -R2018a). Used
 * only to compile the gateways with -fsyntax-only where MATLAB is absent;
 * nothing here is linked or run. */
#ifndef EMC_TEST_MEX_API_H
#define EMC_TEST_MEX_API_H

#include <stddef.h>
#include <stdint.h>

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef int8_t int8_T;
typedef uint8_t uint8_T;
typedef int16_t int16_T;
typedef uint16_t uint16_T;
typedef int32_t int32_T;
typedef uint32_t uint32_T;
typedef int64_t int64_T;
typedef uint64_t uint64_T;
typedef unsigned char mxLogical;
typedef float mxSingle;
typedef double mxDouble;
typedef struct { float real, imag; } mxComplexSingle;

typedef enum {
  mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxVOID_CLASS,
  mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS,
  mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS, mxFUNCTION_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...);
int mexAtExit(void (*ExitFcn)(void));
void* mxMalloc(mwSize n);
void* mxGetData(const mxArray* pm);
mxSingle* mxGetSingles(const mxArray* pa);
mxDouble* mxGetDoubles(const mxArray* pa);
mxComplexSingle* mxGetComplexSingles(const mxArray* pa);
size_t mxGetNumberOfElements(const mxArray* pm);
mwSize mxGetNumberOfDimensions(const mxArray* pm);
const mwSize* mxGetDimensions(const mxArray* pm);
mxClassID mxGetClassID(const mxArray* pm);
int mxIsNumeric(const mxArray* pm);
int mxIsLogical(const mxArray* pm);
int mxIsComplex(const mxArray* pm);
int mxIsSingle(const mxArray* pm);
int mxIsUint64(const mxArray* pm);
int mxIsEmpty(const mxArray* pm);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity flag);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);

#endif
