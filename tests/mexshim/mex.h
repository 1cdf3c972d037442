/*
 * Minimal stand-in for the MATLAB MEX API, for compiling and testing this
 * repository's own MEX gateways (sgldbench_amd/mex) on a machine without
 * MATLAB.  Test infrastructure only: it implements the subset of mxArray
 * calls the gateways use, with the semantics of the documented MATLAB API
 * (column-major dense arrays, cell arrays, char strings, uint64 scalars, a
 * sparse flag), and makes mexErrMsgIdAndTxt throw a C++ exception that the
 * test driver catches (MATLAB longjmps out of the MEX call instead).
 */
#ifndef SGLD_TEST_MEX_H
#define SGLD_TEST_MEX_H
#include <stddef.h>
#include <stdint.h>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef int32_t int32_T;
typedef struct mxArray_tag mxArray;
typedef enum {
    mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxVOID_CLASS,
    mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS,
    mxINT32_CLASS, mxUINT32_CLASS, mxINT64_CLASS, mxUINT64_CLASS
} mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif
double *mxGetPr(const mxArray *a);
void *mxGetData(const mxArray *a);
mwSize mxGetM(const mxArray *a);
mwSize mxGetN(const mxArray *a);
mwSize mxGetNumberOfElements(const mxArray *a);
mwSize mxGetNumberOfDimensions(const mxArray *a);
const mwSize *mxGetDimensions(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsInt32(const mxArray *a);
int mxIsUint8(const mxArray *a);
int mxIsUint64(const mxArray *a);
int mxIsLogical(const mxArray *a);
int mxIsCell(const mxArray *a);
int mxIsChar(const mxArray *a);
int mxIsSparse(const mxArray *a);
int mxIsComplex(const mxArray *a);
mxArray *mxGetCell(const mxArray *a, mwIndex i);
double mxGetScalar(const mxArray *a);
char *mxArrayToString(const mxArray *a);
void mxFree(void *p);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity c);
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
void mexWarnMsgIdAndTxt(const char *id, const char *fmt, ...);
/* the gateway's entry point, C linkage as MATLAB's mex.h declares it */
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);
#ifdef __cplusplus
}
#endif
#endif
