// Ks = Solving_AssembleCmptStencilFromNonFinestLevel(KsPrevious,
// elementUpwardMap, interpolatingKe, localMapping, numProjectNodes) on the
// GPU: the MEX of src/MEXfuncs/Solving_AssembleCmptStencilFromNonFinestLevel.cpp:
// 8-143 (called at src/Solving_AssembleFEAstencil.m:56) over
// sgld_assemble_nonfinest; interpolatingKe / localMapping as for the finest level.
#include "sgld_mex.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    sgldmex::want_args(nrhs, 5, nlhs, 1, "MyToolbox:arrayProduct");
    const mxArray *Kp = prhs[0], *up = prhs[1], *iKe = prhs[2], *lm = prhs[3];
    if (!sgldmex::is_real_double(Kp))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:KsPreviousNotDouble", "KsPrevious must be a real double array.");
    const mwSize nd = mxGetNumberOfDimensions(Kp);
    const mwSize *d = mxGetDimensions(Kp);
    if (nd < 2 || nd > 3 || d[0] != 24 || d[1] != 24)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:KsPreviousDim", "KsPrevious must be 24x24xN.");
    const int64_t numPrev = nd == 3 ? (int64_t)d[2] : 1;
    if (!sgldmex::is_real_int32(up) || (mxGetN(up) != 64 && mxGetN(up) != 8))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix",
                          "Input elementUpwardMap must be an M*64 or M*8 int32 matrix.");
    if (!mxIsSparse(iKe) || !mxIsDouble(iKe))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotSparse", "Input interpolatingKe must be a sparse double matrix.");
    if (!sgldmex::is_real_int32(lm) || mxGetN(lm) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input localMapping must be a column vector.");
    const int64_t nq = (int64_t)mxGetN(up);
    const mwSize nC = mxGetM(up);
    const mwSize dims[3] = {24, 24, nC};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    sgldmex::check(sgld_assemble_nonfinest(mxGetPr(Kp), numPrev, (const int32_t *)mxGetData(up), (int64_t)nC, nq,
                                           mxGetPr(plhs[0])));
}
