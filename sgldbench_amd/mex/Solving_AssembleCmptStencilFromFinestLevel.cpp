// Ks = Solving_AssembleCmptStencilFromFinestLevel(Ke, E, elementUpwardMap,
// interpolatingKe, localMapping, numProjectNodes) on the GPU: the MEX of
// src/MEXfuncs/Solving_AssembleCmptStencilFromFinestLevel.cpp:12-169 (called
// at src/Solving_AssembleFEAstencil.m:33) over sgld_assemble_finest.  The
// trilinear interpolatingKe / localMapping pair is checked as the reference
// checks it; the library derives the same operator from the span
// ((span+1)^3 = numProjectNodes), so its values are not read.
#include "sgld_mex.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    sgldmex::want_args(nrhs, 6, nlhs, 1, "MyToolbox:arrayProduct");
    const mxArray *Ke = prhs[0], *E = prhs[1], *up = prhs[2], *iKe = prhs[3], *lm = prhs[4];
    if (!sgldmex::is_real_double(Ke) || mxGetM(Ke) != 24 || mxGetN(Ke) != 24)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNot24x24", "Input Ke must be a 24x24 double matrix.");
    if (!sgldmex::is_real_int32(up) || (mxGetN(up) != 64 && mxGetN(up) != 8))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix",
                          "Input elementUpwardMap must be an M*64 or M*8 int32 matrix.");
    if (!mxIsSparse(iKe) || !mxIsDouble(iKe))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotSparse", "Input interpolatingKe must be a sparse double matrix.");
    if (!sgldmex::is_real_int32(lm) || mxGetN(lm) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input localMapping must be a column vector.");
    const int64_t nq = (int64_t)mxGetN(up), npn = sgldmex::scalar_count(prhs[5]);
    if ((nq == 8 && npn != 27) || (nq == 64 && npn != 125))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix",
                          "numProjectNodes must be 27 (8 children) or 125 (64 children).");
    const mwSize nC = mxGetM(up);
    const mwSize dims[3] = {24, 24, nC};
    plhs[0] = mxCreateNumericArray(3, dims, mxDOUBLE_CLASS, mxREAL);
    sgldmex::check(sgld_assemble_finest(mxGetPr(Ke), 24, 24, mxGetPr(E), (int64_t)mxGetNumberOfElements(E),
                                        (const int32_t *)mxGetData(up), (int64_t)nC, nq, mxGetPr(plhs[0])));
}
