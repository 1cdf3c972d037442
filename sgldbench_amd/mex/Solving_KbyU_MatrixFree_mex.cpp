// Y = Solving_KbyU_MatrixFree_mex(U, eNodMat, Ke, E, colors) on the GPU: the
// MEX of src/MEXfuncs/Solving_KbyU_MatrixFree_mex.cpp:6-103 (called at
// src/Solving_KbyU_MatrixFree.m:9) over sgld_kbyu, with its checks and
// error identifiers.
#include "sgld_mex.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    sgldmex::want_args(nrhs, 5, nlhs, 1, "MATLAB:mexFunction");
    const mxArray *U = prhs[0], *eN = prhs[1], *Ke = prhs[2], *E = prhs[3];
    if (!sgldmex::is_real_double(U) || mxGetN(U) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input U must be a column vector.");
    if (!sgldmex::is_real_int32(eN) || mxGetN(eN) != 8)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix", "Input eNodMat must be an Mx8 int32 matrix.");
    if (!sgldmex::is_real_double(Ke) || mxGetM(Ke) != 24 || mxGetN(Ke) != 24)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNot24x24", "Input Ke must be a 24x24 double matrix.");
    if (!sgldmex::is_real_double(E) || mxGetN(E) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input E must be a column vector.");
    const sgldmex::Colours c = sgldmex::colours(prhs[4], "Solving_KbyU:colorArray", "Solving_KbyU:colorCell");
    const mwSize n = mxGetM(U);
    plhs[0] = mxCreateDoubleMatrix(n, 1, mxREAL);
    sgldmex::check(sgld_kbyu(mxGetPr(U), (int64_t)n, (const int32_t *)mxGetData(eN), (int64_t)mxGetM(eN),
                             mxGetPr(Ke), 24, 24, mxGetPr(E), (int64_t)mxGetM(E), c.ids.data(), c.off.data(),
                             c.count, mxGetPr(plhs[0])));
}
