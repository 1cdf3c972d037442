// ce = TopOpti_CmptUnitCompliance_mex(U, eNodMat, Ke) on the GPU: the MEX of
// src/MEXfuncs/TopOpti_CmptUnitCompliance_mex.cpp:6-59 (called at
// src/TopOpti_ComputeUnitCompliance.m:7) over sgld_unit_compliance.
#include "sgld_mex.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs != 3) mexErrMsgIdAndTxt("MATLAB:mexFunction:nrhs", "3 inputs required.");
    if (nlhs != 1) mexErrMsgIdAndTxt("MATLAB:mexFunction:nlhs", "One output required.");
    const mxArray *U = prhs[0], *eN = prhs[1], *Ke = prhs[2];
    if (!sgldmex::is_real_double(U) || mxGetN(U) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input U must be a column vector.");
    if (!sgldmex::is_real_int32(eN) || mxGetN(eN) != 8)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix", "Input eNodMat must be an Mx8 int32 matrix.");
    if (!sgldmex::is_real_double(Ke) || mxGetM(Ke) != 24 || mxGetN(Ke) != 24)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNot24x24", "Input Ke must be a 24x24 double matrix.");
    const mwSize ne = mxGetM(eN);
    plhs[0] = mxCreateDoubleMatrix(ne, 1, mxREAL);
    sgldmex::check(sgld_unit_compliance(mxGetPr(U), (int64_t)mxGetM(U), (const int32_t *)mxGetData(eN),
                                        (int64_t)ne, mxGetPr(Ke), 24, 24, mxGetPr(plhs[0])));
}
