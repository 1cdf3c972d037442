// xF = Solving_Interpolation_MatrixFree_mex(xC, transferMat, RI, eNodMat,
// numNodesFine, numNodesFineGhost, solidNodeMapCoarser2Finer,
// transferMatCoeffi, colors) on the GPU: the MEX of
// src/MEXfuncs/Solving_Interpolation_MatrixFree_mex.cpp:6-190 (called at
// src/Solving_InterpolationDeviation.m:7-9) over sgld_interpolate.
#include "sgld_mex.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    sgldmex::want_args(nrhs, 9, nlhs, 1, "MATLAB:mexFunction");
    const mxArray *xC = prhs[0], *tm = prhs[1], *RI = prhs[2], *eN = prhs[3], *map = prhs[6], *co = prhs[7];
    if (!sgldmex::is_real_double(xC) || mxGetN(xC) != 3)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input xCoarser must be a n-by-3 matrix.");
    if (!sgldmex::is_real_int32(tm) || (mxGetM(tm) != 125 && mxGetM(tm) != 27))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix",
                          "Input transferMat must be an 27xM or 125xM int32 matrix.");
    if (!sgldmex::is_real_double(RI) || mxGetN(RI) != 8 || (mxGetM(RI) != 125 && mxGetM(RI) != 27))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMx8", "Input RI must be a 125/27 x8 double matrix.");
    if (!sgldmex::is_real_int32(eN) || mxGetN(eN) != 8)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix", "Input eNodMat must be an Mx8 int32 matrix.");
    if (!sgldmex::is_real_int32(map) || mxGetN(map) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector",
                          "Input solidNodeMapCoarser2Finer must be a column vector.");
    if (!sgldmex::is_real_double(co) || mxGetN(co) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input transferMatCoeffi must be a column vector.");
    const sgldmex::Colours c =
        sgldmex::colours(prhs[8], "Solving_KbyU:colorArray", "Solving_Interpolation:colorCell");
    const int64_t nF = sgldmex::scalar_count(prhs[4]), nG = sgldmex::scalar_count(prhs[5]);
    plhs[0] = mxCreateDoubleMatrix((mwSize)nF, 3, mxREAL);
    sgldmex::check(sgld_interpolate(mxGetPr(xC), (int64_t)mxGetM(xC), (const int32_t *)mxGetData(tm),
                                    (int64_t)mxGetM(tm), mxGetPr(RI), (int64_t)mxGetM(RI), 8,
                                    (const int32_t *)mxGetData(eN), (int64_t)mxGetM(eN), nF, nG,
                                    (const int32_t *)mxGetData(map), mxGetPr(co), c.ids.data(), c.off.data(),
                                    c.count, mxGetPr(plhs[0])));
}
