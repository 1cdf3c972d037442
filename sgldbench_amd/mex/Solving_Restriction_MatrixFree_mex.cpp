// rC = Solving_Restriction_MatrixFree_mex(rF, transferMat, RIT, eNodMat,
// numNodesCoarse, numNodesFineGhost, solidNodeMapCoarser2Finer,
// transferMatCoeffi, colors) on the GPU: the MEX of
// src/MEXfuncs/Solving_Restriction_MatrixFree_mex.cpp:6-191 (called at
// src/Solving_RestrictResidual.m:6-8) over sgld_restrict.
#include "sgld_mex.hpp"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    sgldmex::want_args(nrhs, 9, nlhs, 1, "MATLAB:mexFunction");
    const mxArray *rF = prhs[0], *tm = prhs[1], *RIT = prhs[2], *eN = prhs[3], *map = prhs[6], *co = prhs[7];
    if (!sgldmex::is_real_double(rF) || mxGetN(rF) != 3)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input rFiner must be a n-by-3 matrix.");
    if (!sgldmex::is_real_int32(tm) || (mxGetM(tm) != 125 && mxGetM(tm) != 27))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix",
                          "Input transferMat must be an 27xM or 125xM int32 matrix.");
    if (!sgldmex::is_real_double(RIT) || mxGetM(RIT) != 8 || (mxGetN(RIT) != 125 && mxGetN(RIT) != 27))
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMx8", "Input RIT must be a 8x 125/27 double matrix.");
    if (!sgldmex::is_real_int32(eN) || mxGetN(eN) != 8)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix", "Input eNodMat must be an Mx8 int32 matrix.");
    if (!sgldmex::is_real_int32(map) || mxGetN(map) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector",
                          "Input solidNodeMapCoarser2Finer must be a column vector.");
    if (!sgldmex::is_real_double(co) || mxGetN(co) != 1)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "Input transferMatCoeffi must be a column vector.");
    const sgldmex::Colours c = sgldmex::colours(prhs[8], "Solving_KbyU:colorArray", "Solving_KbyU:colorCell");
    const int64_t nC = sgldmex::scalar_count(prhs[4]), nG = sgldmex::scalar_count(prhs[5]);
    plhs[0] = mxCreateDoubleMatrix((mwSize)nC, 3, mxREAL);
    sgldmex::check(sgld_restrict(mxGetPr(rF), (int64_t)mxGetM(rF), (const int32_t *)mxGetData(tm),
                                 (int64_t)mxGetM(tm), mxGetPr(RIT), 8, (int64_t)mxGetN(RIT),
                                 (const int32_t *)mxGetData(eN), (int64_t)mxGetM(eN), nC, nG,
                                 (const int32_t *)mxGetData(map), mxGetPr(co), c.ids.data(), c.off.data(), c.count,
                                 mxGetPr(plhs[0])));
}
