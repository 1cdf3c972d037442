// sgld_ctx_mex: the device-resident solver context (include/sgldfea.h,
// sgld_ctx_*) as one MEX function with a command string, for the .m shims of
// Solving_CG_GMGS.m, Solving_KbyU_MatrixFree.m, Solving_Vcycle.m,
// Solving_AssembleFEAstencil.m and TopOpti_ComputeUnitCompliance.m.
//
//   h = sgld_ctx_mex('create', voxels, numLevels, nonDyadic, poissonRatio, cellSize)
//       voxels: resY x resX x resZ uint8/logical/double (nonzero = solid);
//       numLevels / nonDyadic as numLevels_ / nonDyadic_ before
//       Solving_BuildingMeshHierarchy.m:12-16 adjusts them; h: uint64 handle
//   sgld_ctx_mex('set_fixed', h, fixedDOFs)     logical mask, or 1-based DOF ids
//   sgld_ctx_mex('set_modulus', h, eleModulus)  numElements values
//   sgld_ctx_mex('assemble', h, weightFactorJacobi)
//   [U, its, relres] = sgld_ctx_mex('pcg', h, F, U0, tol, maxIT[, 'fp64' | 'mixed'])
//   y  = sgld_ctx_mex('kbyu', h, u)             Solving_KbyU_MatrixFree(u)
//   z  = sgld_ctx_mex('vcycle', h, r)           Solving_Vcycle(r)
//   ce = sgld_ctx_mex('unit_compliance', h, U)  TopOpti_ComputeUnitCompliance()
//   info = sgld_ctx_mex('info', h)              levels x 8: resX resY resZ span
//                                               numElements numNodes numDOFs latticeNodes
//   sgld_ctx_mex('destroy', h)
// A pcg that reaches maxIT returns its result with a warning, as
// Solving_PreconditionedConjugateGradientSolver.m:32-35 does.
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "sgld_mex.hpp"

namespace {

std::string command(const mxArray *a) {
    if (!mxIsChar(a)) mexErrMsgIdAndTxt("sgld:ctx:command", "The first argument must be a command string.");
    char *s = mxArrayToString(a);
    std::string out = s ? s : "";
    mxFree(s);
    return out;
}

sgld_ctx *handle(const mxArray *a) {
    if (!mxIsUint64(a) || mxGetNumberOfElements(a) != 1)
        mexErrMsgIdAndTxt("sgld:ctx:handle", "The context handle must be the uint64 scalar 'create' returned.");
    uint64_t v = 0;
    std::memcpy(&v, mxGetData(a), sizeof(v));
    if (!v) mexErrMsgIdAndTxt("sgld:ctx:handle", "Null context handle.");
    return reinterpret_cast<sgld_ctx *>(v);
}

int64_t info_at(sgld_ctx *C, int level, int k) {
    int64_t out[8];
    sgldmex::check(sgld_ctx_info(C, level, out));
    return out[k];
}

std::vector<double> doubles(const mxArray *a, int64_t n, const char *what) {
    if (!(mxIsDouble(a) && !mxIsComplex(a) && !mxIsSparse(a)) || (int64_t)mxGetNumberOfElements(a) != n)
        mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "%s must be a real double vector of %lld entries.",
                          what, (long long)n);
    const double *p = mxGetPr(a);
    return std::vector<double>(p, p + n);
}

void want(int nrhs, int n, const std::string &cmd) {
    if (nrhs != n) mexErrMsgIdAndTxt("MATLAB:mexFunction:nrhs", "'%s' takes %d inputs.", cmd.c_str(), n);
}

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
    if (nrhs < 1) mexErrMsgIdAndTxt("MATLAB:mexFunction:nrhs", "A command string is required.");
    const std::string cmd = command(prhs[0]);
    if (cmd == "create") {
        want(nrhs, 6, cmd);
        const mxArray *v = prhs[1];
        const mwSize nd = mxGetNumberOfDimensions(v);
        const mwSize *d = mxGetDimensions(v);
        if (nd > 3 || !(mxIsUint8(v) || mxIsLogical(v) || mxIsDouble(v)))
            mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotMatrix",
                              "voxels must be a resY x resX x resZ uint8, logical or double array.");
        const int ry = (int)d[0], rx = (int)d[1], rz = nd == 3 ? (int)d[2] : 1;
        const size_t nv = (size_t)rx * ry * rz;
        std::vector<uint8_t> vox(nv);
        if (mxIsDouble(v)) {
            const double *p = mxGetPr(v);
            for (size_t i = 0; i < nv; ++i) vox[i] = p[i] != 0.0;
        } else {
            const uint8_t *p = (const uint8_t *)mxGetData(v);
            for (size_t i = 0; i < nv; ++i) vox[i] = p[i] != 0;
        }
        sgld_ctx *C = nullptr;
        sgldmex::check(sgld_ctx_create(&C, rx, ry, rz, vox.data(), (int)mxGetScalar(prhs[2]),
                                       (int)mxGetScalar(prhs[3]), mxGetScalar(prhs[4]), mxGetScalar(prhs[5])));
        const mwSize one[2] = {1, 1};
        plhs[0] = mxCreateNumericArray(2, one, mxUINT64_CLASS, mxREAL);
        const uint64_t h = reinterpret_cast<uint64_t>(C);
        std::memcpy(mxGetData(plhs[0]), &h, sizeof(h));
        return;
    }
    if (nrhs < 2) mexErrMsgIdAndTxt("MATLAB:mexFunction:nrhs", "'%s' needs the context handle.", cmd.c_str());
    sgld_ctx *C = handle(prhs[1]);
    if (cmd == "destroy") {
        want(nrhs, 2, cmd);
        sgldmex::check(sgld_ctx_destroy(C));
        return;
    }
    const int64_t nDof = info_at(C, 1, 6), nEl = info_at(C, 1, 4);
    if (cmd == "set_fixed") {
        want(nrhs, 3, cmd);
        const mxArray *f = prhs[2];
        std::vector<uint8_t> mask(nDof, 0);
        const int64_t n = (int64_t)mxGetNumberOfElements(f);
        if (mxIsLogical(f)) {
            if (n != nDof)
                mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "A logical fixedDOFs mask needs %lld entries.",
                                  (long long)nDof);
            const uint8_t *p = (const uint8_t *)mxGetData(f);
            for (int64_t i = 0; i < n; ++i) mask[i] = p[i] != 0;
        } else if (mxIsDouble(f) || mxIsInt32(f)) {
            for (int64_t i = 0; i < n; ++i) {
                const int64_t id = mxIsDouble(f) ? (int64_t)mxGetPr(f)[i] : (int64_t)((const int32_t *)mxGetData(f))[i];
                if (id < 1 || id > nDof)
                    mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "fixedDOFs(%lld) = %lld outside 1..%lld.",
                                      (long long)(i + 1), (long long)id, (long long)nDof);
                mask[id - 1] = 1;
            }
        } else {
            mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "fixedDOFs must be logical, double or int32.");
        }
        sgldmex::check(sgld_ctx_set_fixed(C, mask.data()));
        return;
    }
    if (cmd == "set_modulus") {
        want(nrhs, 3, cmd);
        const std::vector<double> E = doubles(prhs[2], nEl, "eleModulus");
        sgldmex::check(sgld_ctx_set_modulus(C, E.data()));
        return;
    }
    if (cmd == "assemble") {
        want(nrhs, 3, cmd);
        sgldmex::check(sgld_ctx_assemble(C, mxGetScalar(prhs[2])));
        return;
    }
    if (cmd == "pcg") {
        if (nrhs != 6 && nrhs != 7) mexErrMsgIdAndTxt("MATLAB:mexFunction:nrhs", "'pcg' takes 6 or 7 inputs.");
        const std::vector<double> F = doubles(prhs[2], nDof, "F");
        std::vector<double> U = mxGetNumberOfElements(prhs[3]) == 0 ? std::vector<double>(nDof, 0.0)
                                                                     : doubles(prhs[3], nDof, "U0");
        int prec = SGLD_PREC_FP64;
        if (nrhs == 7) {
            const std::string p = command(prhs[6]);
            if (p == "mixed") prec = SGLD_PREC_MIXED;
            else if (p != "fp64")
                mexErrMsgIdAndTxt("MATLAB:mexFunction:inputNotVector", "precision must be 'fp64' or 'mixed'.");
        }
        int its = 0;
        double rr = 0.0;
        const int rc = sgld_ctx_pcg(C, F.data(), U.data(), mxGetScalar(prhs[4]), (int)mxGetScalar(prhs[5]), prec,
                                    &its, &rr);
        if (rc == SGLD_ERR_MAXIT)
            mexWarnMsgIdAndTxt("sgld:maxit", "PCG reached maxIT = %d, relative residual %.3e.",
                               (int)mxGetScalar(prhs[5]), rr);
        else
            sgldmex::check(rc);
        plhs[0] = mxCreateDoubleMatrix((mwSize)nDof, 1, mxREAL);
        std::memcpy(mxGetPr(plhs[0]), U.data(), sizeof(double) * nDof);
        if (nlhs > 1) plhs[1] = mxCreateDoubleScalar((double)its);
        if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(rr);
        return;
    }
    if (cmd == "kbyu" || cmd == "vcycle") {
        want(nrhs, 3, cmd);
        const std::vector<double> u = doubles(prhs[2], nDof, cmd == "kbyu" ? "u" : "r");
        plhs[0] = mxCreateDoubleMatrix((mwSize)nDof, 1, mxREAL);
        sgldmex::check(cmd == "kbyu" ? sgld_ctx_kbyu(C, u.data(), mxGetPr(plhs[0]))
                                     : sgld_ctx_vcycle(C, u.data(), mxGetPr(plhs[0])));
        return;
    }
    if (cmd == "unit_compliance") {
        want(nrhs, 3, cmd);
        const std::vector<double> U = doubles(prhs[2], nDof, "U");
        plhs[0] = mxCreateDoubleMatrix((mwSize)nEl, 1, mxREAL);
        sgldmex::check(sgld_ctx_unit_compliance(C, U.data(), mxGetPr(plhs[0])));
        return;
    }
    if (cmd == "info") {
        want(nrhs, 2, cmd);
        const int L = sgld_ctx_info(C, 0, nullptr);
        plhs[0] = mxCreateDoubleMatrix((mwSize)L, 8, mxREAL);
        double *o = mxGetPr(plhs[0]);
        for (int l = 1; l <= L; ++l)
            for (int k = 0; k < 8; ++k) o[(l - 1) + (int64_t)L * k] = (double)info_at(C, l, k);
        return;
    }
    mexErrMsgIdAndTxt("sgld:ctx:command", "Unknown command '%s'.", cmd.c_str());
}
