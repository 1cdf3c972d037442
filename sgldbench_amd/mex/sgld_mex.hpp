// Shared marshalling of the MEX gateways over libsgldfea (include/sgldfea.h).
//
// Each gateway keeps the MATLAB call signature, argument checks and error
// identifiers of the MEX function it replaces, turns MATLAB arrays into the
// C ABI's pointers and sizes (colour cells into concatenated ids plus
// offsets), and forwards a failing status as mexErrMsgIdAndTxt with the
// identifier that sgld_last_error() begins with.  Compile one gateway per
// MEX name: mex -largeArrayDims <gateway>.cpp -I<repo>/include -L<lib dir> -lsgldfea
#pragma once
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mex.h"
#include "sgldfea.h"

namespace sgldmex {

// The C ABI's failure, as a MATLAB error whose identifier is the message's
// prefix up to ": " (e.g. "MATLAB:mexFunction:inputNot24x24").
inline void raise_last(int status) {
    const char *msg = sgld_last_error();
    std::string m = msg ? msg : "";
    const size_t cut = m.find(": ");
    const std::string id = cut != std::string::npos ? m.substr(0, cut) : "sgld:status" + std::to_string(status);
    mexErrMsgIdAndTxt(id.c_str(), "%s", cut != std::string::npos ? m.c_str() + cut + 2 : m.c_str());
}

inline void check(int status) {
    if (status != SGLD_OK) raise_last(status);
}

inline void want_args(int nrhs, int nWant, int nlhs, int nlhsWant, const char *idPrefix) {
    if (nrhs != nWant) {
        const std::string id = std::string(idPrefix) + ":nrhs";
        mexErrMsgIdAndTxt(id.c_str(), "%d inputs required.", nWant);
    }
    if (nlhs != nlhsWant) {
        const std::string id = std::string(idPrefix) + ":nlhs";
        mexErrMsgIdAndTxt(id.c_str(), "%d output required.", nlhsWant);
    }
}

inline bool is_real_double(const mxArray *a) { return mxIsDouble(a) && !mxIsComplex(a) && !mxIsSparse(a); }
inline bool is_real_int32(const mxArray *a) { return mxIsInt32(a) && !mxIsComplex(a); }

// A cell array of int32 element-id vectors (the colours of Solving_Coloring.m)
// as concatenated ids and prefix offsets.  Empty cells are allowed.
struct Colours {
    std::vector<int32_t> ids;
    std::vector<int64_t> off;
    int count = 0;
};

inline Colours colours(const mxArray *cell, const char *arrayId, const char *cellId) {
    if (!mxIsCell(cell)) mexErrMsgIdAndTxt(arrayId, "colorArray must be a cell array.");
    Colours c;
    const mwSize n = mxGetNumberOfElements(cell);
    c.count = (int)n;
    c.off.assign(n + 1, 0);
    for (mwSize i = 0; i < n; ++i) {
        const mxArray *e = mxGetCell(cell, i);
        mwSize k = 0;
        if (e) {
            if (!is_real_int32(e)) mexErrMsgIdAndTxt(cellId, "Each color cell must be a real int32 vector.");
            k = mxGetNumberOfElements(e);
            const int32_t *p = (const int32_t *)mxGetData(e);
            c.ids.insert(c.ids.end(), p, p + k);
        }
        c.off[i + 1] = c.off[i] + (int64_t)k;
    }
    if (c.ids.empty()) c.ids.push_back(0);  // a valid pointer for the zero-length case
    return c;
}

inline int64_t scalar_count(const mxArray *a) { return (int64_t)mxGetScalar(a); }

}  // namespace sgldmex
