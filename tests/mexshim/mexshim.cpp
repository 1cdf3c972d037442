// Implementation of tests/mexshim/mex.h plus a small C interface for the
// Python tests (build arrays, call a gateway's mexFunction, read results).
// Test infrastructure only.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
    mxClassID cls = mxDOUBLE_CLASS;
    std::vector<mwSize> dims{0, 0};
    std::vector<unsigned char> data;
    std::vector<mxArray *> cells;
    bool sparse = false, complex = false;
};

namespace {
size_t elsize(mxClassID c) {
    switch (c) {
        case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
        case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
        case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
        default: return 1;
    }
}
mwSize count(const mxArray *a) {
    mwSize n = 1;
    for (mwSize d : a->dims) n *= d;
    return n;
}
struct MexError : std::runtime_error {
    std::string id;
    MexError(const std::string &i, const std::string &m) : std::runtime_error(m), id(i) {}
};
std::string g_warn;
}  // namespace

extern "C" {
double *mxGetPr(const mxArray *a) { return (double *)a->data.data(); }
void *mxGetData(const mxArray *a) { return (void *)a->data.data(); }
mwSize mxGetM(const mxArray *a) { return a->dims[0]; }
mwSize mxGetN(const mxArray *a) {
    mwSize n = 1;
    for (size_t k = 1; k < a->dims.size(); ++k) n *= a->dims[k];
    return n;
}
mwSize mxGetNumberOfElements(const mxArray *a) { return a->cls == mxCELL_CLASS ? a->cells.size() : count(a); }
mwSize mxGetNumberOfDimensions(const mxArray *a) { return a->dims.size(); }
const mwSize *mxGetDimensions(const mxArray *a) { return a->dims.data(); }
int mxIsDouble(const mxArray *a) { return a->cls == mxDOUBLE_CLASS; }
int mxIsInt32(const mxArray *a) { return a->cls == mxINT32_CLASS; }
int mxIsUint8(const mxArray *a) { return a->cls == mxUINT8_CLASS; }
int mxIsUint64(const mxArray *a) { return a->cls == mxUINT64_CLASS; }
int mxIsLogical(const mxArray *a) { return a->cls == mxLOGICAL_CLASS; }
int mxIsCell(const mxArray *a) { return a->cls == mxCELL_CLASS; }
int mxIsChar(const mxArray *a) { return a->cls == mxCHAR_CLASS; }
int mxIsSparse(const mxArray *a) { return a->sparse; }
int mxIsComplex(const mxArray *a) { return a->complex; }
mxArray *mxGetCell(const mxArray *a, mwIndex i) { return i < a->cells.size() ? a->cells[i] : nullptr; }
double mxGetScalar(const mxArray *a) {
    if (a->data.empty()) return 0.0;
    switch (a->cls) {
        case mxDOUBLE_CLASS: return *(const double *)a->data.data();
        case mxINT32_CLASS: return *(const int32_t *)a->data.data();
        case mxUINT64_CLASS: return (double)*(const uint64_t *)a->data.data();
        case mxUINT8_CLASS: case mxLOGICAL_CLASS: return a->data[0];
        default: return 0.0;
    }
}
char *mxArrayToString(const mxArray *a) {
    const mwSize n = count(a);
    char *s = (char *)std::malloc(n + 1);
    const uint16_t *c = (const uint16_t *)a->data.data();
    for (mwSize i = 0; i < n; ++i) s[i] = (char)c[i];
    s[n] = 0;
    return s;
}
void mxFree(void *p) { std::free(p); }
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID cls, mxComplexity) {
    mxArray *a = new mxArray_tag();
    a->cls = cls;
    a->dims.assign(dims, dims + ndim);
    a->data.assign(count(a) * elsize(cls), 0);
    return a;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c) {
    const mwSize d[2] = {m, n};
    return mxCreateNumericArray(2, d, mxDOUBLE_CLASS, c);
}
mxArray *mxCreateDoubleScalar(double v) {
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    *mxGetPr(a) = v;
    return a;
}
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw MexError(id, buf);
}
void mexWarnMsgIdAndTxt(const char *id, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_warn = std::string(id) + ": " + buf;
}

// ---- driver interface (ctypes) ----
mxArray *shim_array(int cls, int ndim, const int64_t *dims, const void *src) {
    std::vector<mwSize> d(dims, dims + ndim);
    mxArray *a = mxCreateNumericArray((mwSize)ndim, d.data(), (mxClassID)cls, mxREAL);
    if (src && !a->data.empty()) std::memcpy(a->data.data(), src, a->data.size());
    return a;
}
mxArray *shim_sparse_flagged(int64_t m, int64_t n) {
    const mwSize d[2] = {(mwSize)m, (mwSize)n};
    mxArray *a = mxCreateNumericArray(2, d, mxDOUBLE_CLASS, mxREAL);
    a->sparse = true;
    return a;
}
mxArray *shim_cell(int n, mxArray **elems) {
    mxArray *a = new mxArray_tag();
    a->cls = mxCELL_CLASS;
    a->dims = {1, (mwSize)n};
    a->cells.assign(elems, elems + n);
    return a;
}
mxArray *shim_string(const char *s) {
    const size_t n = std::strlen(s);
    const mwSize d[2] = {1, n};
    mxArray *a = mxCreateNumericArray(2, d, mxCHAR_CLASS, mxREAL);
    uint16_t *c = (uint16_t *)a->data.data();
    for (size_t i = 0; i < n; ++i) c[i] = (unsigned char)s[i];
    return a;
}
void shim_free(mxArray *a) {
    if (!a) return;
    for (mxArray *c : a->cells) shim_free(c);
    delete a;
}
int shim_class(const mxArray *a) { return (int)a->cls; }
int shim_dims(const mxArray *a, int64_t *out) {
    for (size_t k = 0; k < a->dims.size(); ++k) out[k] = (int64_t)a->dims[k];
    return (int)a->dims.size();
}
const void *shim_data(const mxArray *a) { return a->data.data(); }
typedef void (*mexfn)(int, mxArray **, int, const mxArray **);
// 0: returned; 1: mexErrMsgIdAndTxt (identifier and message copied out)
int shim_call(void *fn, int nlhs, mxArray **plhs, int nrhs, mxArray **prhs, char *id, char *msg, int cap) {
    g_warn.clear();
    try {
        ((mexfn)fn)(nlhs, plhs, nrhs, (const mxArray **)prhs);
    } catch (const MexError &e) {
        std::snprintf(id, cap, "%s", e.id.c_str());
        std::snprintf(msg, cap, "%s", e.what());
        return 1;
    }
    return 0;
}
const char *shim_last_warning(void) { return g_warn.c_str(); }
}
