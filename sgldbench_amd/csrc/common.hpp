// Shared host/device helpers of libsgldfea: status handling, device buffers,
// element geometry constants.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/sgldfea.h"

namespace sgld {

void set_error(const char *fmt, ...);

#define SGLD_TRY(expr)                       \
    do {                                     \
        int rc_ = (expr);                    \
        if (rc_ != SGLD_OK) return rc_;      \
    } while (0)

#define SGLD_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            sgld::set_error("sgld:hip: %s (%s) at %s:%d", hipGetErrorName(e_), #call,        \
                            __FILE__, __LINE__);                                            \
            return SGLD_ERR_HIP;                                                            \
        }                                                                                   \
    } while (0)

#define SGLD_ARG(cond, ...)                  \
    do {                                     \
        if (!(cond)) {                       \
            sgld::set_error(__VA_ARGS__);    \
            return SGLD_ERR_ARG;             \
        }                                    \
    } while (0)

// Owning device allocation (freed on scope exit).
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        release();
        p = o.p; n = o.n; o.p = nullptr; o.n = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    int alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            set_error("sgld:nomem: device allocation of %zu bytes failed", count * sizeof(T));
            return SGLD_ERR_NOMEM;
        }
        n = count;
        return SGLD_OK;
    }
    int upload(const T *src, size_t count) {
        SGLD_TRY(alloc(count));
        if (count) SGLD_HIP(hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice));
        return SGLD_OK;
    }
    int zero(hipStream_t s = 0) {
        SGLD_HIP(hipMemsetAsync(p, 0, n * sizeof(T), s));
        return SGLD_OK;
    }
};

// Local-node corner bits of the hexahedron, node order of
// src/FEA_GaussianIntegral.m:2-9 (1:(0,0,0) 2:(1,0,0) 3:(1,1,0) 4:(0,1,0), +z).
constexpr int kCornerX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
constexpr int kCornerY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
constexpr int kCornerZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

// Host helpers shared by the stateless and stateful paths.
void element_stiffness(double nu, double h, double *Ke);
void stress_operator(double nu, double h, double *T);  // 48 x 24, row-major
void transfer_table(int span, double *RI);          // (span+1)^3 x 8 col-major
void child_prolongation(int span, int q, double p[8][8]);
void galerkin_child_add(const double p[8][8], const double *K, double *out);

inline int blocks_for(int64_t n, int t) { return (int)((n + t - 1) / t); }
// Grid for a grid-stride kernel: the dispatch packet holds the grid size in
// work-items as 32 bits, so one-item-per-thread launches must stay below 2^32.
inline int bounded_blocks(int64_t n, int t) {
    const int64_t b = (n + t - 1) / t, cap = (int64_t)1 << 20;
    return (int)(b < cap ? b : cap);
}

}  // namespace sgld
