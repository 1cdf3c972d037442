// Density filter of the topology optimisation (the step after the compliance
// sensitivities in every TopOpti iteration, TopOpti_GlobalVolumeConstraint.m:
// 43, 128-129, 173): cone weights w(e1, e2) = max(0, rMin - |c_e1 - c_e2|)
// over the solid voxels in the box of half-width ceil(rMin) - 1,
//   Hs[e1] = sum_e2 w            (TopOpti_SetupDensityFilter_mex.cpp:10-51)
//   Y[e1]  = sum_e2 w x[e2]      (TopOpti_PerformDensityFiltering_mex.cpp:10-52)
// The values are gathered on the voxel lattice (void voxels hold 0, so they
// add exact zeros) in the reference's neighbour order - layer, then column,
// then row - with FP64 products and sums and no contraction, so the results
// equal the reference algorithm bit for bit.  The weights depend only on the
// offset and are tabulated once per call on the host with the same
// expression rMin - sqrt(d^2).
#include <cmath>
#include <vector>

#include "lattice_ctx.hpp"

#pragma clang fp contract(off)

namespace sgld {
namespace {

// xl[v] = x[efwd[v] - 1] on solid voxels, 0 on void ones; x == null: 1 on solid.
__global__ void k_filter_scatter(const double *__restrict__ x, const int32_t *__restrict__ efwd, int64_t nv,
                                 double *__restrict__ xl) {
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e = efwd[v];
        xl[v] = e ? (x ? x[e - 1] : 1.0) : 0.0;
    }
}

// One thread per voxel, rows fastest so a wavefront reads contiguous runs.
__global__ void __launch_bounds__(256) k_density_filter(const double *__restrict__ xl,
                                                        const int32_t *__restrict__ efwd, int nx, int ny,
                                                        int nz, int R, const double *__restrict__ w,
                                                        double *__restrict__ Y) {
    const int64_t nxy = (int64_t)nx * ny, nv = nxy * nz;
    const int W = 2 * R + 1;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (int64_t)gridDim.x * blockDim.x) {
        const int32_t e1 = efwd[v];
        if (!e1) continue;
        const int jj = (int)(v % ny), ii = (int)((v / ny) % nx), kk = (int)(v / nxy);
        const int k0 = max(kk - R, 0), k1 = min(kk + R, nz - 1);
        const int i0 = max(ii - R, 0), i1 = min(ii + R, nx - 1);
        const int j0 = max(jj - R, 0), j1 = min(jj + R, ny - 1);
        double acc = 0.0;
        for (int k2 = k0; k2 <= k1; ++k2)
            for (int i2 = i0; i2 <= i1; ++i2) {
                const double *xrow = xl + k2 * nxy + (int64_t)i2 * ny;
                const double *wrow = w + ((k2 - kk + R) * W + (i2 - ii + R)) * W + (R - jj);
                for (int j2 = j0; j2 <= j1; ++j2) acc += wrow[j2] * xrow[j2];
            }
        Y[e1 - 1] = acc;
    }
}

// LDS-tiled form for R <= kFRmax: a block filters a kFBy x kFBx x kFBz
// (row, column, layer) brick of voxels, one thread each, from its
// neighbourhood staged in LDS (0 outside the lattice).  Every voxel walks the
// full (2R+1)^3 window in the same order as k_density_filter; a neighbour
// outside the lattice adds w * 0 = +0, which leaves the sum unchanged (the sum
// starts at +0 and never becomes -0), so the results are the same bits.  The
// weight offsets are uniform across the block (scalar loads); the per-thread
// gathers come from LDS instead of the cache hierarchy.
constexpr int kFBy = 8, kFBx = 8, kFBz = 4, kFRmax = 4;
template <int R>
__global__ void __launch_bounds__(256) k_density_filter_tiled(const double *__restrict__ xl,
                                                              const int32_t *__restrict__ efwd, int nx, int ny,
                                                              int nz, const double *__restrict__ w,
                                                              double *__restrict__ Y, int bricksY, int bricksX) {
    constexpr int W = 2 * R + 1, SY = kFBy + 2 * R, SX = kFBx + 2 * R, SZ = kFBz + 2 * R;
    __shared__ double sh[SY * SX * SZ];
    const int64_t nxy = (int64_t)nx * ny;
    const int by = blockIdx.x % bricksY, bx = (blockIdx.x / bricksY) % bricksX, bz = blockIdx.x / (bricksY * bricksX);
    const int y0 = by * kFBy - R, x0 = bx * kFBx - R, z0 = bz * kFBz - R;
    for (int t = threadIdx.x; t < SY * SX * SZ; t += blockDim.x) {
        const int ly = t % SY, lx = (t / SY) % SX, lz = t / (SY * SX);
        const int gy = y0 + ly, gx = x0 + lx, gz = z0 + lz;
        const bool in = gy >= 0 && gy < ny && gx >= 0 && gx < nx && gz >= 0 && gz < nz;
        sh[t] = in ? xl[gz * nxy + (int64_t)gx * ny + gy] : 0.0;
    }
    __syncthreads();
    const int ty = threadIdx.x % kFBy, tx = (threadIdx.x / kFBy) % kFBx, tz = threadIdx.x / (kFBy * kFBx);
    const int jj = by * kFBy + ty, ii = bx * kFBx + tx, kk = bz * kFBz + tz;
    if (jj >= ny || ii >= nx || kk >= nz) return;
    const int32_t e1 = efwd[kk * nxy + (int64_t)ii * ny + jj];
    if (!e1) return;
    double acc = 0.0;
    for (int dz = 0; dz < W; ++dz)
#pragma unroll
        for (int dx = 0; dx < W; ++dx) {
            const double *srow = sh + ((tz + dz) * SX + (tx + dx)) * SY + ty;
            const double *wrow = w + (dz * W + dx) * W;
#pragma unroll
            for (int dy = 0; dy < W; ++dy) acc += wrow[dy] * srow[dy];
        }
    Y[e1 - 1] = acc;
}

std::vector<double> cone_weights(double rMin, int R) {
    const int W = 2 * R + 1;
    std::vector<double> w((size_t)W * W * W);
    for (int dz = -R; dz <= R; ++dz)
        for (int dx = -R; dx <= R; ++dx)
            for (int dy = -R; dy <= R; ++dy) {
                const int d2 = dx * dx + dy * dy + dz * dz;
                const double d = rMin - std::sqrt((double)d2);
                w[((size_t)(dz + R) * W + (dx + R)) * W + (dy + R)] = d > 0.0 ? d : 0.0;
            }
    return w;
}

// Filter with the voxel map on the device; x_dev null computes Hs.  With a
// context (`timed`) the lattice copy of x and the weights live in the
// context's work buffer, allocated on first use and kept.
int run_filter(hipStream_t s, const double *x_dev, double rMin, const int32_t *efwd_dev, int resX, int resY,
               int resZ, double *Y_dev, sgld_ctx *timed = nullptr) {
    const int R = (int)std::ceil(rMin) - 1;
    SGLD_ARG(R >= 0 && R <= 32, "sgld:arg: rMin must be in (0, 33]");
    const int64_t nv = (int64_t)resX * resY * resZ;
    std::vector<double> w = cone_weights(rMin, R);
    struct View {
        double *p;
    };
    DevBuf<double> own;
    DevBuf<double> &buf = timed ? timed->filterWork : own;
    if (buf.n < (size_t)nv + w.size()) SGLD_TRY(buf.alloc((size_t)nv + w.size()));
    View xl{buf.p}, dw{buf.p + nv};
    SGLD_HIP(hipMemcpyAsync(dw.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_filter_scatter, dim3(bounded_blocks(nv, 256)), dim3(256), 0, s, x_dev, efwd_dev, nv, xl.p);
    if (timed) SGLD_HIP(hipEventRecord(timed->e0, s));
    if (R <= kFRmax) {
        const int bY = (resY + kFBy - 1) / kFBy, bX = (resX + kFBx - 1) / kFBx, bZ = (resZ + kFBz - 1) / kFBz;
        const dim3 grid((unsigned)((int64_t)bY * bX * bZ)), block(kFBy * kFBx * kFBz);
        switch (R) {
            case 0: hipLaunchKernelGGL(k_density_filter_tiled<0>, grid, block, 0, s, xl.p, efwd_dev, resX, resY, resZ, dw.p, Y_dev, bY, bX); break;
            case 1: hipLaunchKernelGGL(k_density_filter_tiled<1>, grid, block, 0, s, xl.p, efwd_dev, resX, resY, resZ, dw.p, Y_dev, bY, bX); break;
            case 2: hipLaunchKernelGGL(k_density_filter_tiled<2>, grid, block, 0, s, xl.p, efwd_dev, resX, resY, resZ, dw.p, Y_dev, bY, bX); break;
            case 3: hipLaunchKernelGGL(k_density_filter_tiled<3>, grid, block, 0, s, xl.p, efwd_dev, resX, resY, resZ, dw.p, Y_dev, bY, bX); break;
            default: hipLaunchKernelGGL(k_density_filter_tiled<4>, grid, block, 0, s, xl.p, efwd_dev, resX, resY, resZ, dw.p, Y_dev, bY, bX); break;
        }
    } else {
        hipLaunchKernelGGL(k_density_filter, dim3(bounded_blocks(nv, 256)), dim3(256), 0, s, xl.p, efwd_dev, resX,
                           resY, resZ, R, dw.p, Y_dev);
    }
    if (timed) SGLD_HIP(hipEventRecord(timed->e1, s));
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipStreamSynchronize(s));
    if (timed) {
        float ms = 0;
        SGLD_HIP(hipEventElapsedTime(&ms, timed->e0, timed->e1));
        timed->stats[5] = ms;
    }
    return SGLD_OK;
}

// Host-side checks of the MEX arguments: every id in 0..numElements, each
// element id used once.
int check_map(int64_t numElements, const int32_t *efwd, int resX, int resY, int resZ) {
    SGLD_ARG(efwd, "sgld:arg: eleMapForward is null");
    SGLD_ARG(resX > 0 && resY > 0 && resZ > 0, "sgld:arg: resolution must be positive");
    SGLD_ARG(numElements >= 0 && numElements <= (int64_t)resX * resY * resZ,
             "sgld:arg: numElements exceeds the voxel count");
    const int64_t nv = (int64_t)resX * resY * resZ;
    std::vector<uint8_t> seen((size_t)numElements + 1, 0);
    for (int64_t v = 0; v < nv; ++v) {
        const int32_t e = efwd[v];
        SGLD_ARG(e >= 0 && e <= numElements, "sgld:arg: eleMapForward(%lld) = %d outside 0..numElements",
                 (long long)(v + 1), (int)e);
        if (e) {
            SGLD_ARG(!seen[e], "sgld:arg: element %d appears twice in eleMapForward", (int)e);
            seen[e] = 1;
        }
    }
    return SGLD_OK;
}

}  // namespace
}  // namespace sgld

using namespace sgld;

extern "C" int sgld_density_filter_setup(double rMin, int64_t numElements, const int32_t *eleMapForward,
                                         int resX, int resY, int resZ, double *Hs) {
    SGLD_TRY(check_map(numElements, eleMapForward, resX, resY, resZ));
    SGLD_ARG(Hs || numElements == 0, "sgld:arg: Hs is null");
    const int64_t nv = (int64_t)resX * resY * resZ;
    DevBuf<int32_t> dmap;
    DevBuf<double> dY;
    SGLD_TRY(dmap.upload(eleMapForward, nv));
    SGLD_TRY(dY.alloc(numElements));
    SGLD_TRY(run_filter(nullptr, nullptr, rMin, dmap.p, resX, resY, resZ, dY.p));
    if (numElements) SGLD_HIP(hipMemcpy(Hs, dY.p, sizeof(double) * numElements, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_density_filter(const double *x, double rMin, int64_t numElements, const int32_t *eleMapForward,
                                   int resX, int resY, int resZ, double *Y) {
    SGLD_TRY(check_map(numElements, eleMapForward, resX, resY, resZ));
    SGLD_ARG((x && Y) || numElements == 0, "sgld:arg: x or Y is null");
    const int64_t nv = (int64_t)resX * resY * resZ;
    DevBuf<int32_t> dmap;
    DevBuf<double> dx, dY;
    SGLD_TRY(dmap.upload(eleMapForward, nv));
    SGLD_TRY(dx.upload(x, numElements));
    SGLD_TRY(dY.alloc(numElements));
    SGLD_TRY(run_filter(nullptr, dx.p, rMin, dmap.p, resX, resY, resZ, dY.p));
    if (numElements) SGLD_HIP(hipMemcpy(Y, dY.p, sizeof(double) * numElements, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_ctx_density_filter_device(sgld_ctx *C, const double *dx, double rMin, double *dY) {
    SGLD_ARG(C && dY, "sgld:arg: null argument");
    const Grid &g = C->lv[0].g;
    SGLD_TRY(ensure_efwd(C));
    return run_filter(C->s, dx, rMin, C->efwd.p, g.nx, g.ny, g.nz, dY, C);
}
