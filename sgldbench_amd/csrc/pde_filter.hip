// PDE (Helmholtz) density filter of the local-volume-constraint topology
// optimisation (TopOpti_BuildPDEfilter_matrixFree.m,
// TopOpti_PDEFiltering_matrixFree.m): element values are spread to the nodes
// (src/8 per element node), the scalar system (iRmin^2 K + M) x = b is solved
// by Jacobi-preconditioned CG (the reference recurrence,
// Solving_PreconditionedConjugateGradientSolver.m, x0 = 0) with a
// matrix-free product of the 8 x 8 element matrix (Solving_KbyU_MatrixFree8x8_mex.cpp),
// and the nodal solution is averaged back to the elements.  All vectors are
// scalar node lattices of the finest level (void nodes 0); products gather
// node-centrically over the solid elements at each node.
#include <cmath>
#include <cstring>
#include <vector>

#include "lattice_ctx.hpp"

namespace sgld {
namespace {

__device__ __forceinline__ double psum(double v) {
    __shared__ double sh[4];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += sh[w];
    return s;  // thread 0
}

// node n of the lattice: the solid elements around it, their local index j of n
#define FOR_NODE_ELEMENTS(BODY)                                                                    \
    for (int dl = 0; dl < 2; ++dl)                                                                  \
        for (int dc = 0; dc < 2; ++dc)                                                              \
            for (int dr = 0; dr < 2; ++dr) {                                                        \
                const int er = r - 1 + dr, ec = c - 1 + dc, el = l - 1 + dl;                        \
                if (er < 0 || er >= gd.ny || ec < 0 || ec >= gd.nx || el < 0 || el >= gd.nz) continue; \
                const int64_t v = er + (int64_t)gd.ny * ec + (int64_t)gd.ny * gd.nx * el;           \
                if (!efwd[v]) continue;                                                             \
                const int j = (dl ? 0 : 4) + (dr ? ((1 - dc) ? 2 : 3) : ((1 - dc) ? 1 : 0));        \
                BODY                                                                                \
            }

// b = element values spread to nodes (src / 8), dinv = 1 / sum of diag(Ke)
__global__ void __launch_bounds__(256) k_pde_setup(const double *__restrict__ src, const int32_t *__restrict__ efwd,
                                                   Grid gd, const double *__restrict__ Ke, double *__restrict__ b,
                                                   double *__restrict__ dinv, double *__restrict__ partial) {
    const int64_t NY = gd.NY(), NX = gd.NX(), nN = gd.nodes();
    double acc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nN; g += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(g % NY), c = (int)((g / NY) % NX), l = (int)(g / (NY * NX));
        double bv = 0.0, d = 0.0;
        FOR_NODE_ELEMENTS(bv += src[efwd[v] - 1] * 0.125; d += Ke[j * 9];)
        b[g] = bv;
        dinv[g] = d > 0.0 ? 1.0 / d : 0.0;
        acc += bv * bv;
    }
    const double s = psum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// y = K u (scalar 8 x 8 element matrix, column-major), partial u.y
__global__ void __launch_bounds__(256) k_pde_kbyu(const double *__restrict__ u, const int32_t *__restrict__ efwd,
                                                  Grid gd, const double *__restrict__ Ke, double *__restrict__ y,
                                                  double *__restrict__ partial) {
    const int64_t NY = gd.NY(), NX = gd.NX(), nN = gd.nodes(), pl = NY * NX;
    double acc = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < nN; g += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(g % NY), c = (int)((g / NY) % NX), l = (int)(g / pl);
        double yv = 0.0;
        FOR_NODE_ELEMENTS(
            for (int k = 0; k < 8; ++k) {
                const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
                const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
                const int zk = (k >= 4);
                yv += Ke[j + 8 * k] * u[(er + 1 - yk) + NY * (ec + xk) + pl * (el + zk)];
            })
        y[g] = yv;
        acc += u[g] * yv;
    }
    const double s = psum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// x += lambda p; r -= lambda q; z = dinv r; partials r.r and z.r
__global__ void __launch_bounds__(256) k_pde_update(double *__restrict__ x, double *__restrict__ r,
                                                    double *__restrict__ z, const double *__restrict__ p,
                                                    const double *__restrict__ q, const double *__restrict__ dinv,
                                                    double lambda, int64_t n, double *__restrict__ prr,
                                                    double *__restrict__ pzr) {
    double a = 0.0, bz = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        x[i] += lambda * p[i];
        const double rn = r[i] - lambda * q[i];
        r[i] = rn;
        const double zn = dinv[i] * rn;
        z[i] = zn;
        a += rn * rn;
        bz += zn * rn;
    }
    const double s1 = psum(a);
    __syncthreads();
    const double s2 = psum(bz);
    if (threadIdx.x == 0) {
        prr[blockIdx.x] = s1;
        pzr[blockIdx.x] = s2;
    }
}

__global__ void k_pde_axpby(double *__restrict__ p, const double *__restrict__ z, double beta, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = z[i] + beta * p[i];
}

// tar_e = mean of the solution at the element's 8 nodes
__global__ void k_pde_n2e(const double *__restrict__ x, Grid gd, const int32_t *__restrict__ emap, int64_t ne,
                          double *__restrict__ tar) {
    const int64_t NY = gd.NY(), pl = gd.NY() * gd.NX();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = emap ? emap[i] : i;
        const int er = (int)(v % gd.ny), ec = (int)((v / gd.ny) % gd.nx), el = (int)(v / ((int64_t)gd.ny * gd.nx));
        double sum = 0.0;
        for (int k = 0; k < 8; ++k) {
            const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
            const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
            const int zk = (k >= 4);
            sum += x[(er + 1 - yk) + NY * (ec + xk) + pl * (el + zk)];
        }
        tar[i] = sum * 0.125;
    }
}

// TopOpti_BuildPDEfilter_matrixFree.m:7-20 (unit cell): iRmin^2 dShape'dShape + N'diag(w detJ)N
void pde_element(double rHatMin, double *Ke) {
    const double g = 0.577350269189626;
    double K0[64] = {}, K1[64] = {};
    for (int q = 0; q < 8; ++q) {
        const double s = (kCornerX[q] ? 1 : -1) * g, t = (kCornerY[q] ? 1 : -1) * g, p = (kCornerZ[q] ? 1 : -1) * g;
        double N[8], dN[3][8];
        for (int j = 0; j < 8; ++j) {
            const int cs = kCornerX[j] ? 1 : -1, ct = kCornerY[j] ? 1 : -1, cp = kCornerZ[j] ? 1 : -1;
            N[j] = 0.125 * (1 + cs * s) * (1 + ct * t) * (1 + cp * p);
            dN[0][j] = 0.125 * cs * (1 + ct * t) * (1 + cp * p);
            dN[1][j] = 0.125 * ct * (1 + cs * s) * (1 + cp * p);
            dN[2][j] = 0.125 * cp * (1 + cs * s) * (1 + ct * t);
        }
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                for (int d = 0; d < 3; ++d) K0[a + 8 * b] += dN[d][a] * dN[d][b];
                K1[a + 8 * b] += N[a] * (1.0 / 8.0) * N[b];
            }
    }
    const double rr = rHatMin / 2.0 / std::sqrt(3.0);
    for (int i = 0; i < 64; ++i) Ke[i] = rr * rr * K0[i] + K1[i];
}

}  // namespace
}  // namespace sgld

using namespace sgld;

extern "C" int sgld_ctx_pde_filter_device(sgld_ctx *C, const double *dsrc, double rHatMin, double tol, int maxIT,
                                          double *dtar, int *itsOut) {
    SGLD_ARG(C && dsrc && dtar, "sgld:arg: null argument");
    SGLD_ARG(rHatMin > 0 && tol > 0 && maxIT >= 1, "sgld:arg: need rHatMin > 0, tol > 0, maxIT >= 1");
    SGLD_TRY(ensure_efwd(C));
    const Grid &gd = C->lv[0].g;
    const int64_t n = gd.nodes(), ne = C->lv[0].numElements;
    double Ke[64];
    pde_element(rHatMin, Ke);
    DevBuf<double> dKe, b, x, r, z, p, q, dinv, part;
    SGLD_TRY(dKe.upload(Ke, 64));
    for (DevBuf<double> *v : {&b, &x, &r, &z, &p, &q, &dinv}) SGLD_TRY(v->alloc(n));
    const int nb = grid_blocks(n);
    SGLD_TRY(part.alloc(2 * (size_t)nb));
    std::vector<double> hp(2 * (size_t)nb);
    auto reduce = [&](int which, double *out) -> int {
        SGLD_HIP(hipMemcpyAsync(hp.data(), part.p + (size_t)which * nb, sizeof(double) * nb, hipMemcpyDeviceToHost, C->s));
        SGLD_HIP(hipStreamSynchronize(C->s));
        double s = 0;
        for (int i = 0; i < nb; ++i) s += hp[i];
        *out = s;
        return SGLD_OK;
    };
    hipLaunchKernelGGL(k_pde_setup, dim3(nb), dim3(256), 0, C->s, dsrc, C->efwd.p, gd, dKe.p, b.p, dinv.p, part.p);
    SGLD_HIP(hipGetLastError());
    double bb = 0;
    SGLD_TRY(reduce(0, &bb));
    const double normB = std::sqrt(bb);
    // r = b, z = M r, p = z, rho = z.r  (x0 = 0)
    SGLD_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * n, C->s));
    SGLD_HIP(hipMemsetAsync(p.p, 0, sizeof(double) * n, C->s));
    SGLD_HIP(hipMemcpyAsync(r.p, b.p, sizeof(double) * n, hipMemcpyDeviceToDevice, C->s));
    SGLD_HIP(hipMemsetAsync(q.p, 0, sizeof(double) * n, C->s));
    hipLaunchKernelGGL(k_pde_update, dim3(nb), dim3(256), 0, C->s, x.p, r.p, z.p, p.p, q.p, dinv.p, 0.0, n,
                       part.p, part.p + nb);
    SGLD_HIP(hipMemcpyAsync(p.p, z.p, sizeof(double) * n, hipMemcpyDeviceToDevice, C->s));
    double rho = 0;
    SGLD_TRY(reduce(1, &rho));
    int its = 0;
    if (normB > 0.0) {
        for (its = 1; its <= maxIT; ++its) {
            hipLaunchKernelGGL(k_pde_kbyu, dim3(nb), dim3(256), 0, C->s, p.p, C->efwd.p, gd, dKe.p, q.p, part.p);
            double pq = 0;
            SGLD_TRY(reduce(0, &pq));
            const double lambda = rho / pq;
            hipLaunchKernelGGL(k_pde_update, dim3(nb), dim3(256), 0, C->s, x.p, r.p, z.p, p.p, q.p, dinv.p, lambda,
                               n, part.p, part.p + nb);
            double rr = 0, rho2 = 0;
            SGLD_TRY(reduce(0, &rr));
            SGLD_TRY(reduce(1, &rho2));
            if (std::sqrt(rr) / normB < tol) break;
            hipLaunchKernelGGL(k_pde_axpby, dim3(nb), dim3(256), 0, C->s, p.p, (const double *)z.p, rho2 / rho, n);
            rho = rho2;
        }
        if (its > maxIT) its = maxIT;
    }
    hipLaunchKernelGGL(k_pde_n2e, dim3(grid_blocks(ne)), dim3(256), 0, C->s, x.p, gd,
                       C->full ? nullptr : C->eleMap.p, ne, dtar);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipStreamSynchronize(C->s));
    if (itsOut) *itsOut = its;
    return SGLD_OK;
}

extern "C" int sgld_ctx_pde_filter(sgld_ctx *C, const double *src, double rHatMin, double tol, int maxIT,
                                   double *tar, int *its) {
    SGLD_ARG(C && src && tar, "sgld:arg: null argument");
    const int64_t ne = C->lv[0].numElements;
    DevBuf<double> ds, dt;
    SGLD_TRY(ds.upload(src, ne));
    SGLD_TRY(dt.alloc(ne));
    SGLD_TRY(sgld_ctx_pde_filter_device(C, ds.p, rHatMin, tol, maxIT, dt.p, its));
    SGLD_HIP(hipMemcpy(tar, dt.p, sizeof(double) * ne, hipMemcpyDeviceToHost));
    return SGLD_OK;
}
