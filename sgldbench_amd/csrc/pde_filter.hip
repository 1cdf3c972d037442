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

// Device-side scalars of the filter's CG (so a batch of iterations runs
// without host waits); every iteration kernel returns at once after a stop.
struct PCtl {
    double rho, lambda, beta, normb, tol;
    int its, maxit, done, pad;
};
__device__ __forceinline__ bool pstopped(const PCtl *c) { return c && c->done; }

// The 8 x 8 element matrix by value: with the indices known at compile time
// its entries are uniform scalar loads.
struct PdeKe {
    double v[64];  // column-major, Ke[j + 8 k]
};

// y = K u (scalar 8 x 8 element matrix), partial u.y.  A thread owns a node
// column (row, col) and a chunk of kPdeCz layers and marches up it: the 3 x 3
// node plane above is loaded per layer, the two below are kept from the
// previous layers, and the 4 solid flags of the element layer above likewise,
// so a node costs 9 + 4 loads instead of 27 + 8.  Loads are unconditional
// (indices clamped into the lattice; a clamped value only meets elements
// outside it, whose weight is 0) and each element is weighted by 1 (solid) or
// 0 (void or outside), so they issue together.
constexpr int kPdeCz = 8;
__global__ void __launch_bounds__(256) k_pde_kbyu(const double *__restrict__ u, const int32_t *__restrict__ efwd,
                                                  Grid gd, PdeKe K, double *__restrict__ y,
                                                  double *__restrict__ partial, const PCtl *ctl) {
    if (pstopped(ctl)) return;
    const int NY = gd.ny + 1, NX = gd.nx + 1, NZ = gd.nz + 1;
    const int64_t pl = (int64_t)NY * NX, ple = (int64_t)gd.ny * gd.nx;
    const int64_t cols = pl, chunks = (NZ + kPdeCz - 1) / kPdeCz;
    double acc = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < cols * chunks;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = t % cols, ch = t / cols;
        const int r = (int)(col % NY), c = (int)(col / NY);
        const int l0 = (int)ch * kPdeCz, l1 = min(l0 + kPdeCz, NZ);
        int64_t off[3][3];  // in-plane offsets of the 3 x 3 neighbourhood (clamped)
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy)
                off[dx][dy] = min(max(r - 1 + dy, 0), NY - 1) + (int64_t)NY * min(max(c - 1 + dx, 0), NX - 1);
        // element (er, ec) columns around the node: in-plane offset and whether inside
        int64_t eoff[2][2];
        bool ein[2][2];
#pragma unroll
        for (int dc = 0; dc < 2; ++dc)
#pragma unroll
            for (int dr = 0; dr < 2; ++dr) {
                const int er = r - 1 + dr, ec = c - 1 + dc;
                ein[dc][dr] = er >= 0 && er < gd.ny && ec >= 0 && ec < gd.nx;
                eoff[dc][dr] = ein[dc][dr] ? er + (int64_t)gd.ny * ec : 0;
            }
        auto plane = [&](int lz, double (&pv)[3][3]) {
            const int64_t base = pl * min(max(lz, 0), NZ - 1);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) pv[dx][dy] = u[base + off[dx][dy]];
        };
        auto solid = [&](int el, double (&w)[2][2]) {
            const bool lin = el >= 0 && el < gd.nz;
            const int64_t base = lin ? ple * el : 0;
#pragma unroll
            for (int dc = 0; dc < 2; ++dc)
#pragma unroll
                for (int dr = 0; dr < 2; ++dr) {
                    const int32_t ev = efwd[base + eoff[dc][dr]];
                    w[dc][dr] = (lin && ein[dc][dr] && ev) ? 1.0 : 0.0;
                }
        };
        double P[3][3][3];   // node planes l-1, l, l+1
        double Wl[2][2][2];  // solid weights of element layers l-1, l
        plane(l0 - 1, P[0]);
        plane(l0, P[1]);
        solid(l0 - 1, Wl[0]);
        for (int l = l0; l < l1; ++l) {
            plane(l + 1, P[2]);
            solid(l, Wl[1]);
            double yv = 0.0;
#pragma unroll
            for (int dl = 0; dl < 2; ++dl)
#pragma unroll
                for (int dc = 0; dc < 2; ++dc)
#pragma unroll
                    for (int dr = 0; dr < 2; ++dr) {
                        const int j = (dl ? 0 : 4) + (dr ? ((1 - dc) ? 2 : 3) : ((1 - dc) ? 1 : 0));
                        double sEl = 0.0;
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
                            const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
                            const int zk = (k >= 4);
                            sEl += K.v[j + 8 * k] * P[dl + zk][dc + xk][dr + 1 - yk];
                        }
                        yv += Wl[dl][dc][dr] * sEl;
                    }
            y[pl * l + col] = yv;
            acc += P[1][1][1] * yv;
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int dy = 0; dy < 3; ++dy) {
                    P[0][dx][dy] = P[1][dx][dy];
                    P[1][dx][dy] = P[2][dx][dy];
                }
#pragma unroll
            for (int dc = 0; dc < 2; ++dc)
#pragma unroll
                for (int dr = 0; dr < 2; ++dr) Wl[0][dc][dr] = Wl[1][dc][dr];
        }
    }
    const double s = psum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0] = sum of partial[0 .. nb), out[1] = sum of partial[nb .. 2 nb)
// (one block, fixed order), so the host copies two values instead of the
// partials.
__global__ void __launch_bounds__(256) k_pde_fin(const double *__restrict__ partial, int nb, int sets,
                                                 double *__restrict__ out) {
    for (int k = 0; k < sets; ++k) {
        double a = 0.0;
        for (int i = threadIdx.x; i < nb; i += blockDim.x) a += partial[(int64_t)k * nb + i];
        const double s = psum(a);
        if (threadIdx.x == 0) out[k] = s;
        __syncthreads();
    }
}

// Scalar steps of one CG iteration (Solving_PreconditionedConjugateGradientSolver.m:
// 15-29) from the partial sums: op 0, lambda = rho / p.q; op 1, with r.r and
// z.r, the stopping test on ||r|| / ||b|| (done 1; done 2 at maxit), then
// beta = z.r / rho and rho = z.r.
__global__ void __launch_bounds__(256) k_pde_fin_iter(const double *__restrict__ partial, int nb, int op,
                                                      PCtl *ctl) {
    if (ctl->done) return;
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        a += partial[i];
        if (op == 1) b += partial[nb + i];
    }
    const double sa = psum(a);
    __syncthreads();
    const double sb = psum(b);
    if (threadIdx.x != 0) return;
    if (op == 0) {
        ctl->lambda = ctl->rho / sa;
        return;
    }
    ctl->its += 1;
    if (sqrt(sa) / ctl->normb < ctl->tol) {
        ctl->done = 1;
        return;
    }
    if (ctl->its >= ctl->maxit) ctl->done = 2;
    ctl->beta = sb / ctl->rho;
    ctl->rho = sb;
}

// x += lambda p; r -= lambda q; z = dinv r; partials r.r and z.r
__global__ void __launch_bounds__(256) k_pde_update(double *__restrict__ x, double *__restrict__ r,
                                                    double *__restrict__ z, const double *__restrict__ p,
                                                    const double *__restrict__ q, const double *__restrict__ dinv,
                                                    double lambda0, int64_t n, double *__restrict__ prr,
                                                    double *__restrict__ pzr, const PCtl *ctl) {
    if (pstopped(ctl)) return;
    const double lambda = ctl ? ctl->lambda : lambda0;
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

__global__ void k_pde_axpby(double *__restrict__ p, const double *__restrict__ z, int64_t n, const PCtl *ctl) {
    if (pstopped(ctl)) return;
    const double beta = ctl->beta;
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
    PdeKe K;
    std::memcpy(K.v, Ke, sizeof(Ke));
    // work vectors b, x, r, z, p, q, dinv, the element matrix, partial sums
    // and the control block, in one allocation kept by the context (the
    // per-call allocations cost about as much as the iterations)
    const int nb = grid_blocks(n);
    const size_t need = 7 * (size_t)n + 64 + 2 * (size_t)nb + 2 + 8;
    if (C->pdeWork.n < need) SGLD_TRY(C->pdeWork.alloc(need));
    struct View {
        double *p;
    };
    View b{C->pdeWork.p}, x{b.p + n}, r{x.p + n}, z{r.p + n}, p{z.p + n}, q{p.p + n}, dinv{q.p + n};
    View dKe{dinv.p + n}, part{dKe.p + 64};
    PCtl *dctl = reinterpret_cast<PCtl *>(part.p + 2 * (size_t)nb + 2);
    static_assert(sizeof(PCtl) <= 8 * sizeof(double), "control block slot");
    SGLD_HIP(hipMemcpyAsync(dKe.p, Ke, sizeof(Ke), hipMemcpyHostToDevice, C->s));
    double *sums = part.p + 2 * (size_t)nb;  // k_pde_fin's two results
    // the two sums travel through the context's pinned control-block mirror
    // (16 bytes, one wait), not through pageable memory
    double *hs = reinterpret_cast<double *>(C->hctl);
    auto reduce_sets = [&](int first, int sets, double *out0, double *out1) -> int {
        hipLaunchKernelGGL(k_pde_fin, dim3(1), dim3(256), 0, C->s, (const double *)part.p + (size_t)first * nb, nb,
                           sets, sums);
        SGLD_HIP(hipGetLastError());
        SGLD_HIP(hipMemcpyAsync(hs, sums, sizeof(double) * sets, hipMemcpyDeviceToHost, C->s));
        SGLD_HIP(hipStreamSynchronize(C->s));
        *out0 = hs[0];
        if (out1) *out1 = hs[1];
        return SGLD_OK;
    };
    auto reduce = [&](int which, double *out) -> int { return reduce_sets(which, 1, out, nullptr); };
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
                       part.p, part.p + nb, (const PCtl *)nullptr);
    SGLD_HIP(hipMemcpyAsync(p.p, z.p, sizeof(double) * n, hipMemcpyDeviceToDevice, C->s));
    double rho = 0;
    SGLD_TRY(reduce(1, &rho));
    int its = 0;
    if (normB > 0.0) {
        // batches of iterations with the scalars on the device; the host
        // looks at the control block once per batch (through the context's
        // pinned mirror)
        static_assert(sizeof(PCtl) <= sizeof(Ctl), "PCtl travels through the Ctl mirror");
        PCtl *hc = reinterpret_cast<PCtl *>(C->hctl);
        *hc = PCtl{rho, 0.0, 0.0, normB, tol, 0, maxIT, 0, 0};
        PCtl *dcp = dctl;
        SGLD_HIP(hipMemcpyAsync(dcp, hc, sizeof(PCtl), hipMemcpyHostToDevice, C->s));
        constexpr int kBatch = 8;
        for (;;) {
            for (int i = 0; i < kBatch; ++i) {
                hipLaunchKernelGGL(k_pde_kbyu, dim3(nb), dim3(256), 0, C->s, p.p, C->efwd.p, gd, K, q.p, part.p,
                                   (const PCtl *)dcp);
                hipLaunchKernelGGL(k_pde_fin_iter, dim3(1), dim3(256), 0, C->s, (const double *)part.p, nb, 0, dcp);
                hipLaunchKernelGGL(k_pde_update, dim3(nb), dim3(256), 0, C->s, x.p, r.p, z.p, p.p, q.p, dinv.p, 0.0,
                                   n, part.p, part.p + nb, (const PCtl *)dcp);
                hipLaunchKernelGGL(k_pde_fin_iter, dim3(1), dim3(256), 0, C->s, (const double *)part.p, nb, 1, dcp);
                hipLaunchKernelGGL(k_pde_axpby, dim3(nb), dim3(256), 0, C->s, p.p, (const double *)z.p, n,
                                   (const PCtl *)dcp);
            }
            SGLD_HIP(hipGetLastError());
            SGLD_HIP(hipMemcpyAsync(hc, dcp, sizeof(PCtl), hipMemcpyDeviceToHost, C->s));
            SGLD_HIP(hipStreamSynchronize(C->s));
            if (hc->done) break;
        }
        its = hc->its;
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
