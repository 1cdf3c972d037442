// Device-resident voxel FEA solver on the node lattice: K.u, additive
// multigrid V-cycle, PCG (FP64, and FP32-inner with FP64 refinement), unit
// compliance and the Galerkin setup.  See lattice_kernels.hpp for the layout.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "lattice_ctx.hpp"

namespace sgld {
namespace {

// The element matrix is per context: the dense 24 x 24 Ke by pointer to the
// context's device copy, the nine mode-space coefficients (ModeK, see
// mode_forces) by value as a kernel argument.

// Local node j of the corner (x, y, z) bits: node order of FEA_GaussianIntegral.m:2-9.
__host__ __device__ constexpr int corner_index(int x, int y, int z) {
    return (z ? 4 : 0) + (y ? (x ? 2 : 3) : (x ? 1 : 0));
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[16];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    v = (threadIdx.x < nw) ? sh[threadIdx.x] : 0.0;
    if (w == 0)
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;  // valid in thread 0
}

__device__ __forceinline__ void block_sum2(double &a, double &b) {
    __shared__ double sh[2][16];
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        sh[0][w] = a;
        sh[1][w] = b;
    }
    __syncthreads();
    const int nw = (blockDim.x + 63) >> 6;
    a = (threadIdx.x < nw) ? sh[0][threadIdx.x] : 0.0;
    b = (threadIdx.x < nw) ? sh[1][threadIdx.x] : 0.0;
    if (w == 0)
        for (int o = 32; o > 0; o >>= 1) {
            a += __shfl_down(a, o, 64);
            b += __shfl_down(b, o, 64);
        }
}

__device__ __forceinline__ bool stopped(const Ctl *ctl) { return ctl && ctl->done; }

// Finest-level Jacobi weights dinv = 2 omega / diag: FP64 for the FP64 solve;
// for the FP32 inner solve of the mixed mode rounded to bfloat16 (a diagonal
// preconditioner perturbed by <= 2^-9 relative is still SPD, it is used
// consistently in z and rho, and the FP64 reliable updates fix the accuracy),
// which halves the bytes of the two finest passes that read it.
__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite f)
}
template <class T> struct FineW;
// Loads of a vector that a pass streams once: nontemporal (NT) when the
// pass's vectors outgrow the Infinity Cache, so that nothing of them would
// survive to the next pass anyway (measured on 147 M-entry FP32 vectors: a
// direction-update-shaped pass over 64-entry chunks 4.90 -> 5.33 TB/s; on the
// 125 M-DOF ellipsoid solve, mixed 875 -> 840 ms, and 864 ms with p and r,
// which the next kernel reads, left on cached loads), ordinary loads otherwise.
template <bool NT, class V>
__device__ __forceinline__ V ld_stream(const V *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
// The matching store, for outputs of a pass beyond the Infinity Cache that
// the next pass streams from HBM: the caches then keep what the next pass
// re-reads (125 M-DOF ellipsoids, mixed solve 778 -> 764 ms with the direction
// and residual updates' stores nontemporal, of which the FP32 K.u after the
// direction update 498 -> 466 us; the fused residual update's stores
// 525 -> 511 us).  SGLD_NT_STORE=0 / SGLD_NT_STORE_FUSED=0 at build time:
// ordinary stores.
#ifndef SGLD_NT_STORE
#define SGLD_NT_STORE 1
#endif
#ifndef SGLD_NT_STORE_FUSED
#define SGLD_NT_STORE_FUSED 1
#endif
template <bool NT, class V>
__device__ __forceinline__ void st_stream(V *p, V v) {
    if constexpr (NT && SGLD_NT_STORE) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Entry of a finest-level weight or FP32 correction vector: the FP64 solve's
// vectors are lattice vectors (entry i of the lattice at i); the bfloat16
// weights and the FP32 correction of the mixed solve hold only the DOF chunks
// of the span, in span order (lattice entry i at its span position t; t = i
// on a full cuboid), since only streaming passes read them: on an irregular
// domain the passes then touch a third of the address range (125 M-DOF
// ellipsoids: residual update 360 -> 338 us, direction update 633 -> 614 us,
// measured on the same access pattern).
template <class T>
__device__ __forceinline__ int64_t fine_pos(int64_t i, int64_t t) {
    return sizeof(T) == 4 ? t : i;
}

template <> struct FineW<double> {
    using D = double;
    template <bool NT = false>
    __device__ __forceinline__ static double get(const double *p, int64_t i) { return ld_stream<NT>(&p[i]); }
    __device__ __forceinline__ static double value(double raw) { return raw; }
};
template <> struct FineW<float> {
    using D = uint16_t;
    template <bool NT = false>
    __device__ __forceinline__ static float get(const uint16_t *p, int64_t i) {
        return __uint_as_float((uint32_t)ld_stream<NT>(&p[i]) << 16);
    }
    __device__ __forceinline__ static float value(uint16_t raw) { return __uint_as_float((uint32_t)raw << 16); }
};

// Grid-stride passes over the entries of a Span (all entries, or the listed
// chunks of an irregular domain).
#define FOR_SPAN(i, S)                                                                           \
    for (int64_t t_ = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t_ < (S).count;            \
         t_ += (int64_t)gridDim.x * blockDim.x)                                                    \
        if (const int64_t i = (S).at(t_); i < (S).n)

// Block index and grid size a kernel body works with: the hardware ones for
// a launch of one solve; for a launch over many solves (one block range per
// instance, see the multi-instance iteration) the block's index within its
// instance's range and that range's length.
struct VB {
    unsigned b, g;  // unsigned, as blockIdx.x / gridDim.x
    __device__ static VB self() { return VB{blockIdx.x, gridDim.x}; }
};
#define FOR_SPAN_V(i, S, vb)                                                                      \
    for (int64_t t_ = (int64_t)(vb).b * blockDim.x + threadIdx.x; t_ < (S).count;               \
         t_ += (int64_t)(vb).g * blockDim.x)                                                        \
        if (const int64_t i = (S).at(t_); i < (S).n)
// The same, with the entry's position t in the span as well.
#define FOR_SPAN_T(i, t, S)                                                                       \
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (S).count;              \
         t += (int64_t)gridDim.x * blockDim.x)                                                      \
        if (const int64_t i = (S).at(t); i < (S).n)
#define FOR_SPAN_VT(i, t, S, vb)                                                                  \
    for (int64_t t = (int64_t)(vb).b * blockDim.x + threadIdx.x; t < (S).count;                  \
         t += (int64_t)(vb).g * blockDim.x)                                                         \
        if (const int64_t i = (S).at(t); i < (S).n)

// ----------------------------------------------------------------------------
// K.u: y = sum_e E_e Ke u_e at every lattice node, gathered node-centrically
// over the node's 8 incident voxels (no atomics, no colours).  Fixed and void
// DOFs are set to zero (Solving_KbyU_MatrixFree.m:39).  Optional partial of
// u.y for the PCG step length.  Grid-stride over nodes.
// ----------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ void kbyu_body(VB vb, const T *__restrict__ u, T *__restrict__ y,
                                              const T *__restrict__ Ke, const T *__restrict__ E,
                                              const uint8_t *__restrict__ nflag, Grid gd, Span S,
                                              double *__restrict__ partial, const Ctl *ctl) {
    if (stopped(ctl)) return;
    const int64_t NY = gd.NY(), NX = gd.NX();
    const int64_t pl = NY * NX;
    double dotacc = 0.0;
    FOR_SPAN_V(g, S, vb) {
        const int r = (int)(g % NY), c = (int)((g / NY) % NX), l = (int)(g / pl);
        const uint8_t f = nflag[g];
        T acc[3] = {0, 0, 0};
        if (f & kSolid) {
#pragma unroll
            for (int dl = 0; dl < 2; ++dl)
#pragma unroll
                for (int dc = 0; dc < 2; ++dc)
#pragma unroll
                    for (int dr = 0; dr < 2; ++dr) {
                        const int er = r - 1 + dr, ec = c - 1 + dc, el = l - 1 + dl;
                        if (er < 0 || er >= gd.ny || ec < 0 || ec >= gd.nx || el < 0 || el >= gd.nz)
                            continue;
                        // this node is corner (x, y, z) = (1-dc, dr, 1-dl) of the element
                        const int j = corner_index(1 - dc, dr, 1 - dl);
                        const T Ee = E[er + (int64_t)gd.ny * ec + (int64_t)gd.ny * gd.nx * el];
                        T s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
                            const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
                            const int zk = (k >= 4);
                            const int64_t gk = (er + 1 - yk) + NY * (ec + xk) + pl * (el + zk);
                            const T u0 = u[3 * gk], u1 = u[3 * gk + 1], u2 = u[3 * gk + 2];
                            s0 += Ke[(3 * j + 0) + 24 * (3 * k + 0)] * u0 +
                                  Ke[(3 * j + 0) + 24 * (3 * k + 1)] * u1 +
                                  Ke[(3 * j + 0) + 24 * (3 * k + 2)] * u2;
                            s1 += Ke[(3 * j + 1) + 24 * (3 * k + 0)] * u0 +
                                  Ke[(3 * j + 1) + 24 * (3 * k + 1)] * u1 +
                                  Ke[(3 * j + 1) + 24 * (3 * k + 2)] * u2;
                            s2 += Ke[(3 * j + 2) + 24 * (3 * k + 0)] * u0 +
                                  Ke[(3 * j + 2) + 24 * (3 * k + 1)] * u1 +
                                  Ke[(3 * j + 2) + 24 * (3 * k + 2)] * u2;
                        }
                        acc[0] += Ee * s0;
                        acc[1] += Ee * s1;
                        acc[2] += Ee * s2;
                    }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T v = ((f & kSolid) && !(f & (1 << a))) ? acc[a] : T(0);
            y[3 * g + a] = v;
            if (partial) dotacc += (double)u[3 * g + a] * (double)v;
        }
    }
    if (partial) {
        const double s = block_sum(dotacc);
        if (threadIdx.x == 0) partial[vb.b] = s;
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_kbyu(const T *__restrict__ u, T *__restrict__ y,
                                              const T *__restrict__ Ke, const T *__restrict__ E,
                                              const uint8_t *__restrict__ nflag, Grid gd, Span S,
                                              double *__restrict__ partial, const Ctl *ctl) {
    kbyu_body<T>(VB::self(), u, y, Ke, E, nflag, gd, S, partial, ctl);
}


// Mode forces of one hexahedron: uh[a][m] holds the Hadamard modes
// (m = x + 2y + 4z difference bits) of component a on entry, the force modes
// E K' uh on exit (in place).
template <class T>
__device__ __forceinline__ void mode_forces(T (&uh)[3][8], T Ee, const ModeK<T> &km) {
    // K' is block diagonal: modes couple only within {(x,1),(y,2),(z,4)},
    // {(x,6),(y,5),(z,3)}, the pairs {(a,e_b),(b,e_a)} and {(a,e_a+e_c),
    // (b,e_b+e_c)}, and (a,7) alone; mode 0 (translation) carries no force.
    // For the isotropic cube, axis permutations leave K' invariant, so nine
    // numbers define it (km9: own-axis linear, cross-axis linear, bilinear
    // with own axis, bilinear without, trilinear, then the four couplings).
    // E scales the nine coefficients (9 multiplies instead of 21 on the modes)
    // (the 3 x 3 blocks as d x_i + c (x + y + z) with d = own - coupling:
    // one shared sum and product instead of three)
    const T dLo = Ee * (km.v[0] - km.v[5]), dLx = Ee * km.v[1], dBo = Ee * km.v[2], dBx = Ee * (km.v[3] - km.v[8]);
    const T dT = Ee * km.v[4];
    const T c0 = Ee * km.v[5], c1 = Ee * km.v[6], c2 = Ee * km.v[7], c3 = Ee * km.v[8];
    {
        const T x = uh[0][1], yv = uh[1][2], z = uh[2][4];
        const T cs = c0 * (x + yv + z);
        uh[0][1] = dLo * x + cs;
        uh[1][2] = dLo * yv + cs;
        uh[2][4] = dLo * z + cs;
    }
    {
        const T x = uh[0][6], yv = uh[1][5], z = uh[2][3];
        const T cs = c3 * (x + yv + z);
        uh[0][6] = dBx * x + cs;
        uh[1][5] = dBx * yv + cs;
        uh[2][3] = dBx * z + cs;
    }
    {
        const T x = uh[0][2], yv = uh[1][1];
        uh[0][2] = dLx * x + c1 * yv;
        uh[1][1] = dLx * yv + c1 * x;
    }
    {
        const T x = uh[0][4], z = uh[2][1];
        uh[0][4] = dLx * x + c1 * z;
        uh[2][1] = dLx * z + c1 * x;
    }
    {
        const T yv = uh[1][4], z = uh[2][2];
        uh[1][4] = dLx * yv + c1 * z;
        uh[2][2] = dLx * z + c1 * yv;
    }
    {
        const T x = uh[0][5], yv = uh[1][6];
        uh[0][5] = dBo * x + c2 * yv;
        uh[1][6] = dBo * yv + c2 * x;
    }
    {
        const T x = uh[0][3], z = uh[2][6];
        uh[0][3] = dBo * x + c2 * z;
        uh[2][6] = dBo * z + c2 * x;
    }
    {
        const T yv = uh[1][3], z = uh[2][5];
        uh[1][3] = dBo * yv + c2 * z;
        uh[2][5] = dBo * z + c2 * yv;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        uh[a][7] *= dT;
        uh[a][0] = T(0);  // translation carries no force (callers skip it)
    }
}

// ----------------------------------------------------------------------------
// K.u, tiled mode-space form ("element columns").
//
// A block owns a 15 x 15 (row x col) patch of output nodes and a chunk of
// node layers; thread (ei, ej) owns the element column under one face of the
// 16 x 16 element patch around it and marches up through the layers.  The
// per-element Walsh-Hadamard transform is split separably: the (x, y) part is
// taken once per face and node layer (F), the z part combines the faces
// below and above an element; on the way back the z part splits the element
// forces onto its two faces, the lower face completes (its upper-element part
// is carried in registers from the previous layer), and only then is the
// (x, y) part inverted and scattered to the face's four nodes, 12 LDS adds
// per element into an FP64 node-layer accumulator.  With K' = H Ke H^T / 64
// block diagonal (mode_forces), an element costs about 175 adds/multiplies
// instead of the 576 FMA of the dense product.
//
// Persistent, XCD-aware grid: block b runs on XCD b % 8 and walks the work
// items (patch x z-chunk) of that XCD's contiguous share, so patches sharing
// halo nodes run behind the same L2.  Node layers are prefetched two
// iterations ahead into registers and staged in an LDS ring (4 slots with
// one block barrier per element layer, 3 slots with two).  On an
// irregular domain `wlist` holds the (ascending) work items whose output
// nodes touch a solid element; the others would only write zeros over void
// nodes that are zero already.
// ----------------------------------------------------------------------------

// Raw buffer access (CDNA buffer instructions): a load whose byte offset is
// at or past the descriptor's range returns 0 and a store there is dropped,
// so slots outside the lattice need neither a select nor a branch.
constexpr uint32_t kOob = 0x7FFFFFF0u;  // offset past every range used here
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <class T> __device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t off);
template <> __device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ void buf_store(double v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v),
                                          r, (int)off, 0, 0);
}
__device__ __forceinline__ void buf_store(float v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)off, 0, 0);
}

template <class T>
__device__ __forceinline__ void face_modes(const T *__restrict__ lay, int ei, int ej, T (&F)[3][4]) {
    // corners of face (ei, ej): (x0,y0) = node (ei+1, ej), (x1,y0) = (ei+1, ej+1),
    // (x0,y1) = (ei, ej), (x1,y1) = (ei, ej+1)
    const T *n00 = lay + 3 * ((ei + 1) + kTN * ej);
    const T *n10 = lay + 3 * ((ei + 1) + kTN * (ej + 1));
    const T *n01 = lay + 3 * (ei + kTN * ej);
    const T *n11 = lay + 3 * (ei + kTN * (ej + 1));
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const T s0 = n00[a] + n10[a], d0 = n10[a] - n00[a];
        const T s1 = n01[a] + n11[a], d1 = n11[a] - n01[a];
        F[a][0] = s0 + s1;
        F[a][1] = d0 + d1;
        F[a][2] = s1 - s0;
        F[a][3] = d1 - d0;
    }
}

// OneBar selects one block barrier per element layer instead of two: a
// 4-slot node-layer ring and 3 accumulators put every write-after-read pair
// (staging layer el+2 over layer el-2, clearing the accumulator of layer el+1
// last read for layer el-2) one step apart, so the barrier of the step in
// between orders them.
template <class T, bool OneBar>
__global__ void __launch_bounds__(256, sizeof(T) == 8 ? 3 : 4) k_kbyu_tiled(const T *__restrict__ u, T *__restrict__ y,
                                                       const T *__restrict__ E,
                                                       const uint8_t *__restrict__ nflag, Grid gd,
                                                       TileMap tm, const int32_t *__restrict__ wlist,
                                                       int64_t wn, double *__restrict__ partial,
                                                       const Ctl *ctl, ModeK<T> km) {
    if (stopped(ctl)) return;
    constexpr int kSlots = OneBar ? 4 : 3, kAcc = OneBar ? 3 : 2;
    __shared__ T us[kSlots][kTN * kTN * 3];
    // node forces of two (three) layers, FP64 in both precisions: native
    // ds_add_f32 made the FP32 kernel 6.7x slower than ds_add_f64 (337 vs 50 us, cfg 1)
    __shared__ double acc[kAcc][kTN * kTN * 3];
    const int t = threadIdx.x;
    const int NY = gd.ny + 1, NX = gd.nx + 1, NZ = gd.nz + 1;
    const int64_t pl = (int64_t)NY * NX;       // nodes per layer
    const int64_t ple = (int64_t)gd.ny * gd.nx;  // elements per layer
    const int ei = t % kTE, ej = t / kTE;
    const int64_t W = wlist ? wn : tm.work();  // work items: all, or the listed ones
    const int G = gridDim.x;
    const int nxcd = G < kXCD ? G : kXCD;  // XCDs that hold at least one block
    const int xcd = blockIdx.x % nxcd, slotInXcd = blockIdx.x / nxcd;
    const int blocksOnXcd = G / nxcd + (xcd < G % nxcd ? 1 : 0);
    const int64_t wBeg = W * xcd / nxcd, wEnd = W * (xcd + 1) / nxcd;
    for (int k = t; k < kAcc * kTN * kTN * 3; k += 256) (&acc[0][0])[k] = 0.0;
    // LDS offsets (in T) of this face's corners
    const int f00 = 3 * ((ei + 1) + kTN * ej), f10 = 3 * ((ei + 1) + kTN * (ej + 1));
    const int f01 = 3 * (ei + kTN * ej), f11 = 3 * (ei + kTN * (ej + 1));
    // A staged node layer is kTN column runs of 3 kTN contiguous values (the
    // LDS layout 3 (row + kTN col) + a); thread t copies values j = t + 256 q,
    // so every global access of a wavefront is contiguous.  The outputs are
    // the interior of the staged layer, written from the same slots.  Per
    // slot: lOff, the byte offset of its value from the staged corner of a
    // node layer (kOob outside the lattice), and oOff, the same for outputs
    // (kOob unless an output node); node flags sit at oOff / (3 sizeof(T)).
    const FastDiv d3 = FastDiv::make3();
    uint32_t lOff[kStageSlots], oOff[kStageSlots];
    double dotacc = 0.0;
    for (int64_t w = wBeg + slotInXcd; w < wEnd; w += blocksOnXcd) {
        int64_t rem = wlist ? wlist[w] : w;
        const int tz = (int)(rem / ((int64_t)tm.tilesR * tm.tilesC));
        rem -= (int64_t)tz * tm.tilesR * tm.tilesC;
        const int tc = (int)(rem / tm.tilesR), tr = (int)(rem % tm.tilesR);
        const int R0 = tr * kTO, C0 = tc * kTO, Z0 = tz * tm.cz;
        const int Zend = min(Z0 + tm.cz, NZ);
        const int er = R0 - 1 + ei, ec = C0 - 1 + ej;
        const bool einside = er >= 0 && er < gd.ny && ec >= 0 && ec < gd.nx;
#pragma unroll
        for (int q = 0; q < kStageSlots; ++q) {
            const int j = t + 256 * q, col = j / (3 * kTN), off = j - 3 * kTN * col, row = off / 3;
            const int gr = R0 - 1 + row, gcl = C0 - 1 + col;
            const bool ok = j < 3 * kTN * kTN && gr >= 0 && gr < NY && gcl >= 0 && gcl < NX;
            const uint32_t byteOff = (uint32_t)sizeof(T) * (uint32_t)(3 * NY * col + off);
            lOff[q] = ok ? byteOff : kOob;
            oOff[q] = (ok && row >= 1 && row <= kTO && col >= 1 && col <= kTO) ? byteOff : kOob;
        }
        // per node layer: the staged corner (R0 - 1, C0 - 1) and the bytes
        // spanned by the staged columns
        const int64_t sBase = 3 * ((int64_t)(R0 - 1) + (int64_t)NY * (C0 - 1));
        const uint32_t vBytes = (uint32_t)sizeof(T) * (uint32_t)(3 * NY * (kTN - 1) + 3 * kTN);
        const uint32_t fBytes = (uint32_t)(NY * (kTN - 1) + kTN);
        auto uRsrc = [&](int L) { return buf_rsrc(u + sBase + 3 * pl * L, (L >= 0 && L < NZ) ? vBytes : 0u); };
        const T *pE = E + (einside ? er + (int64_t)gd.ny * ec + ple * Z0 : 0);  // layer el+1
        T pf[kStageSlots];
        auto loadL = [&](__amdgpu_buffer_rsrc_t r, T (&q)[kStageSlots]) {
#pragma unroll
            for (int k = 0; k < kStageSlots; ++k) q[k] = buf_load<T>(r, lOff[k]);
        };
        auto stage = [&](T *dst, const T (&q)[kStageSlots]) {
#pragma unroll
            for (int k = 0; k < kStageSlots; ++k)
                if (k < kStageSlots - 1 || t + 256 * k < 3 * kTN * kTN) dst[t + 256 * k] = q[k];
        };
        // layers el, el+1, el+2 (and el-1 with OneBar)
        T *sPrev = us[0], *sCur = us[1], *sNext = us[2], *sFree = us[kSlots - 1];
        double *aCur = acc[0], *aNext = acc[1], *aFree = acc[kAcc - 1];
        {   // stage node layers Z0-1 and Z0; layer Z0+1 stays in flight in pf
            T p0[kStageSlots], p1[kStageSlots];
            loadL(uRsrc(Z0 - 1), p0);
            loadL(uRsrc(Z0), p1);
            loadL(uRsrc(Z0 + 1), pf);
            __syncthreads();  // previous work item done with LDS
            stage(sPrev, p0);
            stage(sCur, p1);
        }
        __syncthreads();
        T Flo[3][4], Alo[3][4];
        face_modes<T>(sPrev, ei, ej, Flo);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q) Alo[a][q] = T(0);
        T Ecur = (einside && Z0 - 1 >= 0) ? pE[-ple] : T(0);
        bool eokCur = true;  // Ecur valid (else the element is outside the lattice)
        // One element layer.  On entry sPrev/sCur hold node layers el and el+1,
        // `ready` holds layer el+2 (loaded during the previous step), Flo are
        // the face modes of layer el and Alo the part of layer el's face
        // forces from element layer el-1.  `spare` receives layer el+3 and
        // stays in flight through the next step; the caller alternates the
        // two buffers so no register copy waits on a load.
        auto step = [&](int el, T (&ready)[kStageSlots], T (&spare)[kStageSlots]) {
            // loads in the order they are consumed: output flags of node layer
            // el (after the barrier), the modulus of element layer el+1 (end of
            // the step), node layer el+3 (two steps later)
            const bool out = el >= Z0;
            const __amdgpu_buffer_rsrc_t fr = buf_rsrc(nflag + sBase / 3 + pl * el, out ? fBytes : 0u);
            uint32_t flv[kStageSlots];
#pragma unroll
            for (int q = 0; q < kStageSlots; ++q)
                flv[q] = __builtin_amdgcn_raw_buffer_load_b8(fr, (int)d3.div(oOff[q] / (uint32_t)sizeof(T)), 0, 0);
            const bool eok = einside && el + 1 < gd.nz;
            const T Enext = *(eok ? pE : E);  // raw; zeroed where it is used
            loadL(uRsrc(el + 3), spare);
            T m[3][8];
            {
                T Fhi[3][4];
                face_modes<T>(sCur, ei, ej, Fhi);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        m[a][q] = Flo[a][q] + Fhi[a][q];
                        m[a][q + 4] = Fhi[a][q] - Flo[a][q];
                        Flo[a][q] = Fhi[a][q];
                    }
            }
            mode_forces<T>(m, eokCur ? Ecur : T(0), km);
            if (el >= Z0) {
                // lower face complete: inverse (x, y) transform, scatter 4 nodes
                // (force mode 0 is zero: written out, since x + 0 is not folded)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T A0 = Alo[a][0] - m[a][4], A1 = Alo[a][1] + (m[a][1] - m[a][5]);
                    const T A2 = Alo[a][2] + (m[a][2] - m[a][6]), A3 = Alo[a][3] + (m[a][3] - m[a][7]);
                    const T t0 = A0 - A2, t1 = A1 - A3, t2 = A0 + A2, t3 = A1 + A3;
                    atomicAdd(aCur + f00 + a, (double)(t0 - t1));
                    atomicAdd(aCur + f10 + a, (double)(t0 + t1));
                    atomicAdd(aCur + f01 + a, (double)(t2 - t3));
                    atomicAdd(aCur + f11 + a, (double)(t2 + t3));
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                Alo[a][0] = m[a][4];
#pragma unroll
                for (int q = 1; q < 4; ++q) Alo[a][q] = m[a][q] + m[a][q + 4];
            }
            Ecur = Enext;
            eokCur = eok;
            if constexpr (OneBar) {
                // layer el+2 into the slot of layer el-2 and a clear accumulator
                // for layer el+1 (both last read before the previous barrier)
                stage(sNext, ready);
                for (int k = t; k < kTN * kTN * 3; k += 256) aNext[k] = 0.0;
            }
            __syncthreads();  // node layer el accumulated
            if (out) {
                const __amdgpu_buffer_rsrc_t yr = buf_rsrc(y + sBase + 3 * pl * el, vBytes);
#pragma unroll
                for (int q = 0; q < kStageSlots; ++q) {
                    if (q == kStageSlots - 1 && t + 256 * q >= 3 * kTN * kTN) continue;
                    const uint32_t vi = oOff[q] / (uint32_t)sizeof(T);
                    const uint32_t a = vi - 3 * d3.div(vi);  // component (any value for kOob: flags are 0 there)
                    const int j = t + 256 * q;
                    const T v = ((flv[q] & kSolid) && !(flv[q] & (1u << a))) ? (T)aCur[j] : T(0);
                    buf_store(v, yr, oOff[q]);
                    dotacc += (double)sPrev[j] * (double)v;
                }
            }
            if constexpr (OneBar) {
                T *tmp = sPrev;
                sPrev = sCur;
                sCur = sNext;
                sNext = sFree;
                sFree = tmp;
                double *ta = aCur;
                aCur = aNext;
                aNext = aFree;
                aFree = ta;
                pE += ple;
            } else {
                // clear the other accumulator (read during the previous step,
                // scattered into during the next); halo nodes get discarded sums
                for (int k = t; k < kTN * kTN * 3; k += 256) aNext[k] = 0.0;
                // layer el+2 into the slot of layer el-1 (free since the barrier above)
                stage(sNext, ready);
                T *tmp = sPrev;
                sPrev = sCur;
                sCur = sNext;
                sNext = tmp;
                double *ta = aCur;
                aCur = aNext;
                aNext = ta;
                pE += ple;
                __syncthreads();  // layer el+2 staged, accumulator clear
            }
        };
        T pn[kStageSlots];
        for (int el = Z0 - 1; el < Zend; el += 2) {
            step(el, pf, pn);
            if (el + 1 < Zend) step(el + 1, pn, pf);
        }
    }
    if (partial) {
        const double s = block_sum(dotacc);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// ----------------------------------------------------------------------------
// K.u, tiled mode-space form with owner-computes node sums (no atomics, the
// same result in every run).
//
// Work decomposition, node-layer staging and element arithmetic are those of
// k_kbyu_tiled.  After the inverse (x, y) transform, element column (ei, ej)
// holds the forces of its lower face's corners A = node (ei+1, ej),
// B = (ei+1, ej+1), C = (ei, ej), D = (ei, ej+1).  Thread (ei, ej) owns node
// (ei, ej) and sums, always in this order,
//   y = (C(ei, ej) + A(ei-1, ej)) + (D(ei, ej-1) + B(ei-1, ej-1)).
// The lanes of one 16-lane DPP row are the rows ei of one patch column, so
// the (ei-1) terms arrive by a row shift; the (ej-1) pair comes through one
// LDS exchange per element layer, double-buffered so the layer's single block
// barrier orders every write and read.  Outputs are stored from registers.
// ----------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T row_shr1(T v);  // lane - 1 of the 16-lane row; 0 in its lane 0
template <> __device__ __forceinline__ float row_shr1<float>(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));
}
template <> __device__ __forceinline__ double row_shr1<double>(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, 0x111, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), 0x111, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// StageOut: node sums go through a double-buffered 16 x 16-node LDS layer and
// are stored one step later with contiguous full-wavefront stores (value
// t + 256 q of the layer); otherwise every thread stores its node's three
// values itself.
template <class T, bool StageOut>
__device__ __forceinline__ void kbyu_owner_body(VB vb, const T *__restrict__ u, T *__restrict__ y,
                                                       const T *__restrict__ E,
                                                       const uint8_t *__restrict__ nflag, Grid gd,
                                                       TileMap tm, const int32_t *__restrict__ wlist,
                                                       int64_t wn, double *__restrict__ partial,
                                                       const Ctl *ctl, ModeK<T> km) {
    if (stopped(ctl)) return;
    constexpr int kSlots = 4;
    __shared__ T us[kSlots][kTN * kTN * 3];
    __shared__ T xch[2][3][256];  // D + B(ei-1) of every element column, two element layers
    constexpr int kON = 3 * kTE * kTE;       // staged output values per node layer
    __shared__ T ost[StageOut ? 2 : 1][StageOut ? kON : 1];
    const int t = threadIdx.x;
    const int NY = gd.ny + 1, NX = gd.nx + 1, NZ = gd.nz + 1;
    const int64_t pl = (int64_t)NY * NX;       // nodes per layer
    const int64_t ple = (int64_t)gd.ny * gd.nx;  // elements per layer
    const int ei = t % kTE, ej = t / kTE;
    const int64_t W = wlist ? wn : tm.work();
    const int G = vb.g;
    const int nxcd = G < kXCD ? G : kXCD;
    const int xcd = vb.b % nxcd, slotInXcd = vb.b / nxcd;
    const int blocksOnXcd = G / nxcd + (xcd < G % nxcd ? 1 : 0);
    const int64_t wBeg = W * xcd / nxcd, wEnd = W * (xcd + 1) / nxcd;
    const int f00 = 3 * ((ei + 1) + kTN * ej), f10 = 3 * ((ei + 1) + kTN * (ej + 1));
    const int f01 = 3 * (ei + kTN * ej), f11 = 3 * (ei + kTN * (ej + 1));
    (void)f00; (void)f10; (void)f11;
    uint32_t lOff[kStageSlots];
    double dotacc = 0.0;
    // every blocksOnXcd-th item of the XCD's range from slotInXcd: the blocks
    // of one XCD work on neighbouring items at any time (tickets from a
    // per-XCD atomic counter instead balanced the irregular domain's uneven
    // items better, 125 M-DOF ellipsoids FP32 481 -> 466 us, but made the
    // p.q partials depend on timing and cost 36 -> 52 us on 256x128x128)
    for (int64_t w = wBeg + slotInXcd; w < wEnd; w += blocksOnXcd) {
        int64_t rem = wlist ? wlist[w] : w;
        const int tz = (int)(rem / ((int64_t)tm.tilesR * tm.tilesC));
        rem -= (int64_t)tz * tm.tilesR * tm.tilesC;
        const int tc = (int)(rem / tm.tilesR), tr = (int)(rem % tm.tilesR);
        const int R0 = tr * kTO, C0 = tc * kTO, Z0 = tz * tm.cz;
        const int Zend = min(Z0 + tm.cz, NZ);
        const int er = R0 - 1 + ei, ec = C0 - 1 + ej;
        const bool einside = er >= 0 && er < gd.ny && ec >= 0 && ec < gd.nx;
#pragma unroll
        for (int q = 0; q < kStageSlots; ++q) {
            const int j = t + 256 * q, col = j / (3 * kTN), off = j - 3 * kTN * col, row = off / 3;
            const int gr = R0 - 1 + row, gcl = C0 - 1 + col;
            const bool ok = j < 3 * kTN * kTN && gr >= 0 && gr < NY && gcl >= 0 && gcl < NX;
            lOff[q] = ok ? (uint32_t)sizeof(T) * (uint32_t)(3 * NY * col + off) : kOob;
        }
        // this thread's output node (ei, ej) of the staged window, if an
        // output node (ei, ej >= 1) inside the lattice
        const bool onode = ei >= 1 && ej >= 1 && er < NY && ec < NX;
        const uint32_t fOff = onode ? (uint32_t)(ei + NY * ej) : kOob;
        const uint32_t yOff = onode ? (uint32_t)sizeof(T) * (uint32_t)(3 * (ei + NY * ej)) : kOob;
        // staged outputs: value j = t + 256 q of a layer is node (row, col) =
        // ((j % 48) / 3, j / 48) of the 16 x 16 window, component j % 3
        uint32_t oOff[kON / 256];
#pragma unroll
        for (int q = 0; q < kON / 256; ++q) {
            const int j = t + 256 * q, col = j / (3 * kTE), off = j - 3 * kTE * col, row = off / 3;
            const bool ok = row >= 1 && col >= 1 && R0 - 1 + row < NY && C0 - 1 + col < NX;
            oOff[q] = ok ? (uint32_t)sizeof(T) * (uint32_t)(3 * NY * col + off) : kOob;
        }
        const int64_t sBase = 3 * ((int64_t)(R0 - 1) + (int64_t)NY * (C0 - 1));
        const uint32_t vBytes = (uint32_t)sizeof(T) * (uint32_t)(3 * NY * (kTN - 1) + 3 * kTN);
        const uint32_t fBytes = (uint32_t)(NY * (kTN - 1) + kTN);
        auto uRsrc = [&](int L) { return buf_rsrc(u + sBase + 3 * pl * L, (L >= 0 && L < NZ) ? vBytes : 0u); };
        const T *pE = E + (einside ? er + (int64_t)gd.ny * ec + ple * Z0 : 0);  // layer el+1
        T pf[kStageSlots];
        auto loadL = [&](__amdgpu_buffer_rsrc_t r, T (&q)[kStageSlots]) {
#pragma unroll
            for (int k = 0; k < kStageSlots; ++k) q[k] = buf_load<T>(r, lOff[k]);
        };
        auto stage = [&](T *dst, const T (&q)[kStageSlots]) {
#pragma unroll
            for (int k = 0; k < kStageSlots; ++k)
                if (k < kStageSlots - 1 || t + 256 * k < 3 * kTN * kTN) dst[t + 256 * k] = q[k];
        };
        T *sPrev = us[0], *sCur = us[1], *sNext = us[2], *sFree = us[3];
        {
            T p0[kStageSlots], p1[kStageSlots];
            loadL(uRsrc(Z0 - 1), p0);
            loadL(uRsrc(Z0), p1);
            loadL(uRsrc(Z0 + 1), pf);
            __syncthreads();  // previous work item done with LDS
            stage(sPrev, p0);
            stage(sCur, p1);
        }
        __syncthreads();
        T Flo[3][4], Alo[3][4];
        face_modes<T>(sPrev, ei, ej, Flo);
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int q = 0; q < 4; ++q) Alo[a][q] = T(0);
        T Ecur = (einside && Z0 - 1 >= 0) ? pE[-ple] : T(0);
        bool eokCur = true;
        T dotw = T(0);  // this work item's share of u.y, in T (one FMA per value)
        // stores staged node layer L (its values are in ost[L & 1])
        auto flush = [&](int L) {
            const __amdgpu_buffer_rsrc_t yr = buf_rsrc(y + sBase + 3 * pl * L, vBytes);
#pragma unroll
            // (nontemporal stores here measured slower: 125 M-DOF ellipsoids
            // FP32 467 -> 579 us)
            for (int q = 0; q < kON / 256; ++q) buf_store(ost[L & 1][t + 256 * q], yr, oOff[q]);
        };
        auto step = [&](int el, T (&ready)[kStageSlots], T (&spare)[kStageSlots]) {
            const bool out = el >= Z0;
            const __amdgpu_buffer_rsrc_t fr = buf_rsrc(nflag + sBase / 3 + pl * el, out ? fBytes : 0u);
            const uint32_t fl = __builtin_amdgcn_raw_buffer_load_b8(fr, (int)fOff, 0, 0);
            const bool eok = einside && el + 1 < gd.nz;
            const T Enext = *(eok ? pE : E);  // raw; zeroed where it is used
            loadL(uRsrc(el + 3), spare);
            T m[3][8];
            {
                T Fhi[3][4];
                face_modes<T>(sCur, ei, ej, Fhi);
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        m[a][q] = Flo[a][q] + Fhi[a][q];
                        m[a][q + 4] = Fhi[a][q] - Flo[a][q];
                        Flo[a][q] = Fhi[a][q];
                    }
            }
            mode_forces<T>(m, eokCur ? Ecur : T(0), km);
            T left[3] = {T(0), T(0), T(0)};
            T *xw = &xch[el & 1][0][0];
            if (out) {
                // lower face complete: inverse (x, y) transform to its corners
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T A0 = Alo[a][0] - m[a][4], A1 = Alo[a][1] + (m[a][1] - m[a][5]);
                    const T A2 = Alo[a][2] + (m[a][2] - m[a][6]), A3 = Alo[a][3] + (m[a][3] - m[a][7]);
                    const T t0 = A0 - A2, t1 = A1 - A3, t2 = A0 + A2, t3 = A1 + A3;
                    const T cA = t0 - t1, cB = t0 + t1, cC = t2 - t3, cD = t2 + t3;
                    left[a] = cC + row_shr1<T>(cA);
                    xw[256 * a + t] = cD + row_shr1<T>(cB);
                }
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                Alo[a][0] = m[a][4];
#pragma unroll
                for (int q = 1; q < 4; ++q) Alo[a][q] = m[a][q] + m[a][q + 4];
            }
            Ecur = Enext;
            eokCur = eok;
            stage(sNext, ready);  // layer el+2 into the slot of layer el-2
            __syncthreads();      // corner sums of layer el exchanged
            if (StageOut && el - 1 >= Z0) flush(el - 1);  // staged in the previous step
            if (out) {
                const __amdgpu_buffer_rsrc_t yr = buf_rsrc(y + sBase + 3 * pl * el, vBytes);
                const uint32_t keep = (fl & kSolid) ? (~fl & 7u) : 0u;  // components kept (solid, free)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T v0 = left[a] + (ej >= 1 ? xw[256 * a + t - kTE] : T(0));
                    const T v = (keep & (1u << a)) ? v0 : T(0);
                    if constexpr (StageOut) ost[el & 1][3 * t + a] = v;
                    else buf_store(v, yr, yOff + (uint32_t)sizeof(T) * (uint32_t)a);
                    dotw += sPrev[f01 + a] * v;
                }
            }
            T *tmp = sPrev;
            sPrev = sCur;
            sCur = sNext;
            sNext = sFree;
            sFree = tmp;
            pE += ple;
        };
        T pn[kStageSlots];
        for (int el = Z0 - 1; el < Zend; el += 2) {
            step(el, pf, pn);
            if (el + 1 < Zend) step(el + 1, pn, pf);
        }
        if constexpr (StageOut) {
            __syncthreads();  // the last layer staged
            if (Zend - 1 >= Z0) flush(Zend - 1);
        }
        dotacc += (double)dotw;
    }
    if (partial) {
        const double s = block_sum(dotacc);
        if (threadIdx.x == 0) partial[vb.b] = s;
    }
}
template <class T, bool StageOut>
__global__ void __launch_bounds__(256, sizeof(T) == 8 ? 3 : 4) k_kbyu_owner(const T *__restrict__ u, T *__restrict__ y,
                                                       const T *__restrict__ E,
                                                       const uint8_t *__restrict__ nflag, Grid gd,
                                                       TileMap tm, const int32_t *__restrict__ wlist,
                                                       int64_t wn, double *__restrict__ partial,
                                                       const Ctl *ctl, ModeK<T> km) {
    kbyu_owner_body<T, StageOut>(VB::self(), u, y, E, nflag, gd, tm, wlist, wn, partial, ctl, km);
}

// ----------------------------------------------------------------------------
// Restriction r_c = P^T r_f with trilinear P of span S: coarse node N gathers
// fine nodes S N + d, |d_i| < S, weight prod (1 - |d_i|/S).
// ----------------------------------------------------------------------------
template <class T, int S>
__device__ __forceinline__ void restrict_body(VB vb, const T *__restrict__ rf, Grid gf,
                                                  T *__restrict__ rc, Grid gc, NodeDiv dc,
                                                  const uint8_t *__restrict__ cflag, int skipVoid,
                                                  const Ctl *ctl) {
    if (stopped(ctl)) return;
    const int64_t G = (int64_t)vb.b * blockDim.x + threadIdx.x;
    if (G >= gc.nodes()) return;
    if (skipVoid && !(cflag[G] & kSolid)) {  // irregular domains: void coarse nodes read nothing
        rc[3 * G] = rc[3 * G + 1] = rc[3 * G + 2] = T(0);
        return;
    }
    int R, Cc, Lc;
    dc.rcl((uint32_t)G, R, Cc, Lc);
    const int64_t NYf = gf.NY(), NXf = gf.NX(), NZf = gf.NZ();
    // Branch-free stencil: a fine index outside the lattice is clamped and
    // gets weight 0, so the loads issue together instead of one by one.
    T a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
    for (int dl = -(S - 1); dl <= S - 1; ++dl) {
        const int64_t lu = (int64_t)S * Lc + dl;
        const int64_t lf = lu < 0 ? 0 : (lu >= NZf ? NZf - 1 : lu);
        const T wz = (lu == lf) ? T(S - (dl < 0 ? -dl : dl)) / T(S) : T(0);
#pragma unroll
        for (int dc = -(S - 1); dc <= S - 1; ++dc) {
            const int64_t cu = (int64_t)S * Cc + dc;
            const int64_t cf = cu < 0 ? 0 : (cu >= NXf ? NXf - 1 : cu);
            const T wx = (cu == cf) ? wz * T(S - (dc < 0 ? -dc : dc)) / T(S) : T(0);
#pragma unroll
            for (int dr = -(S - 1); dr <= S - 1; ++dr) {
                const int64_t ru = (int64_t)S * R + dr;
                const int64_t rr = ru < 0 ? 0 : (ru >= NYf ? NYf - 1 : ru);
                const T w = (ru == rr) ? wx * T(S - (dr < 0 ? -dr : dr)) / T(S) : T(0);
                const T *src = rf + 3 * (rr + NYf * cf + NYf * NXf * lf);
                a0 += w * src[0];
                a1 += w * src[1];
                a2 += w * src[2];
            }
        }
    }
    const bool solid = cflag[G] & kSolid;
    rc[3 * G] = solid ? a0 : T(0);
    rc[3 * G + 1] = solid ? a1 : T(0);
    rc[3 * G + 2] = solid ? a2 : T(0);
}
template <class T, int S>
__global__ void __launch_bounds__(256) k_restrict(const T *__restrict__ rf, Grid gf,
                                                  T *__restrict__ rc, Grid gc, NodeDiv dc,
                                                  const uint8_t *__restrict__ cflag, int skipVoid,
                                                  const Ctl *ctl) {
    restrict_body<T, S>(VB::self(), rf, gf, rc, gc, dc, cflag, skipVoid, ctl);
}

// ----------------------------------------------------------------------------
// Span-2 restriction, LDS-tiled.  A block owns a 32 x 8 patch of coarse node
// columns and a chunk of coarse layers; fine layers stream through LDS (rows
// 2R0-1 .. 2R0+63, cols 2C0-1 .. 2C0+15, loaded with coalesced row runs and
// prefetched one layer ahead).  Each thread reduces its 3 x 3 in-plane
// neighbourhood (weights 1/2, 1, 1/2 per axis) and folds the result into the
// coarse layers above and below (weight 1 for the even fine layer 2L, 1/2 to
// L and L+1 for 2L+1).
// ----------------------------------------------------------------------------
constexpr int kRR = 32, kRC = 8;                    // coarse rows, cols per block
constexpr int kRFR = 2 * kRR + 1, kRFC = 2 * kRC + 1;  // staged fine rows, cols
constexpr int kRFN = kRFR * kRFC;                    // staged fine nodes per layer
constexpr int kRPer = (kRFN + 255) / 256;            // staged nodes per thread

template <class T>
__device__ __forceinline__ void restrict2_tiled_body(VB vb, const T *__restrict__ rf, Grid gf,
                                                         T *__restrict__ rc, Grid gc,
                                                         const uint8_t *__restrict__ cflag,
                                                         int czc, int tilesR, int tilesC,
                                                         const int32_t *__restrict__ tiles,
                                                         const uint8_t *__restrict__ fflag,
                                                         const Ctl *ctl) {
    if (stopped(ctl)) return;
    __shared__ T sh[2][kRFN * 3];
    const int t = threadIdx.x;
    int bid = tiles ? tiles[vb.b] : (int)vb.b;  // listed tiles: those with a solid coarse node
    const int tr = bid % tilesR;
    bid /= tilesR;
    const int tc = bid % tilesC;
    const int tz = bid / tilesC;
    const int NYf = gf.ny + 1, NXf = gf.nx + 1, NZf = gf.nz + 1;
    const int NYc = gc.ny + 1, NXc = gc.nx + 1, NZc = gc.nz + 1;
    const int64_t plf = (int64_t)NYf * NXf, plc = (int64_t)NYc * NXc;
    const int R0 = tr * kRR, C0 = tc * kRC;
    const int L0 = tz * czc, L1 = min(L0 + czc, NZc);  // coarse layers [L0, L1)
    const int ci = t % kRR, cj = t / kRR;
    const int Rc = R0 + ci, Cc = C0 + cj;
    const bool mine = Rc < NYc && Cc < NXc;
    // staged fine node k -> global offset within a layer (or -1)
    int64_t goff[kRPer];
#pragma unroll
    for (int i = 0; i < kRPer; ++i) {
        const int k = t + 256 * i;
        goff[i] = -1;
        if (k < kRFN) {
            const int r = 2 * R0 - 1 + k % kRFR, c = 2 * C0 - 1 + k / kRFR;
            if (r >= 0 && r < NYf && c >= 0 && c < NXf) goff[i] = r + (int64_t)NYf * c;
        }
    }
    // Loads are unconditional (a valid address for staged nodes outside the
    // lattice) and stay in flight; stash() zeroes the invalid ones.  With
    // conditional loads the compiler waits for the prefetch at each branch.
    // With fflag (irregular domains), a void fine node (its residual is 0)
    // loads from the vector's first entries instead, one line every void node
    // shares, so the tile's void parts move no bytes; its flags are loaded a
    // layer before its residual so that the address does not wait for them.
    T pf[kRPer][3];
    bool pok[kRPer];
    uint8_t fl[kRPer];
    auto fetch_flags = [&](int lf) {
#pragma unroll
        for (int i = 0; i < kRPer; ++i)
            fl[i] = (fflag && goff[i] >= 0 && lf >= 0 && lf < NZf) ? fflag[goff[i] + plf * lf] : (uint8_t)kSolid;
    };
    auto fetch = [&](int lf) {
#pragma unroll
        for (int i = 0; i < kRPer; ++i) {
            pok[i] = goff[i] >= 0 && lf >= 0 && lf < NZf && (fl[i] & kSolid);
            const T *src = pok[i] ? rf + 3 * (goff[i] + plf * lf) : rf;
            pf[i][0] = src[0];
            pf[i][1] = src[1];
            pf[i][2] = src[2];
        }
    };
    auto stash = [&](T *dst) {
#pragma unroll
        for (int i = 0; i < kRPer; ++i) {
            const int k = t + 256 * i;
            if (k < kRFN) {
                dst[3 * k] = pok[i] ? pf[i][0] : T(0);
                dst[3 * k + 1] = pok[i] ? pf[i][1] : T(0);
                dst[3 * k + 2] = pok[i] ? pf[i][2] : T(0);
            }
        }
    };
    const int lfBeg = 2 * L0 - 1, lfEnd = 2 * (L1 - 1) + 1;  // fine layers [lfBeg, lfEnd]
    // flag of the coarse node written next (loaded ahead: used where it is
    // loaded, it cost each coarse layer a memory round trip)
    uint8_t cfl = (mine && L0 < L1) ? cflag[Rc + (int64_t)NYc * Cc + plc * L0] : (uint8_t)0;
    fetch_flags(lfBeg);
    fetch(lfBeg);
    stash(sh[0]);
    fetch_flags(lfBeg + 1);
    fetch(lfBeg + 1);
    fetch_flags(lfBeg + 2);
    __syncthreads();
    T acc[3] = {0, 0, 0}, nxt[3] = {0, 0, 0};
    int buf = 0;
    for (int lf = lfBeg; lf <= lfEnd; ++lf) {
        const T *lay = sh[buf];
        T P[3] = {0, 0, 0};
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc)
#pragma unroll
            for (int dr = -1; dr <= 1; ++dr) {
                const T w = T((dc == 0 ? 2 : 1) * (dr == 0 ? 2 : 1)) * T(0.25);
                const T *src = lay + 3 * ((2 * ci + 1 + dr) + kRFR * (2 * cj + 1 + dc));
                P[0] += w * src[0];
                P[1] += w * src[1];
                P[2] += w * src[2];
            }
        // fine layer lf = 2L (weight 1 to L) or 2L+1 (1/2 to L and L+1)
        const bool even = ((lf - lfBeg) & 1) == 1;  // lfBeg = 2L0-1 is odd
        if (even) {
#pragma unroll
            for (int a = 0; a < 3; ++a) acc[a] += P[a];
        } else {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                acc[a] += T(0.5) * P[a];
                nxt[a] = T(0.5) * P[a];
            }
            const int L = (lf - 1) / 2;  // coarse layer completed by fine layer 2L+1
            if (lf > lfBeg && mine && L >= L0 && L < L1) {
                const int64_t G = Rc + (int64_t)NYc * Cc + plc * L;
                const bool solid = cfl & kSolid;
                if (L + 1 < L1) cfl = cflag[G + plc];  // the next coarse layer's flag, two fine layers ahead
#pragma unroll
                for (int a = 0; a < 3; ++a) rc[3 * G + a] = solid ? acc[a] : T(0);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) acc[a] = nxt[a];
        }
        if (lf < lfEnd) {
            stash(sh[buf ^ 1]);  // layer lf+1 (fetched one iteration ago)
            if (lf + 2 <= lfEnd) {
                fetch(lf + 2);       // with the flags loaded one iteration ago
                fetch_flags(lf + 3);
            }
            __syncthreads();
            buf ^= 1;
        }
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_restrict2_tiled(const T *__restrict__ rf, Grid gf,
                                                         T *__restrict__ rc, Grid gc,
                                                         const uint8_t *__restrict__ cflag,
                                                         int czc, int tilesR, int tilesC,
                                                         const int32_t *__restrict__ tiles,
                                                         const uint8_t *__restrict__ fflag,
                                                         const Ctl *ctl) {
    restrict2_tiled_body<T>(VB::self(), rf, gf, rc, gc, cflag, czc, tilesR, tilesC, tiles, fflag, ctl);
}

// Residual update fused with the first restriction (span 2, LDS-tiled as
// restrict2_tiled_body): r' = r - lambda q with the partials of r'.r' and
// sum dinv r'^2 (Solving_PreconditionedConjugateGradientSolver.m:16-17), and
// r_2 = P^T r' (Solving_Vcycle.m:13 on level 2) from the same staged values,
// so the residual is read once per iteration instead of twice.
// Every tile forms r' at all the fine nodes it stages; it writes back, and
// sums, only the nodes it owns: fine node n belongs to the tile of coarse
// node floor(n / 2) along each axis (the staged row, column and layer below
// the tile's own belong to the neighbours).  A solid fine node's owner coarse
// node is a node of a solid coarse element, so the listed tiles own every
// solid fine node and no void one is written.  Neighbouring tiles read the
// shared nodes' r while their owner writes r', so r' goes to a second vector
// (rOut; the iterations alternate the two).  Persistent over the tiles, one
// pair of partials per block.  dinvL: the finest weights in lattice order.
template <class T>
__device__ __forceinline__ void update_restrict_body(VB vb, const T *__restrict__ rf, const T *__restrict__ q,
                                                     T *__restrict__ rOut,
                                                     const typename FineW<T>::D *__restrict__ dinvL, Grid gf,
                                                     T *__restrict__ rc, Grid gc, const uint8_t *__restrict__ cflag,
                                                     int czc, int tilesR, int tilesC, int64_t nTiles,
                                                     const int32_t *__restrict__ tiles,
                                                     const uint8_t *__restrict__ fflag, double *__restrict__ partRR,
                                                     double *__restrict__ partDR, const Ctl *ctl) {
    if (stopped(ctl)) return;
    __shared__ T sh[2][kRFN * 3];
    const T lt = (T)ctl->lambda;
    const int t = threadIdx.x;
    const int NYf = gf.ny + 1, NXf = gf.nx + 1, NZf = gf.nz + 1;
    const int NYc = gc.ny + 1, NXc = gc.nx + 1, NZc = gc.nz + 1;
    const int64_t plf = (int64_t)NYf * NXf, plc = (int64_t)NYc * NXc;
    const int ci = t % kRR, cj = t / kRR;
    double rr = 0.0, dr = 0.0;
    for (int64_t w = vb.b; w < nTiles; w += vb.g) {
        int bid = tiles ? tiles[w] : (int)w;
        const int tr = bid % tilesR;
        bid /= tilesR;
        const int tc = bid % tilesC;
        const int tz = bid / tilesC;
        const int R0 = tr * kRR, C0 = tc * kRC;
        const int L0 = tz * czc, L1 = min(L0 + czc, NZc);  // coarse layers [L0, L1)
        const int Rc = R0 + ci, Cc = C0 + cj;
        const bool mine = Rc < NYc && Cc < NXc;
        int64_t goff[kRPer];
        bool own[kRPer];
#pragma unroll
        for (int i = 0; i < kRPer; ++i) {
            const int k = t + 256 * i;
            goff[i] = -1;
            own[i] = false;
            if (k < kRFN) {
                const int kr = k % kRFR, kc = k / kRFR;
                const int r = 2 * R0 - 1 + kr, c = 2 * C0 - 1 + kc;
                if (r >= 0 && r < NYf && c >= 0 && c < NXf) goff[i] = r + (int64_t)NYf * c;
                own[i] = kr >= 1 && kc >= 1;
            }
        }
        const int lfBeg = 2 * L0 - 1, lfEnd = 2 * (L1 - 1) + 1;  // fine layers [lfBeg, lfEnd]
        // as restrict2_tiled_body: unconditional loads (void and outside
        // nodes read the vectors' first entries), the flags of a layer loaded
        // one fetch before its values
        uint8_t fl[kRPer];
        auto fetch_flags = [&](int lf) {
#pragma unroll
            for (int i = 0; i < kRPer; ++i)
                fl[i] = (fflag && goff[i] >= 0 && lf >= 0 && lf < NZf) ? fflag[goff[i] + plf * lf] : (uint8_t)kSolid;
        };
        // r, q and, with them (a layer ahead, not where they are used), the
        // weights: loaded in the stash they cost each layer a second memory
        // round trip
        using D = typename FineW<T>::D;
        auto fetch = [&](int lf, T (&v)[kRPer][6], D (&wv)[kRPer][3], bool (&ok)[kRPer]) {
#pragma unroll
            for (int i = 0; i < kRPer; ++i) {
                ok[i] = goff[i] >= 0 && lf >= 0 && lf < NZf && lf <= lfEnd && (fl[i] & kSolid);
                const int64_t o = ok[i] ? 3 * (goff[i] + plf * lf) : 0;
                const int64_t ow = ok[i] && own[i] && lf >= 2 * L0 ? o : 0;  // weights of owned nodes only
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    v[i][a] = rf[o + a];
                    v[i][3 + a] = q[o + a];
                    wv[i][a] = dinvL[ow + a];
                }
            }
        };
        // r' of layer lf (fetched into v) into LDS; the owned nodes to rOut
        auto stash = [&](T *dst, int lf, const T (&v)[kRPer][6], const D (&wv)[kRPer][3], const bool (&ok)[kRPer]) {
            const bool ownLayer = lf >= 2 * L0;  // the layer below the tile's belongs to the tile below
#pragma unroll
            for (int i = 0; i < kRPer; ++i) {
                const int k = t + 256 * i;
                if (k < kRFN) {
                    T u[3];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        u[a] = ok[i] ? v[i][a] - lt * v[i][3 + a] : T(0);
                        dst[3 * k + a] = u[a];
                    }
                    if (ok[i] && own[i] && ownLayer) {
                        const int64_t g = goff[i] + plf * lf;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            // nontemporal in FP32 (525 -> 511 us per launch);
                            // in FP64 they wrote 1.41 GB instead of 1.21 GB
                            // (partial lines) and ran slower (1,029 -> 1,042 us)
                            st_stream<SGLD_NT_STORE_FUSED != 0 && sizeof(T) == 4>(&rOut[3 * g + a], u[a]);
                            const double ud = (double)u[a];
                            rr += ud * ud;
                            dr += (double)FineW<T>::value(wv[i][a]) * ud * ud;
                        }
                    }
                }
            }
        };
        // (a second register set fetching two layers ahead measured slower
        // in FP32: 568 -> 679 us per launch on the 125 M-DOF ellipsoids)
        T vA[kRPer][6];
        D wA[kRPer][3];
        bool okA[kRPer];
        uint8_t cfl = (mine && L0 < L1) ? cflag[Rc + (int64_t)NYc * Cc + plc * L0] : (uint8_t)0;
        fetch_flags(lfBeg);
        fetch(lfBeg, vA, wA, okA);
        stash(sh[0], lfBeg, vA, wA, okA);
        fetch_flags(lfBeg + 1);
        fetch(lfBeg + 1, vA, wA, okA);
        fetch_flags(lfBeg + 2);
        __syncthreads();
        T acc[3] = {0, 0, 0};
        int buf = 0;
        // layer lf from LDS; `cur` holds layer lf + 1, which is stashed here,
        // and then receives layer lf + 2
        auto step = [&](int lf, T (&cur)[kRPer][6], D (&wc)[kRPer][3], bool (&okc)[kRPer]) {
            const T *lay = sh[buf];
            T P[3] = {0, 0, 0};
#pragma unroll
            for (int dc = -1; dc <= 1; ++dc)
#pragma unroll
                for (int dr2 = -1; dr2 <= 1; ++dr2) {
                    const T wgt = T((dc == 0 ? 2 : 1) * (dr2 == 0 ? 2 : 1)) * T(0.25);
                    const T *src = lay + 3 * ((2 * ci + 1 + dr2) + kRFR * (2 * cj + 1 + dc));
                    P[0] += wgt * src[0];
                    P[1] += wgt * src[1];
                    P[2] += wgt * src[2];
                }
            const bool even = ((lf - lfBeg) & 1) == 1;  // lfBeg = 2L0-1 is odd
            if (even) {
#pragma unroll
                for (int a = 0; a < 3; ++a) acc[a] += P[a];
            } else {
                T nx[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    acc[a] += T(0.5) * P[a];
                    nx[a] = T(0.5) * P[a];
                }
                const int L = (lf - 1) / 2;
                if (lf > lfBeg && mine && L >= L0 && L < L1) {
                    const int64_t G = Rc + (int64_t)NYc * Cc + plc * L;
                    const bool solid = cfl & kSolid;
                    if (L + 1 < L1) cfl = cflag[G + plc];  // two fine layers ahead
#pragma unroll
                    for (int a = 0; a < 3; ++a) rc[3 * G + a] = solid ? acc[a] : T(0);
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) acc[a] = nx[a];
            }
            if (lf < lfEnd) {
                stash(sh[buf ^ 1], lf + 1, cur, wc, okc);
                if (lf + 2 <= lfEnd) {
                    fetch(lf + 2, cur, wc, okc);
                    fetch_flags(lf + 3);
                }
                __syncthreads();
                buf ^= 1;
            }
        };
        for (int lf = lfBeg; lf <= lfEnd; ++lf) step(lf, vA, wA, okA);
        __syncthreads();  // the next tile restages both buffers
    }
    block_sum2(rr, dr);
    if (threadIdx.x == 0) {
        partRR[vb.b] = rr;
        partDR[vb.b] = dr;
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_update_restrict(const T *__restrict__ rf, const T *__restrict__ q,
                                                         T *__restrict__ rOut,
                                                         const typename FineW<T>::D *__restrict__ dinvL, Grid gf,
                                                         T *__restrict__ rc, Grid gc, const uint8_t *__restrict__ cflag,
                                                         int czc, int tilesR, int tilesC, int64_t nTiles,
                                                         const int32_t *__restrict__ tiles,
                                                         const uint8_t *__restrict__ fflag,
                                                         double *__restrict__ partRR, double *__restrict__ partDR,
                                                         const Ctl *ctl) {
    update_restrict_body<T>(VB::self(), rf, q, rOut, dinvL, gf, rc, gc, cflag, czc, tilesR, tilesC, nTiles, tiles,
                            fflag, partRR, partDR, ctl);
}

// Component a of (P x_c)(n), span S, at any fine node n (the caller zeroes
// void nodes).  Branch-free: per axis the two coarse neighbours are q and
// q + (m > 0) with weights (S - m)/S and m/S, so all eight loads are in the
// coarse lattice and issue together (with data-dependent skips the compiler
// waits for each gather in turn).
template <class T, int S>
__device__ __forceinline__ T interp_dof(const T *__restrict__ xc, Grid gc, const NodeDiv &df, uint32_t n,
                                        int a) {
    int r, c, l;
    df.rcl(n, r, c, l);
    const int64_t NYc = gc.NY(), NXc = gc.NX();
    const int qr = r / S, mr = r % S, qc = c / S, mc = c % S, ql = l / S, ml = l % S;
    const T *base = xc + 3 * (qr + NYc * qc + NYc * NXc * ql) + a;
    const int64_t dr = 3 * (mr > 0), dc = 3 * NYc * (mc > 0), dl = 3 * NYc * NXc * (ml > 0);
    const T wr1 = T(mr) / T(S), wc1 = T(mc) / T(S), wl1 = T(ml) / T(S);
    const T wr0 = T(1) - wr1, wc0 = T(1) - wc1, wl0 = T(1) - wl1;
    const T v000 = base[0], v100 = base[dr], v010 = base[dc], v110 = base[dc + dr];
    const T v001 = base[dl], v101 = base[dl + dr], v011 = base[dl + dc], v111 = base[dl + dc + dr];
    return wl0 * (wc0 * (wr0 * v000 + wr1 * v100) + wc1 * (wr0 * v010 + wr1 * v110)) +
           wl1 * (wc0 * (wr0 * v001 + wr1 * v101) + wc1 * (wr0 * v011 + wr1 * v111));
}

// ----------------------------------------------------------------------------
// Prolongation + both Jacobi terms of one V-cycle level (Solving_Vcycle.m:35-49):
// x_f = P x_c + dinv_f r_f with dinv = 2 omega / diag; on the finest level the
// fixed DOFs are zeroed (Solving_Vcycle.m:50) and the partial of x_f . r_f is
// produced for the next PCG rho.
// ----------------------------------------------------------------------------
template <class T, int S>
__device__ __forceinline__ void prolong_body(VB vb, const T *__restrict__ xc, Grid gc,
                                                 T *__restrict__ xf, NodeDiv df,
                                                 const T *__restrict__ rf,
                                                 const T *__restrict__ dinv,
                                                 const uint8_t *__restrict__ fflag, int finest, Span dofs,
                                                 double *__restrict__ partial, const Ctl *ctl) {
    if (stopped(ctl)) return;
    const FastDiv d3 = FastDiv::make3();
    double dotacc = 0.0;
    FOR_SPAN_V(i, dofs, vb) {  // one DOF per lane: every vector access is contiguous
        const uint32_t n = d3.div((uint32_t)i);
        const int a = (int)((uint32_t)i - 3 * n);
        const T pv = interp_dof<T, S>(xc, gc, df, n, a);
        const T sv = dinv[i] * rf[i];
        const uint8_t f = fflag[n];
        const T val = ((f & kSolid) && !(finest && (f & (1 << a)))) ? pv + sv : T(0);
        xf[i] = val;
        if (partial) dotacc += (double)val * (double)rf[i];
    }
    if (partial) {
        const double s = block_sum(dotacc);
        if (threadIdx.x == 0) partial[vb.b] = s;
    }
}
template <class T, int S>
__global__ void __launch_bounds__(256) k_prolong(const T *__restrict__ xc, Grid gc,
                                                 T *__restrict__ xf, NodeDiv df,
                                                 const T *__restrict__ rf,
                                                 const T *__restrict__ dinv,
                                                 const uint8_t *__restrict__ fflag, int finest, Span dofs,
                                                 double *__restrict__ partial, const Ctl *ctl) {
    prolong_body<T, S>(VB::self(), xc, gc, xf, df, rf, dinv, fflag, finest, dofs, partial, ctl);
}

// The same pass with each accelerator complex die (XCD, one L2 each; block b
// runs on XCD b mod 8) owning one contiguous eighth of the span, in 64-entry
// units, so that values gathered from a shared coarser vector are fetched
// into one L2 rather than into all eight.  Positions of one wavefront still
// share t >> 6 (what Span::at needs for 64-entry chunks).  Needs the grid to
// be a multiple of 8.
#define FOR_SPAN_XCD_VT(i, t, S, vb)                                                                    \
    for (int64_t u_ = ((S).count + 63) >> 6, e_ = min<int64_t>((S).count, (u_ * ((vb).b % 8 + 1) / 8) << 6), \
                 t = ((u_ * ((vb).b % 8) / 8) << 6) + (int64_t)((vb).b / 8) * blockDim.x + threadIdx.x;  \
         t < e_; t += (int64_t)((vb).g / 8) * blockDim.x)                                             \
        if (const int64_t i = (S).at(t); i < (S).n)

// ----------------------------------------------------------------------------
// Finest-level V-cycle output folded into the PCG direction update
// (Solving_Vcycle.m:35-50 followed by
// Solving_PreconditionedConjugateGradientSolver.m:28 and :15): per node
//   z = mask(P x_2 + dinv r),  x += lambda p,  p = z + beta p   (in place)
// with lambda of the previous iteration (its y += lambda p, deferred to here
// so that p is read once) and beta = rho'/rho of this one.  z is never
// stored: rho' = z.r is formed without it as x_2.r_2 + sum dinv r^2, because
// P^T = R, r_2 = R r and r = 0 on fixed DOFs.
// ----------------------------------------------------------------------------
template <class T, int S, bool NT>
__device__ __forceinline__ void prolong_dir_body(VB vb, const T *__restrict__ xc, Grid gc, NodeDiv df,
                                                     const T *__restrict__ rf,
                                                     const typename FineW<T>::D *__restrict__ dinv,
                                                     const uint8_t *__restrict__ fflag, T *__restrict__ x,
                                                     T *__restrict__ p, Span dofs, int xsplit, const Ctl *ctl) {
    if (stopped(ctl)) return;
    const T lam = (T)ctl->lambda, beta = (T)ctl->beta;
    const FastDiv d3 = FastDiv::make3();
    // one DOF per lane: every vector access of a wavefront is contiguous;
    // one contiguous range per XCD keeps the coarse gathers in its L2
    // (125 M-DOF ellipsoids, FP32: 4.12 GB fetched + written per launch with
    // blocks interleaved over the XCDs, against 2.81 GB algorithmic)
    auto body = [&](int64_t i, int64_t t) {
        const int64_t k = fine_pos<T>(i, t);  // dinv and x (span order in FP32)
        const uint32_t n = d3.div((uint32_t)i);
        const int a = (int)((uint32_t)i - 3 * n);
        const T pv = interp_dof<T, S>(xc, gc, df, n, a);
        const T sv = FineW<T>::template get<NT>(dinv, k) * ld_stream<NT>(&rf[i]);
        const uint8_t f = fflag[n];
        const T po = ld_stream<NT>(&p[i]);
        const T z = ((f & kSolid) && !(f & (1 << a))) ? pv + sv : T(0);
        st_stream<NT>(&x[k], ld_stream<NT>(&x[k]) + lam * po);
        st_stream<NT>(&p[i], z + beta * po);
    };
    if (xsplit) {
        FOR_SPAN_XCD_VT(i, t, dofs, vb) body(i, t);
    } else {
        FOR_SPAN_VT(i, t, dofs, vb) body(i, t);
    }
}
template <class T, int S, bool NT>
__global__ void __launch_bounds__(256) k_prolong_dir(const T *__restrict__ xc, Grid gc, NodeDiv df,
                                                     const T *__restrict__ rf,
                                                     const typename FineW<T>::D *__restrict__ dinv,
                                                     const uint8_t *__restrict__ fflag, T *__restrict__ x,
                                                     T *__restrict__ p, Span dofs, int xsplit, const Ctl *ctl) {
    prolong_dir_body<T, S, NT>(VB::self(), xc, gc, df, rf, dinv, fflag, x, p, dofs, xsplit, ctl);
}

// The same pass with the coarse values staged in LDS (span 2, used when the
// coarse column fits): one work item is the pair of fine columns (2k, 2k+1)
// of node layer l, a contiguous DOF range; its interpolation reads coarse
// columns k, k+1 of coarse layers l/2, l/2 + (l odd), which are copied to LDS
// once (4 contiguous runs of 3 NYc values) instead of being gathered from the
// vector cache eight times per DOF.  Clamped runs carry weight 0.
template <class T>
__device__ __forceinline__ void prolong_dir_cols_body(VB vb, const T *__restrict__ xc, Grid gc, Grid gf,
                                                          const T *__restrict__ rf,
                                                          const typename FineW<T>::D *__restrict__ dinv,
                                                          const uint8_t *__restrict__ fflag, T *__restrict__ x,
                                                          T *__restrict__ p, const int32_t *__restrict__ items,
                                                          int64_t nItems, const Ctl *ctl) {
    extern __shared__ unsigned char smem_raw[];
    T *cs = reinterpret_cast<T *>(smem_raw);  // [layer][col][3 NYc]
    if (stopped(ctl)) return;
    const T lam = (T)ctl->lambda, beta = (T)ctl->beta;
    const int NYf = gf.ny + 1, NXf = gf.nx + 1;
    const int NYc = gc.ny + 1, NXc = gc.nx + 1, NZc = gc.nz + 1;
    const int ppl = (NXf + 1) / 2, runLen = 3 * NYc, colLen = 3 * NYf;
    for (int64_t w = vb.b; w < nItems; w += vb.g) {
        const int item = items ? items[w] : (int)w;
        const int l = item / ppl, k = item - l * ppl;
        const int c0 = 2 * k, ncol = (2 * k + 1 < NXf) ? 2 : 1;
        const int ql = l >> 1, ml = l & 1;
        __syncthreads();  // the previous item is done with LDS
        for (int t = threadIdx.x; t < 4 * runLen; t += blockDim.x) {
            const int run = t / runLen, off = t - run * runLen;
            const int cc = min(k + (run & 1), NXc - 1), ll = min(ql + (run >> 1), NZc - 1);
            cs[t] = xc[3 * ((int64_t)NYc * (cc + (int64_t)NXc * ll)) + off];
        }
        __syncthreads();
        const int64_t d0 = 3 * ((int64_t)NYf * (c0 + (int64_t)NXf * l));
        const T wl1 = T(ml) * T(0.5), wl0 = T(1) - wl1;
        for (int t = threadIdx.x; t < ncol * colLen; t += blockDim.x) {
            const int cl = t >= colLen ? 1 : 0;  // fine column c0 + cl; odd columns sit between coarse columns
            const int rem = t - cl * colLen, r = rem / 3, a = rem - 3 * r;
            const int qr = r >> 1, mr = r & 1;
            const T wr1 = T(mr) * T(0.5), wr0 = T(1) - wr1, wc1 = T(cl) * T(0.5), wc0 = T(1) - wc1;
            const T *b0 = cs + 3 * qr + a;  // run (layer 0, col 0)
            const int dr = 3 * mr;
            const T v = wl0 * (wc0 * (wr0 * b0[0] + wr1 * b0[dr]) + wc1 * (wr0 * b0[runLen] + wr1 * b0[runLen + dr])) +
                        wl1 * (wc0 * (wr0 * b0[2 * runLen] + wr1 * b0[2 * runLen + dr]) +
                               wc1 * (wr0 * b0[3 * runLen] + wr1 * b0[3 * runLen + dr]));
            const int64_t i = d0 + t;
            const uint8_t f = fflag[(d0 / 3) + cl * NYf + r];
            const T sv = FineW<T>::get(dinv, i) * rf[i];
            const T z = ((f & kSolid) && !(f & (1 << a))) ? v + sv : T(0);
            const T po = p[i];
            x[i] = x[i] + lam * po;
            p[i] = z + beta * po;
        }
    }
}
template <class T>
__global__ void __launch_bounds__(256) k_prolong_dir_cols(const T *__restrict__ xc, Grid gc, Grid gf,
                                                          const T *__restrict__ rf,
                                                          const typename FineW<T>::D *__restrict__ dinv,
                                                          const uint8_t *__restrict__ fflag, T *__restrict__ x,
                                                          T *__restrict__ p, const int32_t *__restrict__ items,
                                                          int64_t nItems, const Ctl *ctl) {
    prolong_dir_cols_body<T>(VB::self(), xc, gc, gf, rf, dinv, fflag, x, p, items, nItems, ctl);
}

// Sum over the block returned to every thread: the lane-0 wave sums are
// added by thread 0 in wave order and broadcast, so every block summing the
// same values with the same mapping gets the same bits.
__device__ __forceinline__ double block_allsum(double v) {
    __shared__ double sh[17];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) t += sh[i];
        sh[16] = t;
    }
    __syncthreads();
    return sh[16];
}

// ||r||^2 and ||f||^2 of a reliable update from their per-block partials, in
// one fixed order (a 1024-thread block): out[0], out[1] in thread 0.
__device__ __forceinline__ void sum2_partials(const double *__restrict__ partRR, const double *__restrict__ partBB,
                                              int nb, double &rr, double &bb) {
    rr = 0.0;
    bb = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        rr += partRR[i];
        bb += partBB[i];
    }
    block_sum2(rr, bb);
}
__global__ void __launch_bounds__(1024) k_sum2(const double *__restrict__ partRR, const double *__restrict__ partBB,
                                               int nb, double *__restrict__ out) {
    double rr, bb;
    sum2_partials(partRR, partBB, nb, rr, bb);
    if (threadIdx.x == 0) {
        out[0] = rr;
        out[1] = bb;
    }
}

// Solving_PreconditionedConjugateGradientSolver.m:16-17: r -= lambda q with
// partials of r.r (stopping test) and of sum dinv r^2 (the fine smoother's
// share of the next rho).
// lambda = rho / p.q comes from the control block (k_fin FIN_PQ of the K.u
// partials, one block: summing them in every block of this pass cost 8 KB
// of reads per block, about 40 us per merged launch over 32 instances).
template <class T, bool NT>
__device__ __forceinline__ void update_r_body(VB vb, T *__restrict__ r, const T *__restrict__ q,
                                                  const typename FineW<T>::D *__restrict__ dinv, Span S,
                                                  double *__restrict__ partRR, double *__restrict__ partDR,
                                                  Ctl *ctl) {
    if (stopped(ctl)) return;
    const T lt = (T)ctl->lambda;
    double rr = 0.0, dr = 0.0;
    FOR_SPAN_VT(i, t, S, vb) {
        const T rn = ld_stream<NT>(&r[i]) - lt * ld_stream<NT>(&q[i]);
        st_stream<NT>(&r[i], rn);
        const double rd = (double)rn;
        rr += rd * rd;
        dr += (double)FineW<T>::template get<NT>(dinv, fine_pos<T>(i, t)) * rd * rd;
    }
    block_sum2(rr, dr);
    if (threadIdx.x == 0) {
        partRR[vb.b] = rr;
        partDR[vb.b] = dr;
    }
}
template <class T, bool NT>
__global__ void __launch_bounds__(256) k_update_r(T *__restrict__ r, const T *__restrict__ q,
                                                  const typename FineW<T>::D *__restrict__ dinv, Span S,
                                                  double *__restrict__ partRR, double *__restrict__ partDR,
                                                  Ctl *ctl) {
    update_r_body<T, NT>(VB::self(), r, q, dinv, S, partRR, partDR, ctl);
}

// partial of sum dinv r^2 (rho of a residual that k_update_r did not produce)
template <class T>
__global__ void __launch_bounds__(256) k_dinv_rr(const T *__restrict__ r, const typename FineW<T>::D *__restrict__ dinv,
                                                 Span S, double *__restrict__ partial) {
    double acc = 0.0;
    FOR_SPAN_T(i, t, S) {
        const double rd = (double)r[i];
        acc += (double)FineW<T>::get(dinv, fine_pos<T>(i, t)) * rd * rd;
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// r = b - q on free DOFs, 0 on fixed DOFs (DOF i = 3 node + a).  The
// reference's loads never sit on fixed DOFs (FEA_ApplyBoundaryCondition.m:9-10
// releases a loaded node's fixation), and K.u is zero there, so r = 0 on
// fixed DOFs from then on.
__global__ void k_residual_masked(const double *__restrict__ b, const double *__restrict__ q,
                                  double *__restrict__ r, const uint8_t *__restrict__ nflag, Span S) {
    FOR_SPAN(i, S) {
        const uint32_t node = (uint32_t)(i / 3), a = (uint32_t)(i - 3 * (int64_t)node);
        r[i] = (nflag[node] & (1u << a)) ? 0.0 : b[i] - q[i];
    }
}

// Reliable update of the mixed mode: x += scale (x32 + lambda p32), x32 = 0
// (the FP32 correction, in span order, with the stopped iteration's deferred
// lambda p32, folded into the FP64 solution).
__global__ void k_fold_mixed(double *__restrict__ x, float *__restrict__ x32, const float *__restrict__ p32,
                             double scale, Span S, const Ctl *ctl) {
    const float lam = (float)ctl->lambda;
    FOR_SPAN_T(i, t, S) {
        x[i] += scale * (double)(x32[t] + lam * p32[i]);
        x32[t] = 0.0f;
    }
}

// r = mask(b - q) (zero on fixed DOFs, as k_residual_masked) with partials of
// r.r into partRR and b.b into partBB.
__global__ void __launch_bounds__(256) k_residual_norms(const double *__restrict__ b, const double *__restrict__ q,
                                                        double *__restrict__ r, const uint8_t *__restrict__ nflag,
                                                        Span S, double *__restrict__ partRR,
                                                        double *__restrict__ partBB) {
    double rr = 0.0, bb = 0.0;
    FOR_SPAN(i, S) {
        const uint32_t node = (uint32_t)(i / 3), a = (uint32_t)(i - 3 * (int64_t)node);
        const double bi = b[i];
        const double ri = (nflag[node] & (1u << a)) ? 0.0 : bi - q[i];
        r[i] = ri;
        rr += ri * ri;
        bb += bi * bi;
    }
    block_sum2(rr, bb);
    if (threadIdx.x == 0) {
        partRR[blockIdx.x] = rr;
        partBB[blockIdx.x] = bb;
    }
}

// The same for a reliable update of the mixed solve: the residual goes
// straight to the FP32 vector, scaled by invScale (the solve's fixed power of
// two), and only its norm is taken in FP64.
__global__ void __launch_bounds__(256) k_residual_norms32(const double *__restrict__ b, const double *__restrict__ q,
                                                          float *__restrict__ r32, double invScale,
                                                          const uint8_t *__restrict__ nflag, Span S,
                                                          double *__restrict__ partRR, double *__restrict__ partBB) {
    double rr = 0.0, bb = 0.0;
    FOR_SPAN(i, S) {
        const uint32_t node = (uint32_t)(i / 3), a = (uint32_t)(i - 3 * (int64_t)node);
        const double bi = b[i];
        const double ri = (nflag[node] & (1u << a)) ? 0.0 : bi - q[i];
        r32[i] = (float)(ri * invScale);
        rr += ri * ri;
        bb += bi * bi;
    }
    block_sum2(rr, bb);
    if (threadIdx.x == 0) {
        partRR[blockIdx.x] = rr;
        partBB[blockIdx.x] = bb;
    }
}

// x += lambda p with lambda of the last executed iteration (the deferred
// update of a solve that stopped, or before a reliable update).
template <class T>
__global__ void k_axpy_lambda(T *__restrict__ x, const T *__restrict__ p, Span S, const Ctl *ctl) {
    const T lam = (T)ctl->lambda;
    FOR_SPAN(i, S) x[i] = x[i] + lam * p[i];
}

// ----------------------------------------------------------------------------
// Small-level tail of the V-cycle in two launches.  Below level t0 the
// restrictions and prolongations are latency-bound (a few microseconds per
// launch whatever the size), so they are collapsed: linear interpolation on
// nested grids composes exactly, P(t0 <- t0+1) ... P(t0+m-1 <- t0+m) is the
// trilinear interpolation of span 2^m, and R = P^T.  Hence
//   r_{t0+m} = R_{2^m} r_{t0}                                    (m = 1..M)
//   x_{t0}   = D_{t0} r_{t0} + sum_{m<M} P_{2^m} D_{t0+m} r_{t0+m} + P_{2^M} x_{t0+M}
// which equals Solving_Vcycle.m:12-26 and :35-49 on these levels (D = 2 omega
// / diag; void coarse nodes: a solid node's stencil touches only corners of
// non-empty coarse elements, and void entries are zero).
// ----------------------------------------------------------------------------
constexpr int kTailMax = 3;  // fused levels below t0: spans 2, 4, 8

template <class T>
struct TailArgs {
    int M;                          // fused levels t0+1 .. t0+M (t0+M is the coarsest)
    Grid g0;                        // level t0
    NodeDiv d0;
    const T *r0, *dinv0;
    const uint8_t *flag0;
    T *x0;
    Grid g[kTailMax];
    NodeDiv dv[kTailMax];
    T *r[kTailMax];
    const T *dinv[kTailMax];
    const uint8_t *flag[kTailMax];
    const T *xc;                    // x of the coarsest level
    int blk[kTailMax + 1];          // restriction blocks: level t0+1+m uses [blk[m], blk[m+1])
};

// lanes per target node of level t0+1+m: one thread (27 taps), one wavefront
// (343 taps), one block (3375 taps)
__host__ __device__ constexpr int tail_lanes(int m) { return m == 0 ? 1 : (m == 1 ? 64 : 256); }

// r of one node of level t0+1+m (span S = 2^(m+1) from level t0), gathered by
// P lanes; spans and lane counts are compile-time so the tap decode is cheap.
template <class T, int S, int P>
__device__ __forceinline__ void tail_restrict_node(const TailArgs<T> &A, int m, double (*sh)[4], unsigned vbb) {
    constexpr int Wd = 2 * S - 1, taps = Wd * Wd * Wd;
    const int lane = threadIdx.x % P;
    const int64_t G = (int64_t)(vbb - A.blk[m]) * (256 / P) + threadIdx.x / P;
    const Grid gc = A.g[m];
    const bool live = G < gc.nodes();
    int R = 0, Cc = 0, Lc = 0;
    if (live) A.dv[m].rcl((uint32_t)G, R, Cc, Lc);
    const int64_t NYf = A.g0.NY(), NXf = A.g0.NX(), NZf = A.g0.NZ();
    T a0 = 0, a1 = 0, a2 = 0;
    if (live) {
        // branch-free taps: indices clamped into the lattice, weight 0 outside
#pragma unroll
        for (int k0 = 0; k0 < taps; k0 += P) {
            const int k = k0 + lane;
            const int kk = k < taps ? k : taps - 1;
            const int dr = kk % Wd - (S - 1), dc = (kk / Wd) % Wd - (S - 1), dl = kk / (Wd * Wd) - (S - 1);
            const int64_t ru = (int64_t)S * R + dr, cu = (int64_t)S * Cc + dc, lu = (int64_t)S * Lc + dl;
            const int64_t rr = ru < 0 ? 0 : (ru >= NYf ? NYf - 1 : ru);
            const int64_t cf = cu < 0 ? 0 : (cu >= NXf ? NXf - 1 : cu);
            const int64_t lf = lu < 0 ? 0 : (lu >= NZf ? NZf - 1 : lu);
            const bool in = k < taps && ru == rr && cu == cf && lu == lf;
            const T w = in ? T((S - (dr < 0 ? -dr : dr)) * (S - (dc < 0 ? -dc : dc)) * (S - (dl < 0 ? -dl : dl))) /
                                 T(S * S * S)
                           : T(0);
            const T *src = A.r0 + 3 * (rr + NYf * cf + NYf * NXf * lf);
            a0 += w * src[0];
            a1 += w * src[1];
            a2 += w * src[2];
        }
    }
    double s0 = a0, s1 = a1, s2 = a2;
    if (P >= 64) {
        for (int o = 32; o > 0; o >>= 1) {
            s0 += __shfl_down(s0, o, 64);
            s1 += __shfl_down(s1, o, 64);
            s2 += __shfl_down(s2, o, 64);
        }
    }
    if (P == 256) {
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            sh[0][w] = s0;
            sh[1][w] = s1;
            sh[2][w] = s2;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s0 = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
            s1 = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
            s2 = sh[2][0] + sh[2][1] + sh[2][2] + sh[2][3];
        }
    }
    if (live && lane == 0) {
        const bool solid = A.flag[m][G] & kSolid;
        A.r[m][3 * G] = solid ? (T)s0 : T(0);
        A.r[m][3 * G + 1] = solid ? (T)s1 : T(0);
        A.r[m][3 * G + 2] = solid ? (T)s2 : T(0);
    }
}

template <class T>
__device__ __forceinline__ void tail_restrict_body(VB vb, TailArgs<T> A, const Ctl *ctl) {
    __shared__ double sh[3][4];
    if (stopped(ctl)) return;
    int m = 0;
    while (m + 1 < A.M && (int)vb.b >= A.blk[m + 1]) ++m;  // uniform per block
    if (m == 0) tail_restrict_node<T, 2, tail_lanes(0)>(A, 0, sh, vb.b);
    else if (m == 1) tail_restrict_node<T, 4, tail_lanes(1)>(A, 1, sh, vb.b);
    else tail_restrict_node<T, 8, tail_lanes(2)>(A, 2, sh, vb.b);
}
template <class T>
__global__ void __launch_bounds__(256) k_tail_restrict(TailArgs<T> A, const Ctl *ctl) {
    tail_restrict_body<T>(VB::self(), A, ctl);
}

// component a of sum_corners w(n) v[G] for the span-S interpolation from
// level grid gc at level-t0 node (r, c, l); `dv` null: plain v, else dv * v.
template <class T, int S>
__device__ __forceinline__ T tail_interp(const T *__restrict__ v, const T *__restrict__ dv, Grid gc, int r,
                                         int c, int l, int a) {
    const int64_t NYc = gc.NY(), NXc = gc.NX();
    const int qr = r / S, mr = r % S, qc = c / S, mc = c % S, ql = l / S, ml = l % S;
    const int64_t base = 3 * (qr + NYc * qc + NYc * NXc * ql) + a;
    const int64_t dr = 3 * (mr > 0), dc = 3 * NYc * (mc > 0), dl = 3 * NYc * NXc * (ml > 0);
    const T wr1 = T(mr) / T(S), wc1 = T(mc) / T(S), wl1 = T(ml) / T(S);
    const T wr0 = T(1) - wr1, wc0 = T(1) - wc1, wl0 = T(1) - wl1;
    auto at = [&](int64_t o) { return dv ? dv[base + o] * v[base + o] : v[base + o]; };
    return wl0 * (wc0 * (wr0 * at(0) + wr1 * at(dr)) + wc1 * (wr0 * at(dc) + wr1 * at(dc + dr))) +
           wl1 * (wc0 * (wr0 * at(dl) + wr1 * at(dl + dr)) + wc1 * (wr0 * at(dl + dc) + wr1 * at(dl + dc + dr)));
}

template <class T, int M>
__device__ __forceinline__ void tail_prolong_body(VB vb, TailArgs<T> A, double *__restrict__ partial, const Ctl *ctl) {
    if (stopped(ctl)) return;
    const FastDiv d3 = FastDiv::make3();
    double dotacc = 0.0;
    FOR_SPAN_V(i, Span::all(3 * A.g0.nodes()), vb) {
        const uint32_t n = d3.div((uint32_t)i);
        const int a = (int)((uint32_t)i - 3 * n);
        int r, c, l;
        A.d0.rcl(n, r, c, l);
        T v = A.dinv0[i] * A.r0[i];
        if (M > 1) v += tail_interp<T, 2>(A.r[0], A.dinv[0], A.g[0], r, c, l, a);
        if (M > 2) v += tail_interp<T, 4>(A.r[1], A.dinv[1], A.g[1], r, c, l, a);
        v += tail_interp<T, (2 << (M - 1))>(A.xc, (const T *)nullptr, A.g[M - 1], r, c, l, a);
        const T val = (A.flag0[n] & kSolid) ? v : T(0);
        A.x0[i] = val;
        if (partial) dotacc += (double)val * (double)A.r0[i];
    }
    if (partial) {
        const double s = block_sum(dotacc);
        if (threadIdx.x == 0) partial[vb.b] = s;
    }
}
template <class T, int M>
__global__ void __launch_bounds__(256) k_tail_prolong(TailArgs<T> A, double *__restrict__ partial, const Ctl *ctl) {
    tail_prolong_body<T, M>(VB::self(), A, partial, ctl);
}

// Coarsest level (Solving_Vcycle.m:29-31): x(free) = L^-T L^-1 r(free) with
// K_L(free, free) = L L^T, applied as two triangular products with the
// inverse factor W = L^-1 (the device-parallel form of the two triangular
// solves; the applied operator W^T W stays symmetric positive definite in
// floating point).  Accumulated in FP64; W in FP64, or for the FP32 inner
// solve an FP32 copy once W outgrows the Infinity Cache (the preconditioner
// W32^T W32 is as symmetric positive definite as W^T W, and the product reads
// half the bytes).  One wavefront per row.
// y = W r(free): row i of W holds entries 0..i.
// With Lds (n <= kCoarseLds), a block first gathers the entries r(free)
// its rows need into LDS, so the dot products read W and LDS only (the
// free-DOF gather inside the loop made each step two dependent loads).
constexpr int64_t kCoarseLds = 4096;
template <class T, class TW, bool Lds>
__device__ __forceinline__ void coarse_lower_body(VB vb, const TW *__restrict__ Wrow,
                                                      const int32_t *__restrict__ freeIdx, int64_t n,
                                                      const T *__restrict__ r, double *__restrict__ y,
                                                      const Ctl *ctl) {
    if (stopped(ctl)) return;
    const int64_t row = (int64_t)vb.b * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    __shared__ double rs[Lds ? kCoarseLds : 1];
    if constexpr (Lds) {
        const int64_t jEnd = min<int64_t>(n, (int64_t)(vb.b + 1) * (blockDim.x >> 6));
        for (int64_t j = threadIdx.x; j < jEnd; j += blockDim.x) rs[j] = (double)r[freeIdx[j]];
        __syncthreads();
    }
    if (row >= n) return;
    const TW *w = Wrow + tri_row_off(row);
    double acc = 0.0;
    for (int64_t j = lane; j <= row; j += 64)
        acc += (double)w[j] * (Lds ? rs[j] : (double)r[freeIdx[j]]);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) y[row] = acc;
}
template <class T, class TW, bool Lds>
__global__ void __launch_bounds__(256) k_coarse_lower(const TW *__restrict__ Wrow,
                                                      const int32_t *__restrict__ freeIdx, int64_t n,
                                                      const T *__restrict__ r, double *__restrict__ y,
                                                      const Ctl *ctl) {
    coarse_lower_body<T, TW, Lds>(VB::self(), Wrow, freeIdx, n, r, y, ctl);
}
// x(free) = W^T y: row j of W^T (column j of W) holds entries j..n-1.  Fixed
// and void DOFs of x are never written (zero since the assembly).
template <class T, class TW>
__device__ __forceinline__ void coarse_upper_body(VB vb, const TW *__restrict__ Wcol,
                                                      const int32_t *__restrict__ freeIdx, int64_t n,
                                                      const double *__restrict__ y, T *__restrict__ x,
                                                      const Ctl *ctl) {
    if (stopped(ctl)) return;
    const int64_t j = (int64_t)vb.b * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    const TW *w = Wcol + tri_col_off(j, n);
    double acc = 0.0;
    for (int64_t t = lane; t < n - j; t += 64) acc += (double)w[t] * y[j + t];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) x[freeIdx[j]] = (T)acc;
}
template <class T, class TW>
__global__ void __launch_bounds__(256) k_coarse_upper(const TW *__restrict__ Wcol,
                                                      const int32_t *__restrict__ freeIdx, int64_t n,
                                                      const double *__restrict__ y, T *__restrict__ x,
                                                      const Ctl *ctl) {
    coarse_upper_body<T, TW>(VB::self(), Wcol, freeIdx, n, y, x, ctl);
}

// ----------------------------- PCG vector kernels ---------------------------
template <class T>
__device__ __forceinline__ void dot_body(VB vb, const T *__restrict__ a, const T *__restrict__ b, Span S,
                                             double *__restrict__ partial, const Ctl *ctl) {
    if (stopped(ctl)) return;
    double acc = 0.0;
    FOR_SPAN_V(i, S, vb) acc += (double)a[i] * (double)b[i];
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[vb.b] = s;
}
template <class T>
__global__ void __launch_bounds__(256) k_dot(const T *__restrict__ a, const T *__restrict__ b, Span S,
                                             double *__restrict__ partial, const Ctl *ctl) {
    dot_body<T>(VB::self(), a, b, S, partial, ctl);
}

// r = b - q (q may be null: r = b).
template <class T>
__global__ void k_residual(const T *__restrict__ b, const T *__restrict__ q, T *__restrict__ r, Span S) {
    FOR_SPAN(i, S) r[i] = q ? b[i] - q[i] : b[i];
}

// Solving_PreconditionedConjugateGradientSolver.m:15-18:
// y += lambda p; r -= lambda z; partial r.r.
template <class T>
__global__ void __launch_bounds__(256) k_update_xr(T *__restrict__ x, T *__restrict__ r,
                                                   const T *__restrict__ p,
                                                   const T *__restrict__ q, Span S,
                                                   double *__restrict__ partial, const Ctl *ctl) {
    if (stopped(ctl)) return;
    const double lam = ctl->lambda;
    const T lt = (T)lam;
    double acc = 0.0;
    FOR_SPAN(i, S) {
        x[i] = x[i] + lt * p[i];
        const T rn = r[i] - lt * q[i];
        r[i] = rn;
        acc += (double)rn * (double)rn;
    }
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// :28 p = z + (x2/x1) p, into the other p buffer
template <class T>
__global__ void k_update_p_into(T *__restrict__ pnew, const T *__restrict__ z,
                                const T *__restrict__ pold, Span S, const Ctl *ctl) {
    if (stopped(ctl)) return;
    const T beta = (T)ctl->beta;
    FOR_SPAN(i, S) pnew[i] = z[i] + beta * pold[i];
}

template <class T>
__global__ void k_copy(T *__restrict__ dst, const T *__restrict__ src, Span S) {
    FOR_SPAN(i, S) dst[i] = src[i];
}

// dst = (To)(src * scale)
template <class To, class Ti>
__global__ void k_convert(To *__restrict__ dst, const Ti *__restrict__ src, double scale, Span S) {
    FOR_SPAN(i, S) dst[i] = (To)((double)src[i] * scale);
}

template <class T>
__global__ void k_fill(T *__restrict__ v, T val, Span S) {
    FOR_SPAN(i, S) v[i] = val;
}

enum FinOp { FIN_NORMB = 0, FIN_RHO0, FIN_RR, FIN_RHO, FIN_PQ };

// Second reduction stage (fixed order, deterministic) and the scalar updates
// of the PCG recurrence.
__device__ __forceinline__ void fin_body(const double *__restrict__ partial, int nb,
                                         const double *__restrict__ partial2, int nb2, int op, Ctl *ctl) {
    if (op != FIN_NORMB && op != FIN_RHO0 && ctl->done) return;
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[i];
    for (int i = threadIdx.x; i < nb2; i += blockDim.x) acc += partial2[i];
    const double s = block_sum(acc);
    if (threadIdx.x != 0) return;
    switch (op) {
        case FIN_NORMB: ctl->normb = sqrt(s); break;
        case FIN_RHO0:
            ctl->rho = s;
            ctl->beta = 0.0;
            break;
        case FIN_RR:
            ctl->rr = s;
            ctl->res = sqrt(s) / ctl->normb;
            ctl->its += 1;
            if (ctl->res < ctl->tol && (ctl->flags & kCtlSoft)) ctl->flags |= kCtlWant;
            if (ctl->res < ctl->tol && !(ctl->flags & kCtlSoft)) ctl->done = 1;
            else if (ctl->its >= ctl->maxit) ctl->done = 2;
            break;
        case FIN_RHO:
            ctl->beta = s / ctl->rho;
            ctl->rho = s;
            break;
        case FIN_PQ:  // Solving_PreconditionedConjugateGradientSolver.m:15
            ctl->pq = s;
            ctl->lambda = ctl->rho / s;
            break;
    }
}
__global__ void __launch_bounds__(1024) k_fin(const double *__restrict__ partial, int nb,
                                              const double *__restrict__ partial2, int nb2, int op,
                                              Ctl *ctl) {
    fin_body(partial, nb, partial2, nb2, op, ctl);
}

// ------------------------- multi-instance launches --------------------------
// One launch of an iteration kernel over many independent solves: instance i
// runs virtual blocks [start[i], start[i + 1]) of the grid with its own
// arguments, as its own launch of start[i + 1] - start[i] blocks would (same
// block indices, so per-block partial sums and their order are the same).  An
// instance whose solve has stopped returns at once, as in its own launch.
template <class A>
struct MTab {
    const A *args;
    const int *start;  // n + 1 entries
    int n;
};
__device__ __forceinline__ int mt_inst(const int *start, int n) {
    const int b = (int)blockIdx.x;
    int lo = 0, hi = n - 1;  // the last i with start[i] <= b (uniform per block)
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Kernel descriptors: the argument block A, the body run(vb, a), the launch
// of the single-solve kernel with the same arguments, and the launch bounds
// (kOcc 0: none beyond the block size).
template <class T> struct DirA {
    const T *xc; Grid gc; NodeDiv df; const T *rf; const typename FineW<T>::D *dinv; const uint8_t *fflag;
    T *x; T *p; Span dofs; int xsplit; const Ctl *ctl;
};
template <class T, int S, bool NT> struct OpDir {
    using A = DirA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        prolong_dir_body<T, S, NT>(vb, a.xc, a.gc, a.df, a.rf, a.dinv, a.fflag, a.x, a.p, a.dofs, a.xsplit, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_prolong_dir<T, S, NT>), dim3(grid), dim3(256), shm, s, a.xc, a.gc, a.df, a.rf, a.dinv,
                           a.fflag, a.x, a.p, a.dofs, a.xsplit, a.ctl);
    }
};
template <class T> struct DirColsA {
    const T *xc; Grid gc, gf; const T *rf; const typename FineW<T>::D *dinv; const uint8_t *fflag; T *x; T *p;
    const int32_t *items; int64_t nItems; const Ctl *ctl;
};
template <class T> struct OpDirCols {
    using A = DirColsA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        prolong_dir_cols_body<T>(vb, a.xc, a.gc, a.gf, a.rf, a.dinv, a.fflag, a.x, a.p, a.items, a.nItems, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_prolong_dir_cols<T>, dim3(grid), dim3(256), shm, s, a.xc, a.gc, a.gf, a.rf, a.dinv,
                           a.fflag, a.x, a.p, a.items, a.nItems, a.ctl);
    }
};
template <class T> struct KuOwnerA {
    const T *u; T *y; const T *E; const uint8_t *nflag; Grid gd; TileMap tm; const int32_t *wlist; int64_t wn;
    double *partial; const Ctl *ctl; ModeK<T> km;
};
template <class T, bool StageOut> struct OpKuOwner {
    using A = KuOwnerA<T>;
    static constexpr int kThreads = 256, kOcc = sizeof(T) == 8 ? 3 : 4;
    __device__ static void run(VB vb, const A &a) {
        kbyu_owner_body<T, StageOut>(vb, a.u, a.y, a.E, a.nflag, a.gd, a.tm, a.wlist, a.wn, a.partial, a.ctl, a.km);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_kbyu_owner<T, StageOut>), dim3(grid), dim3(256), shm, s, a.u, a.y, a.E, a.nflag, a.gd,
                           a.tm, a.wlist, a.wn, a.partial, a.ctl, a.km);
    }
};
template <class T> struct KuDenseA {
    const T *u; T *y; const T *Ke; const T *E; const uint8_t *nflag; Grid gd; Span S; double *partial;
    const Ctl *ctl;
};
template <class T> struct OpKuDense {
    using A = KuDenseA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        kbyu_body<T>(vb, a.u, a.y, a.Ke, a.E, a.nflag, a.gd, a.S, a.partial, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_kbyu<T>, dim3(grid), dim3(256), shm, s, a.u, a.y, a.Ke, a.E, a.nflag, a.gd, a.S,
                           a.partial, a.ctl);
    }
};
template <class T> struct UpdA {
    T *r; const T *q; const typename FineW<T>::D *dinv; Span S; double *partRR; double *partDR; Ctl *ctl;
};
template <class T, bool NT> struct OpUpd {
    using A = UpdA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        update_r_body<T, NT>(vb, a.r, a.q, a.dinv, a.S, a.partRR, a.partDR, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_update_r<T, NT>), dim3(grid), dim3(256), shm, s, a.r, a.q, a.dinv, a.S, a.partRR,
                           a.partDR, a.ctl);
    }
};
struct FinA {
    const double *partial; int nb; const double *partial2; int nb2; int op; Ctl *ctl;
};
struct OpFin {
    using A = FinA;
    static constexpr int kThreads = 1024, kOcc = 0;
    __device__ static void run(VB, const A &a) { fin_body(a.partial, a.nb, a.partial2, a.nb2, a.op, a.ctl); }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_fin, dim3(grid), dim3(1024), shm, s, a.partial, a.nb, a.partial2, a.nb2, a.op, a.ctl);
    }
};
template <class T> struct ResA {
    const T *rf; Grid gf; T *rc; Grid gc; NodeDiv dc; const uint8_t *cflag; int skipVoid; const Ctl *ctl;
};
template <class T, int S> struct OpRes {
    using A = ResA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        restrict_body<T, S>(vb, a.rf, a.gf, a.rc, a.gc, a.dc, a.cflag, a.skipVoid, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_restrict<T, S>), dim3(grid), dim3(256), shm, s, a.rf, a.gf, a.rc, a.gc, a.dc, a.cflag,
                           a.skipVoid, a.ctl);
    }
};
template <class T> struct Res2A {
    const T *rf; Grid gf; T *rc; Grid gc; const uint8_t *cflag; int czc, tilesR, tilesC; const int32_t *tiles;
    const uint8_t *fflag; const Ctl *ctl;
};
template <class T> struct OpRes2 {
    using A = Res2A<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        restrict2_tiled_body<T>(vb, a.rf, a.gf, a.rc, a.gc, a.cflag, a.czc, a.tilesR, a.tilesC, a.tiles, a.fflag,
                                a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_restrict2_tiled<T>, dim3(grid), dim3(256), shm, s, a.rf, a.gf, a.rc, a.gc, a.cflag,
                           a.czc, a.tilesR, a.tilesC, a.tiles, a.fflag, a.ctl);
    }
};
template <class T> struct UpdResA {
    const T *rf; const T *q; T *rOut; const typename FineW<T>::D *dinvL; Grid gf; T *rc; Grid gc;
    const uint8_t *cflag; int czc, tilesR, tilesC; int64_t nTiles; const int32_t *tiles; const uint8_t *fflag;
    double *partRR; double *partDR; const Ctl *ctl;
};
template <class T> struct OpUpdRes {
    using A = UpdResA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        update_restrict_body<T>(vb, a.rf, a.q, a.rOut, a.dinvL, a.gf, a.rc, a.gc, a.cflag, a.czc, a.tilesR,
                                a.tilesC, a.nTiles, a.tiles, a.fflag, a.partRR, a.partDR, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_update_restrict<T>, dim3(grid), dim3(256), shm, s, a.rf, a.q, a.rOut, a.dinvL, a.gf,
                           a.rc, a.gc, a.cflag, a.czc, a.tilesR, a.tilesC, a.nTiles, a.tiles, a.fflag, a.partRR,
                           a.partDR, a.ctl);
    }
};
template <class T> struct ProA {
    const T *xc; Grid gc; T *xf; NodeDiv df; const T *rf; const T *dinv; const uint8_t *fflag; int finest;
    Span dofs; double *partial; const Ctl *ctl;
};
template <class T, int S> struct OpPro {
    using A = ProA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        prolong_body<T, S>(vb, a.xc, a.gc, a.xf, a.df, a.rf, a.dinv, a.fflag, a.finest, a.dofs, a.partial, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_prolong<T, S>), dim3(grid), dim3(256), shm, s, a.xc, a.gc, a.xf, a.df, a.rf, a.dinv,
                           a.fflag, a.finest, a.dofs, a.partial, a.ctl);
    }
};
template <class T> struct TailRA {
    TailArgs<T> A; const Ctl *ctl;
};
template <class T> struct OpTailR {
    using A = TailRA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) { tail_restrict_body<T>(vb, a.A, a.ctl); }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_tail_restrict<T>, dim3(grid), dim3(256), shm, s, a.A, a.ctl);
    }
};
template <class T> struct TailPA {
    TailArgs<T> A; double *partial; const Ctl *ctl;
};
template <class T, int M> struct OpTailP {
    using A = TailPA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) { tail_prolong_body<T, M>(vb, a.A, a.partial, a.ctl); }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_tail_prolong<T, M>), dim3(grid), dim3(256), shm, s, a.A, a.partial, a.ctl);
    }
};
template <class T, class TW> struct CLoA {
    const TW *Wrow; const int32_t *freeIdx; int64_t n; const T *r; double *y; const Ctl *ctl;
};
template <class T, class TW, bool Lds> struct OpCLo {
    using A = CLoA<T, TW>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        coarse_lower_body<T, TW, Lds>(vb, a.Wrow, a.freeIdx, a.n, a.r, a.y, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_coarse_lower<T, TW, Lds>), dim3(grid), dim3(256), shm, s, a.Wrow, a.freeIdx, a.n, a.r,
                           a.y, a.ctl);
    }
};
template <class T, class TW> struct CUpA {
    const TW *Wcol; const int32_t *freeIdx; int64_t n; const double *y; T *x; const Ctl *ctl;
};
template <class T, class TW> struct OpCUp {
    using A = CUpA<T, TW>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        coarse_upper_body<T, TW>(vb, a.Wcol, a.freeIdx, a.n, a.y, a.x, a.ctl);
    }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL((k_coarse_upper<T, TW>), dim3(grid), dim3(256), shm, s, a.Wcol, a.freeIdx, a.n, a.y, a.x,
                           a.ctl);
    }
};
template <class T> struct DotA {
    const T *a; const T *b; Span S; double *partial; const Ctl *ctl;
};
template <class T> struct OpDot {
    using A = DotA<T>;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) { dot_body<T>(vb, a.a, a.b, a.S, a.partial, a.ctl); }
    static void single(const A &a, int grid, size_t shm, hipStream_t s) {
        hipLaunchKernelGGL(k_dot<T>, dim3(grid), dim3(256), shm, s, a.a, a.b, a.S, a.partial, a.ctl);
    }
};

template <class Op>
__global__ void __launch_bounds__(Op::kThreads) k_multi(MTab<typename Op::A> t) {
    const int i = mt_inst(t.start, t.n);
    const VB vb{(unsigned)((int)blockIdx.x - t.start[i]), (unsigned)(t.start[i + 1] - t.start[i])};
    Op::run(vb, t.args[i]);
}
template <class Op>
__global__ void __launch_bounds__(Op::kThreads, Op::kOcc > 0 ? Op::kOcc : 1) k_multi_occ(MTab<typename Op::A> t) {
    const int i = mt_inst(t.start, t.n);
    const VB vb{(unsigned)((int)blockIdx.x - t.start[i]), (unsigned)(t.start[i + 1] - t.start[i])};
    Op::run(vb, t.args[i]);
}
template <class Op>
void multi_launch(const void *args, const int *start, int n, int grid, size_t shm, hipStream_t s) {
    const MTab<typename Op::A> t{(const typename Op::A *)args, start, n};
    if constexpr (Op::kOcc > 0) hipLaunchKernelGGL(k_multi_occ<Op>, dim3(grid), dim3(Op::kThreads), shm, s, t);
    else hipLaunchKernelGGL(k_multi<Op>, dim3(grid), dim3(Op::kThreads), shm, s, t);
}

// ---- reliable updates of the mixed solve on the device (merged solves) ----
// Per instance: the solve's fixed numbers and update state (MixCtl), and a
// gate block whose `done` is 0 only while an update runs, so that the
// update's kernels (the same V-cycle kernels, gated by it) do nothing for
// instances that are iterating.  The steps and decisions are those of the
// host path (mixed_update_begin / mixed_update_end).
struct MixCtl {
    double scale, normbS, delta, tolS, res;
    int itsAtUpdate, maxStretch, updates, finished, keep, doneCode;
};
struct MixCheckA {
    Ctl *ctl; Ctl *gate; MixCtl *mix;
};
struct OpMixCheck {
    using A = MixCheckA;
    static constexpr int kThreads = 64, kOcc = 0;
    __device__ static void run(VB, const A &a) {
        if (threadIdx.x != 0) return;
        const bool need = !a.mix->finished && (a.ctl->done != 0 || (a.ctl->flags & kCtlWant) ||
                                               a.ctl->its - a.mix->itsAtUpdate >= a.mix->maxStretch);
        a.gate->done = need ? 0 : 1;
        if (need) a.mix->itsAtUpdate = a.ctl->its;
    }
    static void single(const A &, int, size_t, hipStream_t) {}  // merged solves only
};
// The gated update passes (fold, norms, weighted rho) are launched with a
// fraction of the blocks of a single-solve pass: launches in which no
// instance updates (most of them) then cost a few microseconds instead of
// about 30 us for the exits of 57 k blocks (32 uneven cuboids).  Each
// launched block runs the virtual blocks b, b + g, ... of the single-solve
// grid of vbs blocks, so the per-block partial sums, and with them every
// iterate, stay those of the single solve.
#define FOR_VBLOCKS(v, vb, vbs) for (unsigned v = (vb).b; v < (unsigned)(vbs); v += (vb).g)
struct MixFoldA {
    double *x; float *x32; const float *p32; Span S; const Ctl *ctl; const Ctl *gate; const MixCtl *mix; int vbs;
};
struct OpMixFold {
    using A = MixFoldA;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        if (a.gate->done) return;
        const float lam = (float)a.ctl->lambda;
        const double scale = a.mix->scale;
        FOR_VBLOCKS(v, vb, a.vbs) {
            const VB w{v, (unsigned)a.vbs};
            FOR_SPAN_VT(i, t, a.S, w) {
                a.x[i] += scale * (double)(a.x32[t] + lam * a.p32[i]);
                a.x32[t] = 0.0f;
            }
        }
    }
    static void single(const A &, int, size_t, hipStream_t) {}
};
struct MixNormsA {
    const double *b; const double *q; float *r32; const uint8_t *nflag; Span S; double *partRR; double *partBB;
    const Ctl *gate; const MixCtl *mix; int vbs;
};
struct OpMixNorms {
    using A = MixNormsA;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        if (a.gate->done) return;
        const double invScale = 1.0 / a.mix->scale;
        FOR_VBLOCKS(v, vb, a.vbs) {
            const VB w{v, (unsigned)a.vbs};
            double rr = 0.0, bb = 0.0;
            FOR_SPAN_V(i, a.S, w) {
                const uint32_t node = (uint32_t)(i / 3), c = (uint32_t)(i - 3 * (int64_t)node);
                const double bi = a.b[i];
                const double ri = (a.nflag[node] & (1u << c)) ? 0.0 : bi - a.q[i];
                a.r32[i] = (float)(ri * invScale);
                rr += ri * ri;
                bb += bi * bi;
            }
            block_sum2(rr, bb);
            if (threadIdx.x == 0) {
                a.partRR[v] = rr;
                a.partBB[v] = bb;
            }
        }
    }
    static void single(const A &, int, size_t, hipStream_t) {}
};
struct MixDecideA {
    const double *partRR; const double *partBB; int nb; Ctl *ctl; Ctl *gate; MixCtl *mix;
};
struct OpMixDecide {
    using A = MixDecideA;
    static constexpr int kThreads = 1024, kOcc = 0;
    __device__ static void run(VB, const A &a) {
        if (a.gate->done) return;
        double rr, bb;
        sum2_partials(a.partRR, a.partBB, a.nb, rr, bb);  // as k_sum2 on the single-solve path
        if (threadIdx.x != 0) return;
        Ctl &c = *a.ctl;
        MixCtl &m = *a.mix;
        const double normr = sqrt(rr), res = bb > 0.0 ? sqrt(rr / bb) : 0.0;
        m.updates += 1;
        m.res = res;
        if (res < m.tolS || c.done == 2 || !(normr > 0.0)) {
            m.finished = 1;
            m.doneCode = res < m.tolS ? 1 : 2;
            c.done = 1;
            a.gate->done = 1;  // no rho: the solve has stopped
            return;
        }
        // the direction's rho, as mixed_update_end
        double rhoDir = c.rho;
        int keep = 1;
        if (!c.done) {
            keep = c.beta > 0.0;
            if (keep) rhoDir = c.rho / c.beta;
        }
        m.keep = keep;
        c.done = 0;
        c.flags &= ~kCtlWant;
        c.rho = rhoDir;
        c.lambda = 0.0;
        c.normb = normr / m.scale;
        c.tol = fmax(m.delta, 0.5 * m.tolS * m.normbS / normr);
    }
    static void single(const A &, int, size_t, hipStream_t) {}
};
struct MixDinvRRA {
    const float *r; const uint16_t *dinv; Span S; double *partial; const Ctl *gate; int vbs;
};
struct OpMixDinvRR {
    using A = MixDinvRRA;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) {
        if (a.gate->done) return;
        FOR_VBLOCKS(v, vb, a.vbs) {
            const VB w{v, (unsigned)a.vbs};
            double acc = 0.0;
            FOR_SPAN_VT(i, t, a.S, w) {
                const double rd = (double)a.r[i];
                acc += (double)FineW<float>::get(a.dinv, t) * rd * rd;
            }
            const double s = block_sum(acc);
            if (threadIdx.x == 0) a.partial[v] = s;
        }
    }
    static void single(const A &, int, size_t, hipStream_t) {}
};
struct MixFinRhoA {
    const double *partial; int nb; const double *partial2; int nb2; Ctl *ctl; const Ctl *gate; const MixCtl *mix;
};
struct OpMixFinRho {
    using A = MixFinRhoA;
    static constexpr int kThreads = 1024, kOcc = 0;
    __device__ static void run(VB, const A &a) {
        if (a.gate->done) return;
        fin_body(a.partial, a.nb, a.partial2, a.nb2, a.mix->keep ? FIN_RHO : FIN_RHO0, a.ctl);
    }
    static void single(const A &, int, size_t, hipStream_t) {}
};

// Copies each instance's control block into one array (one host copy per batch).
__global__ void k_gather_ctl(Ctl *const *ctls, int n, Ctl *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *ctls[i];
}

// ------------------------------- setup kernels ------------------------------
// Finest diagonal: diag = sum_e E_e Ke_jj (Solving_AssembleFEAstencil.m:102-121);
// dinv = 2 omega / diag on solid nodes.
__global__ void k_diag0(const double *__restrict__ Ke, const double *__restrict__ E, Grid gd,
                        const uint8_t *__restrict__ nflag,
                        double omega, double *__restrict__ dinv, float *__restrict__ dinv32) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gd.nodes()) return;
    const int64_t NY = gd.NY(), NX = gd.NX();
    const int r = (int)(g % NY), c = (int)((g / NY) % NX), l = (int)(g / (NY * NX));
    double d[3] = {0, 0, 0};
    for (int dl = 0; dl < 2; ++dl)
        for (int dc = 0; dc < 2; ++dc)
            for (int dr = 0; dr < 2; ++dr) {
                const int er = r - 1 + dr, ec = c - 1 + dc, el = l - 1 + dl;
                if (er < 0 || er >= gd.ny || ec < 0 || ec >= gd.nx || el < 0 || el >= gd.nz) continue;
                const int j = corner_index(1 - dc, dr, 1 - dl);
                const double Ee = E[er + (int64_t)gd.ny * ec + (int64_t)gd.ny * gd.nx * el];
                for (int a = 0; a < 3; ++a) d[a] += Ke[(3 * j + a) * 25] * Ee;
            }
    const bool solid = nflag[g] & kSolid;
    for (int a = 0; a < 3; ++a) {
        const double v = solid ? 2.0 * omega / d[a] : 0.0;
        dinv[3 * g + a] = v;
        dinv32[3 * g + a] = (float)v;
    }
}

// The mixed solve's bfloat16 weights in span order (entry i of the lattice at
// its span position t).
__global__ void k_pack_w16_lattice(const float *__restrict__ dinv32, uint16_t *__restrict__ dinv16, Span S) {
    FOR_SPAN(i, S) dinv16[i] = to_bf16(dinv32[i]);
}
__global__ void k_pack_w16(const float *__restrict__ dinv32, uint16_t *__restrict__ dinv16, Span S) {
    FOR_SPAN_T(i, t, S) dinv16[t] = to_bf16(dinv32[i]);
}

// Coarse diagonal from element matrices (Solving_AssembleFEAstencil.m:82-99).
__global__ void k_diagk(const double *__restrict__ Ks, Grid gd, const uint8_t *__restrict__ nflag,
                        double omega, double *__restrict__ dinv, float *__restrict__ dinv32) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= gd.nodes()) return;
    const int64_t NY = gd.NY(), NX = gd.NX();
    const int r = (int)(g % NY), c = (int)((g / NY) % NX), l = (int)(g / (NY * NX));
    double d[3] = {0, 0, 0};
    for (int dl = 0; dl < 2; ++dl)
        for (int dc = 0; dc < 2; ++dc)
            for (int dr = 0; dr < 2; ++dr) {
                const int er = r - 1 + dr, ec = c - 1 + dc, el = l - 1 + dl;
                if (er < 0 || er >= gd.ny || ec < 0 || ec >= gd.nx || el < 0 || el >= gd.nz) continue;
                const int j = corner_index(1 - dc, dr, 1 - dl);
                const int64_t e = er + (int64_t)gd.ny * ec + (int64_t)gd.ny * gd.nx * el;
                for (int a = 0; a < 3; ++a) d[a] += Ks[576 * e + (3 * j + a) * 25];
            }
    const bool solid = nflag[g] & kSolid;
    for (int a = 0; a < 3; ++a) {
        const double v = solid ? 2.0 * omega / d[a] : 0.0;
        dinv[3 * g + a] = v;
        dinv32[3 * g + a] = (float)v;
    }
}

// Level-2 element matrices from the fine moduli and the span-s templates
// H_q = P_q^T Ke P_q (Solving_AssembleCmptStencilFromFinestLevel.cpp:71-169).
__global__ void k_galerkin_from_E(const double *__restrict__ H, int s, const double *__restrict__ E,
                                  Grid gf, Grid gc, double *__restrict__ Ks) {
    const int64_t total = gc.elems() * 576;
    const int nq = s * s * s;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / 576;
        const int t = (int)(idx % 576);
        const int er = (int)(e % gc.ny), ec = (int)((e / gc.ny) % gc.nx), el = (int)(e / ((int64_t)gc.ny * gc.nx));
        double acc = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int qr = q % s, qc = (q / s) % s, ql = q / (s * s);
            const int64_t v = (int64_t)(s * er + qr) + (int64_t)gf.ny * (s * ec + qc) +
                              (int64_t)gf.ny * gf.nx * (s * el + ql);
            acc += E[v] * H[576 * q + t];
        }
        Ks[idx] = acc;
    }
}

// Level >= 3: Ks_e = sum_q P_q^T Ks_prev(child q) P_q
// (Solving_AssembleCmptStencilFromNonFinestLevel.cpp:78-143).  One block of 576
// threads per coarse lattice element.
__global__ void __launch_bounds__(576) k_galerkin_from_Ks(const double *__restrict__ Kf, Grid gf,
                                                          Grid gc, int s,
                                                          const double *__restrict__ P,
                                                          double *__restrict__ Ks) {
    __shared__ double sK[576], sT[576], sP[64];
    const int t = threadIdx.x;
    const int row = t % 24, col = t / 24;
    const int nq = s * s * s;
    for (int64_t e = blockIdx.x; e < gc.elems(); e += gridDim.x) {
        const int er = (int)(e % gc.ny), ec = (int)((e / gc.ny) % gc.nx), el = (int)(e / ((int64_t)gc.ny * gc.nx));
        double out = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int qr = q % s, qc = (q / s) % s, ql = q / (s * s);
            const int64_t v = (int64_t)(s * er + qr) + (int64_t)gf.ny * (s * ec + qc) +
                              (int64_t)gf.ny * gf.nx * (s * el + ql);
            __syncthreads();
            sK[t] = Kf[576 * v + t];
            if (t < 64) sP[t] = P[64 * q + t];
            __syncthreads();
            {
                const int J = col / 3, a = col % 3;
                double acc = 0.0;
                for (int j = 0; j < 8; ++j) {
                    const double w = sP[j * 8 + J];
                    if (w != 0.0) acc += w * sK[row + 24 * (3 * j + a)];
                }
                sT[t] = acc;
            }
            __syncthreads();
            {
                const int I = row / 3, a = row % 3;
                double acc = 0.0;
                for (int i = 0; i < 8; ++i) {
                    const double w = sP[i * 8 + I];
                    if (w != 0.0) acc += w * sT[3 * i + a + 24 * col];
                }
                out += acc;
            }
        }
        Ks[576 * e + t] = out;
    }
}

// Dense FP64 Cholesky K = L L^T of the coarsest free-DOF matrix and its
// inverse factor W = L^-1, blocked in 64 x 64 tiles (column-major n x n, lower
// triangle; L overwrites A).  Per 64-column panel k0: factor the diagonal tile
// in LDS (k_potrf_tile), solve the rows below against it (k_trsm_panel), then
// subtract the panel's outer product from the trailing lower triangle
// (k_syrk_tiles).  W is formed block row by block row: W_II = L_II^-1 for every
// diagonal tile (k_trinv_tiles), then W_IJ = -W_II sum_{J<=K<I} L_IK W_KJ for
// all J < I of block row I in one launch (k_trinv_blockrow).  A pivot that is
// not positive is reported (1-based) in *bad.
constexpr int kCB = 64;  // tile edge

// 4 x 4 outputs per thread of a 64 x 64 tile product over 64 inner indices:
// acc[a][b] += X[k][tr + 16 a] * Y[k][tc + 16 b] (k-major LDS tiles).
__device__ __forceinline__ void tile_fma(const double (*X)[kCB + 1], const double (*Y)[kCB + 1], int tr, int tc,
                                         double acc[4][4]) {
#pragma unroll 4
    for (int k = 0; k < kCB; ++k) {
        double x[4], y[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            x[q] = X[k][tr + 16 * q];
            y[q] = Y[k][tc + 16 * q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int w = 0; w < 4; ++w) acc[q][w] = fma(x[q], y[w], acc[q][w]);
    }
}

// Diagonal tile k0 (kb = min(64, n - k0) columns), one block of 256 threads.
__global__ void __launch_bounds__(256) k_potrf_tile(double *__restrict__ A, int64_t n, int64_t k0,
                                                    unsigned long long *bad) {
    __shared__ double t[kCB][kCB + 1];  // t[col][row]
    const int kb = (int)min<int64_t>(kCB, n - k0);
    for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
        const int i = e % kb, j = e / kb;
        t[j][i] = i >= j ? A[(k0 + i) + n * (k0 + j)] : 0.0;
    }
    for (int k = 0; k < kb; ++k) {
        __syncthreads();
        const double piv = t[k][k];
        if (threadIdx.x == 0 && !(piv > 0.0)) atomicCAS(bad, 0ull, (unsigned long long)(k0 + k + 1));
        const double d = sqrt(piv);
        __syncthreads();
        for (int i = k + threadIdx.x; i < kb; i += blockDim.x) t[k][i] = i == k ? d : t[k][i] / d;
        __syncthreads();
        for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
            const int i = e % kb, j = e / kb;
            if (j > k && i >= j) t[j][i] -= t[k][i] * t[k][j];
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kb * kb; e += blockDim.x) {
        const int i = e % kb, j = e / kb;
        if (i >= j) A[(k0 + i) + n * (k0 + j)] = t[j][i];
    }
}

// Rows i >= k0 + 64 of panel k0: x L_kk^T = a (forward substitution), one
// row per thread; the panel tile L_kk is read from LDS by every lane at once.
__global__ void __launch_bounds__(64) k_trsm_panel(double *__restrict__ A, int64_t n, int64_t k0) {
    __shared__ double l[kCB][kCB + 1];  // l[row][col] of L_kk
    for (int e = threadIdx.x; e < kCB * kCB; e += blockDim.x) {
        const int i = e % kCB, j = e / kCB;
        l[i][j] = i >= j ? A[(k0 + i) + n * (k0 + j)] : 0.0;
    }
    __syncthreads();
    const int64_t i = k0 + kCB + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x[kCB];
#pragma unroll
    for (int j = 0; j < kCB; ++j) x[j] = A[i + n * (k0 + j)];
#pragma unroll
    for (int j = 0; j < kCB; ++j) {
        double v = x[j];
#pragma unroll
        for (int m = 0; m < j; ++m) v = fma(-x[m], l[j][m], v);
        x[j] = v / l[j][j];
    }
#pragma unroll
    for (int j = 0; j < kCB; ++j) A[i + n * (k0 + j)] = x[j];
}

// Linear id -> (ti, tj), ti >= tj, of the lower-triangular tile pairs.
__device__ __forceinline__ void lower_pair(int64_t id, int64_t &ti, int64_t &tj) {
    int64_t r = (int64_t)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while (r * (r + 1) / 2 > id) --r;
    while ((r + 1) * (r + 2) / 2 <= id) ++r;
    ti = r;
    tj = id - r * (r + 1) / 2;
}

// Trailing update of panel k0: A(R_i, R_j) -= P(R_i) P(R_j)^T for the lower
// tile pairs (i >= j) of rows/columns >= k0 + 64, P = A(:, k0 .. k0 + 63).
__global__ void __launch_bounds__(256) k_syrk_tiles(double *__restrict__ A, int64_t n, int64_t k0) {
    __shared__ double Pi[kCB][kCB + 1], Pj[kCB][kCB + 1];  // [k][row]
    int64_t ti, tj;
    lower_pair(blockIdx.x, ti, tj);
    const int64_t base = k0 + kCB, r0 = base + kCB * ti, c0 = base + kCB * tj;
    for (int e = threadIdx.x; e < kCB * kCB; e += blockDim.x) {
        const int r = e % kCB, k = e / kCB;
        Pi[k][r] = r0 + r < n ? A[(r0 + r) + n * (k0 + k)] : 0.0;
        Pj[k][r] = c0 + r < n ? A[(c0 + r) + n * (k0 + k)] : 0.0;
    }
    __syncthreads();
    const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
    double acc[4][4] = {};
    tile_fma(Pi, Pj, tr, tc, acc);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int64_t r = r0 + tr + 16 * q, c = c0 + tc + 16 * w;
            if (r < n && c < n && r >= c) A[r + n * c] -= acc[q][w];
        }
}

// W_II = L_II^-1 for diagonal tile I = blockIdx.x: column c of the tile per
// thread, by forward substitution of e_c in LDS (x[r][c], each thread its own
// column; the upper part of the tile comes out 0).
__global__ void __launch_bounds__(64) k_trinv_tiles(const double *__restrict__ L, double *__restrict__ W,
                                                    int64_t n) {
    __shared__ double l[kCB][kCB + 1], x[kCB][kCB + 1];  // l[row][col], x[row][col]
    const int64_t i0 = kCB * (int64_t)blockIdx.x;
    const int kb = (int)min<int64_t>(kCB, n - i0);
    for (int e = threadIdx.x; e < kCB * kCB; e += blockDim.x) {
        const int i = e % kCB, j = e / kCB;
        l[i][j] = (i < kb && j < kb && i >= j) ? L[(i0 + i) + n * (i0 + j)] : (i == j ? 1.0 : 0.0);
    }
    __syncthreads();
    const int c = threadIdx.x;
    for (int r = 0; r < kCB; ++r) {
        double v = r == c ? 1.0 : 0.0;
        for (int m = c; m < r; ++m) v = fma(-l[r][m], x[m][c], v);  // x[m][c] = 0 for m < c
        x[r][c] = r < c ? 0.0 : v / l[r][r];
    }
    if (c >= kb) return;
    for (int r = 0; r < kb; ++r) W[(i0 + r) + n * (i0 + c)] = x[r][c];
}

// Block row I of W: W_IJ = -W_II S, S = sum_{K=J}^{I-1} L_IK W_KJ, for J =
// blockIdx.x < I (rows of the last block row beyond n read as 0).
__global__ void __launch_bounds__(256) k_trinv_blockrow(const double *__restrict__ L, double *__restrict__ W,
                                                        int64_t n, int64_t I) {
    __shared__ double X[kCB][kCB + 1], Y[kCB][kCB + 1];  // [k][row], [k][col]
    const int64_t J = blockIdx.x, i0 = kCB * I, j0 = kCB * J;
    const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
    double acc[4][4] = {};
    for (int64_t K = J; K < I; ++K) {
        const int64_t k0 = kCB * K;
        __syncthreads();
        for (int e = threadIdx.x; e < kCB * kCB; e += blockDim.x) {
            const int a = e % kCB, b = e / kCB;
            X[b][a] = i0 + a < n ? L[(i0 + a) + n * (k0 + b)] : 0.0;  // L_IK(a, b)
            Y[a][b] = W[(k0 + a) + n * (j0 + b)];                      // W_KJ(a, b)
        }
        __syncthreads();
        tile_fma(X, Y, tr, tc, acc);
    }
    // S to LDS (Y[k][col] = S(k, col)), W_II to X ([k][row] = W_II(row, k))
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int w = 0; w < 4; ++w) Y[tr + 16 * q][tc + 16 * w] = acc[q][w];
    for (int e = threadIdx.x; e < kCB * kCB; e += blockDim.x) {
        const int a = e % kCB, b = e / kCB;
        X[b][a] = (i0 + a < n && i0 + b < n) ? W[(i0 + a) + n * (i0 + b)] : 0.0;
    }
    __syncthreads();
    double out[4][4] = {};
    tile_fma(X, Y, tr, tc, out);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int64_t r = i0 + tr + 16 * q;
            if (r < n) W[r + n * (j0 + tc + 16 * w)] = -out[q][w];
        }
}

// Packed rows and columns of the lower-triangular W (column-major n x n;
// tri_row_off / tri_col_off).
__global__ void k_tri_pack(const double *__restrict__ W, int64_t n, double *__restrict__ Wrow,
                           double *__restrict__ Wcol) {
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n * n;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % n, j = idx / n;
        if (j > i) continue;
        Wrow[tri_row_off(i) + j] = W[idx];
        Wcol[tri_col_off(j, n) + (i - j)] = W[idx];
    }
}

// Compressed <-> lattice.
template <class T>
__global__ void k_scatter_nodes(const T *__restrict__ src, const int32_t *__restrict__ map,
                                int64_t n, T *__restrict__ dst) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t g = map[j];
    dst[3 * g] = src[3 * j];
    dst[3 * g + 1] = src[3 * j + 1];
    dst[3 * g + 2] = src[3 * j + 2];
}
template <class T>
__global__ void k_gather_nodes(const T *__restrict__ src, const int32_t *__restrict__ map,
                               int64_t n, T *__restrict__ dst) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t g = map[j];
    dst[3 * j] = src[3 * g];
    dst[3 * j + 1] = src[3 * g + 1];
    dst[3 * j + 2] = src[3 * g + 2];
}
__global__ void k_scatter_elems(const double *__restrict__ src, const int32_t *__restrict__ map,
                                int64_t n, double *__restrict__ dst, float *__restrict__ dst32) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t v = map ? map[j] : j;
    dst[v] = src[j];
    dst32[v] = (float)src[j];
}

// Unit compliance ce_e = u_e^T Ke u_e per lattice element
// (TopOpti_CmptUnitCompliance_mex.cpp:61-107), then gathered to compressed order.
__global__ void __launch_bounds__(256) k_compliance(const double *__restrict__ Ke, const double *__restrict__ u, Grid gd,
                                                    const int32_t *__restrict__ emap, int64_t ne,
                                                    double *__restrict__ ce) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    const int64_t v = emap ? emap[i] : i;
    const int er = (int)(v % gd.ny), ec = (int)((v / gd.ny) % gd.nx), el = (int)(v / ((int64_t)gd.ny * gd.nx));
    const int64_t NY = gd.NY(), pl = gd.NY() * gd.NX();
    double ue[24];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
        const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
        const int zk = (k >= 4);
        const int64_t g = (er + 1 - yk) + NY * (ec + xk) + pl * (el + zk);
        ue[3 * k] = u[3 * g];
        ue[3 * k + 1] = u[3 * g + 1];
        ue[3 * k + 2] = u[3 * g + 2];
    }
    double q = 0.0;
#pragma unroll
    for (int r = 0; r < 24; ++r) {
        double y = 0.0;
#pragma unroll
        for (int c = 0; c < 24; ++c) y += Ke[r + 24 * c] * ue[c];
        q += ue[r] * y;
    }
    ce[i] = q;
}

// Element energy u_e^T Ke u_e of one hexahedron from its 24 DOFs (corner k at
// (x, y, z) = (kCornerX, kCornerY, kCornerZ)[k]).  With the mode-space
// coefficients of the isotropic cube (modes != 0) it is v^T K' v with v the
// per-component Walsh-Hadamard modes of the corners (Ke = H^T K' H per
// component, since K' = H Ke H^T / 64 and H^T H = 8 I): 72 adds for the
// transforms, about 60 operations in mode_forces and a 21-term dot product,
// instead of the 576 multiply-adds of the dense form, which the other
// element matrices use.
__device__ __forceinline__ double element_energy(const double (&ue)[24], const double *__restrict__ Ke,
                                                 const ModeK<double> &km, int modes) {
    if (modes) {
        double v[3][8], w[3][8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
            const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
            const int zk = (k >= 4);
#pragma unroll
            for (int a = 0; a < 3; ++a) v[a][xk + 2 * yk + 4 * zk] = ue[3 * k + a];
        }
        // mode m: bit d set -> +1 at coordinate 1, -1 at coordinate 0 along axis d
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (!(i & bit)) {
                        const double lo = v[a][i], hi = v[a][i | bit];
                        v[a][i] = lo + hi;
                        v[a][i | bit] = hi - lo;
                    }
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int m = 0; m < 8; ++m) w[a][m] = v[a][m];
        mode_forces<double>(w, 1.0, km);
        double q = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int m = 1; m < 8; ++m) q += v[a][m] * w[a][m];
        return q;
    }
    double q = 0.0;
#pragma unroll
    for (int r = 0; r < 24; ++r) {
        double y = 0.0;
#pragma unroll
        for (int c = 0; c < 24; ++c) y += Ke[r + 24 * c] * ue[c];
        q += ue[r] * y;
    }
    return q;
}

// Compliance sensitivity (TopOpti_GlobalVolumeConstraint.m:113-118 with
// TopOpti_DeMaterialInterpolation.m:5): ce = u_e^T Ke u_e,
// dc_e = -penal (E0 - Emin) x_e^(penal-1) ce * scale, and the partial of the
// compliance sum_e E_e ce (E_e the modulus on the lattice).
struct SensA {
    const double *Ke; const double *u; Grid gd; const int32_t *emap; int64_t ne; const double *x; const double *E;
    double penal, dE, scale; double *dc; double *partial; ModeK<double> km; int modes;
};
__device__ __forceinline__ void sensitivity_body(VB vb, const SensA &a) {
    const Grid gd = a.gd;
    const int64_t NY = gd.NY(), pl = gd.NY() * gd.NX();
    double cacc = 0.0;
    for (int64_t i = (int64_t)vb.b * blockDim.x + threadIdx.x; i < a.ne; i += (int64_t)vb.g * blockDim.x) {
        const int64_t v = a.emap ? a.emap[i] : i;
        const int er = (int)(v % gd.ny), ec = (int)((v / gd.ny) % gd.nx), el = (int)(v / ((int64_t)gd.ny * gd.nx));
        double ue[24];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
            const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
            const int zk = (k >= 4);
            const int64_t g = (er + 1 - yk) + NY * (ec + xk) + pl * (el + zk);
            ue[3 * k] = a.u[3 * g];
            ue[3 * k + 1] = a.u[3 * g + 1];
            ue[3 * k + 2] = a.u[3 * g + 2];
        }
        const double q = element_energy(ue, a.Ke, a.km, a.modes);
        a.dc[i] = -a.penal * a.dE * pow(a.x[i], a.penal - 1.0) * q * a.scale;
        cacc += a.E[v] * q;
    }
    const double s = block_sum(cacc);
    if (threadIdx.x == 0) a.partial[vb.b] = s * a.scale;
}
struct OpSens {
    using A = SensA;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB vb, const A &a) { sensitivity_body(vb, a); }
    static void single(const A &, int, size_t, hipStream_t) {}  // merged launches only
};
// One block per instance: the sum of its nb partials (in a fixed order) into *out.
struct SumA {
    const double *partial; int nb; double *out;
};
struct OpSum {
    using A = SumA;
    static constexpr int kThreads = 256, kOcc = 0;
    __device__ static void run(VB, const A &a) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < a.nb; i += blockDim.x) acc += a.partial[i];
        const double s = block_sum(acc);
        if (threadIdx.x == 0) *a.out = s;
    }
    static void single(const A &, int, size_t, hipStream_t) {}  // merged launches only
};

// Nodal stresses (FEA_StressAnalysis.m:33-59), node-centric: each solid node
// gathers, from each solid element around it, E_e T_j u_e (T_j: the 6 rows of
// the element nodal-stress operator for the node's local index j), averages
// over those elements, and forms the von Mises stress.  stress: nn x 6
// column-major in compressed node order.
__global__ void __launch_bounds__(256) k_stress(const double *__restrict__ u, const double *__restrict__ E,
                                                const int32_t *__restrict__ efwd, Grid gd,
                                                const double *__restrict__ T, const int32_t *__restrict__ nodMap,
                                                int64_t nn, double *__restrict__ stress, double *__restrict__ vm) {
    const int64_t NY = gd.NY(), NX = gd.NX(), pl = NY * NX;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = nodMap ? nodMap[i] : i;
        const int r = (int)(g % NY), c = (int)((g / NY) % NX), l = (int)(g / pl);
        double acc[6] = {0, 0, 0, 0, 0, 0};
        int cnt = 0;
        for (int dl = 0; dl < 2; ++dl)
            for (int dc = 0; dc < 2; ++dc)
                for (int dr = 0; dr < 2; ++dr) {
                    const int er = r - 1 + dr, ec = c - 1 + dc, el = l - 1 + dl;
                    if (er < 0 || er >= gd.ny || ec < 0 || ec >= gd.nx || el < 0 || el >= gd.nz) continue;
                    const int64_t v = er + (int64_t)gd.ny * ec + (int64_t)gd.ny * gd.nx * el;
                    if (!efwd[v]) continue;
                    const int j = corner_index(1 - dc, dr, 1 - dl);
                    const double *Tj = T + 6 * j * 24;
                    double s6[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int xk = (k == 1 || k == 2 || k == 5 || k == 6);
                        const int yk = (k == 2 || k == 3 || k == 6 || k == 7);
                        const int zk = (k >= 4);
                        const int64_t gk = (er + 1 - yk) + NY * (ec + xk) + pl * (el + zk);
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const double uv = u[3 * gk + a];
#pragma unroll
                            for (int q = 0; q < 6; ++q) s6[q] += Tj[q * 24 + 3 * k + a] * uv;
                        }
                    }
                    const double Ee = E[v];
#pragma unroll
                    for (int q = 0; q < 6; ++q) acc[q] += Ee * s6[q];
                    ++cnt;
                }
        double sv[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            sv[q] = cnt ? acc[q] / cnt : 0.0;
            stress[i + nn * q] = sv[q];
        }
        vm[i] = sqrt(0.5 * ((sv[0] - sv[1]) * (sv[0] - sv[1]) + (sv[1] - sv[2]) * (sv[1] - sv[2]) +
                            (sv[2] - sv[0]) * (sv[2] - sv[0])) +
                     3.0 * (sv[5] * sv[5] + sv[3] * sv[3] + sv[4] * sv[4]));
    }
}

}  // namespace
}  // namespace sgld

// ============================== host context ================================
using namespace sgld;

namespace {

// Positions of the launches in one PCG iteration (enqueue_iteration): every
// instance's launches appear in increasing order, so the multi-instance
// iteration can merge the instances' launches step by step.
enum Step {
    ST_DIR = 100, ST_KU = 200, ST_FIN_PQ = 250, ST_UPD = 300, ST_FIN_RR = 400,
    ST_RES = 500,     // + level
    ST_TAILR = 600, ST_CLO = 700, ST_CUP = 710, ST_TAILP = 800,
    ST_PRO = 1000,    // - level
    ST_DOT = 1100, ST_FIN_RHO = 1200, ST_OTHER = 1300
};

// Launch kernel Op with argument block a on the context's stream, or, while
// the context records its iteration, append it to the recording.
template <class Op>
int emit(sgld_ctx *C, int step, int grid, size_t shm, const typename Op::A &a) {
    if (!C->rec) {
        Op::single(a, grid, shm, C->s);
        SGLD_HIP(hipGetLastError());
        return SGLD_OK;
    }
    RecOp o{C->rec->base + step, &multi_launch<Op>, grid, shm, std::vector<char>(sizeof(a))};
    std::memcpy(o.args.data(), &a, sizeof(a));
    C->rec->ops.push_back(std::move(o));
    return SGLD_OK;
}

template <class T> T *lvl_r(Level &l);
template <> double *lvl_r<double>(Level &l) { return l.r.p; }
template <> float *lvl_r<float>(Level &l) { return l.r32.p; }
template <class T> T *lvl_x(Level &l);
template <> double *lvl_x<double>(Level &l) { return l.x.p; }
template <> float *lvl_x<float>(Level &l) { return l.x32.p; }
template <class T> const T *lvl_dinv(Level &l);
template <> const double *lvl_dinv<double>(Level &l) { return l.dinv.p; }
template <> const float *lvl_dinv<float>(Level &l) { return l.dinv32.p; }
// finest-level Jacobi weights of the PCG passes (FP64, or bfloat16 for FP32)
template <class T> const typename FineW<T>::D *fine_dinv(sgld_ctx *C);
template <> const double *fine_dinv<double>(sgld_ctx *C) { return C->lv[0].dinv.p; }
template <> const uint16_t *fine_dinv<float>(sgld_ctx *C) { return C->lv[0].dinv16.p; }
template <class T> const T *ctx_E(sgld_ctx *C);
template <> const double *ctx_E<double>(sgld_ctx *C) { return C->E.p; }
template <> const float *ctx_E<float>(sgld_ctx *C) { return C->E32.p; }
template <class T> const T *ctx_Ke(sgld_ctx *C);
template <> const double *ctx_Ke<double>(sgld_ctx *C) { return C->KeDev.p; }
template <> const float *ctx_Ke<float>(sgld_ctx *C) { return C->KeDev32.p; }
template <class T> ModeK<T> ctx_km(const sgld_ctx *C);
template <> ModeK<double> ctx_km<double>(const sgld_ctx *C) { return C->km64; }
template <> ModeK<float> ctx_km<float>(const sgld_ctx *C) { return C->km32; }

double *slot(sgld_ctx *C, int k) { return C->partial.p + (int64_t)k * kRedBlocks; }

// Kernel timer (KernelTimer): events around one launch on the context stream.
enum { KT_DIR64 = 0, KT_DIR32, KT_KU64, KT_KU32, KT_UPD64, KT_UPD32, KT_RES64, KT_RES32 };
int kt_begin(sgld_ctx *C, int id, const char *name) {
    KernelTimer &T = C->kt;
    if (!T.on || !T.inSolve) return SGLD_OK;
    T.name[id] = name;
    KernelTimer::Rec r{id, T.curIter, nullptr, nullptr};
    for (hipEvent_t *e : {&r.a, &r.b}) {
        if (!T.pool.empty()) {
            *e = T.pool.back();
            T.pool.pop_back();
        } else {
            SGLD_HIP(hipEventCreate(e));
        }
    }
    SGLD_HIP(hipEventRecord(r.a, C->s));
    T.pending.push_back(r);
    return SGLD_OK;
}
int kt_end(sgld_ctx *C) {
    KernelTimer &T = C->kt;
    if (!T.on || !T.inSolve) return SGLD_OK;
    SGLD_HIP(hipEventRecord(T.pending.back().b, C->s));
    return SGLD_OK;
}
// After the stream has drained: launches of batch iterations >= `worked`
// ran past the stop flag.
int kt_harvest(sgld_ctx *C, int worked) {
    KernelTimer &T = C->kt;
    for (const KernelTimer::Rec &r : T.pending) {
        float ms = 0.0f;
        SGLD_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        T.ms[r.id] += ms;
        T.n[r.id] += 1;
        if (r.iter < worked) {
            T.msWork[r.id] += ms;
            T.nWork[r.id] += 1;
        }
        T.pool.push_back(r.a);
        T.pool.push_back(r.b);
    }
    T.pending.clear();
    return SGLD_OK;
}

// Kernel names of the streaming passes: kind 0 direction update (span 2),
// 1 direction update (span 4), 2 residual update; NT: nontemporal loads.
const char *stream_name(bool f64, bool nt, int kind) {
    static const char *const n[12] = {
        "k_prolong_dir<float, 2, false>",  "k_prolong_dir<float, 2, true>",  "k_prolong_dir<float, 4, false>",
        "k_prolong_dir<float, 4, true>",   "k_update_r<float, false>",       "k_update_r<float, true>",
        "k_prolong_dir<double, 2, false>", "k_prolong_dir<double, 2, true>", "k_prolong_dir<double, 4, false>",
        "k_prolong_dir<double, 4, true>",  "k_update_r<double, false>",      "k_update_r<double, true>"};
    return n[6 * f64 + 2 * kind + nt];
}

// (the FP32 owner K.u compiled for 5 waves per SIMD, 96 VGPRs with 16-25
// spilled, measured slower than at 4: 125 M-DOF ellipsoids 504 against 483 us)
const char *owner_name(bool f64, bool stage) {
    static const char *const n[4] = {"k_kbyu_owner<float, false>", "k_kbyu_owner<float, true>",
                                     "k_kbyu_owner<double, false>", "k_kbyu_owner<double, true>"};
    return n[2 * f64 + stage];
}

// y = K u; with zin != null first pout = zin + beta u (the PCG direction
// update of the previous iteration) and y = K pout.
template <class T>
int enqueue_kbyu(sgld_ctx *C, const T *u, T *y, double *partial, const Ctl *ctl, int *nbOut,
                 const T *zin = nullptr, T *pout = nullptr) {
    const Level &l0 = C->lv[0];
    int nb;
    if (zin) {  // the direction update runs as its own bandwidth-bound pass
        const Span S = dof_span(C);
        hipLaunchKernelGGL(k_update_p_into<T>, dim3(grid_blocks(S.count)), dim3(256), 0, C->s, pout, zin, u, S,
                           ctl);
        SGLD_HIP(hipGetLastError());
        u = pout;
    }
    constexpr bool f64 = sizeof(T) == 8;
    SGLD_TRY(kt_begin(C, f64 ? KT_KU64 : KT_KU32,
                      C->kbyuVariant != 1 ? (f64 ? "k_kbyu<double>" : "k_kbyu<float>")
                      : C->kbyuOwner ? owner_name(f64, C->kbyuStageOut[f64 ? 0 : 1])
                      : C->kbyuOneBar ? (f64 ? "k_kbyu_tiled<double, true>" : "k_kbyu_tiled<float, true>")
                                      : (f64 ? "k_kbyu_tiled<double, false>" : "k_kbyu_tiled<float, false>")));
    if (C->kbyuVariant == 1) {
        const int q = f64 ? 0 : 1;
        const TiledSchedule &ts = C->ksched[q];
        nb = ts.grid;
        if (C->kbyuOwner) {
            const KuOwnerA<T> a{u, y, ctx_E<T>(C), l0.nflag.p, l0.g, ts.tm,
                                C->full ? nullptr : (const int32_t *)C->kwork[q].p, C->nKwork[q], partial, ctl,
                                ctx_km<T>(C)};
            SGLD_TRY(C->kbyuStageOut[q] ? (emit<OpKuOwner<T, true>>(C, ST_KU, nb, 0, a))
                                        : (emit<OpKuOwner<T, false>>(C, ST_KU, nb, 0, a)));
        }
        else if (C->rec) {
            set_error("sgld:arg: the multi-instance iteration needs the owner-computes or the dense K.u");
            return SGLD_ERR_ARG;
        } else if (C->kbyuOneBar)
            hipLaunchKernelGGL((k_kbyu_tiled<T, true>), dim3(nb), dim3(256), 0, C->s, u, y, ctx_E<T>(C),
                               l0.nflag.p, l0.g, ts.tm, C->full ? nullptr : (const int32_t *)C->kwork[q].p,
                               C->nKwork[q], partial, ctl, ctx_km<T>(C));
        else
            hipLaunchKernelGGL((k_kbyu_tiled<T, false>), dim3(nb), dim3(256), 0, C->s, u, y, ctx_E<T>(C),
                               l0.nflag.p, l0.g, ts.tm, C->full ? nullptr : (const int32_t *)C->kwork[q].p,
                               C->nKwork[q], partial, ctl, ctx_km<T>(C));
    } else {
        const Span S = node_span(C);
        nb = grid_blocks(S.count);
        SGLD_TRY(emit<OpKuDense<T>>(C, ST_KU, nb, 0,
                                    KuDenseA<T>{u, y, ctx_Ke<T>(C), ctx_E<T>(C), l0.nflag.p, l0.g, S, partial, ctl}));
    }
    SGLD_HIP(hipGetLastError());
    SGLD_TRY(kt_end(C));
    if (nbOut) *nbOut = nb;
    return SGLD_OK;
}

// Mode-space coefficients K' = H Ke H^T / 64 (see mode_forces).  Returns
// false, and the dense kernel is used, unless Ke has the block-diagonal mode
// structure and the axis-permutation symmetry of an isotropic cube; km9
// receives the nine defining values.
bool mode_coefficients(const double *Ke, double *km9) {
    double H[8][8];
    for (int m = 0; m < 8; ++m)
        for (int j = 0; j < 8; ++j) {
            const int c[3] = {kCornerX[j], kCornerY[j], kCornerZ[j]};
            double f = 1;
            for (int d = 0; d < 3; ++d)
                if ((m >> d) & 1) f *= c[d] ? 1.0 : -1.0;
            H[m][j] = f;
        }
    std::vector<std::vector<double>> Kp(24, std::vector<double>(24));
    double maxabs = 0;
    for (int a = 0; a < 3; ++a)
        for (int m = 0; m < 8; ++m)
            for (int b = 0; b < 3; ++b)
                for (int mp = 0; mp < 8; ++mp) {
                    double acc = 0;
                    for (int j = 0; j < 8; ++j)
                        for (int k = 0; k < 8; ++k)
                            acc += H[m][j] * Ke[(3 * j + a) + 24 * (3 * k + b)] * H[mp][k];
                    Kp[8 * a + m][8 * b + mp] = acc / 64.0;
                    maxabs = std::max(maxabs, std::fabs(acc / 64.0));
                }
    const int eb[3] = {1, 2, 4};
    // classes: (row, col) entries that must share one value
    std::vector<std::vector<std::pair<int, int>>> cls(9);
    bool used[24][24] = {};
    for (int a = 0; a < 3; ++a) {
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        auto D = [&](int k, int m) { cls[k].push_back({8 * a + m, 8 * a + m}); };
        D(0, eb[a]);
        D(1, eb[b]); D(1, eb[c]);
        D(2, eb[a] + eb[b]); D(2, eb[a] + eb[c]);
        D(3, eb[b] + eb[c]);
        D(4, 7);
        for (int bb = 0; bb < 3; ++bb) {
            if (bb == a) continue;
            const int cc = 3 - a - bb;
            cls[5].push_back({8 * a + eb[a], 8 * bb + eb[bb]});
            cls[6].push_back({8 * a + eb[bb], 8 * bb + eb[a]});
            cls[7].push_back({8 * a + eb[a] + eb[cc], 8 * bb + eb[bb] + eb[cc]});
            cls[8].push_back({8 * a + eb[bb] + eb[cc], 8 * bb + eb[a] + eb[cc]});
        }
    }
    for (int k = 0; k < 9; ++k) {
        const double v = Kp[cls[k][0].first][cls[k][0].second];
        km9[k] = v;
        for (auto &e : cls[k]) {
            if (std::fabs(Kp[e.first][e.second] - v) > 1e-13 * maxabs) return false;
            used[e.first][e.second] = true;
        }
    }
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j)
            if (!used[i][j] && std::fabs(Kp[i][j]) > 1e-13 * maxabs) return false;
    return true;
}

// Whether the restriction into level k runs the LDS-tiled span-2 kernel.
template <class T>
bool restrict_tiled(const sgld_ctx *C, int k) {
    const Level &Cc = C->lv[k];
    // (restrictTiled 2: tiled into level 1 only, node-centric below, one
    // kernel per level for every instance of a merged multi-instance solve)
    const bool tiled = C->restrictTiled < 0
                           ? Cc.g.nodes() >= (sizeof(T) == 8 ? C->restrictTiledMin : C->restrictTiledMin32)
                           : C->restrictTiled == 2 ? k == 1 : C->restrictTiled != 0;
    return Cc.span == 2 && tiled;
}

template <class T>
int enqueue_restrict(sgld_ctx *C, int k, const T *rf, T *rc, const Ctl *ctl) {
    Level &F = C->lv[k - 1], &Cc = C->lv[k];
    const int nb = blocks_for(Cc.g.nodes(), 256);
    // by size and precision: the tiled kernel marches 2 czc + 1 fine layers per
    // block, which pays off with enough coarse nodes to fill the device, in
    // FP32 only beyond the Infinity Cache (256x128x128 first level, FP64 /
    // FP32: tiled 31 / 17 us, node-centric 35 / 13 us; 512^3 ellipsoids FP32
    // node-centric 222 us for 0.56 GB)
    const NodeDiv dc = NodeDiv::make(Cc.g);
    if (restrict_tiled<T>(C, k)) {
        const int tilesR = (int)((Cc.g.NY() + kRR - 1) / kRR), tilesC = (int)((Cc.g.NX() + kRC - 1) / kRC);
        const int czc = C->restrictCz, tilesZ = (int)((Cc.g.NZ() + czc - 1) / czc);
        // irregular domains: only the tiles holding a solid coarse node (the
        // others stay zero from set-up)
        const bool list = !C->full && Cc.rtilesCz == czc;
        const int nbt = list ? (int)Cc.nRtiles : tilesR * tilesC * tilesZ;
        if (nbt > 0)
            SGLD_TRY(emit<OpRes2<T>>(C, ST_RES + k, nbt, 0,
                                     Res2A<T>{rf, F.g, rc, Cc.g, Cc.nflag.p, czc, tilesR, tilesC,
                                              list ? (const int32_t *)Cc.rtiles.p : nullptr,
                                              C->full || !C->restrictSkipVoid ? nullptr : F.nflag.p, ctl}));
        return SGLD_OK;
    }
    const ResA<T> a{rf, F.g, rc, Cc.g, dc, Cc.nflag.p, (int)!C->full, ctl};
    return Cc.span == 2 ? emit<OpRes<T, 2>>(C, ST_RES + k, nb, 0, a) : emit<OpRes<T, 4>>(C, ST_RES + k, nb, 0, a);
}

template <class T>
int enqueue_prolong(sgld_ctx *C, int k, const T *xc, T *xf, const T *rf, double *partial,
                    const Ctl *ctl, int *nbOut) {
    Level &F = C->lv[k - 1], &Cc = C->lv[k];
    const int finest = (k == 1);
    // irregular domains: only the level's DOF chunks holding a solid
    // node (the rest stays zero from set-up)
    const Span S = finest ? dof_span(C)
                          : (C->full || !F.dofChunks.p ? Span::all(3 * F.g.nodes())
                                                       : Span::listed(F.dofChunks.p, 3 * F.g.nodes(), F.nDofChunks,
                                                                      C->chunkLog));
    const int nb = grid_blocks(S.count);
    const NodeDiv df = NodeDiv::make(F.g);
    const ProA<T> a{xc, Cc.g, xf, df, rf, lvl_dinv<T>(F), F.nflag.p, finest, S, partial, ctl};
    SGLD_TRY(Cc.span == 2 ? (emit<OpPro<T, 2>>(C, ST_PRO - k, nb, 0, a)) : (emit<OpPro<T, 4>>(C, ST_PRO - k, nb, 0, a)));
    if (nbOut) *nbOut = nb;
    return SGLD_OK;
}

// z = M r (Solving_Vcycle.m), partial z.r into `partial` when non-null.
template <class T>
int enqueue_vcycle_coarse(sgld_ctx *C, const T *r, double *partialX2, int *nbX2, const Ctl *ctl, int kFirst = 1);

template <class T>
int enqueue_vcycle(sgld_ctx *C, const T *r, T *z, double *partial, const Ctl *ctl, int *nbOut) {
    SGLD_TRY(enqueue_vcycle_coarse<T>(C, r, nullptr, nullptr, ctl));
    SGLD_TRY(enqueue_prolong<T>(C, 1, lvl_x<T>(C->lv[1]), z, r, partial, ctl, nbOut));
    return SGLD_OK;
}

int fin(sgld_ctx *C, const double *partial, int nb, int op, const double *partial2 = nullptr, int nb2 = 0) {
    const int step = op == FIN_RR ? ST_FIN_RR : op == FIN_RHO ? ST_FIN_RHO : op == FIN_PQ ? ST_FIN_PQ : ST_OTHER;
    return emit<OpFin>(C, step, 1, 0, FinA{partial, nb, partial2, nb2, op, C->ctl.p});
}

// Top t0 of the fused small-level tail (k_tail_restrict / k_tail_prolong),
// or -1: the first level k >= 1 with at most 2^17 nodes, raised so that at
// most kTailMax levels lie below it; none when no level lies below.
int tail_top(const sgld_ctx *C) {
    if (!C->tailOn) return -1;
    const int L = C->L;
    int k = 1;
    while (k < L && C->lv[k].g.nodes() > ((int64_t)1 << 17)) ++k;
    k = std::max(k, L - 1 - kTailMax);
    return k < L - 1 ? k : -1;
}

template <class T>
TailArgs<T> tail_args(sgld_ctx *C, int t0) {
    TailArgs<T> A{};
    Level &l0 = C->lv[t0];
    A.M = C->L - 1 - t0;
    A.g0 = l0.g;
    A.d0 = NodeDiv::make(l0.g);
    A.r0 = lvl_r<T>(l0);
    A.dinv0 = lvl_dinv<T>(l0);
    A.flag0 = l0.nflag.p;
    A.x0 = lvl_x<T>(l0);
    int b = 0;
    for (int m = 0; m < A.M; ++m) {
        Level &lm = C->lv[t0 + 1 + m];
        A.g[m] = lm.g;
        A.dv[m] = NodeDiv::make(lm.g);
        A.r[m] = lvl_r<T>(lm);
        A.dinv[m] = lvl_dinv<T>(lm);
        A.flag[m] = lm.nflag.p;
        A.blk[m] = b;
        const int perBlock = 256 / tail_lanes(m);
        b += (int)((lm.g.nodes() + perBlock - 1) / perBlock);
    }
    A.blk[A.M] = b;
    A.xc = lvl_x<T>(C->lv[C->L - 1]);
    return A;
}

// The V-cycle below the finest level for residual r (Solving_Vcycle.m:12-49
// for levels >= 2): restrictions down to the coarsest, its dense solve, and
// prolongation + smoothing back up to level 2, whose x_2 stays in lv[1]; the
// small levels run as the fused tail.  partialX2 (if not null) receives the
// partials of x_2 . r_2 (count in *nbX2).  kFirst 2: level 2's residual is
// already there (fused with the residual update), r is not read.
template <class T>
int enqueue_vcycle_coarse(sgld_ctx *C, const T *r, double *partialX2, int *nbX2, const Ctl *ctl, int kFirst) {
    const int L = C->L, t0 = tail_top(C);
    const int kLast = t0 > 0 ? t0 : L - 1;  // last level reached by single restrictions
    const T *rf = kFirst > 1 ? lvl_r<T>(C->lv[kFirst - 1]) : r;
    for (int k = kFirst; k <= kLast; ++k) {
        if (k == 1) SGLD_TRY(kt_begin(C, sizeof(T) == 8 ? KT_RES64 : KT_RES32, "first restriction"));
        SGLD_TRY(enqueue_restrict<T>(C, k, rf, lvl_r<T>(C->lv[k]), ctl));
        if (k == 1) SGLD_TRY(kt_end(C));
        rf = lvl_r<T>(C->lv[k]);
    }
    TailArgs<T> A{};
    if (t0 > 0) {
        A = tail_args<T>(C, t0);
        SGLD_TRY(emit<OpTailR<T>>(C, ST_TAILR, A.blk[A.M], 0, TailRA<T>{A, ctl}));
    }
    Level &Lc = C->lv[L - 1];
    const int64_t n = C->nCoarseFree;
    const int nbc = blocks_for(n, 4);
    if (sizeof(T) == 4 && C->cWrow32.p) {
        SGLD_TRY((emit<OpCLo<T, float, false>>(C, ST_CLO, nbc, 0,
                                               CLoA<T, float>{(const float *)C->cWrow32.p, C->cFree.p, n,
                                                              (const T *)lvl_r<T>(Lc), C->cY.p, ctl})));
        SGLD_TRY((emit<OpCUp<T, float>>(C, ST_CUP, nbc, 0,
                                        CUpA<T, float>{(const float *)C->cWcol32.p, C->cFree.p, n,
                                                       (const double *)C->cY.p, lvl_x<T>(Lc), ctl})));
    } else {
        const CLoA<T, double> lo{(const double *)C->cWrow.p, C->cFree.p, n, (const T *)lvl_r<T>(Lc), C->cY.p, ctl};
        SGLD_TRY(n <= kCoarseLds ? (emit<OpCLo<T, double, true>>(C, ST_CLO, nbc, 0, lo))
                                 : (emit<OpCLo<T, double, false>>(C, ST_CLO, nbc, 0, lo)));
        SGLD_TRY((emit<OpCUp<T, double>>(C, ST_CUP, nbc, 0,
                                         CUpA<T, double>{(const double *)C->cWcol.p, C->cFree.p, n,
                                                         (const double *)C->cY.p, lvl_x<T>(Lc), ctl})));
    }
    if (t0 > 0) {
        const int nb = grid_blocks(3 * A.g0.nodes());
        double *part = t0 == 1 ? partialX2 : nullptr;
        const TailPA<T> a{A, part, ctl};
        if (A.M == 1) SGLD_TRY((emit<OpTailP<T, 1>>(C, ST_TAILP, nb, 0, a)));
        else if (A.M == 2) SGLD_TRY((emit<OpTailP<T, 2>>(C, ST_TAILP, nb, 0, a)));
        else SGLD_TRY((emit<OpTailP<T, 3>>(C, ST_TAILP, nb, 0, a)));
        if (part && nbX2) *nbX2 = nb;
    }
    for (int k = kLast; k >= 2; --k)
        SGLD_TRY(enqueue_prolong<T>(C, k, lvl_x<T>(C->lv[k]), lvl_x<T>(C->lv[k - 1]), lvl_r<T>(C->lv[k - 1]),
                                    k == 2 ? partialX2 : nullptr, ctl, k == 2 ? nbX2 : nullptr));
    if (L == 2 && partialX2) {  // level 2 is the coarsest: x_2 . r_2 by its own pass
        const Span S2 = Span::all(3 * C->lv[1].g.nodes());
        *nbX2 = grid_blocks(S2.count);
        SGLD_TRY(emit<OpDot<T>>(C, ST_DOT, *nbX2, 0,
                                DotA<T>{(const T *)lvl_x<T>(Lc), (const T *)lvl_r<T>(Lc), S2, partialX2, ctl}));
    }
    return SGLD_OK;
}

// rho = z.r for a residual r set outside the iteration (start of a solve,
// reliable update), z = M r: sum dinv r^2 on the finest level plus x_2 . r_2
// from the coarse V-cycle; op FIN_RHO0 (restart, beta = 0) or FIN_RHO.
template <class T>
int enqueue_rho(sgld_ctx *C, const T *r, int op) {
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    hipLaunchKernelGGL(k_dinv_rr<T>, dim3(nbv), dim3(256), 0, C->s, r, fine_dinv<T>(C), V, slot(C, 3));
    SGLD_HIP(hipGetLastError());
    int nb2 = 0;
    SGLD_TRY(enqueue_vcycle_coarse<T>(C, r, slot(C, 2), &nb2, C->ctl.p));
    return fin(C, slot(C, 2), nb2, op, slot(C, 3), nbv);
}

// One PCG iteration (Solving_PreconditionedConjugateGradientSolver.m:14-29,
// with z = Solving_Vcycle(r)), rotated: it starts with the finest level of the
// previous iteration's V-cycle, fused with that iteration's deferred
// y += lambda p and this one's p = z + beta p (k_prolong_dir), then
// q = K p, lambda = rho / p.q, r -= lambda q with the stopping test on the
// recursive residual, and the coarse V-cycle of the new r with rho' and beta.
// Finest-level traffic: 6 vector passes + 1/8 for the direction update, K.u,
// 4 for the residual update, 1 + 1/8 for the first restriction.
// Weights of the fused residual update and restriction: the finest Jacobi
// weights (FP64, or the FP32 inner solve's bfloat16 ones) in lattice order.
template <class T> const typename FineW<T>::D *fine_dinv_lattice(sgld_ctx *C);
template <> const double *fine_dinv_lattice<double>(sgld_ctx *C) { return C->lv[0].dinv.p; }
template <> const uint16_t *fine_dinv_lattice<float>(sgld_ctx *C) {
    return C->full ? C->lv[0].dinv16.p : C->dinv16L.p;
}

// Whether the iterations fuse the residual update with the first restriction:
// a span-2 LDS-tiled first restriction, and not while a multi-instance
// iteration is recorded (it replays one iteration, the fused form alternates
// two residual vectors).
template <class T>
bool fuse_res_applies(const sgld_ctx *C) {
    return (sizeof(T) == 8 ? C->fuseRes : C->fuseRes32) && !C->rec && C->L >= 2 && restrict_tiled<T>(C, 1);
}
template <class T>
bool fuse_res(const sgld_ctx *C) {  // and its lattice-order weights exist
    return fuse_res_applies<T>(C) && (sizeof(T) == 8 || C->full || C->dinv16L.p);
}

template <class T>
int enqueue_iteration(sgld_ctx *C, T *x, T *r, T *p, T *q, T *rOut = nullptr) {
    const Span S = dof_span(C);
    Ctl *ctl = C->ctl.p;
    Level &l0 = C->lv[0], &l1 = C->lv[1];
    const NodeDiv d0 = NodeDiv::make(l0.g);
    const int nbv = (int)std::max<int64_t>(1, std::min<int64_t>(C->vecBlocks, (S.count + 255) / 256));
    const int nbx = std::max(kXCD, nbv / kXCD * kXCD);  // XCD-split direction update: a multiple of 8
    constexpr bool f64 = sizeof(T) == 8;
    const bool nt = C->streamNT[f64 ? 0 : 1];
    // LDS-staged coarse columns where the gathers bound the pass: FP64 on a
    // full cuboid (measured 146 -> 105 us on 256x128x128).  FP32 is HBM-bound
    // either way, and on irregular domains whole column pairs cost more than
    // the 64-entry chunks of the DOF pass (cfg 2: 1.86 -> 2.51 s).
    const size_t colBytes = sizeof(T) * 4 * 3 * (size_t)l1.g.NY();  // staged coarse runs
    const bool cols = l1.span == 2 && C->dirCols && f64 && C->full && colBytes <= 48 * 1024;
    SGLD_TRY(kt_begin(C, f64 ? KT_DIR64 : KT_DIR32,
                      cols ? "k_prolong_dir_cols<double>"
                      : stream_name(f64, nt, l1.span == 2 ? 0 : 1)));
    if (cols) {
        const int64_t items = ((l0.g.NX() + 1) / 2) * l0.g.NZ();
        const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(items, 65536));
        SGLD_TRY(emit<OpDirCols<T>>(C, ST_DIR, nb, colBytes,
                                    DirColsA<T>{(const T *)lvl_x<T>(l1), l1.g, l0.g, (const T *)r, fine_dinv<T>(C),
                                                l0.nflag.p, x, p, (const int32_t *)nullptr, items,
                                                (const Ctl *)ctl}));
    } else {
        const DirA<T> a{lvl_x<T>(l1), l1.g, d0, (const T *)r, fine_dinv<T>(C), l0.nflag.p, x, p, S,
                        C->dirXcd ? 1 : 0, (const Ctl *)ctl};
        const int nb = C->dirXcd ? nbx : nbv;
        if (l1.span == 2) SGLD_TRY(nt ? (emit<OpDir<T, 2, true>>(C, ST_DIR, nb, 0, a))
                                      : (emit<OpDir<T, 2, false>>(C, ST_DIR, nb, 0, a)));
        else SGLD_TRY(nt ? (emit<OpDir<T, 4, true>>(C, ST_DIR, nb, 0, a))
                         : (emit<OpDir<T, 4, false>>(C, ST_DIR, nb, 0, a)));
    }
    SGLD_TRY(kt_end(C));
    int nb = 0;
    SGLD_TRY(enqueue_kbyu<T>(C, p, q, slot(C, 0), ctl, &nb));
    SGLD_TRY(fin(C, slot(C, 0), nb, FIN_PQ));
    if (rOut && fuse_res<T>(C)) {
        // r' = r - lambda q into rOut, fused with its restriction to level 2
        Level &F = C->lv[0];
        const int czc = C->restrictCz;
        const int tilesR = (int)((l1.g.NY() + kRR - 1) / kRR), tilesC = (int)((l1.g.NX() + kRC - 1) / kRC);
        const int tilesZ = (int)((l1.g.NZ() + czc - 1) / czc);
        const bool list = !C->full && l1.rtilesCz == czc;
        const int64_t nTiles = list ? l1.nRtiles : (int64_t)tilesR * tilesC * tilesZ;
        const int nbf = (int)std::max<int64_t>(1, std::min<int64_t>(nTiles, kRedBlocks));
        SGLD_TRY(kt_begin(C, f64 ? KT_UPD64 : KT_UPD32,
                          f64 ? "k_update_restrict<double>" : "k_update_restrict<float>"));
        SGLD_TRY(emit<OpUpdRes<T>>(C, ST_UPD, nbf, 0,
                                   UpdResA<T>{(const T *)r, (const T *)q, rOut, fine_dinv_lattice<T>(C), F.g,
                                              lvl_r<T>(l1), l1.g, l1.nflag.p, czc, tilesR, tilesC, nTiles,
                                              list ? (const int32_t *)l1.rtiles.p : nullptr,
                                              C->full || !C->restrictSkipVoid ? nullptr : F.nflag.p, slot(C, 1),
                                              slot(C, 3), ctl}));
        SGLD_TRY(kt_end(C));
        SGLD_TRY(fin(C, slot(C, 1), nbf, FIN_RR));
        int nb2 = 0;
        SGLD_TRY(enqueue_vcycle_coarse<T>(C, rOut, slot(C, 2), &nb2, ctl, 2));
        return fin(C, slot(C, 2), nb2, FIN_RHO, slot(C, 3), nbf);
    }
    SGLD_TRY(kt_begin(C, f64 ? KT_UPD64 : KT_UPD32, stream_name(f64, nt, 2)));
    const UpdA<T> ua{r, (const T *)q, fine_dinv<T>(C), S, slot(C, 1), slot(C, 3), ctl};
    SGLD_TRY(nt ? (emit<OpUpd<T, true>>(C, ST_UPD, nbv, 0, ua)) : (emit<OpUpd<T, false>>(C, ST_UPD, nbv, 0, ua)));
    SGLD_TRY(kt_end(C));
    SGLD_TRY(fin(C, slot(C, 1), nbv, FIN_RR));
    int nb2 = 0;
    SGLD_TRY(enqueue_vcycle_coarse<T>(C, r, slot(C, 2), &nb2, ctl));
    SGLD_TRY(fin(C, slot(C, 2), nb2, FIN_RHO, slot(C, 3), nbv));
    return SGLD_OK;
}

int true_residual(sgld_ctx *C, double *res, double *normr);
int true_residual_enqueue(sgld_ctx *C);
void true_residual_sums(const sgld_ctx *C, double *res, double *normr);
int residual_to32(sgld_ctx *C, double invScale, double *rr, double *bb);
int residual_to32_enqueue(sgld_ctx *C, double invScale);
void residual_to32_sums(const sgld_ctx *C, double *rr, double *bb);

// A solve as a resumable state machine: pcg_begin enqueues and runs the set-up,
// pcg_enqueue_batch queues `batch` iterations plus the copy of the device
// control block to its host mirror (stream-capturable: kernels and one
// memcpy, fixed buffers), pcg_after_batch runs the host decisions once the
// batch has completed.  A single solve loops enqueue / sync / after; the
// multi-instance driver captures each instance's batch once as a graph.
struct SolveState {
    int precision = SGLD_PREC_FP64;
    double tol = 1e-8;
    int maxit = 500, batch = 8;
    bool finished = false;
    int its = 0, done = 0, updates = 0;
    double res = 0.0;
    // mixed precision
    double scale = 1.0, normb = 0.0, delta = 0.1;
    int itsAtUpdate = 0;
    int maxStretch = 96;  // force a reliable update after this many FP32 iterations
    // FP32 stretches that reach their tolerance keep iterating until the
    // reliable update (Ctl kCtlSoft): for merged multi-instance solves, whose
    // updates run every few iterations
    bool soft = false;
};

// The set-up of a solve in two halves: the first queues work up to the first
// value the host needs, the second (once the stream has passed it) decides
// and queues the rest; pcg_begin runs both with a wait between them.
int fp64_begin_enqueue(sgld_ctx *C, SolveState &S) {
    // Solving_PreconditionedConjugateGradientSolver.m:7-12 on lattice vectors (x holds x0)
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    *C->hctl = Ctl{};
    C->hctl->tol = S.tol;
    C->hctl->maxit = S.maxit;
    SGLD_HIP(hipMemcpyAsync(C->ctl.p, C->hctl, sizeof(Ctl), hipMemcpyHostToDevice, C->s));
    hipLaunchKernelGGL(k_dot<double>, dim3(nbv), dim3(256), 0, C->s, C->b.p, C->b.p, V, slot(C, 3), nullptr);
    SGLD_HIP(hipGetLastError());
    SGLD_TRY(fin(C, slot(C, 3), nbv, FIN_NORMB));
    SGLD_TRY(enqueue_kbyu<double>(C, C->x.p, C->q.p, nullptr, nullptr, nullptr));
    hipLaunchKernelGGL(k_residual_masked, dim3(nbv), dim3(256), 0, C->s, C->b.p, C->q.p, C->r.p,
                       C->lv[0].nflag.p, V);
    SGLD_HIP(hipGetLastError());
    SGLD_TRY(enqueue_rho<double>(C, C->r.p, FIN_RHO0));  // rho = z.r, beta = 0: first p = z
    hipLaunchKernelGGL(k_fill<double>, dim3(nbv), dim3(256), 0, C->s, C->p.p, 0.0, V);  // beta 0 x p
    SGLD_HIP(hipMemcpyAsync(C->hctl, C->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, C->s));
    return SGLD_OK;
}
int fp64_begin_finish(sgld_ctx *C, SolveState &S) {
    if (!(C->hctl->normb > 0.0)) {  // F = 0: solution 0, nothing to iterate
        S.finished = true;
        S.its = 0;
        S.res = 0.0;
        S.done = 1;
    }
    return SGLD_OK;
}

int mixed_begin_enqueue(sgld_ctx *C, SolveState &) { return true_residual_enqueue(C); }
int mixed_begin_finish(sgld_ctx *C, SolveState &S) {
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    double res = 0, normr = 0;
    true_residual_sums(C, &res, &normr);
    S.its = 0;
    S.updates = 0;
    S.res = res;
    if (!(normr > 0.0) || res < S.tol) {
        S.finished = true;
        S.done = 1;
        return SGLD_OK;
    }
    S.normb = normr / res;
    // Reliable update when the FP32 residual has dropped by delta = 0.1.
    // Measured, mixed solve at delta 0.01 / 0.1: 256x128x128 cuboid 29.2 /
    // 31.2 ms, part_R256 44.5 / 42.3 ms (its long FP32 stretches drift: true /
    // recursive residual 1.8 at delta 0.01, 1.05 at 0.1), 125 M-DOF
    // ellipsoids 0.893 / 0.884 s.  Starting at 0.1 and narrowing after a
    // clean first stretch gave 30.0-31.0 ms, 42.7-43.0 ms and 0.902 s: no
    // signal in the first stretch tells the cuboid from the ellipsoids.
    S.delta = 0.1;
    if (const char *env = std::getenv("SGLD_MIXED_DELTA")) S.delta = std::atof(env);
    if (const char *env = std::getenv("SGLD_MIXED_STRETCH")) S.maxStretch = std::max(1, std::atoi(env));
    // The FP32 vectors hold values divided by one power of two near the
    // initial residual norm for the whole solve (exact, and FP32's range is
    // ample for the 1e-8 decrease after it), so a reliable update replaces r
    // without rescaling p, rho or the correction.
    S.scale = std::ldexp(1.0, (int)std::lround(std::log2(normr)));
    S.itsAtUpdate = 0;
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nbv), dim3(256), 0, C->s, C->r32.p, C->r.p,
                       1.0 / S.scale, V);
    // x32 in span order: positions t <= lattice entry i < nDof
    hipLaunchKernelGGL(k_fill<float>, dim3(nbv), dim3(256), 0, C->s, C->x32.p, 0.0f,
                       Span::all(std::min<int64_t>(V.count, C->nDof())));
    SGLD_HIP(hipGetLastError());
    *C->hctl = Ctl{};
    // the FP32 recursion's stop: its residual, relative to the stretch's
    // start (normb, in scaled units), below delta, or at the solve's tolerance
    C->hctl->tol = std::max(S.delta, 0.5 * S.tol * S.normb / normr);
    C->hctl->maxit = S.maxit;
    C->hctl->normb = normr / S.scale;
    C->hctl->flags = S.soft ? kCtlSoft : 0;
    SGLD_HIP(hipMemcpyAsync(C->ctl.p, C->hctl, sizeof(Ctl), hipMemcpyHostToDevice, C->s));
    SGLD_TRY(enqueue_rho<float>(C, C->r32.p, FIN_RHO0));
    hipLaunchKernelGGL(k_fill<float>, dim3(nbv), dim3(256), 0, C->s, C->p32.p, 0.0f, V);  // beta 0 x p
    SGLD_HIP(hipGetLastError());
    return SGLD_OK;
}

int ensure_fused_vectors(sgld_ctx *C, int precision);
int pcg_begin_enqueue(sgld_ctx *C, SolveState &S) {
    S.finished = false;
    SGLD_TRY(ensure_fused_vectors(C, S.precision));
    return S.precision == SGLD_PREC_FP64 ? fp64_begin_enqueue(C, S) : mixed_begin_enqueue(C, S);
}
int pcg_begin_finish(sgld_ctx *C, SolveState &S) {
    return S.precision == SGLD_PREC_FP64 ? fp64_begin_finish(C, S) : mixed_begin_finish(C, S);
}
int pcg_begin(sgld_ctx *C, SolveState &S) {
    SGLD_TRY(pcg_begin_enqueue(C, S));
    SGLD_HIP(hipStreamSynchronize(C->s));
    return pcg_begin_finish(C, S);
}

// `batch` iterations.  The last iteration's y += lambda p stays pending (it is
// folded into the next iteration's first pass); pcg_after_batch applies it
// when the solve stops or before a reliable update.
// With the fused residual update (fuse_res) iteration i reads the residual
// from r and writes it to rB for even i, the other way round for odd i; the
// batch is even, so a batch whose iterations all ran leaves it in r (a solve
// that stops, or whose FP32 stretch stops, inside a batch does not go on from
// the recursive residual: it has finished, or a reliable update rewrites r).
int ensure_fused_vectors(sgld_ctx *C, int precision) {
    const bool f64 = precision == SGLD_PREC_FP64;
    if (!(f64 ? fuse_res_applies<double>(C) : fuse_res_applies<float>(C))) return SGLD_OK;
    if (!f64 && !C->full && !C->dinv16L.p) {  // lattice-order bfloat16 weights (set-up once, kept)
        const Span V = dof_span(C);
        SGLD_TRY(C->dinv16L.alloc(C->nDof()));
        hipLaunchKernelGGL(k_pack_w16_lattice, dim3(grid_blocks(V.count)), dim3(256), 0, C->s,
                           (const float *)C->lv[0].dinv32.p, C->dinv16L.p, V);
        SGLD_HIP(hipGetLastError());
    }
    if (f64 && !C->rB.p) {
        SGLD_TRY(C->rB.alloc(C->nDof()));
        SGLD_TRY(C->rB.zero(C->s));  // void entries are never written
    }
    if (!f64 && !C->r32B.p) {
        SGLD_TRY(C->r32B.alloc(C->nDof()));
        SGLD_HIP(hipMemsetAsync(C->r32B.p, 0, sizeof(float) * C->r32B.n, C->s));
    }
    return SGLD_OK;
}

int pcg_enqueue_batch(sgld_ctx *C, const SolveState &S) {
    for (int i = 0; i < S.batch; ++i) {
        C->kt.curIter = i;
        const bool odd = i & 1;
        if (S.precision == SGLD_PREC_FP64) {
            double *rIn = odd && C->rB.p ? C->rB.p : C->r.p;
            double *rOut = C->rB.p ? (odd ? C->r.p : C->rB.p) : nullptr;
            SGLD_TRY(enqueue_iteration<double>(C, C->x.p, rIn, C->p.p, C->q.p, rOut));
        } else {
            float *rIn = odd && C->r32B.p ? C->r32B.p : C->r32.p;
            float *rOut = C->r32B.p ? (odd ? C->r32.p : C->r32B.p) : nullptr;
            SGLD_TRY(enqueue_iteration<float>(C, C->x32.p, rIn, C->p32.p, C->q32.p, rOut));
        }
    }
    C->kt.curIter = -1;
    SGLD_HIP(hipMemcpyAsync(C->hctl, C->ctl.p, sizeof(Ctl), hipMemcpyDeviceToHost, C->s));
    return SGLD_OK;
}

int mixed_update_begin(sgld_ctx *C, SolveState &S, bool *update);
int mixed_update_end(sgld_ctx *C, SolveState &S);

// Host step after a completed batch (the control block is in C->hctl).
int pcg_after_batch(sgld_ctx *C, SolveState &S) {
    const Ctl st = *C->hctl;
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    if (S.precision == SGLD_PREC_FP64) {
        if (st.done) {  // apply the stopped iteration's y += lambda p
            hipLaunchKernelGGL(k_axpy_lambda<double>, dim3(nbv), dim3(256), 0, C->s, C->x.p, (const double *)C->p.p,
                               V, (const Ctl *)C->ctl.p);
            SGLD_HIP(hipGetLastError());
            S.finished = true;
            S.its = st.its;
            S.res = st.res;
            S.done = st.done;
        }
        return SGLD_OK;
    }
    // Mixed: the FP32 Krylov iteration (values divided by the solve's fixed
    // scale); when its recursive residual has dropped by delta since the
    // stretch began (or after maxStretch iterations), fold the FP32
    // correction into the FP64 solution and replace the residual by the FP64
    // true residual (reliable update).
    bool update = false;
    SGLD_TRY(mixed_update_begin(C, S, &update));
    if (!update) return SGLD_OK;
    SGLD_HIP(hipStreamSynchronize(C->s));
    return mixed_update_end(C, S);
}

// Reliable update of the mixed solve, in two halves so that a host driving
// many solves can queue every solve's first half before it waits: the first
// folds the FP32 correction into x and queues r = f - K x (FP64 K.u) with
// the copy of its norms' partials; the second, once the stream has passed
// that copy, decides and queues the continuation.
int mixed_update_begin(sgld_ctx *C, SolveState &S, bool *update) {
    const Ctl st = *C->hctl;
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    S.its = st.its;
    *update = st.done || (st.flags & kCtlWant) || S.its - S.itsAtUpdate >= S.maxStretch;
    if (!*update) return SGLD_OK;
    S.itsAtUpdate = S.its;
    hipLaunchKernelGGL(k_fold_mixed, dim3(nbv), dim3(256), 0, C->s, C->x.p, C->x32.p, (const float *)C->p32.p,
                       S.scale, V, (const Ctl *)C->ctl.p);  // with the pending x32 += lambda p32
    SGLD_HIP(hipGetLastError());
    return residual_to32_enqueue(C, 1.0 / S.scale);
}

int mixed_update_end(sgld_ctx *C, SolveState &S) {
    const Ctl st = *C->hctl;
    double rr = 0, bb = 0;
    residual_to32_sums(C, &rr, &bb);
    const double normr = std::sqrt(rr), res = bb > 0 ? std::sqrt(rr / bb) : 0.0;
    ++S.updates;
    S.res = res;
    if (res < S.tol || st.done == 2 || !(normr > 0.0)) {
        S.finished = true;
        S.done = res < S.tol ? 1 : 2;
        return SGLD_OK;
    }
    // rho of the iteration whose direction p holds, for beta = rho' / rho of
    // the replaced residual: after a stop that iteration's own rho; after a
    // forced update (stretch limit, every iteration of the batch complete)
    // the last iteration has already formed rho_{i+1} and beta_{i+1} =
    // rho_{i+1} / rho_i, so rho_i = rho_{i+1} / beta_{i+1}.  Continuing the
    // direction there, instead of restarting it, keeps slow stretches from
    // compounding (125 M-DOF ellipsoids at delta 0.01: 665 -> 406 iterations).
    double rhoDir = st.rho;
    bool keep = true;
    if (!st.done) {
        keep = st.beta > 0.0;
        if (keep) rhoDir = st.rho / st.beta;
    }
    *C->hctl = st;
    C->hctl->done = 0;
    C->hctl->flags &= ~kCtlWant;
    C->hctl->rho = rhoDir;
    C->hctl->lambda = 0.0;  // applied above
    C->hctl->normb = normr / S.scale;  // the new stretch's start
    C->hctl->tol = std::max(S.delta, 0.5 * S.tol * S.normb / normr);
    SGLD_HIP(hipMemcpyAsync(C->ctl.p, C->hctl, sizeof(Ctl), hipMemcpyHostToDevice, C->s));
    // rho = z.r and beta for the replaced residual, then continue from p
    // (restart, beta = 0, only when the direction has just been restarted)
    SGLD_TRY(enqueue_rho<float>(C, C->r32.p, keep ? FIN_RHO : FIN_RHO0));
    return SGLD_OK;
}

// The context's element matrix on the device (dense FP64 and FP32 copies)
// and its mode-space coefficients, set once when the context is created.
int element_data(sgld_ctx *C) {
    SGLD_TRY(C->KeDev.upload(C->Ke, 576));
    float k32[576];
    for (int i = 0; i < 576; ++i) k32[i] = (float)C->Ke[i];
    SGLD_TRY(C->KeDev32.upload(k32, 576));
    double km[9];
    C->haveModes = mode_coefficients(C->Ke, km) ? 1 : 0;
    if (C->haveModes)
        for (int i = 0; i < 9; ++i) {
            C->km64.v[i] = km[i];
            C->km32.v[i] = (float)km[i];
        }
    return SGLD_OK;
}

int upload_flags(Level &l) {
    if (!l.nflag.p) SGLD_TRY(l.nflag.alloc(l.g.nodes()));
    SGLD_HIP(hipMemcpy(l.nflag.p, l.hflag.data(), l.hflag.size(), hipMemcpyHostToDevice));
    return SGLD_OK;
}

// One solve on lattice vectors b -> x (x holds x0).  Up to S.batch
// iterations are queued between host checks.  Once two checks within one
// stretch (FP64: the solve; mixed: between reliable updates) have measured
// the per-iteration decrease f of the recursive residual, the next batch is
// cut to the iterations predicted to reach that stretch's stop plus one, so
// few iterations run past the device-side stop flag (their ~20 launches
// return at once but still cost the dispatch).
int pcg_solve_loop(sgld_ctx *C, SolveState &S) {
    const int full = S.batch;
    SGLD_TRY(pcg_begin(C, S));
    double res0 = -1.0;
    int its0 = 0, upd0 = -1;
    while (!S.finished) {
        const int before = C->hctl->its;
        SGLD_TRY(pcg_enqueue_batch(C, S));
        SGLD_HIP(hipStreamSynchronize(C->s));
        const Ctl st = *C->hctl;
        SGLD_TRY(kt_harvest(C, st.its - before));
        SGLD_TRY(pcg_after_batch(C, S));
        if (S.finished) break;
        S.batch = full;
        if (upd0 == S.updates && res0 > 0.0 && st.res > 0.0 && st.its > its0 && st.tol > 0.0) {
            const double f = std::pow(st.res / res0, 1.0 / (st.its - its0));
            if (f > 0.0 && f < 1.0) {
                const double k = std::ceil(std::log(st.tol / st.res) / std::log(f));
                if (k >= 0.0 && k + 1.0 < full) S.batch = std::max(1, (int)k + 1);
                // the fused residual update alternates two vectors: keep batches even
                const bool pingPong = S.precision == SGLD_PREC_FP64 ? C->rB.p != nullptr : C->r32B.p != nullptr;
                if (pingPong) S.batch = std::min(full, (S.batch + 1) & ~1);
            }
        }
        // a reliable update starts a new stretch (new reference norm)
        res0 = st.res;
        its0 = st.its;
        upd0 = S.updates;
        if (S.precision != SGLD_PREC_FP64 && C->hctl->normb != st.normb) upd0 = -1;
    }
    SGLD_HIP(hipStreamSynchronize(C->s));
    return kt_harvest(C, 1 << 30);
}

int pcg_solve(sgld_ctx *C, SolveState &S) {
    C->kt.inSolve = true;
    const int rc = pcg_solve_loop(C, S);
    C->kt.inSolve = false;
    C->kt.curIter = -1;
    return rc;
}

// Reliable update of the mixed solve: r32 = (b - K x) * invScale (FP64 K.u
// and difference, FP32 store) with rr = ||b - K x||^2 and bb = ||b||^2;
// host-synchronous.
// r = f - K x (FP64 K.u) written scaled to FP32, with ||r||^2 and ||f||^2
// (summed on the device, k_sum2) copied to the context's pinned buffer;
// residual_to32_sums reads them once the stream has passed the copy.
int residual_to32_enqueue(sgld_ctx *C, double invScale) {
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    if (!C->hpart) SGLD_HIP(hipHostMalloc((void **)&C->hpart, sizeof(double) * 2, hipHostMallocDefault));
    SGLD_TRY(enqueue_kbyu<double>(C, C->x.p, C->q.p, nullptr, nullptr, nullptr));
    hipLaunchKernelGGL(k_residual_norms32, dim3(nbv), dim3(256), 0, C->s, C->b.p, (const double *)C->q.p, C->r32.p,
                       invScale, C->lv[0].nflag.p, V, slot(C, 0), slot(C, 1));
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(1024), 0, C->s, (const double *)slot(C, 0), (const double *)slot(C, 1),
                       nbv, slot(C, 2));
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpyAsync(C->hpart, slot(C, 2), sizeof(double) * 2, hipMemcpyDeviceToHost, C->s));
    return SGLD_OK;
}
void residual_to32_sums(const sgld_ctx *C, double *rr, double *bb) {
    *rr = C->hpart[0];
    *bb = C->hpart[1];
}
int residual_to32(sgld_ctx *C, double invScale, double *rr, double *bb) {
    SGLD_TRY(residual_to32_enqueue(C, invScale));
    SGLD_HIP(hipStreamSynchronize(C->s));
    residual_to32_sums(C, rr, bb);
    return SGLD_OK;
}

// True FP64 relative residual ||b - K x|| / ||b||; leaves b - Kx in C->r.
int true_residual_enqueue(sgld_ctx *C) { return residual64_enqueue(C, C->b.p, C->x.p, C->r.p); }
void true_residual_sums(const sgld_ctx *C, double *res, double *normr) {
    double rr = 0, bb = 0;
    residual64_sums(C, &rr, &bb);
    *normr = std::sqrt(rr);
    *res = bb > 0 ? std::sqrt(rr / bb) : 0.0;
}
int true_residual(sgld_ctx *C, double *res, double *normr) {
    SGLD_TRY(true_residual_enqueue(C));
    SGLD_HIP(hipStreamSynchronize(C->s));
    true_residual_sums(C, res, normr);
    return SGLD_OK;
}

}  // namespace

namespace sgld {

// r = mask(b - K x) with ||r||^2 and ||b||^2 summed on the device (k_sum2,
// fixed order) and copied to the context's pinned pair; residual64_sums
// reads them once the stream has passed the copy (so that a driver of many
// solves can queue every solve's residual before it waits).
int residual64_enqueue(sgld_ctx *C, const double *b, const double *x, double *r) {
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);
    if (!C->hpart) SGLD_HIP(hipHostMalloc((void **)&C->hpart, sizeof(double) * 2, hipHostMallocDefault));
    SGLD_TRY(enqueue_kbyu<double>(C, x, C->q.p, nullptr, nullptr, nullptr));
    hipLaunchKernelGGL(k_residual_norms, dim3(nbv), dim3(256), 0, C->s, b, (const double *)C->q.p, r,
                       C->lv[0].nflag.p, V, slot(C, 0), slot(C, 1));
    hipLaunchKernelGGL(k_sum2, dim3(1), dim3(1024), 0, C->s, (const double *)slot(C, 0), (const double *)slot(C, 1),
                       nbv, slot(C, 2));
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpyAsync(C->hpart, slot(C, 2), sizeof(double) * 2, hipMemcpyDeviceToHost, C->s));
    return SGLD_OK;
}
void residual64_sums(const sgld_ctx *C, double *rr, double *bb) {
    *rr = C->hpart[0];
    *bb = C->hpart[1];
}
int residual64(sgld_ctx *C, const double *b, const double *x, double *r, double *rr, double *bb) {
    SGLD_TRY(residual64_enqueue(C, b, x, r));
    SGLD_HIP(hipStreamSynchronize(C->s));
    residual64_sums(C, rr, bb);
    return SGLD_OK;
}

int set_kchunk(sgld_ctx *C, int q, int cz) {
    TiledSchedule &ts = C->ksched[q];
    const Grid &g = C->lv[0].g;
    ts.set_chunk((int)g.NZ(), cz);
    if (C->full) return SGLD_OK;
    // Work item w = tz P + (tr + tilesR tc) writes node layers [tz cz, tz cz + cz)
    // of its patch; the elements at those nodes are in layers [tz cz - 1, tz cz + cz).
    // z-slab by z-slab: the blocks one XCD runs at a time work on neighbouring
    // patches of one slab (listing each patch's z-chunks together instead was
    // slower: 125 M-DOF ellipsoids FP32 483 -> 590 us)
    const int64_t P = (int64_t)ts.tm.tilesR * ts.tm.tilesC;
    std::vector<int32_t> items;
    for (int tz = 0; tz < ts.tm.tilesZ; ++tz) {
        const int l0 = std::max(tz * cz - 1, 0), l1 = std::min(tz * cz + cz, g.nz);
        for (int64_t p = 0; p < P; ++p) {
            bool any = false;
            for (int l = l0; l < l1 && !any; ++l) any = C->hPatchLayer[l * P + p] != 0;
            if (any) items.push_back((int32_t)(tz * P + p));
        }
    }
    C->nKwork[q] = (int64_t)items.size();
    SGLD_TRY(C->kwork[q].upload(items.data(), items.size()));
    ts.grid = (int)std::max<int64_t>(1, std::min<int64_t>(C->nKwork[q], ts.cap));
    return SGLD_OK;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int to_lattice(sgld_ctx *C, const double *dsrc, double *ddst) {
    const int64_t n = C->nDof();
    if (C->full) {
        SGLD_HIP(hipMemcpyAsync(ddst, dsrc, sizeof(double) * n, hipMemcpyDeviceToDevice, C->s));
        return SGLD_OK;
    }
    SGLD_HIP(hipMemsetAsync(ddst, 0, sizeof(double) * n, C->s));
    const int64_t nn = C->lv[0].numNodes;
    hipLaunchKernelGGL(k_scatter_nodes<double>, dim3(blocks_for(nn, 256)), dim3(256), 0, C->s, dsrc,
                       C->nodMap.p, nn, ddst);
    SGLD_HIP(hipGetLastError());
    return SGLD_OK;
}

int from_lattice(sgld_ctx *C, const double *dsrc, double *ddst) {
    if (C->full) {
        SGLD_HIP(hipMemcpyAsync(ddst, dsrc, sizeof(double) * C->nDof(), hipMemcpyDeviceToDevice, C->s));
        return SGLD_OK;
    }
    const int64_t nn = C->lv[0].numNodes;
    hipLaunchKernelGGL(k_gather_nodes<double>, dim3(blocks_for(nn, 256)), dim3(256), 0, C->s, dsrc,
                       C->nodMap.p, nn, ddst);
    SGLD_HIP(hipGetLastError());
    return SGLD_OK;
}

int check_ready(sgld_ctx *C) {
    SGLD_ARG(C, "sgld:arg: null context");
    if (!C->assembled) {
        set_error("sgld:state: call sgld_ctx_set_fixed, sgld_ctx_set_modulus and sgld_ctx_assemble first");
        return SGLD_ERR_STATE;
    }
    return SGLD_OK;
}

}  // namespace sgld

// ================================= C ABI ====================================
extern "C" int sgld_ctx_create(sgld_ctx **out, int resX, int resY, int resZ, const uint8_t *voxels,
                               int numLevels, int nonDyadic, double nu, double h) {
    SGLD_ARG(out, "sgld:arg: ctx out-pointer is null");
    *out = nullptr;
    SGLD_ARG(resX > 0 && resY > 0 && resZ > 0, "sgld:arg: resolution must be positive");
    SGLD_ARG(numLevels >= 2 && numLevels <= 12, "sgld:arg: numLevels must be in 2..12");
    SGLD_ARG(h > 0 && nu > -1 && nu < 0.5, "sgld:arg: need cellSize > 0, -1 < nu < 0.5");
    const int64_t nvox = (int64_t)resX * resY * resZ;
    SGLD_ARG((int64_t)(resX + 1) * (resY + 1) * (resZ + 1) * 3 < ((int64_t)1 << 31),
             "sgld:arg: lattice too large for 32-bit DOF ids");
    std::unique_ptr<sgld_ctx> C(new sgld_ctx());
    // Solving_BuildingMeshHierarchy.m:12-16
    if (nonDyadic == 1 && numLevels >= 4) numLevels -= 1; else nonDyadic = 0;
    C->L = numLevels;
    if (const char *env = std::getenv("SGLD_RESTRICT_TILED_MIN")) C->restrictTiledMin = std::atoll(env);
    if (const char *env = std::getenv("SGLD_RESTRICT_TILED_MIN32")) C->restrictTiledMin32 = std::atoll(env);
    if (const char *env = std::getenv("SGLD_TAIL")) C->tailOn = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_DIR_COLS")) C->dirCols = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_BRESTRICT_TILED")) C->bRestrictTiled = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_RESTRICT_CZ")) C->restrictCz = std::max(1, std::min(16, std::atoi(env)));
    if (const char *env = std::getenv("SGLD_CHUNK_LOG")) C->chunkLog = std::max(4, std::min(6, std::atoi(env)));
    C->nonDyadic = nonDyadic;
    C->nu = nu;
    C->h = h;
    element_stiffness(nu, h, C->Ke);
    C->lv.resize(numLevels);
    Level &l0 = C->lv[0];
    l0.g = Grid{resX, resY, resZ};
    l0.esolid.assign(nvox, 1);
    if (voxels) for (int64_t v = 0; v < nvox; ++v) l0.esolid[v] = voxels[v] != 0;
    for (int k = 1; k < numLevels; ++k) {
        const int s = (k == 1 && nonDyadic) ? 4 : 2;
        const Grid &gf = C->lv[k - 1].g;
        SGLD_ARG(gf.nx % s == 0 && gf.ny % s == 0 && gf.nz % s == 0,
                 "sgld:arg: level %d resolution %dx%dx%d not divisible by span %d", k, gf.nx, gf.ny,
                 gf.nz, s);
        Level &lc = C->lv[k];
        lc.span = s;
        lc.g = Grid{gf.nx / s, gf.ny / s, gf.nz / s};
        lc.esolid.assign(lc.g.elems(), 0);
        const auto &ef = C->lv[k - 1].esolid;
        // Solving_BuildingMeshHierarchy.m:79 (coarse element exists iff a child is solid)
        for (int64_t v = 0; v < gf.elems(); ++v) {
            if (!ef[v]) continue;
            const int64_t r = v % gf.ny, c = (v / gf.ny) % gf.nx, l = v / ((int64_t)gf.ny * gf.nx);
            lc.esolid[(r / s) + (int64_t)lc.g.ny * (c / s) + (int64_t)lc.g.ny * lc.g.nx * (l / s)] = 1;
        }
    }
    for (int k = 0; k < numLevels; ++k) {
        Level &l = C->lv[k];
        const Grid &g = l.g;
        l.hflag.assign(g.nodes(), 0);
        const int64_t NY = g.NY(), NX = g.NX();
        int64_t ne = 0;
        for (int64_t v = 0; v < g.elems(); ++v) {
            if (!l.esolid[v]) continue;
            ++ne;
            const int64_t r = v % g.ny, c = (v / g.ny) % g.nx, ll = v / ((int64_t)g.ny * g.nx);
            for (int dz = 0; dz < 2; ++dz)
                for (int dx = 0; dx < 2; ++dx)
                    for (int dy = 0; dy < 2; ++dy) l.hflag[(r + dy) + NY * (c + dx) + NY * NX * (ll + dz)] = kSolid;
        }
        l.numElements = ne;
        int64_t nn = 0;
        for (auto f : l.hflag) nn += (f & kSolid) ? 1 : 0;
        l.numNodes = nn;
    }
    Level &f0 = C->lv[0];
    C->full = (f0.numElements == f0.g.elems());
    if (!C->full) {
        C->hNodMap.reserve(f0.numNodes);
        for (int64_t g = 0; g < f0.g.nodes(); ++g)
            if (f0.hflag[g] & kSolid) C->hNodMap.push_back((int32_t)g);
        C->hEleMap.reserve(f0.numElements);
        for (int64_t v = 0; v < f0.g.elems(); ++v)
            if (f0.esolid[v]) C->hEleMap.push_back((int32_t)v);
        SGLD_TRY(C->nodMap.upload(C->hNodMap.data(), C->hNodMap.size()));
        SGLD_TRY(C->eleMap.upload(C->hEleMap.data(), C->hEleMap.size()));
        // chunks (2^chunkLog entries) of the finest node and DOF vectors that
        // hold a solid node
        const Grid &g0 = f0.g;
        const int64_t nNodes = g0.nodes(), cs = (int64_t)1 << C->chunkLog;
        std::vector<int32_t> nodeCh, dofCh;
        for (int64_t c = 0; cs * c < nNodes; ++c) {
            bool any = false;
            for (int64_t g = cs * c; g < std::min(cs * c + cs, nNodes) && !any; ++g) any = f0.hflag[g] & kSolid;
            if (any) nodeCh.push_back((int32_t)c);
        }
        for (int64_t c = 0; cs * c < 3 * nNodes; ++c) {
            bool any = false;
            for (int64_t g = cs * c / 3; g <= std::min(cs * c + cs - 1, 3 * nNodes - 1) / 3 && !any; ++g)
                any = f0.hflag[g] & kSolid;
            if (any) dofCh.push_back((int32_t)c);
        }
        C->nNodeChunks = (int64_t)nodeCh.size();
        C->nDofChunks = (int64_t)dofCh.size();
        nodeCh.insert(nodeCh.end(), 3, 0);  // padding entries (see Span::at)
        dofCh.insert(dofCh.end(), 3, 0);
        SGLD_TRY(C->nodeChunks.upload(nodeCh.data(), nodeCh.size()));
        SGLD_TRY(C->dofChunks.upload(dofCh.data(), dofCh.size()));
        // element layers of each tiled K.u patch window (element rows R0-1 .. R0+kTO-1
        // and columns C0-1 .. C0+kTO-1 of patch (tr, tc), R0 = kTO tr, C0 = kTO tc)
        // that hold a solid element
        const int tilesR = (int)((g0.NY() + kTO - 1) / kTO), tilesC = (int)((g0.NX() + kTO - 1) / kTO);
        const int64_t P = (int64_t)tilesR * tilesC;
        C->hPatchLayer.assign(P * g0.nz, 0);
        for (int64_t v = 0; v < g0.elems(); ++v) {
            if (!f0.esolid[v]) continue;
            const int r = (int)(v % g0.ny), c = (int)((v / g0.ny) % g0.nx);
            const int64_t l = v / ((int64_t)g0.ny * g0.nx);
            for (int tc = c / kTO; tc <= std::min(tilesC - 1, (c + 1) / kTO); ++tc)
                for (int tr = r / kTO; tr <= std::min(tilesR - 1, (r + 1) / kTO); ++tr)
                    C->hPatchLayer[l * P + tr + (int64_t)tilesR * tc] = 1;
        }
    }
    SGLD_HIP(hipStreamCreateWithFlags(&C->s, hipStreamNonBlocking));
    SGLD_HIP(hipEventCreate(&C->e0));
    SGLD_HIP(hipEventCreate(&C->e1));
    SGLD_HIP(hipHostMalloc((void **)&C->hctl, sizeof(Ctl), hipHostMallocDefault));
    // device buffers
    const int64_t nd = C->nDof();
    for (int k = 0; k < numLevels; ++k) {
        Level &l = C->lv[k];
        const int64_t n = 3 * l.g.nodes();
        SGLD_TRY(upload_flags(l));
        SGLD_TRY(l.dinv.alloc(n));
        SGLD_TRY(l.dinv32.alloc(n));
        if (k == 0) SGLD_TRY(l.dinv16.alloc(n));
        if (k > 0) {
            SGLD_TRY(l.r.alloc(n)); SGLD_TRY(l.x.alloc(n));
            SGLD_TRY(l.r32.alloc(n)); SGLD_TRY(l.x32.alloc(n));
            // void entries are never written on irregular domains: start at 0
            SGLD_TRY(l.r.zero(C->s)); SGLD_TRY(l.r32.zero(C->s));
            SGLD_TRY(l.x.zero(C->s)); SGLD_TRY(l.x32.zero(C->s));
            if (!C->full && l.span == 2) {
                // tiles of the tiled restriction that hold a solid coarse node
                const int czc = C->restrictCz;
                const int tilesR = (int)((l.g.NY() + kRR - 1) / kRR), tilesC = (int)((l.g.NX() + kRC - 1) / kRC);
                const int tilesZ = (int)((l.g.NZ() + czc - 1) / czc);
                std::vector<uint8_t> act((size_t)tilesR * tilesC * tilesZ, 0);
                const int64_t NYn = l.g.NY(), NXn = l.g.NX();
                for (int64_t g = 0; g < l.g.nodes(); ++g) {
                    if (!(l.hflag[g] & kSolid)) continue;
                    const int64_t r = g % NYn, c = (g / NYn) % NXn, z = g / (NYn * NXn);
                    act[(r / kRR) + (int64_t)tilesR * ((c / kRC) + (int64_t)tilesC * (z / czc))] = 1;
                }
                std::vector<int32_t> list;
                for (size_t i = 0; i < act.size(); ++i)
                    if (act[i]) list.push_back((int32_t)i);
                l.nRtiles = (int64_t)list.size();
                SGLD_TRY(l.rtiles.upload(list.data(), list.size()));
                l.rtilesCz = czc;
            }
            if (!C->full) {
                // chunks of this level's DOF vectors with a solid node
                std::vector<int32_t> ch;
                const int64_t cs = (int64_t)1 << C->chunkLog;
                for (int64_t c = 0; cs * c < n; ++c) {
                    bool any = false;
                    for (int64_t g = cs * c / 3; g <= std::min(cs * c + cs - 1, n - 1) / 3 && !any; ++g)
                        any = l.hflag[g] & kSolid;
                    if (any) ch.push_back((int32_t)c);
                }
                l.nDofChunks = (int64_t)ch.size();
                ch.insert(ch.end(), 3, 0);  // padding entries (see Span::at)
                SGLD_TRY(l.dofChunks.upload(ch.data(), ch.size()));
            }
        }
    }
    SGLD_TRY(C->E.alloc(f0.g.elems()));
    SGLD_TRY(C->E32.alloc(f0.g.elems()));
    SGLD_TRY(C->E.zero(C->s));
    SGLD_HIP(hipMemsetAsync(C->E32.p, 0, sizeof(float) * f0.g.elems(), C->s));
    for (DevBuf<double> *v : {&C->b, &C->x, &C->r, &C->z, &C->p, &C->p2, &C->q}) {
        SGLD_TRY(v->alloc(nd));
        SGLD_TRY(v->zero(C->s));
    }
    for (DevBuf<float> *v : {&C->b32, &C->r32, &C->z32, &C->p32, &C->p32b, &C->q32, &C->x32}) {
        SGLD_TRY(v->alloc(nd));
        SGLD_HIP(hipMemsetAsync(v->p, 0, sizeof(float) * nd, C->s));
    }
    SGLD_TRY(C->partial.alloc((int64_t)kRedSlots * kRedBlocks));
    SGLD_TRY(C->ctl.alloc(1));
    SGLD_TRY(element_data(C.get()));
    // Streaming passes with nontemporal loads and stores once one DOF vector
    // outgrows half the 256 MB Infinity Cache (a pass streams 3-4 of them, so
    // nothing it reads survives to the next pass); FP64 from a quarter of it
    // (measured: 256x128x128, 103 MB per FP64 vector, FP64 solve 46.9 -> 45.9
    // ms; part_R256, 122 MiB, 58.1 -> 55.6 ms; in FP32 the same forms gained
    // on the cuboid's 51 MB vectors, 29.9 -> 29.1 ms, but lost on
    // part_R256's 61 MiB ones, 37.4 -> 41.0 ms).  SGLD_STREAM_NT=0/1 forces
    // either.
    C->streamNT[0] = (double)nd * 8.0 > 64.0 * (1 << 20);
    C->streamNT[1] = (double)nd * 4.0 > 128.0 * (1 << 20);
    if (const char *env = std::getenv("SGLD_STREAM_NT")) C->streamNT[0] = C->streamNT[1] = std::atoi(env) != 0;
    // The residual update fused with the first restriction saves a read of r
    // where r comes from HBM and the separate passes run over chunk and tile
    // lists: irregular domains with FP64 vectors beyond the same share (125
    // M-DOF ellipsoids: FP64 solve 1529 -> 1468 ms).  Within the share the
    // two passes read from the Infinity Cache and the fused tile march is
    // slower (256x128x128: 46.6 -> 49.7 ms), and on a full 256^3 cuboid it
    // gained nothing (16 FP64 solves 3.025 -> 3.055 s).
    // The FP32 form gains less (bfloat16 weights, half the bytes per node):
    // 575 -> 533 us per iteration on the same workload, with its weights and
    // the coarse flags loaded a layer ahead.
    C->fuseRes = !C->full && (double)nd * 8.0 > 128.0 * (1 << 20);
    C->fuseRes32 = !C->full && (double)nd * 4.0 > 128.0 * (1 << 20);
    if (const char *env = std::getenv("SGLD_FUSE_RES")) C->fuseRes = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_FUSE_RES32")) C->fuseRes32 = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_DIR_XCD")) C->dirXcd = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_RESTRICT_SKIP_VOID")) C->restrictSkipVoid = std::atoi(env) != 0;
    if (const char *env = std::getenv("SGLD_VEC_BLOCKS"))
        C->vecBlocks = std::max(64, std::min(kRedBlocks, std::atoi(env)));
    double km[9];
    C->kbyuVariant = mode_coefficients(C->Ke, km) ? 1 : 0;
    if (const char *env = std::getenv("SGLD_KBYU")) {
        if (!std::strcmp(env, "dense")) C->kbyuVariant = 0;
    }
    if (C->kbyuVariant == 1) {
        const Grid &g0 = f0.g;
        int dev = 0, ncu = 256;
        SGLD_HIP(hipGetDevice(&dev));
        SGLD_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        int per[2] = {0, 0};
        // one barrier per element layer: measured 71-74 vs 75 us (FP64) and 52 vs
        // 54 us (FP32) on 256x128x128; SGLD_KU_BARRIERS=2 selects the two-barrier step
        if (const char *env = std::getenv("SGLD_KU_BARRIERS")) C->kbyuOneBar = std::atoi(env) != 2;
        // owner-computes node sums (no LDS atomics, run-to-run identical);
        // SGLD_KU_OWNER=0 selects the atomic-accumulation kernel
        if (const char *env = std::getenv("SGLD_KU_OWNER")) C->kbyuOwner = std::atoi(env) != 0;
        // Outputs staged through LDS once the output vector outgrows the 256 MB
        // Infinity Cache share (its partial-line stores then reach HBM):
        // 256x128x128 direct / staged FP64 57.6 / 67.5 us, FP32 36.1 / 38.2 us;
        // 125 M-DOF ellipsoids 914 / 867 and 560 / 483 us (the LDS atomic kernel:
        // 63.5, 47.0, 885, 513 us)
        for (int q = 0; q < 2; ++q) C->kbyuStageOut[q] = (double)nd * (q ? 4.0 : 8.0) > 128.0 * (1 << 20);
        if (const char *env = std::getenv("SGLD_KU_STAGE_OUT"))
            C->kbyuStageOut[0] = C->kbyuStageOut[1] = std::atoi(env) != 0;
        if (C->kbyuOwner) {
            SGLD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per[0], C->kbyuStageOut[0] ? k_kbyu_owner<double, true> : k_kbyu_owner<double, false>,
                256, 0));
            SGLD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per[1], C->kbyuStageOut[1] ? k_kbyu_owner<float, true> : k_kbyu_owner<float, false>,
                256, 0));
        } else if (C->kbyuOneBar) {
            SGLD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[0], k_kbyu_tiled<double, true>, 256, 0));
            SGLD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[1], k_kbyu_tiled<float, true>, 256, 0));
        } else {
            SGLD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[0], k_kbyu_tiled<double, false>, 256, 0));
            SGLD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per[1], k_kbyu_tiled<float, false>, 256, 0));
        }
        C->numCU = ncu;
        // Blocks per compute unit of the persistent grid.  A sweep of 2..6 on the
        // 256x128x128 cuboid measured 3 (FP64: 168 VGPRs, 3 waves per SIMD) and
        // 4 (FP32: 114 VGPRs) fastest, one more than the occupancy query gives.
        int measured[2] = {3, 4};
        if (const char *env = std::getenv("SGLD_KU_BPCU64")) measured[0] = std::max(1, std::atoi(env));
        if (const char *env = std::getenv("SGLD_KU_BPCU32")) measured[1] = std::max(1, std::atoi(env));
        for (int q = 0; q < 2; ++q) {
            C->kOcc[q] = std::getenv(q ? "SGLD_KU_BPCU32" : "SGLD_KU_BPCU64") ? measured[q]
                                                                                : std::max(per[q], measured[q]);
            TiledSchedule &ts = C->ksched[q];
            ts.tm.tilesR = (int)((g0.NY() + kTO - 1) / kTO);
            ts.tm.tilesC = (int)((g0.NX() + kTO - 1) / kTO);
            ts.cap = (int)std::min<int64_t>((int64_t)ncu * C->kOcc[q], kRedBlocks);
            // z-chunks of 12 node layers, measured (FP64 / FP32 us, chunk 8 / 12 / 16):
            // 256x128x128 72 / 65 / 73 and 50 / 47 / 52; 256^3 320 / 301 / 295 and
            // 196 / 188 / 184; 125 M-DOF ellipsoids 918 / 887 / 894 and 538 / 517 / 521
            int cz = 12;
            if (const char *env = std::getenv("SGLD_KU_CZ")) cz = std::max(2, std::min(256, std::atoi(env)));
            SGLD_TRY(set_kchunk(C.get(), q, cz));
        }
    }
    SGLD_HIP(hipStreamSynchronize(C->s));
    *out = C.release();
    return SGLD_OK;
}

extern "C" int sgld_ctx_destroy(sgld_ctx *C) {
    if (!C) return SGLD_OK;
    if (C->s) (void)hipStreamSynchronize(C->s);
    if (C->hctl) (void)hipHostFree(C->hctl);
    if (C->hpart) (void)hipHostFree(C->hpart);
    if (C->e0) (void)hipEventDestroy(C->e0);
    if (C->e1) (void)hipEventDestroy(C->e1);
    for (hipEvent_t e : C->kt.pool) (void)hipEventDestroy(e);
    for (const KernelTimer::Rec &r : C->kt.pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    batch_free(C->batch);
    hipStream_t s = C->s;
    delete C;
    if (s) (void)hipStreamDestroy(s);
    return SGLD_OK;
}

extern "C" int sgld_ctx_info(const sgld_ctx *C, int level, int64_t *out) {
    SGLD_ARG(C, "sgld:arg: null context");
    if (level == 0) return C->L;
    SGLD_ARG(level >= 1 && level <= C->L && out, "sgld:arg: level out of range");
    const Level &l = C->lv[level - 1];
    out[0] = l.g.nx; out[1] = l.g.ny; out[2] = l.g.nz; out[3] = l.span;
    out[4] = l.numElements; out[5] = l.numNodes; out[6] = 3 * l.numNodes; out[7] = l.g.nodes();
    return SGLD_OK;
}

extern "C" int sgld_ctx_set_fixed(sgld_ctx *C, const uint8_t *fixed) {
    SGLD_ARG(C && fixed, "sgld:arg: null argument");
    Level &l0 = C->lv[0];
    for (auto &f : l0.hflag) f &= kSolid;
    const int64_t nn = l0.numNodes;
    std::vector<double> ind(C->nDof(), 0.0);
    for (int64_t j = 0; j < nn; ++j) {
        const int64_t g = C->full ? j : C->hNodMap[j];
        for (int a = 0; a < 3; ++a)
            if (fixed[3 * j + a]) {
                l0.hflag[g] |= (uint8_t)(1 << a);
                ind[3 * g + a] = 1.0;
            }
    }
    SGLD_TRY(upload_flags(l0));
    // FEA_SetupVoxelBased.m:13-26: restrict the indicator, fixed iff != 0.
    SGLD_HIP(hipMemcpy(C->r.p, ind.data(), sizeof(double) * ind.size(), hipMemcpyHostToDevice));
    const double *rf = C->r.p;
    for (int k = 1; k < C->L; ++k) {
        Level &lc = C->lv[k];
        SGLD_TRY(enqueue_restrict<double>(C, k, rf, lc.r.p, nullptr));
        std::vector<double> hv(3 * lc.g.nodes());
        SGLD_HIP(hipMemcpyAsync(hv.data(), lc.r.p, sizeof(double) * hv.size(), hipMemcpyDeviceToHost, C->s));
        SGLD_HIP(hipStreamSynchronize(C->s));
        for (int64_t g = 0; g < lc.g.nodes(); ++g) {
            lc.hflag[g] &= kSolid;
            for (int a = 0; a < 3; ++a)
                if (hv[3 * g + a] != 0.0) lc.hflag[g] |= (uint8_t)(1 << a);
        }
        SGLD_TRY(upload_flags(lc));
        rf = lc.r.p;
    }
    C->haveFixed = true;
    C->assembled = false;
    return SGLD_OK;
}

extern "C" int sgld_ctx_set_modulus(sgld_ctx *C, const double *E) {
    SGLD_ARG(C && E, "sgld:arg: null argument");
    const int64_t ne = C->lv[0].numElements;
    for (int64_t i = 0; i < ne; ++i)
        SGLD_ARG(std::isfinite(E[i]) && E[i] > 0.0,
                 "sgld:arg: eleModulus[%lld] = %g must be finite and > 0", (long long)i, E[i]);
    DevBuf<double> tmp;
    SGLD_TRY(tmp.upload(E, ne));
    hipLaunchKernelGGL(k_scatter_elems, dim3(blocks_for(ne, 256)), dim3(256), 0, C->s, tmp.p,
                       C->full ? nullptr : C->eleMap.p, ne, C->E.p, C->E32.p);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipStreamSynchronize(C->s));
    C->haveModulus = true;
    C->assembled = false;
    return SGLD_OK;
}

namespace {
int assemble_impl(sgld_ctx *C, double omega) {
    SGLD_ARG(C, "sgld:arg: null context");
    SGLD_ARG(omega > 0 && omega < 2, "sgld:arg: omega must be in (0, 2)");
    if (!C->haveFixed || !C->haveModulus) {
        set_error("sgld:state: set fixed DOFs and modulus before assembling");
        return SGLD_ERR_STATE;
    }
    const double t0 = now_s();
    C->omega = omega;
    Level &l0 = C->lv[0];
    hipLaunchKernelGGL(k_diag0, dim3(blocks_for(l0.g.nodes(), 256)), dim3(256), 0, C->s, C->KeDev.p, C->E.p, l0.g,
                       l0.nflag.p, omega, l0.dinv.p, l0.dinv32.p);
    SGLD_HIP(hipGetLastError());
    {
        const Span V = dof_span(C);
        hipLaunchKernelGGL(k_pack_w16, dim3(grid_blocks(V.count)), dim3(256), 0, C->s, (const float *)l0.dinv32.p,
                           l0.dinv16.p, V);
        SGLD_HIP(hipGetLastError());
        if (C->dinv16L.p) {  // the same weights in lattice order (fused residual update), refreshed
            hipLaunchKernelGGL(k_pack_w16_lattice, dim3(grid_blocks(V.count)), dim3(256), 0, C->s,
                               (const float *)l0.dinv32.p, C->dinv16L.p, V);
            SGLD_HIP(hipGetLastError());
        }
    }
    DevBuf<double> KsPrev, KsCur;
    for (int k = 1; k < C->L; ++k) {
        Level &lf = C->lv[k - 1], &lc = C->lv[k];
        const int s = lc.span, nq = s * s * s;
        SGLD_TRY(KsCur.alloc(576 * lc.g.elems()));
        if (k == 1) {
            std::vector<double> H(576 * nq, 0.0);
            for (int q = 0; q < nq; ++q) {
                double p[8][8];
                child_prolongation(s, q, p);
                galerkin_child_add(p, C->Ke, H.data() + 576 * q);
            }
            DevBuf<double> dH;
            SGLD_TRY(dH.upload(H.data(), H.size()));
            hipLaunchKernelGGL(k_galerkin_from_E, dim3(bounded_blocks(576 * lc.g.elems(), 256)), dim3(256), 0,
                               C->s, dH.p, s, C->E.p, lf.g, lc.g, KsCur.p);
            SGLD_HIP(hipGetLastError());
            SGLD_HIP(hipStreamSynchronize(C->s));
        } else {
            std::vector<double> P(64 * nq);
            for (int q = 0; q < nq; ++q) {
                double p[8][8];
                child_prolongation(s, q, p);
                std::memcpy(P.data() + 64 * q, p, sizeof(p));
            }
            DevBuf<double> dP;
            SGLD_TRY(dP.upload(P.data(), P.size()));
            hipLaunchKernelGGL(k_galerkin_from_Ks, dim3((unsigned)std::min<int64_t>(lc.g.elems(), 1 << 20)),
                               dim3(576), 0, C->s, KsPrev.p, lf.g, lc.g, s, dP.p, KsCur.p);
            SGLD_HIP(hipGetLastError());
            SGLD_HIP(hipStreamSynchronize(C->s));
        }
        if (k < C->L - 1) {
            hipLaunchKernelGGL(k_diagk, dim3(blocks_for(lc.g.nodes(), 256)), dim3(256), 0, C->s, KsCur.p,
                               lc.g, lc.nflag.p, omega, lc.dinv.p, lc.dinv32.p);
            SGLD_HIP(hipGetLastError());
        }
        KsPrev = std::move(KsCur);
    }
    // Coarsest: K_L on its free DOFs (solid, not fixed), dense FP64 Cholesky
    // K_L = L L^T on the device and the inverse factor W = L^-1 for the two
    // triangular products of the V-cycle (Solving_AssembleFEAstencil.m:123-135,
    // Solving_Vcycle.m:29-31).
    Level &lc = C->lv[C->L - 1];
    const int64_t ne = lc.g.elems(), nn = lc.g.nodes();
    std::vector<int32_t> freeIdx;
    std::vector<int64_t> pos(3 * nn, -1);
    for (int64_t g = 0; g < nn; ++g)
        for (int a = 0; a < 3; ++a)
            if ((lc.hflag[g] & kSolid) && !(lc.hflag[g] & (1 << a))) {
                pos[3 * g + a] = (int64_t)freeIdx.size();
                freeIdx.push_back((int32_t)(3 * g + a));
            }
    const int64_t nf = (int64_t)freeIdx.size();
    if (nf > kCoarseMaxFree) {
        set_error("sgld:arg: the coarsest level (level %d, %dx%dx%d elements) has %lld free DOFs; its dense "
                  "Cholesky factor supports at most %lld: use more levels", C->L, lc.g.nx, lc.g.ny, lc.g.nz,
                  (long long)nf, (long long)kCoarseMaxFree);
        return SGLD_ERR_ARG;
    }
    SGLD_ARG(nf > 0, "sgld:arg: the coarsest level has no free DOF");
    std::vector<double> Ks(576 * ne);
    SGLD_HIP(hipMemcpyAsync(Ks.data(), KsPrev.p, sizeof(double) * Ks.size(), hipMemcpyDeviceToHost, C->s));
    SGLD_HIP(hipStreamSynchronize(C->s));
    KsPrev.release();
    std::vector<double> A((size_t)nf * nf, 0.0);
    const int64_t NY = lc.g.NY(), NX = lc.g.NX();
    for (int64_t e = 0; e < ne; ++e) {
        if (!lc.esolid[e]) continue;
        const int64_t er = e % lc.g.ny, ec = (e / lc.g.ny) % lc.g.nx, el = e / ((int64_t)lc.g.ny * lc.g.nx);
        int64_t dof[24];
        for (int k = 0; k < 8; ++k) {
            const int64_t g = (er + 1 - kCornerY[k]) + NY * (ec + kCornerX[k]) + NY * NX * (el + kCornerZ[k]);
            for (int a = 0; a < 3; ++a) dof[3 * k + a] = pos[3 * g + a];
        }
        for (int c = 0; c < 24; ++c) {
            if (dof[c] < 0) continue;
            for (int r = 0; r < 24; ++r)
                if (dof[r] >= 0) A[dof[r] + nf * dof[c]] += Ks[576 * e + r + 24 * c];
        }
    }
    DevBuf<double> dA, dW;
    DevBuf<unsigned long long> dBad;
    SGLD_TRY(dA.upload(A.data(), A.size()));
    SGLD_TRY(dBad.alloc(1));
    SGLD_TRY(dBad.zero(C->s));
    const int64_t T = (nf + kCB - 1) / kCB;  // tiles per edge
    for (int64_t K = 0; K < T; ++K) {
        const int64_t k0 = kCB * K;
        hipLaunchKernelGGL(k_potrf_tile, dim3(1), dim3(256), 0, C->s, dA.p, nf, k0, dBad.p);
        const int64_t rows = nf - k0 - kCB;  // > 0 only below a full tile
        if (rows <= 0) continue;
        hipLaunchKernelGGL(k_trsm_panel, dim3(blocks_for(rows, 64)), dim3(64), 0, C->s, dA.p, nf, k0);
        const int64_t tt = (rows + kCB - 1) / kCB;
        hipLaunchKernelGGL(k_syrk_tiles, dim3((unsigned)(tt * (tt + 1) / 2)), dim3(256), 0, C->s, dA.p, nf, k0);
    }
    SGLD_HIP(hipGetLastError());
    unsigned long long bad = 0;
    SGLD_HIP(hipMemcpyAsync(&bad, dBad.p, sizeof(bad), hipMemcpyDeviceToHost, C->s));
    SGLD_HIP(hipStreamSynchronize(C->s));
    if (bad) {
        const int64_t dof = freeIdx[bad - 1], g = dof / 3;
        set_error("sgld:notspd: coarsest-level matrix is not positive definite (free DOFs %lld; pivot %lld "
                  "= lattice node (%lld,%lld,%lld) comp %lld, diag %.3e)",
                  (long long)nf, (long long)(bad - 1), (long long)(g % NY), (long long)((g / NY) % NX),
                  (long long)(g / (NY * NX)), (long long)(dof % 3), A[(bad - 1) * (nf + 1)]);
        return SGLD_ERR_NOT_SPD;
    }
    std::vector<double>().swap(A);
    SGLD_TRY(dW.alloc((size_t)nf * nf));
    hipLaunchKernelGGL(k_trinv_tiles, dim3((unsigned)T), dim3(64), 0, C->s, (const double *)dA.p, dW.p, nf);
    for (int64_t I = 1; I < T; ++I)
        hipLaunchKernelGGL(k_trinv_blockrow, dim3((unsigned)I), dim3(256), 0, C->s, (const double *)dA.p, dW.p,
                           nf, I);
    SGLD_HIP(hipGetLastError());
    dA.release();
    SGLD_TRY(C->cWrow.alloc(tri_row_off(nf)));
    SGLD_TRY(C->cWcol.alloc(tri_row_off(nf)));
    hipLaunchKernelGGL(k_tri_pack, dim3(bounded_blocks(nf * nf, 256)), dim3(256), 0, C->s, dW.p, nf, C->cWrow.p,
                       C->cWcol.p);
    // FP32 copies for the FP32 inner solve once the FP64 factor outgrows the
    // Infinity Cache share (8 n^2 bytes > 128 MB, n > 4096)
    C->cWrow32.release();
    C->cWcol32.release();
    if (8.0 * (double)nf * (double)nf > 128.0 * (1 << 20) && !std::getenv("SGLD_COARSE_W64")) {
        const int64_t nw = tri_row_off(nf);
        SGLD_TRY(C->cWrow32.alloc(nw));
        SGLD_TRY(C->cWcol32.alloc(nw));
        const Span all = Span::all(nw);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_blocks(nw)), dim3(256), 0, C->s, C->cWrow32.p,
                           (const double *)C->cWrow.p, 1.0, all);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_blocks(nw)), dim3(256), 0, C->s, C->cWcol32.p,
                           (const double *)C->cWcol.p, 1.0, all);
    }
    SGLD_HIP(hipGetLastError());
    SGLD_TRY(C->cFree.upload(freeIdx.data(), freeIdx.size()));
    SGLD_TRY(C->cY.alloc(16 * nf));
    C->nCoarseFree = nf;
    // the coarsest x is written at free DOFs only: clear the rest (a new mask
    // may have fixed DOFs that were free before)
    SGLD_TRY(lc.x.zero(C->s));
    SGLD_HIP(hipMemsetAsync(lc.x32.p, 0, sizeof(float) * lc.x32.n, C->s));
    SGLD_TRY(batch_clear_coarse(C));
    SGLD_HIP(hipStreamSynchronize(C->s));
    C->stats[4] = now_s() - t0;
    C->assembled = true;
    return SGLD_OK;
}
}  // namespace

extern "C" int sgld_ctx_assemble(sgld_ctx *C, double omega) {
    try {
        if (C) C->assembled = false;
        return assemble_impl(C, omega);
    } catch (const std::bad_alloc &) {
        set_error("sgld:nomem: host allocation failed during assembly");
        return SGLD_ERR_NOMEM;
    }
}

extern "C" int sgld_ctx_kbyu(sgld_ctx *C, const double *u, double *y) {
    SGLD_TRY(check_ready(C));
    SGLD_ARG(u && y, "sgld:arg: null vector");
    const int64_t nc = 3 * C->lv[0].numNodes;
    DevBuf<double> du, dy;
    SGLD_TRY(du.upload(u, nc));
    SGLD_TRY(dy.alloc(nc));
    SGLD_TRY(to_lattice(C, du.p, C->z.p));
    SGLD_TRY(enqueue_kbyu<double>(C, C->z.p, C->q.p, nullptr, nullptr, nullptr));
    SGLD_TRY(from_lattice(C, C->q.p, dy.p));
    SGLD_HIP(hipMemcpyAsync(y, dy.p, sizeof(double) * nc, hipMemcpyDeviceToHost, C->s));
    SGLD_HIP(hipStreamSynchronize(C->s));
    return SGLD_OK;
}

extern "C" int sgld_ctx_vcycle(sgld_ctx *C, const double *r, double *z) {
    SGLD_TRY(check_ready(C));
    SGLD_ARG(r && z, "sgld:arg: null vector");
    const int64_t nc = 3 * C->lv[0].numNodes;
    DevBuf<double> dr, dz;
    SGLD_TRY(dr.upload(r, nc));
    SGLD_TRY(dz.alloc(nc));
    SGLD_TRY(to_lattice(C, dr.p, C->r.p));
    SGLD_TRY(enqueue_vcycle<double>(C, C->r.p, C->z.p, nullptr, nullptr, nullptr));
    SGLD_TRY(from_lattice(C, C->z.p, dz.p));
    SGLD_HIP(hipMemcpyAsync(z, dz.p, sizeof(double) * nc, hipMemcpyDeviceToHost, C->s));
    SGLD_HIP(hipStreamSynchronize(C->s));
    return SGLD_OK;
}

extern "C" int sgld_ctx_pcg_device(sgld_ctx *C, const double *dF, double *dU, double tol, int maxIT,
                                   int precision, int *itsOut, double *relres) {
    SGLD_TRY(check_ready(C));
    SGLD_ARG(dF && dU, "sgld:arg: null vector");
    SGLD_ARG(tol > 0 && maxIT >= 1, "sgld:arg: need tol > 0 and maxIT >= 1");
    SGLD_ARG(precision == SGLD_PREC_FP64 || precision == SGLD_PREC_MIXED, "sgld:arg: unknown precision mode");
    SGLD_HIP(hipDeviceSynchronize());
    const double t0 = now_s();
    SGLD_TRY(to_lattice(C, dF, C->b.p));
    SGLD_TRY(to_lattice(C, dU, C->x.p));
    SolveState S;
    S.precision = precision;
    S.tol = tol;
    S.maxit = maxIT;
    SGLD_TRY(pcg_solve(C, S));
    const int its = S.its, done = S.done, outer = S.updates;
    double tres = 0, normr = 0;
    SGLD_TRY(from_lattice(C, C->x.p, dU));
    SGLD_HIP(hipStreamSynchronize(C->s));
    C->stats[0] = now_s() - t0;
    SGLD_TRY(true_residual(C, &tres, &normr));
    C->stats[1] = its;
    C->stats[2] = tres;
    C->stats[3] = outer;
    if (itsOut) *itsOut = its;
    if (relres) *relres = tres;
    if (done != 1) {
        set_error("sgld:maxit: PCG stopped at maxIT = %d with relative residual %.3e", maxIT, tres);
        return SGLD_ERR_MAXIT;
    }
    return SGLD_OK;
}

extern "C" int sgld_ctx_pcg(sgld_ctx *C, const double *F, double *U, double tol, int maxIT,
                            int precision, int *its, double *relres) {
    SGLD_TRY(check_ready(C));
    SGLD_ARG(F && U, "sgld:arg: null vector");
    const int64_t nc = 3 * C->lv[0].numNodes;
    DevBuf<double> dF, dU;
    SGLD_TRY(dF.upload(F, nc));
    SGLD_TRY(dU.upload(U, nc));
    const int rc = sgld_ctx_pcg_device(C, dF.p, dU.p, tol, maxIT, precision, its, relres);
    if (rc != SGLD_OK && rc != SGLD_ERR_MAXIT) return rc;
    SGLD_HIP(hipMemcpy(U, dU.p, sizeof(double) * nc, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int sgld_ctx_unit_compliance(sgld_ctx *C, const double *U, double *ce) {
    SGLD_ARG(C && U && ce, "sgld:arg: null argument");
    const int64_t nc = 3 * C->lv[0].numNodes, ne = C->lv[0].numElements;
    DevBuf<double> dU, dce;
    SGLD_TRY(dU.upload(U, nc));
    SGLD_TRY(dce.alloc(ne));
    SGLD_TRY(to_lattice(C, dU.p, C->z.p));
    hipLaunchKernelGGL(k_compliance, dim3(blocks_for(ne, 256)), dim3(256), 0, C->s, C->KeDev.p, C->z.p, C->lv[0].g,
                       C->full ? nullptr : C->eleMap.p, ne, dce.p);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpyAsync(ce, dce.p, sizeof(double) * ne, hipMemcpyDeviceToHost, C->s));
    SGLD_HIP(hipStreamSynchronize(C->s));
    return SGLD_OK;
}

extern "C" int sgld_ctx_kbyu_device(sgld_ctx *C, const double *du, double *dy, int precision) {
    SGLD_ARG(C && du && dy, "sgld:arg: null argument");
    SGLD_ARG(C->haveModulus && C->haveFixed, "sgld:state: set fixed DOFs and modulus first");
    SGLD_TRY(to_lattice(C, du, C->z.p));
    if (precision == SGLD_PREC_FP64) {
        SGLD_HIP(hipEventRecord(C->e0, C->s));
        SGLD_TRY(enqueue_kbyu<double>(C, C->z.p, C->q.p, nullptr, nullptr, nullptr));
        SGLD_HIP(hipEventRecord(C->e1, C->s));
        SGLD_TRY(from_lattice(C, C->q.p, dy));
    } else {
        const Span V = dof_span(C);
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_blocks(V.count)), dim3(256), 0, C->s, C->z32.p,
                           C->z.p, 1.0, V);
        SGLD_HIP(hipEventRecord(C->e0, C->s));
        SGLD_TRY(enqueue_kbyu<float>(C, C->z32.p, C->q32.p, nullptr, nullptr, nullptr));
        SGLD_HIP(hipEventRecord(C->e1, C->s));
        hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_blocks(V.count)), dim3(256), 0, C->s, C->q.p,
                           C->q32.p, 1.0, V);
        SGLD_TRY(from_lattice(C, C->q.p, dy));
    }
    SGLD_HIP(hipStreamSynchronize(C->s));
    float ms = 0;
    SGLD_HIP(hipEventElapsedTime(&ms, C->e0, C->e1));
    C->stats[precision == SGLD_PREC_FP64 ? 6 : 7] = ms;
    return SGLD_OK;
}

extern "C" int sgld_ctx_bench_kbyu(sgld_ctx *C, int precision, int reps, double *msPerLaunch) {
    SGLD_ARG(C && msPerLaunch && reps > 0, "sgld:arg: bad argument");
    SGLD_ARG(C->haveModulus && C->haveFixed, "sgld:state: set fixed DOFs and modulus first");
    SGLD_HIP(hipStreamSynchronize(C->s));
    SGLD_HIP(hipEventRecord(C->e0, C->s));
    for (int i = 0; i < reps; ++i) {
        if (precision == SGLD_PREC_FP64)
            SGLD_TRY(enqueue_kbyu<double>(C, C->p.p, C->q.p, nullptr, nullptr, nullptr));
        else
            SGLD_TRY(enqueue_kbyu<float>(C, C->p32.p, C->q32.p, nullptr, nullptr, nullptr));
    }
    SGLD_HIP(hipEventRecord(C->e1, C->s));
    SGLD_HIP(hipEventSynchronize(C->e1));
    float ms = 0;
    SGLD_HIP(hipEventElapsedTime(&ms, C->e0, C->e1));
    *msPerLaunch = ms / reps;
    return SGLD_OK;
}

extern "C" int sgld_ctx_stats(const sgld_ctx *C, double *out) {
    SGLD_ARG(C && out, "sgld:arg: null argument");
    for (int i = 0; i < 8; ++i) out[i] = C->stats[i];
    return SGLD_OK;
}

extern "C" int sgld_ctx_kernel_timing(sgld_ctx *C, int on) {
    SGLD_ARG(C, "sgld:arg: null context");
    C->kt.on = on != 0;
    C->kt.reset();
    return SGLD_OK;
}

extern "C" const char *sgld_ctx_kernel_name(const sgld_ctx *C, int id) {
    if (!C || id < 0 || id >= KernelTimer::kN || !C->kt.name[id]) return "";
    return C->kt.name[id];
}

extern "C" int sgld_ctx_kernel_times(const sgld_ctx *C, double *out) {
    SGLD_ARG(C && out, "sgld:arg: null argument");
    for (int k = 0; k < KernelTimer::kN; ++k) {
        out[4 * k + 0] = (double)C->kt.n[k];
        out[4 * k + 1] = C->kt.ms[k];
        out[4 * k + 2] = (double)C->kt.nWork[k];
        out[4 * k + 3] = C->kt.msWork[k];
    }
    return SGLD_OK;
}

// ------------------------- sensitivities, autotuning ------------------------
SensA sens_args(sgld_ctx *C, const double *u, const double *dx, double penal, double E0, double Emin,
                double scale, double *ddc) {
    return SensA{C->KeDev.p, u, C->lv[0].g, C->full ? nullptr : C->eleMap.p, C->lv[0].numElements, dx, C->E.p,
                 penal, E0 - Emin, scale, ddc, slot(C, 3), C->km64, C->haveModes};
}

extern "C" int sgld_multi_sensitivity_device(sgld_ctx *const *ctxs, int n, const double *const *dU,
                                             const double *const *dx, double penal, double E0, double Emin,
                                             double scale, double *const *ddc, double *compliance);

// One instance through the merged launch: the kernel the merged call runs,
// so that its results and the merged call's agree bit for bit (two kernels
// built from one body may round differently).
extern "C" int sgld_ctx_sensitivity_device(sgld_ctx *C, const double *dU, const double *dx, double penal,
                                           double E0, double Emin, double scale, double *ddc, double *compliance) {
    SGLD_ARG(C && dU && dx && ddc, "sgld:arg: null argument");
    SGLD_ARG(C->haveModulus, "sgld:state: call sgld_ctx_set_modulus first");
    SGLD_ARG(penal >= 1.0, "sgld:arg: penal must be >= 1");
    double c = 0.0;
    SGLD_TRY(sgld_multi_sensitivity_device(&C, 1, &dU, &dx, penal, E0, Emin, scale, &ddc, &c));
    if (compliance) *compliance = c;
    return SGLD_OK;
}

extern "C" int sgld_multi_sensitivity_device(sgld_ctx *const *ctxs, int n, const double *const *dU,
                                             const double *const *dx, double penal, double E0, double Emin,
                                             double scale, double *const *ddc, double *compliance) {
    SGLD_ARG(ctxs && dU && dx && ddc && compliance && n >= 1, "sgld:arg: null argument or n < 1");
    SGLD_ARG(penal >= 1.0, "sgld:arg: penal must be >= 1");
    for (int i = 0; i < n; ++i) {
        SGLD_ARG(ctxs[i] && dU[i] && dx[i] && ddc[i], "sgld:arg: null argument of instance %d", i);
        SGLD_ARG(ctxs[i]->haveModulus, "sgld:state: instance %d: call sgld_ctx_set_modulus first", i);
        for (int k = 0; k < i; ++k) SGLD_ARG(ctxs[k] != ctxs[i], "sgld:arg: instance %d repeats a context", i);
    }
    // full cuboids read U where it lies; the others scatter it onto their
    // lattice first (on their own streams, which are drained before the
    // merged launch on the first context's stream)
    std::vector<SensA> sa(n);
    std::vector<SumA> su(n);
    std::vector<int> start(n + 1, 0), iota(n + 1);
    DevBuf<double> out;
    SGLD_TRY(out.alloc(n));
    for (int i = 0; i < n; ++i) {
        sgld_ctx *C = ctxs[i];
        const double *u = dU[i];
        if (!C->full) {
            SGLD_TRY(to_lattice(C, dU[i], C->z.p));
            u = C->z.p;
        }
        const int nb = grid_blocks(C->lv[0].numElements);
        sa[i] = sens_args(C, u, dx[i], penal, E0, Emin, scale, ddc[i]);
        su[i] = SumA{slot(C, 3), nb, out.p + i};
        start[i + 1] = start[i] + nb;
        iota[i] = i;
    }
    iota[n] = n;
    for (int i = 0; i < n; ++i) SGLD_HIP(hipStreamSynchronize(ctxs[i]->s));
    DevBuf<SensA> dsa;
    DevBuf<SumA> dsu;
    DevBuf<int> dstart, diota;
    SGLD_TRY(dsa.upload(sa.data(), n));
    SGLD_TRY(dsu.upload(su.data(), n));
    SGLD_TRY(dstart.upload(start.data(), n + 1));
    SGLD_TRY(diota.upload(iota.data(), n + 1));
    hipStream_t s = ctxs[0]->s;
    multi_launch<OpSens>(dsa.p, dstart.p, n, start[n], 0, s);
    SGLD_HIP(hipGetLastError());
    multi_launch<OpSum>(dsu.p, diota.p, n, n, 0, s);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpyAsync(compliance, out.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    SGLD_HIP(hipStreamSynchronize(s));
    return SGLD_OK;
}

extern "C" int sgld_ctx_sensitivity(sgld_ctx *C, const double *U, const double *x, double penal, double E0,
                                    double Emin, double scale, double *dc, double *compliance) {
    SGLD_ARG(C && U && x && dc, "sgld:arg: null argument");
    const int64_t nc = 3 * C->lv[0].numNodes, ne = C->lv[0].numElements;
    DevBuf<double> dU, dx, ddc;
    SGLD_TRY(dU.upload(U, nc));
    SGLD_TRY(dx.upload(x, ne));
    SGLD_TRY(ddc.alloc(ne));
    SGLD_TRY(sgld_ctx_sensitivity_device(C, dU.p, dx.p, penal, E0, Emin, scale, ddc.p, compliance));
    SGLD_HIP(hipMemcpy(dc, ddc.p, sizeof(double) * ne, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_ctx_autotune(sgld_ctx *C, int reps, double *report) {
    SGLD_ARG(C && report && reps > 0, "sgld:arg: bad arguments");
    SGLD_ARG(C->haveModulus && C->haveFixed, "sgld:state: call sgld_ctx_set_fixed and sgld_ctx_set_modulus first");
    const int variant0 = C->kbyuVariant;
    const int czs[4] = {8, 12, 16, 32};
    double tiled[2] = {1e300, 1e300};
    int bestCz[2] = {C->ksched[0].tm.cz, C->ksched[1].tm.cz};
    if (variant0 == 1) {
        // each precision's z-chunk on its own (persistent grid = resident blocks)
        for (int q = 0; q < 2; ++q)
            for (int cz : czs) {
                SGLD_TRY(set_kchunk(C, q, cz));
                double t = 0;
                SGLD_TRY(sgld_ctx_bench_kbyu(C, q == 0 ? SGLD_PREC_FP64 : SGLD_PREC_MIXED, reps, &t));
                if (t < tiled[q]) {
                    tiled[q] = t;
                    bestCz[q] = cz;
                }
            }
        for (int q = 0; q < 2; ++q) SGLD_TRY(set_kchunk(C, q, bestCz[q]));
    }
    // against the node-centric dense K.u; the mixed solver spends its
    // iterations in FP32 K.u and one FP64 K.u per reliable update
    C->kbyuVariant = 0;
    double dense64 = 0, dense32 = 0;
    SGLD_TRY(sgld_ctx_bench_kbyu(C, SGLD_PREC_FP64, reps, &dense64));
    SGLD_TRY(sgld_ctx_bench_kbyu(C, SGLD_PREC_MIXED, reps, &dense32));
    const bool useTiled = variant0 == 1 && tiled[1] + tiled[0] / 8.0 <= dense32 + dense64 / 8.0;
    C->kbyuVariant = useTiled ? 1 : 0;
    // span-2 restriction variant, timed on FP32 V-cycles (needs the assembled hierarchy)
    int bestR = C->restrictTiled;
    double tv[2] = {0, 0};
    if (C->assembled) {
        for (int rt = 0; rt < 2; ++rt) {
            C->restrictTiled = rt;
            int nb = 0;
            SGLD_TRY(enqueue_vcycle<float>(C, C->r32.p, C->z32.p, slot(C, 2), nullptr, &nb));
            SGLD_HIP(hipEventRecord(C->e0, C->s));
            for (int i = 0; i < reps; ++i) SGLD_TRY(enqueue_vcycle<float>(C, C->r32.p, C->z32.p, slot(C, 2), nullptr, &nb));
            SGLD_HIP(hipEventRecord(C->e1, C->s));
            SGLD_HIP(hipEventSynchronize(C->e1));
            float ms = 0;
            SGLD_HIP(hipEventElapsedTime(&ms, C->e0, C->e1));
            tv[rt] = ms / reps;
        }
        bestR = tv[1] < tv[0] ? 1 : 0;
    }
    C->restrictTiled = bestR;
    report[0] = useTiled ? 1 : 0;
    report[1] = useTiled ? bestCz[0] : 0;
    report[2] = useTiled ? tiled[0] : dense64;
    report[3] = useTiled ? tiled[1] : dense32;
    report[4] = bestR;
    report[5] = tv[0];
    report[6] = tv[1];
    report[7] = useTiled ? bestCz[1] : 0;
    return SGLD_OK;
}

// --------------------- many instances, one replayed graph --------------------
namespace {

// The reliable update of the mixed solve as device launches, recorded for
// the merged solve ahead of each iteration (steps before the iteration's):
// check (gate open when the FP32 stretch has stopped or hit its length
// limit), fold, r = f - K x in FP64, decide (stop, or set up the stretch),
// and rho = z.r of the new residual through the V-cycle, all gated.
int enqueue_mixed_update_device(sgld_ctx *C, MixCtl *mix, Ctl *gate) {
    const Span V = dof_span(C);
    const int nbv = grid_blocks(V.count);  // virtual blocks: the single-solve grid of these passes
    static const int coarsen = [] {
        const char *env = std::getenv("SGLD_MULTI_UPDATE_COARSEN");
        return env ? std::max(1, std::atoi(env)) : 4;
    }();
    const int nbu = (nbv + coarsen - 1) / coarsen;  // launched blocks
    SGLD_TRY(emit<OpMixCheck>(C, 1, 1, 0, MixCheckA{C->ctl.p, gate, mix}));
    SGLD_TRY(emit<OpMixFold>(C, 2, nbu, 0, MixFoldA{C->x.p, C->x32.p, C->p32.p, V, C->ctl.p, gate, mix, nbv}));
    SGLD_TRY(enqueue_kbyu<double>(C, C->x.p, C->q.p, nullptr, gate, nullptr));  // step ST_KU
    SGLD_TRY(emit<OpMixNorms>(C, ST_KU + 10, nbu, 0,
                              MixNormsA{C->b.p, C->q.p, C->r32.p, C->lv[0].nflag.p, V, slot(C, 0), slot(C, 1), gate,
                                        mix, nbv}));
    SGLD_TRY(emit<OpMixDecide>(C, ST_KU + 20, 1, 0, MixDecideA{slot(C, 0), slot(C, 1), nbv, C->ctl.p, gate, mix}));
    SGLD_TRY(emit<OpMixDinvRR>(C, ST_KU + 30, nbu, 0,
                               MixDinvRRA{C->r32.p, C->lv[0].dinv16.p, V, slot(C, 3), gate, nbv}));
    int nb2 = 0;
    SGLD_TRY(enqueue_vcycle_coarse<float>(C, C->r32.p, slot(C, 2), &nb2, gate));
    return emit<OpMixFinRho>(C, ST_FIN_RHO, 1, 0,
                             MixFinRhoA{slot(C, 2), nb2, slot(C, 3), nbv, C->ctl.p, gate, mix});
}

// The multi-instance iteration: each instance's PCG iteration is recorded
// once (enqueue_iteration with the context's recorder set), and the recorded
// launches of all instances are merged step by step into launches that cover
// every instance (k_multi: instance i runs blocks [start[i], start[i+1])).
// Launches of one step with different kernel instantiations (span, staged
// outputs, coarse variant, tail depth) stay separate launches.
struct MultiIteration {
    struct Launch {
        void (*multi)(const void *, const int *, int, int, size_t, hipStream_t);
        int n, grid, step;
        size_t shm;
        DevBuf<char> args;
        DevBuf<int> start;
    };
    std::vector<Launch> launches;
    // the iteration's launches (steps >= kIterBase), or the update block ahead of it
    int enqueue(hipStream_t s, bool update, bool iteration) const {
        for (const Launch &L : launches) {
            if (L.step >= kIterBase ? !iteration : !update) continue;
            L.multi(L.args.p, L.start.p, L.n, L.grid, L.shm, s);
            SGLD_HIP(hipGetLastError());
        }
        return SGLD_OK;
    }
    static constexpr int kIterBase = 100000;
};

// mix / gate (mixed precision): per live instance, the device update state;
// its update block is recorded ahead of the iteration.
int build_multi_iteration(sgld_ctx *const *ctxs, const std::vector<int> &live, int precision, MultiIteration &M,
                          MixCtl *mix = nullptr, Ctl *gate = nullptr) {
    std::vector<OpRec> rec(live.size());
    for (size_t k = 0; k < live.size(); ++k) {
        sgld_ctx *C = ctxs[live[k]];
        C->rec = &rec[k];
        int rc = SGLD_OK;
        if (precision != SGLD_PREC_FP64 && mix) rc = enqueue_mixed_update_device(C, mix + k, gate + k);
        rec[k].base = MultiIteration::kIterBase;  // the iteration's launches, in both precisions
        if (rc == SGLD_OK)
            rc = precision == SGLD_PREC_FP64 ? enqueue_iteration<double>(C, C->x.p, C->r.p, C->p.p, C->q.p)
                                             : enqueue_iteration<float>(C, C->x32.p, C->r32.p, C->p32.p, C->q32.p);
        C->rec = nullptr;
        if (rc != SGLD_OK) return rc;
    }
    struct E {
        int step;
        const RecOp *op;
    };
    std::vector<E> all;
    for (size_t k = 0; k < live.size(); ++k)
        for (const RecOp &o : rec[k].ops) all.push_back(E{o.step, &o});
    // by step, then kernel instantiation; instances keep their order (stable)
    std::stable_sort(all.begin(), all.end(), [](const E &a, const E &b) {
        if (a.step != b.step) return a.step < b.step;
        return (uintptr_t)a.op->multi < (uintptr_t)b.op->multi;
    });
    M.launches.clear();
    M.launches.reserve(all.size());
    for (size_t i = 0; i < all.size();) {
        size_t j = i;
        while (j < all.size() && all[j].step == all[i].step && all[j].op->multi == all[i].op->multi) ++j;
        const size_t sz = all[i].op->args.size();
        std::vector<char> args(sz * (j - i));
        std::vector<int> start(j - i + 1, 0);
        size_t shm = 0;
        for (size_t q = i; q < j; ++q) {
            std::memcpy(args.data() + sz * (q - i), all[q].op->args.data(), sz);
            start[q - i + 1] = start[q - i] + all[q].op->grid;
            shm = std::max(shm, all[q].op->shm);
        }
        M.launches.emplace_back();
        MultiIteration::Launch &L = M.launches.back();
        L.multi = all[i].op->multi;
        L.step = all[i].step;
        L.n = (int)(j - i);
        L.grid = start.back();
        L.shm = shm;
        SGLD_TRY(L.args.upload(args.data(), args.size()));
        SGLD_TRY(L.start.upload(start.data(), start.size()));
        i = j;
    }
    return SGLD_OK;
}

// Multi-instance solve on merged launches (see MultiIteration): one graph
// holds `batch` iterations of every live instance and the copy of all their
// control blocks; between graphs each instance's host step runs on its own
// stream (convergence; reliable updates in mixed precision).
int multi_pcg_merged(sgld_ctx *const *ctxs, int n, std::vector<SolveState> &S, int precision, int batch,
                     double t0, double *latency) {
    std::vector<int> live;
    for (int i = 0; i < n; ++i)
        if (!S[i].finished) live.push_back(i);
    if (live.empty()) return SGLD_OK;
    static const bool times = std::getenv("SGLD_MULTI_TIMES") != nullptr;
    const double tb0 = now_s();
    const int nl = (int)live.size();
    const bool mixed = precision != SGLD_PREC_FP64;
    // mixed precision: the reliable updates run on the device (see
    // enqueue_mixed_update_device), seeded from each solve's begin
    DevBuf<MixCtl> mix;
    DevBuf<Ctl> gate;
    std::vector<MixCtl> hmix(nl);
    if (mixed) {
        std::vector<Ctl> hg(nl);
        for (int k = 0; k < nl; ++k) {
            const SolveState &st = S[live[k]];
            hmix[k] = MixCtl{st.scale, st.normb, st.delta, st.tol, st.res, 0, st.maxStretch, 0, 0, 1, 0};
            hg[k] = Ctl{};
            hg[k].done = 1;
        }
        SGLD_TRY(mix.upload(hmix.data(), nl));
        SGLD_TRY(gate.upload(hg.data(), nl));
    }
    MultiIteration M;
    SGLD_TRY(build_multi_iteration(ctxs, live, precision, M, mixed ? mix.p : nullptr, mixed ? gate.p : nullptr));
    std::vector<Ctl *> ptrs(nl);
    for (int k = 0; k < nl; ++k) ptrs[k] = ctxs[live[k]]->ctl.p;
    DevBuf<Ctl *> dptrs;
    DevBuf<Ctl> gathered;
    SGLD_TRY(dptrs.upload(ptrs.data(), nl));
    SGLD_TRY(gathered.alloc(nl));
    struct Res {
        hipStream_t s = nullptr;
        Ctl *host = nullptr;
        MixCtl *hostMix = nullptr;
        hipGraph_t g = nullptr;
        hipGraphExec_t x = nullptr;
        ~Res() {
            if (x) (void)hipGraphExecDestroy(x);
            if (g) (void)hipGraphDestroy(g);
            if (host) (void)hipHostFree(host);
            if (hostMix) (void)hipHostFree(hostMix);
            if (s) (void)hipStreamDestroy(s);
        }
    } R;
    SGLD_HIP(hipStreamCreateWithFlags(&R.s, hipStreamNonBlocking));
    SGLD_HIP(hipHostMalloc((void **)&R.host, sizeof(Ctl) * nl, hipHostMallocDefault));
    SGLD_HIP(hipHostMalloc((void **)&R.hostMix, sizeof(MixCtl) * nl, hipHostMallocDefault));
    SGLD_HIP(hipStreamBeginCapture(R.s, hipStreamCaptureModeThreadLocal));
    int rc = SGLD_OK;
    // mixed: the device update block ahead of every `every`-th iteration (a
    // stopped stretch waits at most every - 1 iterations for its update).  Its
    // launches cost even when no instance updates (their grids are sized for
    // an update), so a block per iteration measured slower: 32 uneven cuboids
    // (31.9 M DOF) 85.7 ms of graphs at every = 1, 77.0 at 2, 74.3 at 4 (67.3
    // with the owner K.u throughout).
    static const int every = [] {
        const char *env = std::getenv("SGLD_MULTI_UPDATE_EVERY");
        return env ? std::max(1, std::atoi(env)) : 4;
    }();
    for (int b = 0; b < batch && rc == SGLD_OK; ++b) rc = M.enqueue(R.s, mixed && b % every == 0, true);
    if (rc == SGLD_OK) {
        hipLaunchKernelGGL(k_gather_ctl, dim3((nl + 63) / 64), dim3(64), 0, R.s, (Ctl *const *)dptrs.p, nl,
                           gathered.p);
        if (hipGetLastError() != hipSuccess) rc = SGLD_ERR_HIP;
    }
    if (rc == SGLD_OK && hipMemcpyAsync(R.host, gathered.p, sizeof(Ctl) * nl, hipMemcpyDeviceToHost, R.s) != hipSuccess)
        rc = SGLD_ERR_HIP;
    if (rc == SGLD_OK && mixed &&
        hipMemcpyAsync(R.hostMix, mix.p, sizeof(MixCtl) * nl, hipMemcpyDeviceToHost, R.s) != hipSuccess)
        rc = SGLD_ERR_HIP;
    const hipError_t e = hipStreamEndCapture(R.s, &R.g);
    if (rc != SGLD_OK) return rc;
    SGLD_HIP(e);
    SGLD_HIP(hipGraphInstantiate(&R.x, R.g, nullptr, nullptr, 0));
    const double tb1 = now_s();
    double tGraph = 0.0, tHost = 0.0;
    int nBatches = 0;
    int remaining = nl;
    std::vector<char> counted(nl, 0);
    // every solve stops within maxit iterations (done = 2); in mixed precision
    // a stopped stretch idles at most `every` iterations for its update
    int maxit = 0;
    for (int k = 0; k < nl; ++k) maxit = std::max(maxit, S[live[k]].maxit);
    const int maxBatches = (every + 1) * (maxit / batch + 2) + 8;
    while (remaining > 0) {
        if (nBatches >= maxBatches) {
            set_error("sgld:internal: the merged multi-instance solve did not stop within %d batches", maxBatches);
            return SGLD_ERR_HIP;
        }
        const double tg = now_s();
        SGLD_HIP(hipGraphLaunch(R.x, R.s));
        SGLD_HIP(hipStreamSynchronize(R.s));
        tGraph += now_s() - tg;
        ++nBatches;
        const double th = now_s();
        // host steps: FP64 convergence (and the stopped iteration's deferred
        // update); mixed: the device state says which solves have finished
        for (int k = 0; k < nl; ++k) {
            const int i = live[k];
            if (S[i].finished) continue;
            sgld_ctx *C = ctxs[i];
            *C->hctl = R.host[k];
            if (!mixed) {
                SGLD_TRY(pcg_after_batch(C, S[i]));
            } else if (R.hostMix[k].finished) {
                const MixCtl &m = R.hostMix[k];
                S[i].finished = true;
                S[i].its = R.host[k].its;
                S[i].updates = m.updates;
                S[i].res = m.res;
                S[i].done = m.doneCode;
            }
        }
        for (int k = 0; k < nl; ++k) {
            const int i = live[k];
            if (S[i].finished && !counted[k]) {
                counted[k] = 1;
                --remaining;
                if (latency) latency[i] = now_s() - t0;
            }
        }
        if (!mixed)
            for (int k = 0; k < nl; ++k) SGLD_HIP(hipStreamSynchronize(ctxs[live[k]]->s));
        tHost += now_s() - th;
    }
    if (times) {
        std::string steps;
        for (const MultiIteration::Launch &L : M.launches)
            steps += " " + std::to_string(L.step) + ":" + std::to_string(L.n) + "/" + std::to_string(L.grid);
        std::fprintf(stderr, "[multi] launches (step:instances/blocks):%s\n", steps.c_str());
    }
    if (times)
        std::fprintf(stderr, "[multi] since start %.2f ms: build+capture %.2f ms, %d batches: graphs %.2f ms, host steps %.2f ms, %zu launches per iteration\n",
                     1e3 * (tb0 - t0), 1e3 * (tb1 - tb0), nBatches, 1e3 * tGraph, 1e3 * tHost, M.launches.size());
    return SGLD_OK;
}

}  // namespace

extern "C" int sgld_multi_pcg_device(sgld_ctx *const *ctxs, int n, const double *const *dF, double *const *dU,
                                     double tol, int maxIT, int precision, int batch, int *its, double *relres,
                                     double *latency) {
    SGLD_ARG(ctxs && dF && dU && its && relres && n >= 1, "sgld:arg: null argument or n < 1");
    SGLD_ARG(tol > 0 && maxIT >= 1, "sgld:arg: need tol > 0 and maxIT >= 1");
    SGLD_ARG(precision == SGLD_PREC_FP64 || precision == SGLD_PREC_MIXED, "sgld:arg: unknown precision mode");
    SGLD_ARG(batch >= 2 && batch <= 64 && batch % 2 == 0, "sgld:arg: batch must be even, 2..64");
    for (int i = 0; i < n; ++i) {
        SGLD_TRY(check_ready(ctxs[i]));
        SGLD_ARG(dF[i] && dU[i], "sgld:arg: null vector of instance %d", i);
        for (int k = 0; k < i; ++k) SGLD_ARG(ctxs[k] != ctxs[i], "sgld:arg: instance %d repeats a context", i);
    }
    SGLD_HIP(hipDeviceSynchronize());
    static const bool perInstanceGraphs = [] {
        const char *env = std::getenv("SGLD_MULTI_GRAPHS");
        return env && std::atoi(env) != 0;
    }();
    // The merged launches amortise launch cost over the instances, so the
    // V-cycle runs level by level there: the fused small-level tail trades
    // launches for work (its direct restrictions to the coarsest levels read
    // each fine node 27 + 343 + 3375 times for three fused levels instead of
    // 27 per level), which a merged launch pays for every instance.
    struct TailGuard {
        std::vector<std::pair<sgld_ctx *, bool>> saved;
        ~TailGuard() {
            for (auto &c : saved) c.first->tailOn = c.second;
        }
    } tailGuard;
    if (!perInstanceGraphs)
        for (int i = 0; i < n; ++i) {
            tailGuard.saved.push_back({ctxs[i], ctxs[i]->tailOn});
            ctxs[i]->tailOn = false;
        }
    // A merged launch streams every instance's vectors: its streaming passes
    // take nontemporal loads and stores once the four or more vectors a pass
    // streams over all instances outgrow the 256 MB Infinity Cache (64 MiB
    // per vector), whatever each instance's own size (32 uneven cuboids,
    // 122 MiB per FP32 vector in all: mixed 69 -> 67 ms, FP64 90 -> 87 ms).
    // Cache policy only: each instance's results are unchanged.
    struct NtGuard {
        std::vector<std::pair<sgld_ctx *, std::pair<bool, bool>>> saved;
        ~NtGuard() {
            for (auto &c : saved) {
                c.first->streamNT[0] = c.second.first;
                c.first->streamNT[1] = c.second.second;
            }
        }
    } ntGuard;
    if (!perInstanceGraphs && !std::getenv("SGLD_STREAM_NT")) {
        double total = 0.0;
        for (int i = 0; i < n; ++i) total += (double)ctxs[i]->nDof();
        const bool nt64 = total * 8.0 > 64.0 * (1 << 20), nt32 = total * 4.0 > 64.0 * (1 << 20);
        for (int i = 0; i < n; ++i) {
            ntGuard.saved.push_back({ctxs[i], {ctxs[i]->streamNT[0], ctxs[i]->streamNT[1]}});
            ctxs[i]->streamNT[0] = ctxs[i]->streamNT[0] || nt64;
            ctxs[i]->streamNT[1] = ctxs[i]->streamNT[1] || nt32;
        }
    }
    // A mixed instance whose FP32 stretch has reached its tolerance keeps
    // iterating until the next update block instead of idling there (the
    // blocks run every few iterations): 32 cuboids, merged iterations for the
    // slowest instance and solve time below.  SGLD_MULTI_SOFT=0 idles, as a
    // single solve does, so that each instance repeats its single solve bit for bit.
    const char *softEnv = std::getenv("SGLD_MULTI_SOFT");
    const bool soft = !perInstanceGraphs && (!softEnv || std::atoi(softEnv) != 0);
    const double t0 = now_s();
    std::vector<SolveState> S(n);
    for (int i = 0; i < n; ++i) {
        sgld_ctx *C = ctxs[i];
        S[i].precision = precision;
        S[i].tol = tol;
        S[i].maxit = maxIT;
        S[i].batch = batch;
        S[i].soft = soft;
        SGLD_TRY(to_lattice(C, dF[i], C->b.p));
        SGLD_TRY(to_lattice(C, dU[i], C->x.p));
        SGLD_TRY(pcg_begin_enqueue(C, S[i]));  // every instance's set-up queued before the first wait
    }
    for (int i = 0; i < n; ++i) {
        SGLD_HIP(hipStreamSynchronize(ctxs[i]->s));
        SGLD_TRY(pcg_begin_finish(ctxs[i], S[i]));
    }
    for (int i = 0; i < n; ++i) {
        SGLD_HIP(hipStreamSynchronize(ctxs[i]->s));
        if (latency) latency[i] = S[i].finished ? now_s() - t0 : -1.0;
    }
    bool merged = !perInstanceGraphs;
    if (merged) {
        const int rc = multi_pcg_merged(ctxs, n, S, precision, batch, t0, latency);
        if (rc == SGLD_ERR_ARG) merged = false;  // a kernel variant without a merged launch: per-instance graphs
        else if (rc != SGLD_OK) return rc;
    }
    // each unfinished instance's batch of iterations is captured once as a
    // graph; the host re-launches it on the instance's own stream as soon as
    // the previous batch of that instance has completed (the streams map onto
    // the device's hardware queues, so instances overlap)
    struct Graphs {
        std::vector<hipGraph_t> g;
        std::vector<hipGraphExec_t> x;
        ~Graphs() {
            for (auto e : x)
                if (e) (void)hipGraphExecDestroy(e);
            for (auto h : g)
                if (h) (void)hipGraphDestroy(h);
        }
    } G;
    G.g.assign(n, nullptr);
    G.x.assign(n, nullptr);
    for (int i = 0; i < n && !merged; ++i) {
        if (S[i].finished) continue;
        sgld_ctx *C = ctxs[i];
        SGLD_HIP(hipStreamBeginCapture(C->s, hipStreamCaptureModeThreadLocal));
        const int rc = pcg_enqueue_batch(C, S[i]);
        const hipError_t e = hipStreamEndCapture(C->s, &G.g[i]);
        if (rc != SGLD_OK) return rc;
        SGLD_HIP(e);
        SGLD_HIP(hipGraphInstantiate(&G.x[i], G.g[i], nullptr, nullptr, 0));
    }
    int live = 0;
    for (int i = 0; i < n && !merged; ++i) {
        if (S[i].finished) continue;
        SGLD_HIP(hipGraphLaunch(G.x[i], ctxs[i]->s));
        ++live;
    }
    while (live) {
        bool progressed = false;
        for (int i = 0; i < n; ++i) {
            if (S[i].finished) continue;
            const hipError_t q = hipStreamQuery(ctxs[i]->s);
            if (q == hipErrorNotReady) continue;
            SGLD_HIP(q);
            progressed = true;
            SGLD_TRY(pcg_after_batch(ctxs[i], S[i]));
            if (S[i].finished) {
                if (latency) latency[i] = now_s() - t0;
                --live;
            } else {
                SGLD_HIP(hipGraphLaunch(G.x[i], ctxs[i]->s));  // ordered after any update work
            }
        }
        if (!progressed) std::this_thread::yield();
    }
    for (int i = 0; i < n; ++i) SGLD_HIP(hipStreamSynchronize(ctxs[i]->s));
    int allConv = 1;
    for (int i = 0; i < n; ++i) {  // every instance's copy-back and true residual queued before the first wait
        sgld_ctx *C = ctxs[i];
        SGLD_TRY(from_lattice(C, C->x.p, dU[i]));
        SGLD_TRY(true_residual_enqueue(C));
    }
    for (int i = 0; i < n; ++i) {
        sgld_ctx *C = ctxs[i];
        SGLD_HIP(hipStreamSynchronize(C->s));
        double tres = 0, normr = 0;
        true_residual_sums(C, &tres, &normr);
        its[i] = S[i].its;
        relres[i] = tres;
        C->stats[1] = S[i].its;
        C->stats[2] = tres;
        C->stats[3] = S[i].updates;
        if (S[i].done != 1) allConv = 0;
    }
    for (int i = 0; i < n; ++i) ctxs[i]->stats[0] = now_s() - t0;
    if (std::getenv("SGLD_MULTI_TIMES")) std::fprintf(stderr, "[multi] total %.2f ms\n", 1e3 * (now_s() - t0));
    if (!allConv) {
        set_error("sgld:maxit: not every instance reached tol %.3e within %d iterations", tol, maxIT);
        return SGLD_ERR_MAXIT;
    }
    return SGLD_OK;
}

extern "C" int sgld_ctx_set_schedule(sgld_ctx *C, int kbyuVariant, int zChunk, int restrictTiled,
                                     int blocksPerCU) {
    SGLD_ARG(C, "sgld:arg: null context");
    SGLD_ARG(blocksPerCU >= 0 && blocksPerCU <= 8, "sgld:arg: blocksPerCU must be in 0..8 (0 = occupancy)");
    SGLD_ARG(kbyuVariant == 0 || kbyuVariant == 1, "sgld:arg: kbyuVariant must be 0 (dense) or 1 (tiled)");
    SGLD_ARG(restrictTiled >= -1 && restrictTiled <= 2, "sgld:arg: restrictTiled must be -1, 0, 1 or 2");
    if (kbyuVariant == 1) {
        double km[9];
        SGLD_ARG(mode_coefficients(C->Ke, km), "sgld:arg: the tiled K.u needs the isotropic-cube Ke structure");
        SGLD_ARG(zChunk == 0 || (zChunk >= 2 && zChunk <= 256), "sgld:arg: zChunk must be 0 or in 2..256");
        for (int q = 0; q < 2; ++q) {
            const int per = blocksPerCU > 0 ? blocksPerCU : std::max(1, C->kOcc[q]);
            C->ksched[q].cap = (int)std::min<int64_t>((int64_t)C->numCU * per, kRedBlocks);
            SGLD_TRY(set_kchunk(C, q, zChunk > 0 ? zChunk : C->ksched[q].tm.cz));
        }
    }
    C->kbyuVariant = kbyuVariant;
    C->restrictTiled = restrictTiled;
    return SGLD_OK;
}


// ------------------------------ stress recovery ------------------------------
namespace sgld {
int ensure_efwd(sgld_ctx *C) {
    if (C->efwd.p) return SGLD_OK;
    const Grid &g = C->lv[0].g;
    std::vector<int32_t> m(g.elems(), 0);
    if (C->full)
        for (int64_t v = 0; v < g.elems(); ++v) m[v] = (int32_t)(v + 1);
    else
        for (size_t i = 0; i < C->hEleMap.size(); ++i) m[C->hEleMap[i]] = (int32_t)(i + 1);
    return C->efwd.upload(m.data(), m.size());
}
}  // namespace sgld

extern "C" int sgld_ctx_stress_device(sgld_ctx *C, const double *dU, double *dstress, double *dvm) {
    SGLD_ARG(C && dU && dstress && dvm, "sgld:arg: null argument");
    SGLD_ARG(C->haveModulus, "sgld:state: call sgld_ctx_set_modulus first");
    SGLD_TRY(ensure_efwd(C));
    if (!C->stressT.p) {
        std::vector<double> T(48 * 24);
        stress_operator(C->nu, C->h, T.data());
        SGLD_TRY(C->stressT.upload(T.data(), T.size()));
    }
    const int64_t nn = C->lv[0].numNodes;
    SGLD_TRY(to_lattice(C, dU, C->z.p));
    hipLaunchKernelGGL(k_stress, dim3(bounded_blocks(nn, 256)), dim3(256), 0, C->s, C->z.p, C->E.p, C->efwd.p,
                       C->lv[0].g, C->stressT.p, C->full ? nullptr : C->nodMap.p, nn, dstress, dvm);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipStreamSynchronize(C->s));
    return SGLD_OK;
}

extern "C" int sgld_ctx_stress(sgld_ctx *C, const double *U, double *stress, double *vonMises) {
    SGLD_ARG(C && U && stress && vonMises, "sgld:arg: null argument");
    const int64_t nc = 3 * C->lv[0].numNodes, nn = C->lv[0].numNodes;
    DevBuf<double> dU, ds, dv;
    SGLD_TRY(dU.upload(U, nc));
    SGLD_TRY(ds.alloc(6 * nn));
    SGLD_TRY(dv.alloc(nn));
    SGLD_TRY(sgld_ctx_stress_device(C, dU.p, ds.p, dv.p));
    SGLD_HIP(hipMemcpy(stress, ds.p, sizeof(double) * 6 * nn, hipMemcpyDeviceToHost));
    SGLD_HIP(hipMemcpy(vonMises, dv.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
    return SGLD_OK;
}
