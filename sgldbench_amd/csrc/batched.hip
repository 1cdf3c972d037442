// Batched multi-RHS MG-PCG: up to 16 load cases on one mesh hierarchy, solved
// together in mixed precision.  The reference solves load cases one at a time
// (Solving_CG_GMGS.m per F_); here the Krylov vectors of all cases share one
// lattice array with the 16 cases interleaved per DOF, so every pass over the
// mesh (K.u, transfers, smoothing, updates) serves all of them.
//
// Layout of a batched vector: v[(3 g + a) * 16 + j] (FP32) for lattice node
// g, component a, case j.  K.u runs on v_mfma_f32_32x32x16_bf16 with BF16
// hi/lo split operands (x = hi + lo, hi = bf16(x), lo = bf16(x - hi): 16
// significant bits) and FP32 accumulation:
//   q_e = E_e (Ke_hi d_hi + Ke_hi d_lo + Ke_lo d_hi),  d = u_e - u_e,node0
// Subtracting the element's node-0 displacement (a rigid translation, in the
// null space of Ke) before the split leaves only strain-sized differences to
// round, so smooth directions - the usual case in a multigrid-preconditioned
// CG - do not lose their K.u to cancellation against the displacement
// magnitude (without it the 16-bit operands cost 3x more iterations).
//
// Precision: as the single-RHS mixed solver (lattice.hip), the iteration runs
// on FP32 residuals scaled to unit norm, with FP64 dot products and
// alpha/beta per case, and reliable updates: when a case's recursive residual
// has dropped by delta, its FP32 correction is folded into its FP64 solution,
// its FP64 true residual f - K u (FP64 K.u) replaces the recursive one, and p
// and rho are rescaled so the direction carries on.  Each case stops on its
// own FP64 true residual.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "lattice_ctx.hpp"

namespace sgld {
namespace {

constexpr int NR = 16;  // cases (right-hand sides) per batched vector

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// Device-side per-case PCG state.
struct BCtl {
    double rho[NR], pq[NR], rr[NR], alpha[NR], beta[NR], res[NR];
    int active[NR], restart[NR], its[NR], pad[NR];
};

// BF16 hi/lo split of 8 values into MFMA operand vectors, pairwise: one
// packed conversion gives both hi parts (round to nearest even), their FP32
// values come back by a shift and a mask, one more packed conversion gives
// the lo parts (3 instructions per value).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split8p(const float (&v)[8], bf16x8 &hi, bf16x8 &lo) {
    uint32_t h[4], l[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bf16x2 hp = {(__bf16)v[2 * i], (__bf16)v[2 * i + 1]};
        h[i] = __builtin_bit_cast(uint32_t, hp);
        const float h0 = __uint_as_float(h[i] << 16), h1 = __uint_as_float(h[i] & 0xffff0000u);
        const bf16x2 lp = {(__bf16)(v[2 * i] - h0), (__bf16)(v[2 * i + 1] - h1)};
        l[i] = __builtin_bit_cast(uint32_t, lp);
    }
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 hv = {h[0], h[1], h[2], h[3]}, lv = {l[0], l[1], l[2], l[3]};
    hi = __builtin_bit_cast(bf16x8, hv);
    lo = __builtin_bit_cast(bf16x8, lv);
}

// Sum v over the threads of a block that share threadIdx.x % 16 (the case
// lane of grid-stride loops over batched vectors); thread j < 16 writes
// out[j].
__device__ __forceinline__ void block_sum16(double v, double *__restrict__ out) {
    __shared__ double sh[16][NR];
    v += __shfl_down(v, 32, 64);
    v += __shfl_down(v, 16, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < NR) sh[w][lane] = v;
    __syncthreads();
    if (threadIdx.x < NR) {
        double s = 0.0;
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) s += sh[i][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// ----------------------------------------------------------------------------
// q = K p for 16 cases on the BF16 matrix cores.
//
// A block owns a patch of kPO_R x kPO_C output node columns (15 x 7) and a
// chunk of cz node layers, walked persistently and XCD-aware like the
// single-RHS kernel, and marches the element layers of the chunk.  Node
// layers of p are staged in LDS (coalesced float4 loads, prefetched one layer
// ahead in registers) for the 17 x 9 patch nodes; the 16 x 8 element columns
// touching the patch are MFMA column pairs (2 elements x 16 cases = 32
// columns):
//   B[k][col]: k 0-7 x of nodes 0-7, 8-15 y of nodes 0-7, 16-19 z of nodes
//              0-3, 24-27 z of nodes 4-7 (rest 0), as p - p_node0 split
//              hi/lo;
//   A = Ke with rows i = 8 a + n (component a, element-local node n) and the
//       same column order;
//   D = A B, lane l holds rows 8 (r/4) + 4 (l/32) + r%4: lanes 0-31 the
//       forces of the lower face (nodes 0-3), lanes 32-63 the upper face.
// Schedule (owner-accumulated):
// * each wavefront w owns a fixed block of 2 x 4 element columns of the patch
//   (element rows 2 (w % 8) + {0,1}, columns 4 (w / 8) + {0..3}) and adds its
//   elements' face forces into a private LDS accumulator of the block's 3 x 5
//   nodes (two node layers, the lower and the upper face), so no colour passes
//   or block barriers separate the element pairs; a wavefront's four pairs per
//   element layer are (r, c) and (r, c + 2) of its block, which share no node;
// * the modulus scales D after the product (B = p_e - p_e,node0 split hi/lo);
// * once an element layer is done (one block barrier), node layer el is
//   complete: every owned value sums the 1, 2 or 4 private copies of its node
//   in a fixed order (deterministic), clears them, is masked and stored; the
//   stage slot of that layer is refilled in the same phase, and a second
//   barrier ends the layer;
// * LDS strides are padded so that the two elements of a pair (and the two
//   faces) fall in different banks.
// ----------------------------------------------------------------------------
constexpr int kBT = 1024;                 // threads per block (16 wavefronts)
constexpr int kPE_R = 16, kPE_C = 8;      // element columns per patch (rows, cols)
constexpr int kPN_R = kPE_R + 1, kPN_C = kPE_C + 1;  // staged patch nodes
constexpr int kPO_R = kPE_R - 1, kPO_C = kPE_C - 1;  // owned output nodes
constexpr int kPN = kPN_R * kPN_C;        // 153 staged nodes
constexpr int kNV = 3 * NR;               // floats per node (3 components x 16 cases)
constexpr int kLay = kPN * kNV;           // floats per staged / accumulated layer
constexpr int kLd4 = kLay / 4;            // float4 per layer
constexpr int kLdPer = (kLd4 + kBT - 1) / kBT;
static_assert(kPE_R * kPE_C <= kBT && kPN <= kBT, "staging threads");

__device__ __forceinline__ int bcol_k(int k) {  // B/A column k -> element DOF (3 node + comp), -1 = zero
    if (k < 8) return 3 * k;
    if (k < 16) return 3 * (k - 8) + 1;
    if (k < 20) return 3 * (k - 16) + 2;
    if (k >= 24 && k < 28) return 3 * (k - 20) + 2;
    return -1;
}

constexpr int kACS = 3 * kNV + 8;              // private accumulator: floats per node column (3 nodes + pad)
constexpr int kAW = 5 * kACS;                  // floats per wavefront block (5 node columns)
constexpr int kAZero = (kBT / 64) * kAW;       // 16 floats of zeros after the blocks (never accumulated into)
constexpr int kASlot = kAZero + 32;            // floats per node layer
constexpr int kOItems = kPO_R * kPO_C * 3 * 4;  // owned (node, component, 4 cases) items per node layer (1260)
constexpr int kOPer = (kOItems + kBT - 1) / kBT;  // per thread (2)

// Sum over the threads of a block of four per-thread values that belong to
// cases 4 (t % 4) + i; thread j < 16 writes out[j].
__device__ __forceinline__ void block_sum16q(const double (&v)[4], double *__restrict__ out) {
    __shared__ double sh[kBT / 64][NR];
    double d[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        d[i] = v[i];
        d[i] += __shfl_xor(d[i], 32, 64);
        d[i] += __shfl_xor(d[i], 16, 64);
        d[i] += __shfl_xor(d[i], 8, 64);
        d[i] += __shfl_xor(d[i], 4, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < 4)
#pragma unroll
        for (int i = 0; i < 4; ++i) sh[w][4 * lane + i] = d[i];
    __syncthreads();
    if (threadIdx.x < NR) {
        double s = 0.0;
        for (int i = 0; i < kBT / 64; ++i) s += sh[i][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

template <int CS>  // staged floats per patch node column (17 nodes x 48 + pad)
__global__ void __launch_bounds__(kBT, 1) k_bkbyu(const float *__restrict__ p, float *__restrict__ q,
                                                      const float *__restrict__ E,
                                                      const uint8_t *__restrict__ nflag, Grid gd, TileMap tm,
                                                      const float *__restrict__ Ke, double *__restrict__ partial) {
    constexpr int kSL = kPN_C * CS;  // floats per staged node layer
    static_assert(CS % 4 == 0 && 2 * kASlot < (1 << 16) && kSL < (1 << 13), "packed LDS indices");
    __shared__ __attribute__((aligned(16))) float stage[2 * kSL];
    __shared__ __attribute__((aligned(16))) float acc[2 * kASlot];
    __shared__ float Es[2][kPE_R * kPE_C];
    __shared__ uint8_t Fs[2][kPN + 3];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int g = lane >> 5;  // 0: lower-face rows of D (nodes 0-3); 1: upper face (4-7)
    const int col = lane & 31;
    const int j = col & 15, e = col >> 4;
    const int wr = wave & 7, wc = wave >> 3;
    const int NY = gd.ny + 1, NX = gd.nx + 1, NZ = gd.nz + 1;
    const int64_t pl = (int64_t)NY * NX;
    const int64_t ple = (int64_t)gd.ny * gd.nx;

    bf16x8 Ahi[2], Alo[2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kc = bcol_k(16 * s + 8 * g + i);
            const float a = (col < 24 && kc >= 0) ? Ke[(3 * (col & 7) + (col >> 3)) + 24 * kc] : 0.0f;
            const __bf16 hi = (__bf16)a;
            Ahi[s][i] = hi;
            Alo[s][i] = (__bf16)(a - (float)hi);
        }
    // lane bases: stage (node (2 wr, 4 wc + 2 e) of the patch, case j), private
    // accumulator (local node (0, 2 e)), modulus of element (2 wr, 4 wc + 2 e)
    const int sLane = (4 * wc + 2 * e) * CS + (2 * wr) * kNV + j;
    const int aLane = wave * kAW + (2 * e) * kACS + j;
    const int eLane = 2 * wr + kPE_R * (4 * wc + 2 * e);
    // Owned items of this thread: item it = t + kBT i is node-component
    // it / 4 = 3 node + comp of the 15 x 7 owned nodes (oi, oj) (node =
    // oi + 15 oj, staged patch node (oi + 1, oj + 1)) and cases 4 (t % 4) +
    // 0..3.  The node lies in the block of wavefront (rp / 2, cp / 4) and, for
    // even rp, also in the one below (local row 2), for cp = 4 also in the one
    // to its left (local column 4).  Packed per item: oA = primary | row
    // partner << 16, oC = column partner | both << 16 (absent partners point
    // at zeros), oS = stage index | staged node << 13 | comp << 21 | owned << 23
    // | oi << 24 | oj << 28, oG = global float offset from the patch's first
    // owned node.
    uint32_t oA[kOPer], oC[kOPer], oS[kOPer], oG[kOPer];
    constexpr int kRowP = -kAW + 2 * kNV, kColP = -8 * kAW + 4 * kACS;
#pragma unroll
    for (int i = 0; i < kOPer; ++i) {
        const int it = t + kBT * i;
        const int nc = it >> 2, quad = it & 3;
        const int node = nc / 3, comp = nc - 3 * node;
        const int oi = node % kPO_R, oj = node / kPO_R;
        const int rp = oi + 1, cp = oj + 1;
        const int w = (rp >> 1) + 8 * (cp >> 2), lr = rp & 1, lc = cp & 3;
        const int within = comp * NR + 4 * quad;
        const int o0 = w * kAW + lc * kACS + lr * kNV + within;
        const int o1 = lr == 0 ? o0 + kRowP : kAZero, o2 = lc == 0 ? o0 + kColP : kAZero;
        const int o3 = (lr == 0 && lc == 0) ? o0 + kRowP + kColP : kAZero;
        oA[i] = (uint32_t)o0 | (uint32_t)o1 << 16;
        oC[i] = (uint32_t)o2 | (uint32_t)o3 << 16;
        oS[i] = (uint32_t)(cp * CS + rp * kNV + within) | (uint32_t)(rp + kPN_R * cp) << 13 | (uint32_t)comp << 21 |
                (uint32_t)(it < kOItems) << 23 | (uint32_t)oi << 24 | (uint32_t)oj << 28;
        oG[i] = (uint32_t)((oi + NY * oj) * kNV + within);
    }

    const int64_t W = tm.work();
    const int G = gridDim.x;
    const int nxcd = G < kXCD ? G : kXCD;
    const int xcd = blockIdx.x % nxcd, slotInXcd = blockIdx.x / nxcd;
    const int blocksOnXcd = G / nxcd + (xcd < G % nxcd ? 1 : 0);
    const int64_t wBeg = W * xcd / nxcd, wEnd = W * (xcd + 1) / nxcd;
    double dacc[4] = {0.0, 0.0, 0.0, 0.0};  // cases 4 (t % 4) + i
    const float4 *p4 = reinterpret_cast<const float4 *>(p);

    for (int64_t w = wBeg + slotInXcd; w < wEnd; w += blocksOnXcd) {
        int64_t rem = w;
        const int tz = (int)(rem / ((int64_t)tm.tilesR * tm.tilesC));
        rem -= (int64_t)tz * tm.tilesR * tm.tilesC;
        const int tc = (int)(rem / tm.tilesR), tr = (int)(rem % tm.tilesR);
        const int R0 = tr * kPO_R, C0 = tc * kPO_C, Z0 = tz * tm.cz;
        const int Zend = min(Z0 + tm.cz, NZ);
        int src[kLdPer];  // float4 index within a node layer (a layer has < 2^31 / 12 nodes)
        int dst[kLdPer];
#pragma unroll
        for (int i = 0; i < kLdPer; ++i) {
            const int f = t + i * kBT;
            src[i] = -1;
            dst[i] = -1;
            if (f < kLd4) {
                const int node = f / (kNV / 4), w4 = f % (kNV / 4);
                const int rp = node % kPN_R, cp = node / kPN_R;
                const int r = R0 - 1 + rp, c = C0 - 1 + cp;
                dst[i] = cp * CS + rp * kNV + 4 * w4;
                if (r >= 0 && r < NY && c >= 0 && c < NX) src[i] = (r + NY * c) * (kNV / 4) + w4;
            }
        }
        auto load_layer = [&](int lay, float4 (&v)[kLdPer]) {
            const float4 *pl4 = p4 + pl * (int64_t)lay * (kNV / 4);
#pragma unroll
            for (int i = 0; i < kLdPer; ++i)
                v[i] = (src[i] >= 0 && lay >= 0 && lay < NZ) ? pl4[src[i]] : make_float4(0.f, 0.f, 0.f, 0.f);
        };
        auto put_layer = [&](float *sl, const float4 (&v)[kLdPer]) {
#pragma unroll
            for (int i = 0; i < kLdPer; ++i)
                if (dst[i] >= 0) *reinterpret_cast<float4 *>(sl + dst[i]) = v[i];
        };
        uint32_t inRange = 0;  // owned item i inside the lattice
#pragma unroll
        for (int i = 0; i < kOPer; ++i) {
            const int oi = (oS[i] >> 24) & 15, oj = (int)(oS[i] >> 28);
            if (((oS[i] >> 23) & 1) && R0 + oi < NY && C0 + oj < NX) inRange |= 1u << i;
        }
        float *const qPatch = q + ((int64_t)R0 + (int64_t)NY * C0) * kNV;
        float4 pf[kLdPer];
        int eSrc = -1, fSrc = -1;
        if (t < kPE_R * kPE_C) {
            const int er = R0 - 1 + t % kPE_R, ec = C0 - 1 + t / kPE_R;
            if (er >= 0 && er < gd.ny && ec >= 0 && ec < gd.nx) eSrc = er + gd.ny * ec;
        }
        if (t < kPN) {
            const int r = R0 - 1 + t % kPN_R, c = C0 - 1 + t / kPN_R;
            if (r >= 0 && r < NY && c >= 0 && c < NX) fSrc = r + NY * c;
        }
        auto load_E = [&](int el) {
            return (eSrc >= 0 && el >= 0 && el < gd.nz) ? E[ple * el + eSrc] : 0.0f;
        };
        auto load_F = [&](int lay) {
            return (fSrc >= 0 && lay >= 0 && lay < NZ) ? nflag[pl * lay + fSrc] : (uint8_t)0;
        };
        __syncthreads();  // the previous work item is done with LDS
        for (int i = t; i < 2 * kASlot / 4; i += kBT) reinterpret_cast<float4 *>(acc)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        load_layer(Z0 - 1, pf);
        put_layer(stage + ((Z0 - 1) & 1) * kSL, pf);
        load_layer(Z0, pf);
        put_layer(stage + (Z0 & 1) * kSL, pf);
        load_layer(Z0 + 1, pf);  // in flight
        if (t < kPE_R * kPE_C) Es[(Z0 - 1) & 1][t] = load_E(Z0 - 1);
        float eNext = load_E(Z0);
        uint8_t fNext = load_F(Z0);  // flags of node layer el + 1, staged during layer el
        __syncthreads();
        for (int el = Z0 - 1; el < Zend; ++el) {
            const int s0 = el & 1;
            const float *sb = stage + s0 * kSL, *st = stage + (s0 ^ 1) * kSL;
            const float *sz = stage + (s0 ^ g) * kSL + 2 * NR;  // z of this lane's face
            float *af = acc + (s0 ^ g) * kASlot;                // this lane's face layer
            // the lower face of element layer Z0 - 1 is node layer Z0 - 1, not an output
            const float fo = (g == 1 || el >= Z0) ? 1.0f : 0.0f;
            // the wavefront's 4 element pairs: rows tr2 of its block, columns (tc2, tc2 + 2)
#pragma unroll 1
            for (int tr2 = 0; tr2 < 2; ++tr2)
#pragma unroll
                for (int tc2 = 0; tc2 < 2; ++tc2) {
                    const int b = sLane + tr2 * kNV + tc2 * CS;
                    // local node n: (dr, dc) = (1,0), (1,1), (0,1), (0,0)
                    const int so[4] = {kNV, kNV + CS, CS, 0};
                    float va[8], vb[4];
#pragma unroll
                    for (int n = 0; n < 4; ++n) {
                        va[n] = sb[b + so[n] + NR * g];
                        va[n + 4] = st[b + so[n] + NR * g];
                        vb[n] = sz[b + so[n]];
                    }
                    const float ref_a = va[0], ref_b = sb[b + so[0] + 2 * NR];
                    float da[8], db[8];
                    da[0] = 0.0f;
#pragma unroll
                    for (int n = 1; n < 8; ++n) da[n] = va[n] - ref_a;
#pragma unroll
                    for (int n = 0; n < 4; ++n) {
                        db[n] = vb[n] - ref_b;
                        db[n + 4] = 0.0f;
                    }
                    bf16x8 Bh0, Bl0, Bh1, Bl1;
                    split8p(da, Bh0, Bl0);
                    split8p(db, Bh1, Bl1);
                    f32x16 D = {};
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi[0], Bh0, D, 0, 0, 0);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi[0], Bl0, D, 0, 0, 0);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Alo[0], Bh0, D, 0, 0, 0);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi[1], Bh1, D, 0, 0, 0);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Ahi[1], Bl1, D, 0, 0, 0);
                    D = __builtin_amdgcn_mfma_f32_32x32x16_bf16(Alo[1], Bh1, D, 0, 0, 0);
                    const float Ee = fo * Es[s0][eLane + tr2 + kPE_R * tc2];
                    float *a = af + aLane + tr2 * kNV + tc2 * kACS;
                    const int ao[4] = {kNV, kNV + kACS, kACS, 0};
#pragma unroll
                    for (int c = 0; c < 3; ++c)
#pragma unroll
                        for (int m = 0; m < 4; ++m) a[ao[m] + c * NR] = __builtin_fmaf(Ee, D[4 * c + m], a[ao[m] + c * NR]);
                }
            // the p values this thread's owned items are dotted with (node layer el)
            float4 pv[kOPer];
#pragma unroll
            for (int i = 0; i < kOPer; ++i)
                pv[i] = (inRange >> i) & 1 ? *reinterpret_cast<const float4 *>(sb + (oS[i] & 0x1fff))
                                           : make_float4(0.f, 0.f, 0.f, 0.f);
            __syncthreads();  // element layer el done: node layer el complete, stage slot s0 free
            if (el >= Z0) {
                float *a0 = acc + s0 * kASlot;
                const uint8_t *fl = Fs[s0];
                float *const qL = qPatch + pl * (int64_t)el * kNV;
#pragma unroll
                for (int i = 0; i < kOPer; ++i) {
                    if (!((inRange >> i) & 1)) continue;
                    float4 *x0 = reinterpret_cast<float4 *>(a0 + (oA[i] & 0xffff));
                    float4 *x1 = reinterpret_cast<float4 *>(a0 + (oA[i] >> 16));
                    float4 *x2 = reinterpret_cast<float4 *>(a0 + (oC[i] & 0xffff));
                    float4 *x3 = reinterpret_cast<float4 *>(a0 + (oC[i] >> 16));
                    const float4 v0 = *x0, v1 = *x1, v2 = *x2, v3 = *x3;
                    const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
                    *x0 = z;
                    *x1 = z;
                    *x2 = z;
                    *x3 = z;
                    const uint8_t f8 = fl[(oS[i] >> 13) & 0xff];
                    const int comp = (oS[i] >> 21) & 3;
                    const float ok = ((f8 & kSolid) && !(f8 & (1 << comp))) ? 1.0f : 0.0f;
                    float4 v;
                    v.x = ok * ((v0.x + v1.x) + (v2.x + v3.x));
                    v.y = ok * ((v0.y + v1.y) + (v2.y + v3.y));
                    v.z = ok * ((v0.z + v1.z) + (v2.z + v3.z));
                    v.w = ok * ((v0.w + v1.w) + (v2.w + v3.w));
                    *reinterpret_cast<float4 *>(qL + oG[i]) = v;
                    dacc[0] += (double)v.x * (double)pv[i].x;
                    dacc[1] += (double)v.y * (double)pv[i].y;
                    dacc[2] += (double)v.z * (double)pv[i].z;
                    dacc[3] += (double)v.w * (double)pv[i].w;
                }
            }
            put_layer(stage + s0 * kSL, pf);  // node layer el + 2
            load_layer(el + 3, pf);
            if (t < kPN) Fs[s0 ^ 1][t] = fNext;  // node layer el + 1
            fNext = load_F(el + 2);
            if (t < kPE_R * kPE_C) Es[s0 ^ 1][t] = eNext;  // element layer el + 1
            eNext = load_E(el + 2);
            __syncthreads();
        }
    }
    block_sum16q(dacc, partial + (int64_t)blockIdx.x * NR);
}

// ------------------------------ vector kernels ------------------------------
// Grid-stride over n16 = 16 x lattice DOFs; a thread's case is t % 16.
__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t gstride() { return (int64_t)gridDim.x * blockDim.x; }

// Component c of a float4.
__device__ __forceinline__ float &f4c(float4 &v, int c) { return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w; }
__device__ __forceinline__ float f4c(const float4 &v, int c) { return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w; }

// r -= alpha q per case; partials r.r and sum dinv r^2 per case (the finest
// smoother's share of the next rho; dinv is per DOF, shared by the cases).
// (A float4 form, four cases per thread, measured no faster: 2.19 vs 2.14 ms
// on 256^3 x 16, the arrays being far beyond the Infinity Cache.)
__global__ void __launch_bounds__(256) k_bupdate_r(float *__restrict__ r, const float *__restrict__ q,
                                                   const float *__restrict__ dinv, int64_t n16,
                                                   const BCtl *__restrict__ ctl, double *__restrict__ partRR,
                                                   double *__restrict__ partDR) {
    const float al = (float)ctl->alpha[threadIdx.x & 15];
    double rr = 0.0, dr = 0.0;
    for (int64_t i = gtid(); i < n16; i += gstride()) {
        const float rn = r[i] - al * q[i];
        r[i] = rn;
        const double rd = (double)rn;
        rr += rd * rd;
        dr += (double)dinv[i >> 4] * rd * rd;
    }
    block_sum16(rr, partRR + (int64_t)blockIdx.x * NR);
    block_sum16(dr, partDR + (int64_t)blockIdx.x * NR);
}

// partials of sum dinv r^2 per case (a residual replaced by reliable updates)
__global__ void __launch_bounds__(256) k_bdinv_rr(const float *__restrict__ r, const float *__restrict__ dinv,
                                                  int64_t n16, double *__restrict__ partDR) {
    double dr = 0.0;
    for (int64_t i = gtid(); i < n16; i += gstride()) {
        const double rd = (double)r[i];
        dr += (double)dinv[i >> 4] * rd * rd;
    }
    block_sum16(dr, partDR + (int64_t)blockIdx.x * NR);
}

// partials of a.b per case over n16 interleaved entries
__global__ void __launch_bounds__(256) k_bdot(const float *__restrict__ a, const float *__restrict__ b, int64_t n16,
                                              double *__restrict__ part) {
    double acc = 0.0;
    for (int64_t i = gtid(); i < n16; i += gstride()) acc += (double)a[i] * (double)b[i];
    block_sum16(acc, part + (int64_t)blockIdx.x * NR);
}

enum BFin { BFIN_ALPHA = 0, BFIN_RR, BFIN_RHO };

// Per-case reduction of partial[nb][16] and the scalar step of the PCG
// recurrence (Solving_PreconditionedConjugateGradientSolver.m:15-28).
// 1024 threads, 64 per case, four independent sums each so the loads of a
// thread issue back to back (fixed order: deterministic).
constexpr int kBFinGroups = 64;
__device__ __forceinline__ double bfin_sum(const double *__restrict__ part, int nb, int grp, int j) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = grp;
    for (; b + 3 * kBFinGroups < nb; b += 4 * kBFinGroups) {
        s0 += part[(int64_t)b * NR + j];
        s1 += part[(int64_t)(b + kBFinGroups) * NR + j];
        s2 += part[(int64_t)(b + 2 * kBFinGroups) * NR + j];
        s3 += part[(int64_t)(b + 3 * kBFinGroups) * NR + j];
    }
    for (; b < nb; b += kBFinGroups) s0 += part[(int64_t)b * NR + j];
    return (s0 + s1) + (s2 + s3);
}

__global__ void __launch_bounds__(1024) k_bfin(const double *__restrict__ partial, int nb,
                                               const double *__restrict__ partial2, int nb2, int op,
                                               BCtl *__restrict__ ctl) {
    __shared__ double sh[kBFinGroups][NR];
    const int j = threadIdx.x & 15, grp = threadIdx.x >> 4;
    double s = bfin_sum(partial, nb, grp, j);
    if (partial2) s += bfin_sum(partial2, nb2, grp, j);
    sh[grp][j] = s;
    __syncthreads();
    if (threadIdx.x >= NR) return;
    double tot = 0.0;
    for (int i = 0; i < kBFinGroups; ++i) tot += sh[i][j];
    const bool act = ctl->active[j] != 0;
    if (op == BFIN_ALPHA) {
        ctl->pq[j] = tot;
        ctl->alpha[j] = (act && tot > 0.0) ? ctl->rho[j] / tot : 0.0;
    } else if (op == BFIN_RR) {
        ctl->rr[j] = tot;
        ctl->res[j] = sqrt(tot);
        if (act) ctl->its[j] += 1;
    } else {
        const double rho0 = ctl->rho[j];
        ctl->beta[j] = (act && !ctl->restart[j] && rho0 != 0.0) ? tot / rho0 : 0.0;
        ctl->rho[j] = tot;
        ctl->restart[j] = 0;
    }
}

// ------------------------------ batched V-cycle -----------------------------
template <int S>
__global__ void __launch_bounds__(256) k_brestrict(const float *__restrict__ rf, Grid gf, float *__restrict__ rc,
                                                   Grid gc, NodeDiv dc, const uint8_t *__restrict__ cflag) {
    const int j = threadIdx.x & 15;
    const int64_t nN = gc.nodes();
    const int64_t NYf = gf.NY(), NXf = gf.NX(), NZf = gf.NZ();
    for (int64_t G = gtid() >> 4; G < nN; G += gstride() >> 4) {
        int R, Cc, Lc;
        dc.rcl((uint32_t)G, R, Cc, Lc);
        // branch-free stencil (clamped index, weight 0 outside): the loads
        // issue together instead of one after another
        float a0 = 0, a1 = 0, a2 = 0;
#pragma unroll
        for (int dl = -(S - 1); dl <= S - 1; ++dl) {
            const int64_t lu = (int64_t)S * Lc + dl;
            const int64_t lf = lu < 0 ? 0 : (lu >= NZf ? NZf - 1 : lu);
            const float wz = lu == lf ? float(S - (dl < 0 ? -dl : dl)) / float(S) : 0.0f;
#pragma unroll
            for (int dcc = -(S - 1); dcc <= S - 1; ++dcc) {
                const int64_t cu = (int64_t)S * Cc + dcc;
                const int64_t cf = cu < 0 ? 0 : (cu >= NXf ? NXf - 1 : cu);
                const float wx = cu == cf ? wz * float(S - (dcc < 0 ? -dcc : dcc)) / float(S) : 0.0f;
#pragma unroll
                for (int dr = -(S - 1); dr <= S - 1; ++dr) {
                    const int64_t ru = (int64_t)S * R + dr;
                    const int64_t rr = ru < 0 ? 0 : (ru >= NYf ? NYf - 1 : ru);
                    const float w = ru == rr ? wx * float(S - (dr < 0 ? -dr : dr)) / float(S) : 0.0f;
                    const float *pn = rf + 3 * (rr + NYf * cf + NYf * NXf * lf) * NR + j;
                    a0 += w * pn[0];
                    a1 += w * pn[NR];
                    a2 += w * pn[2 * NR];
                }
            }
        }
        const bool solid = cflag[G] & kSolid;
        float *po = rc + 3 * G * NR + j;
        po[0] = solid ? a0 : 0.0f;
        po[NR] = solid ? a1 : 0.0f;
        po[2 * NR] = solid ? a2 : 0.0f;
    }
}

// Span-2 batched restriction, LDS-tiled and marching up the layers.  A block
// owns kBRR x kBRC coarse node columns (all NR cases, one thread per node and
// case) and walks their coarse layers; for coarse layer L it stages fine
// layers 2L and 2L+1 of the (2 kBRR + 1) x (2 kBRC + 1) fine window (whole
// 192-byte nodes as float4, coalesced along the fine rows), forms each coarse
// node's in-plane 3 x 3 sums (weights 1/2, 1, 1/2 per axis) from LDS, and
// carries the half of fine layer 2L+1 that belongs to layer L+1 in registers.
// Each fine value is read from HBM once per block (plus the window's rim),
// where the node-centric gather re-reads the fine layer of each z-neighbour.
constexpr int kBRR = 8, kBRC = 2;                       // coarse rows, cols per block
constexpr int kBFR = 2 * kBRR + 1, kBFC = 2 * kBRC + 1;  // fine window rows, cols
constexpr int kBFN = kBFR * kBFC;                        // fine nodes per staged layer
constexpr int kBNode4 = 3 * NR / 4;                      // float4 per node (12)
constexpr int kBPad4 = kBNode4 + 1;                      // LDS node stride in float4 (fewer bank conflicts)
constexpr int kBStage = (kBFN * kBNode4 + 255) / 256;    // float4 staged per thread and layer

__global__ void __launch_bounds__(256) k_brestrict2_tiled(const float *__restrict__ rf, Grid gf,
                                                          float *__restrict__ rc, Grid gc,
                                                          const uint8_t *__restrict__ cflag, int tilesR) {
    __shared__ float4 sh[2][kBFN * kBPad4];
    const int t = threadIdx.x, j = t & (NR - 1), cr = (t / NR) % kBRR, cc = t / (NR * kBRR);
    const int tr = blockIdx.x % tilesR, tc = blockIdx.x / tilesR;
    const int NYf = gf.ny + 1, NXf = gf.nx + 1, NZf = gf.nz + 1;
    const int NYc = gc.ny + 1, NXc = gc.nx + 1, NZc = gc.nz + 1;
    const int64_t plf = (int64_t)NYf * NXf, plc = (int64_t)NYc * NXc;
    const int R0 = tr * kBRR, C0 = tc * kBRC, R = R0 + cr, Cn = C0 + cc;
    const bool mine = R < NYc && Cn < NXc;
    // staged float4 q = t + 256 i: window node k = q / 12 (row-major in the
    // fine column), part q % 12; offset of the node within a fine layer (or -1)
    int64_t goff[kBStage];
    int part[kBStage], lds[kBStage];
#pragma unroll
    for (int i = 0; i < kBStage; ++i) {
        const int q = t + 256 * i, k = q / kBNode4;
        part[i] = q - k * kBNode4;
        lds[i] = k * kBPad4 + part[i];
        goff[i] = -1;
        if (k < kBFN) {
            const int rfw = 2 * R0 - 1 + k % kBFR, cfw = 2 * C0 - 1 + k / kBFR;
            if (rfw >= 0 && rfw < NYf && cfw >= 0 && cfw < NXf) goff[i] = rfw + (int64_t)NYf * cfw;
        } else {
            lds[i] = -1;
        }
    }
    const float4 *rf4 = reinterpret_cast<const float4 *>(rf);
    const float *shf0 = reinterpret_cast<const float *>(sh[0]), *shf1 = reinterpret_cast<const float *>(sh[1]);
    float carry[3] = {0.f, 0.f, 0.f};  // half of fine layer 2L-1
    for (int L = 0; L < NZc; ++L) {
        float4 v0[kBStage], v1[kBStage];
#pragma unroll
        for (int i = 0; i < kBStage; ++i) {
            const bool ok0 = goff[i] >= 0 && 2 * L < NZf, ok1 = goff[i] >= 0 && 2 * L + 1 < NZf;
            const int64_t o0 = ok0 ? (goff[i] + plf * (2 * L)) * kBNode4 + part[i] : 0;
            const int64_t o1 = ok1 ? (goff[i] + plf * (2 * L + 1)) * kBNode4 + part[i] : 0;
            v0[i] = rf4[o0];
            v1[i] = rf4[o1];
            if (!ok0) v0[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!ok1) v1[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();  // the previous layer's reads of LDS are done
#pragma unroll
        for (int i = 0; i < kBStage; ++i)
            if (lds[i] >= 0) {
                sh[0][lds[i]] = v0[i];
                sh[1][lds[i]] = v1[i];
            }
        __syncthreads();
        float P0[3] = {0.f, 0.f, 0.f}, P1[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int dc = -1; dc <= 1; ++dc)
#pragma unroll
            for (int dr = -1; dr <= 1; ++dr) {
                const float w = float((dc == 0 ? 2 : 1) * (dr == 0 ? 2 : 1)) * 0.25f;
                const int k = (2 * cr + 1 + dr) + kBFR * (2 * cc + 1 + dc);
                const int o = 4 * k * kBPad4 + j;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    P0[a] += w * shf0[o + a * NR];
                    P1[a] += w * shf1[o + a * NR];
                }
            }
        if (mine) {
            const int64_t G = R + (int64_t)NYc * Cn + plc * L;
            const bool solid = cflag[G] & kSolid;
            float *po = rc + 3 * G * NR + j;
#pragma unroll
            for (int a = 0; a < 3; ++a) po[a * NR] = solid ? carry[a] + P0[a] + 0.5f * P1[a] : 0.0f;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) carry[a] = 0.5f * P1[a];
    }
}

// x_f = P x_c + dinv_f r_f; finest: fixed DOFs 0 and partial x_f.r_f per case.
// Branch-free interpolation: per axis the coarse neighbours q and q + (m > 0)
// with weights (S - m)/S and m/S, all inside the coarse lattice.
template <int S>
__global__ void __launch_bounds__(256) k_bprolong(const float *__restrict__ xc, Grid gc, float *__restrict__ xf,
                                                  Grid gf, NodeDiv df, const float *__restrict__ rf,
                                                  const float *__restrict__ dinv,
                                                  const uint8_t *__restrict__ fflag, int finest,
                                                  double *__restrict__ partial) {
    const int j = threadIdx.x & 15;
    const int64_t nN = gf.nodes();
    const int64_t NYc = gc.NY(), NXc = gc.NX();
    double dotacc = 0.0;
    for (int64_t n = gtid() >> 4; n < nN; n += gstride() >> 4) {
        int r, c, l;
        df.rcl((uint32_t)n, r, c, l);
        const int qr = r / S, mr = r % S, qc = c / S, mc = c % S, ql = l / S, ml = l % S;
        const float *base = xc + 3 * (qr + NYc * qc + NYc * NXc * ql) * NR + j;
        const int64_t dr = 3 * NR * (mr > 0), dcn = 3 * NR * NYc * (mc > 0), dl = 3 * NR * NYc * NXc * (ml > 0);
        const float wr1 = float(mr) / float(S), wc1 = float(mc) / float(S), wl1 = float(ml) / float(S);
        const float wr0 = 1.0f - wr1, wc0 = 1.0f - wc1, wl0 = 1.0f - wl1;
        float v[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float *b = base + a * NR;
            v[a] = wl0 * (wc0 * (wr0 * b[0] + wr1 * b[dr]) + wc1 * (wr0 * b[dcn] + wr1 * b[dcn + dr])) +
                   wl1 * (wc0 * (wr0 * b[dl] + wr1 * b[dl + dr]) + wc1 * (wr0 * b[dl + dcn] + wr1 * b[dl + dcn + dr]));
        }
        const uint8_t f = fflag[n];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const int64_t idx = (3 * n + a) * NR + j;
            const float rv = rf[idx];
            float val = v[a] + dinv[3 * n + a] * rv;
            if (!(f & kSolid) || (finest && (f & (1 << a)))) val = 0.0f;
            xf[idx] = val;
            dotacc += (double)val * (double)rv;
        }
    }
    if (partial) block_sum16(dotacc, partial + (int64_t)blockIdx.x * NR);
}

// Finest level of the V-cycle folded into the direction update, per case j:
// z = mask(P x_2 + dinv r), xc += alpha_j p (the previous iteration's
// deferred update), p = z + beta_j p in place; z is never stored (rho comes
// from x_2.r_2 + sum dinv r^2 because R = P^T and r = 0 on fixed DOFs).
// A thread owns cases 4g..4g+3 of a fine node (float4 accesses, g = threadIdx.x & 3).
template <int S>
__global__ void __launch_bounds__(256) k_bprolong_dir(const float *__restrict__ xc2, Grid gc, Grid gf, NodeDiv df,
                                                      const float *__restrict__ rf, const float *__restrict__ dinv,
                                                      const uint8_t *__restrict__ fflag, float *__restrict__ xc,
                                                      float *__restrict__ p, const BCtl *__restrict__ ctl) {
    constexpr int V4 = NR / 4;  // float4 per DOF
    const int g = threadIdx.x & 3;
    float al[4], be[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        al[c] = (float)ctl->alpha[4 * g + c];
        be[c] = (float)ctl->beta[4 * g + c];
    }
    const int64_t nN = gf.nodes();
    const int64_t NYc = gc.NY(), NXc = gc.NX();
    const float4 *x4 = reinterpret_cast<const float4 *>(xc2);
    const float4 *r4 = reinterpret_cast<const float4 *>(rf);
    float4 *c4 = reinterpret_cast<float4 *>(xc);
    float4 *p4 = reinterpret_cast<float4 *>(p);
    for (int64_t n = gtid() >> 2; n < nN; n += gstride() >> 2) {
        int r, c, l;
        df.rcl((uint32_t)n, r, c, l);
        const int qr = r / S, mr = r % S, qc = c / S, mc = c % S, ql = l / S, ml = l % S;
        const float4 *base = x4 + 3 * (qr + NYc * qc + NYc * NXc * ql) * V4 + g;
        const int64_t dr = 3 * V4 * (mr > 0), dcn = 3 * V4 * NYc * (mc > 0), dl = 3 * V4 * NYc * NXc * (ml > 0);
        const float wr1 = float(mr) / float(S), wc1 = float(mc) / float(S), wl1 = float(ml) / float(S);
        const float wr0 = 1.0f - wr1, wc0 = 1.0f - wc1, wl0 = 1.0f - wl1;
        const uint8_t f = fflag[n];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float4 *b = base + a * V4;
            const float4 b0 = b[0], b1 = b[dr], b2 = b[dcn], b3 = b[dcn + dr];
            const float4 b4 = b[dl], b5 = b[dl + dr], b6 = b[dl + dcn], b7 = b[dl + dcn + dr];
            const int64_t idx = (3 * n + a) * V4 + g;
            const float4 rv = r4[idx];
            float4 po = p4[idx], xo = c4[idx];
            const float d = dinv[3 * n + a];
            const bool keep = (f & kSolid) && !(f & (1 << a));
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float v = wl0 * (wc0 * (wr0 * f4c(b0, k) + wr1 * f4c(b1, k)) +
                                       wc1 * (wr0 * f4c(b2, k) + wr1 * f4c(b3, k))) +
                                wl1 * (wc0 * (wr0 * f4c(b4, k) + wr1 * f4c(b5, k)) +
                                       wc1 * (wr0 * f4c(b6, k) + wr1 * f4c(b7, k)));
                const float sv = d * f4c(rv, k);
                const float z = keep ? v + sv : 0.0f;
                const float pk = f4c(po, k);
                f4c(xo, k) += al[k] * pk;
                f4c(po, k) = z + be[k] * pk;
            }
            c4[idx] = xo;
            p4[idx] = po;
        }
    }
}

// Coarsest level for 16 cases: y = W r(free), then x(free) = W^T y with the
// packed inverse Cholesky factor (see k_coarse_lower / k_coarse_upper); one
// block per row, threads split the row 16 ways x 16 cases.
__global__ void __launch_bounds__(256) k_bcoarse_lower(const double *__restrict__ Wrow,
                                                       const int32_t *__restrict__ freeIdx, int64_t n,
                                                       const float *__restrict__ r, double *__restrict__ y) {
    __shared__ double sh[16][NR];
    const int j = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int64_t row = blockIdx.x;
    const double *w = Wrow + tri_row_off(row);
    double acc = 0.0;
    for (int64_t k = grp; k <= row; k += 16) acc += w[k] * (double)r[(int64_t)freeIdx[k] * NR + j];
    sh[grp][j] = acc;
    __syncthreads();
    if (threadIdx.x < NR) {
        double s = 0.0;
        for (int i = 0; i < 16; ++i) s += sh[i][j];
        y[row * NR + j] = s;
    }
}
__global__ void __launch_bounds__(256) k_bcoarse_upper(const double *__restrict__ Wcol,
                                                       const int32_t *__restrict__ freeIdx, int64_t n,
                                                       const double *__restrict__ y, float *__restrict__ x) {
    __shared__ double sh[16][NR];
    const int j = threadIdx.x & 15, grp = threadIdx.x >> 4;
    const int64_t col = blockIdx.x;
    const double *w = Wcol + tri_col_off(col, n);
    double acc = 0.0;
    for (int64_t t = grp; t < n - col; t += 16) acc += w[t] * y[(col + t) * NR + j];
    sh[grp][j] = acc;
    __syncthreads();
    if (threadIdx.x < NR) {
        double s = 0.0;
        for (int i = 0; i < 16; ++i) s += sh[i][j];
        x[(int64_t)freeIdx[col] * NR + j] = (float)s;
    }
}

// ---------------------------- reliable updates ------------------------------
// For the cases in `mask`: x64_j += scale_j (xc_j + alpha_j p_j), xc_j = 0
// (with the deferred update of the case's last iteration).
__global__ void __launch_bounds__(256) k_bfold(double *__restrict__ x64, float *__restrict__ xc,
                                               const float *__restrict__ p, int64_t nd,
                                               const double *__restrict__ scale, uint32_t mask,
                                               const BCtl *__restrict__ ctl) {
    // A block takes 256 DOFs at a time: the interleaved FP32 values are read
    // as float4 (thread: cases 4g..4g+3, g = threadIdx.x & 3), transposed
    // through LDS, and each case's FP64 run is updated with coalesced accesses.
    __shared__ float sh[NR][257];
    const int t = threadIdx.x, g = t & 3;
    float al[4];
    bool in[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        al[c] = (float)ctl->alpha[4 * g + c];
        in[c] = (mask >> (4 * g + c)) & 1u;
    }
    float4 *c4 = reinterpret_cast<float4 *>(xc);
    const float4 *p4 = reinterpret_cast<const float4 *>(p);
    for (int64_t c0 = (int64_t)blockIdx.x * 256; c0 < nd; c0 += (int64_t)gridDim.x * 256) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int dof = (t + 256 * k) >> 2;
            const int64_t i = c0 + dof;
            if (i < nd) {
                float4 xv = c4[i * (NR / 4) + g];
                const float4 pv = p4[i * (NR / 4) + g];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    sh[4 * g + c][dof] = f4c(xv, c) + al[c] * f4c(pv, c);
                    if (in[c]) f4c(xv, c) = 0.0f;
                }
                c4[i * (NR / 4) + g] = xv;
            }
        }
        __syncthreads();
        const int64_t i = c0 + t;
        if (i < nd)
            for (int j = 0; j < NR; ++j)
                if ((mask >> j) & 1u) x64[(int64_t)j * nd + i] += scale[j] * (double)sh[j][t];
        __syncthreads();
    }
}

// For the cases in `mask`: r_j = r64_j * inv_j (FP32), p_j *= f_j; for the
// cases in `zmask`: r_j = p_j = 0.  Same 256-DOF blocking and LDS transpose
// as k_bfold.
__global__ void __launch_bounds__(256) k_breplace(float *__restrict__ r, float *__restrict__ p,
                                                  const double *__restrict__ r64, int64_t nd,
                                                  const double *__restrict__ inv, const double *__restrict__ f,
                                                  uint32_t mask, uint32_t zmask) {
    __shared__ float sh[NR][257];
    const int t = threadIdx.x, g = t & 3;
    float fj[4];
    bool m[4], z[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        fj[c] = (float)f[4 * g + c];
        m[c] = (mask >> (4 * g + c)) & 1u;
        z[c] = (zmask >> (4 * g + c)) & 1u;
    }
    float4 *r4 = reinterpret_cast<float4 *>(r);
    float4 *p4 = reinterpret_cast<float4 *>(p);
    for (int64_t c0 = (int64_t)blockIdx.x * 256; c0 < nd; c0 += (int64_t)gridDim.x * 256) {
        const int64_t i = c0 + t;
        if (i < nd)
            for (int j = 0; j < NR; ++j)
                if ((mask >> j) & 1u) sh[j][t] = (float)(r64[(int64_t)j * nd + i] * inv[j]);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int dof = (t + 256 * k) >> 2;
            const int64_t id = c0 + dof;
            if (id < nd) {
                float4 rv = r4[id * (NR / 4) + g], pv = p4[id * (NR / 4) + g];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    if (m[c]) {
                        f4c(rv, c) = sh[4 * g + c][dof];
                        f4c(pv, c) = fj[c] * f4c(pv, c);
                    } else if (z[c]) {
                        f4c(rv, c) = 0.0f;
                        f4c(pv, c) = 0.0f;
                    }
                }
                r4[id * (NR / 4) + g] = rv;
                p4[id * (NR / 4) + g] = pv;
            }
        }
        __syncthreads();
    }
}

// FP64 case-major lattice vectors <-> batched FP32 (tests and K.u).
__global__ void __launch_bounds__(256) k_bpack(const double *__restrict__ src, int64_t nd, int nrhs,
                                               float *__restrict__ dst) {
    const int j = threadIdx.x & 15;
    for (int64_t i = gtid() >> 4; i < nd; i += gstride() >> 4)
        dst[i * NR + j] = j < nrhs ? (float)src[(int64_t)j * nd + i] : 0.0f;
}
__global__ void __launch_bounds__(256) k_bunpack(const float *__restrict__ src, int64_t nd, int nrhs,
                                                 double *__restrict__ dst) {
    const int j = threadIdx.x & 15;
    if (j >= nrhs) return;
    for (int64_t i = gtid() >> 4; i < nd; i += gstride() >> 4) dst[(int64_t)j * nd + i] = (double)src[i * NR + j];
}

}  // namespace

// ================================ host side =================================
struct BatchState {
    int64_t nd = 0;                      // lattice DOFs of level 0
    DevBuf<float> r, q, xc, p0;          // level 0, 16 cases
    std::vector<DevBuf<float>> lr, lx;   // levels >= 1
    DevBuf<double> b64, x64, r64;        // 16 case-major FP64 lattice vectors each
    DevBuf<double> part, part2, part3;   // partial sums [blocks][16] (three reductions in flight)
    DevBuf<double> scal;                 // per-case scalars
    DevBuf<float> Kc;                    // Ke (24 x 24, column-major) in FP32
    DevBuf<BCtl> ctl;
    BCtl *hctl = nullptr;                // pinned mirror
    int kGrid = 1;                       // persistent grid of k_bkbyu
    TileMap tiles{1, 1, 1, 32};
    ~BatchState() {
        if (hctl) (void)hipHostFree(hctl);
    }
};

void batch_free(BatchState *b) { delete b; }

int batch_clear_coarse(sgld_ctx *C) {
    if (!C->batch) return SGLD_OK;
    DevBuf<float> &x = C->batch->lx[C->L - 1];
    SGLD_HIP(hipMemsetAsync(x.p, 0, sizeof(float) * x.n, C->s));
    return SGLD_OK;
}

namespace {

constexpr int kBRedBlocks = 4096;  // partial slots per reduction (>= every reducing grid)

int batch_init(sgld_ctx *C) {
    if (C->batch) return SGLD_OK;
    std::unique_ptr<BatchState> B(new BatchState());
    const int64_t nd = C->nDof();
    B->nd = nd;
    const size_t n16 = (size_t)nd * NR;
    SGLD_TRY(B->r.alloc(n16));
    SGLD_TRY(B->q.alloc(n16));
    SGLD_TRY(B->xc.alloc(n16));
    SGLD_TRY(B->p0.alloc(n16));
    B->lr.resize(C->L);
    B->lx.resize(C->L);
    for (int k = 1; k < C->L; ++k) {
        const size_t nk = (size_t)3 * C->lv[k].g.nodes() * NR;
        SGLD_TRY(B->lr[k].alloc(nk));
        SGLD_TRY(B->lx[k].alloc(nk));
    }
    // the coarsest x is written at free DOFs only
    SGLD_HIP(hipMemsetAsync(B->lx[C->L - 1].p, 0, sizeof(float) * B->lx[C->L - 1].n, C->s));
    SGLD_TRY(B->b64.alloc(n16));
    SGLD_TRY(B->x64.alloc(n16));
    SGLD_TRY(B->r64.alloc(n16));
    SGLD_TRY(B->part.alloc((size_t)kBRedBlocks * NR));
    SGLD_TRY(B->part2.alloc((size_t)kBRedBlocks * NR));
    SGLD_TRY(B->part3.alloc((size_t)kBRedBlocks * NR));
    SGLD_TRY(B->scal.alloc(4 * NR));
    SGLD_TRY(B->ctl.alloc(1));
    SGLD_HIP(hipHostMalloc((void **)&B->hctl, sizeof(BCtl), hipHostMallocDefault));
    std::vector<float> kc(C->Ke, C->Ke + 576);
    SGLD_TRY(B->Kc.upload(kc.data(), kc.size()));
    const Grid &g = C->lv[0].g;
    const int tilesR = (int)((g.NY() + kPO_R - 1) / kPO_R), tilesC = (int)((g.NX() + kPO_C - 1) / kPO_C);
    const int cz = 32, tilesZ = (int)((g.NZ() + cz - 1) / cz);
    B->tiles = TileMap{tilesR, tilesC, tilesZ, cz};
    int dev = 0, cus = 0;
    SGLD_HIP(hipGetDevice(&dev));
    SGLD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    B->kGrid = (int)std::min<int64_t>(std::min<int64_t>(B->tiles.work(), cus), kBRedBlocks);
    C->batch = B.release();
    return SGLD_OK;
}

int bgrid(int64_t n16) { return (int)std::min<int64_t>(kBRedBlocks, std::max<int64_t>(1, (n16 + 255) / 256)); }

int bfin(sgld_ctx *C, int nb, int op, const double *part = nullptr, const double *part2 = nullptr, int nb2 = 0) {
    hipLaunchKernelGGL(k_bfin, dim3(1), dim3(16 * kBFinGroups), 0, C->s, part ? part : C->batch->part.p, nb, part2, nb2, op,
                       C->batch->ctl.p);
    SGLD_HIP(hipGetLastError());
    return SGLD_OK;
}

int enqueue_bkbyu(sgld_ctx *C, const float *p, float *q, int *nbOut) {
    BatchState *B = C->batch;
    const Level &l0 = C->lv[0];
    hipLaunchKernelGGL(k_bkbyu<kPN_R * kNV + 8>, dim3(B->kGrid), dim3(kBT), 0, C->s, p, q, C->E32.p,
                       l0.nflag.p, l0.g, B->tiles, B->Kc.p, B->part.p);
    SGLD_HIP(hipGetLastError());
    *nbOut = B->kGrid;
    return SGLD_OK;
}

// The batched V-cycle below the finest level (restrictions, coarsest solve,
// prolongation + smoothing up to level 2); partials of x_2.r_2 per case into
// partX2 when not null (count in *nbX2).
int enqueue_bvcycle_coarse(sgld_ctx *C, const float *r, double *partX2, int *nbX2) {
    BatchState *B = C->batch;
    const int L = C->L;
    const float *rf = r;
    for (int k = 1; k < L; ++k) {
        const Level &F = C->lv[k - 1], &Cc = C->lv[k];
        const int nb = bgrid(Cc.g.nodes() * NR);
        const NodeDiv dc = NodeDiv::make(Cc.g);
        if (Cc.span == 2 && C->bRestrictTiled) {
            const int tilesR = (int)((Cc.g.NY() + kBRR - 1) / kBRR), tilesC = (int)((Cc.g.NX() + kBRC - 1) / kBRC);
            hipLaunchKernelGGL(k_brestrict2_tiled, dim3(tilesR * tilesC), dim3(256), 0, C->s, rf, F.g, B->lr[k].p,
                               Cc.g, Cc.nflag.p, tilesR);
        } else if (Cc.span == 2)
            hipLaunchKernelGGL(k_brestrict<2>, dim3(nb), dim3(256), 0, C->s, rf, F.g, B->lr[k].p, Cc.g, dc,
                               Cc.nflag.p);
        else
            hipLaunchKernelGGL(k_brestrict<4>, dim3(nb), dim3(256), 0, C->s, rf, F.g, B->lr[k].p, Cc.g, dc,
                               Cc.nflag.p);
        SGLD_HIP(hipGetLastError());
        rf = B->lr[k].p;
    }
    const int64_t nf = C->nCoarseFree;
    hipLaunchKernelGGL(k_bcoarse_lower, dim3((int)nf), dim3(256), 0, C->s, C->cWrow.p, C->cFree.p, nf,
                       (const float *)B->lr[L - 1].p, C->cY.p);
    hipLaunchKernelGGL(k_bcoarse_upper, dim3((int)nf), dim3(256), 0, C->s, C->cWcol.p, C->cFree.p, nf,
                       (const double *)C->cY.p, B->lx[L - 1].p);
    SGLD_HIP(hipGetLastError());
    for (int k = L - 1; k >= 2; --k) {
        const Level &F = C->lv[k - 1], &Cc = C->lv[k];
        const int nb = bgrid(F.g.nodes() * NR);
        double *part = k == 2 ? partX2 : nullptr;
        const NodeDiv df = NodeDiv::make(F.g);
        if (Cc.span == 2)
            hipLaunchKernelGGL(k_bprolong<2>, dim3(nb), dim3(256), 0, C->s, B->lx[k].p, Cc.g, B->lx[k - 1].p, F.g, df,
                               B->lr[k - 1].p, F.dinv32.p, F.nflag.p, 0, part);
        else
            hipLaunchKernelGGL(k_bprolong<4>, dim3(nb), dim3(256), 0, C->s, B->lx[k].p, Cc.g, B->lx[k - 1].p, F.g, df,
                               B->lr[k - 1].p, F.dinv32.p, F.nflag.p, 0, part);
        SGLD_HIP(hipGetLastError());
        if (part && nbX2) *nbX2 = nb;
    }
    if (L == 2 && partX2) {  // level 2 is the coarsest
        const int64_t n1 = 3 * C->lv[1].g.nodes() * NR;
        const int nb = bgrid(n1);
        hipLaunchKernelGGL(k_bdot, dim3(nb), dim3(256), 0, C->s, B->lx[1].p, B->lr[1].p, n1, partX2);
        SGLD_HIP(hipGetLastError());
        if (nbX2) *nbX2 = nb;
    }
    return SGLD_OK;
}

// The finest level of the V-cycle folded into the direction update (see
// k_bprolong_dir): p = mask(P x_2 + dinv r) + beta p, xc += alpha p.
int enqueue_bdir(sgld_ctx *C, const float *r, float *p) {
    BatchState *B = C->batch;
    const Level &F = C->lv[0], &Cc = C->lv[1];
    const int nb = bgrid(F.g.nodes() * NR);
    const NodeDiv df = NodeDiv::make(F.g);
    if (Cc.span == 2)
        hipLaunchKernelGGL(k_bprolong_dir<2>, dim3(nb), dim3(256), 0, C->s, B->lx[1].p, Cc.g, F.g, df, r,
                           F.dinv32.p, F.nflag.p, B->xc.p, p, B->ctl.p);
    else
        hipLaunchKernelGGL(k_bprolong_dir<4>, dim3(nb), dim3(256), 0, C->s, B->lx[1].p, Cc.g, F.g, df, r,
                           F.dinv32.p, F.nflag.p, B->xc.p, p, B->ctl.p);
    SGLD_HIP(hipGetLastError());
    return SGLD_OK;
}

// rho = z.r per case for a residual set outside the iteration: sum dinv r^2
// (into part2) plus x_2.r_2 from the coarse V-cycle (into part3).
int enqueue_brho(sgld_ctx *C, const float *r, int64_t n16) {
    BatchState *B = C->batch;
    const int nbv = bgrid(n16);
    hipLaunchKernelGGL(k_bdinv_rr, dim3(nbv), dim3(256), 0, C->s, r, C->lv[0].dinv32.p, n16, B->part2.p);
    SGLD_HIP(hipGetLastError());
    int nbX2 = 0;
    SGLD_TRY(enqueue_bvcycle_coarse(C, r, B->part3.p, &nbX2));
    return bfin(C, nbX2, BFIN_RHO, B->part3.p, B->part2.p, nbv);
}

int upload_ctl(sgld_ctx *C) {
    SGLD_HIP(hipMemcpyAsync(C->batch->ctl.p, C->batch->hctl, sizeof(BCtl), hipMemcpyHostToDevice, C->s));
    return SGLD_OK;
}
int download_ctl(sgld_ctx *C) {
    SGLD_HIP(hipMemcpyAsync(C->batch->hctl, C->batch->ctl.p, sizeof(BCtl), hipMemcpyDeviceToHost, C->s));
    SGLD_HIP(hipStreamSynchronize(C->s));
    return SGLD_OK;
}

// Batched mixed solve of nrhs cases; dF, dU compressed FP64 (numDOF x nrhs,
// case-major), dU holds the initial guesses.
int solve_batched(sgld_ctx *C, int nrhs, const double *dF, double *dU, double tol, int maxit, int *itsOut,
                  double *relres, int *updatesOut) {
    SGLD_TRY(batch_init(C));
    BatchState *B = C->batch;
    const int64_t nd = B->nd, n16 = nd * NR;
    const int64_t nc = 3 * C->lv[0].numNodes;
    double *b64 = B->b64.p, *x64 = B->x64.p, *r64 = B->r64.p;
    SGLD_HIP(hipMemsetAsync(B->xc.p, 0, sizeof(float) * n16, C->s));
    SGLD_HIP(hipMemsetAsync(B->p0.p, 0, sizeof(float) * n16, C->s));
    SGLD_HIP(hipMemsetAsync(B->r.p, 0, sizeof(float) * n16, C->s));
    std::vector<double> scale(NR, 0.0), normb(NR, 0.0), itol(NR, 0.0), fct(NR, 1.0), inv(NR, 0.0);
    std::vector<int> done(NR, 1), itsAtUpdate(NR, 0);
    double delta = 0.1;
    if (const char *env = std::getenv("SGLD_MIXED_DELTA")) delta = std::atof(env);
    const int maxStretch = 96;
    BCtl &h = *B->hctl;
    std::memset(&h, 0, sizeof(BCtl));
    int updates = 0;
    uint32_t mask = 0;
    for (int j = 0; j < nrhs; ++j) {
        SGLD_TRY(to_lattice(C, dF + j * nc, b64 + j * nd));
        SGLD_TRY(to_lattice(C, dU + j * nc, x64 + j * nd));
        double rr = 0, bb = 0;
        SGLD_TRY(residual64(C, b64 + j * nd, x64 + j * nd, r64 + j * nd, &rr, &bb));
        normb[j] = std::sqrt(bb);
        relres[j] = bb > 0 ? std::sqrt(rr / bb) : 0.0;
        itsOut[j] = 0;
        if (!(bb > 0) || relres[j] < tol || !(rr > 0)) continue;
        done[j] = 0;
        scale[j] = std::sqrt(rr);
        inv[j] = 1.0 / scale[j];
        itol[j] = std::max(delta, 0.5 * tol * normb[j] / scale[j]);
        h.active[j] = 1;
        h.restart[j] = 1;
        mask |= 1u << j;
    }
    if (!mask) return SGLD_OK;
    const int nbv = bgrid(n16);
    SGLD_HIP(hipMemcpyAsync(B->scal.p, inv.data(), sizeof(double) * NR, hipMemcpyHostToDevice, C->s));
    SGLD_HIP(hipMemcpyAsync(B->scal.p + NR, fct.data(), sizeof(double) * NR, hipMemcpyHostToDevice, C->s));
    hipLaunchKernelGGL(k_breplace, dim3(nbv), dim3(256), 0, C->s, B->r.p, B->p0.p, r64, nd, B->scal.p,
                       B->scal.p + NR, mask, 0u);
    SGLD_HIP(hipGetLastError());
    SGLD_TRY(upload_ctl(C));
    // Rotated iteration (as the single-RHS solver): the finest level of the
    // previous V-cycle, its deferred xc += alpha p and p = z + beta p run as
    // one pass; p starts at 0 so that the first direction is z.
    float *p = B->p0.p;
    int nb = 0;
    SGLD_TRY(enqueue_brho(C, B->r.p, n16));  // restart: beta = 0
    while (true) {
        SGLD_TRY(enqueue_bdir(C, B->r.p, p));
        SGLD_TRY(enqueue_bkbyu(C, p, B->q.p, &nb));
        SGLD_TRY(bfin(C, nb, BFIN_ALPHA));
        hipLaunchKernelGGL(k_bupdate_r, dim3(nbv), dim3(256), 0, C->s, B->r.p, (const float *)B->q.p,
                           C->lv[0].dinv32.p, n16, B->ctl.p, B->part.p, B->part2.p);
        SGLD_HIP(hipGetLastError());
        SGLD_TRY(bfin(C, nbv, BFIN_RR));
        SGLD_TRY(download_ctl(C));
        uint32_t flag = 0, forced = 0;
        for (int j = 0; j < NR; ++j) {
            if (done[j] || !h.active[j]) continue;
            const bool conv = h.res[j] < itol[j];
            const bool stretch = h.its[j] - itsAtUpdate[j] >= maxStretch;
            if (conv || stretch || h.its[j] >= maxit || !(h.res[j] == h.res[j])) flag |= 1u << j;
            if (!conv && stretch) forced |= 1u << j;
        }
        // An update event sweeps the interleaved vectors whatever the number
        // of cases it serves, so cases already past sqrt(delta) of their
        // drop join it (an early reliable update only adds accuracy).
        if (flag)
            for (int j = 0; j < NR; ++j)
                if (!done[j] && h.active[j] && !((flag >> j) & 1u) && h.res[j] < std::sqrt(itol[j]))
                    flag |= 1u << j;
        if (flag) {
            std::vector<double> sc(NR, 0.0);
            for (int j = 0; j < NR; ++j) sc[j] = scale[j];
            SGLD_HIP(hipMemcpyAsync(B->scal.p + 2 * NR, sc.data(), sizeof(double) * NR, hipMemcpyHostToDevice,
                                    C->s));
            hipLaunchKernelGGL(k_bfold, dim3(nbv), dim3(256), 0, C->s, x64, B->xc.p, (const float *)p, nd,
                               B->scal.p + 2 * NR, flag, B->ctl.p);
            SGLD_HIP(hipGetLastError());
            uint32_t keep = 0, zero = 0;
            for (int j = 0; j < NR; ++j) {
                if (!((flag >> j) & 1u)) continue;
                h.alpha[j] = 0.0;  // the deferred alpha p is in x64 now
                double rr = 0, bb = 0;
                SGLD_TRY(residual64(C, b64 + j * nd, x64 + j * nd, r64 + j * nd, &rr, &bb));
                ++updates;
                relres[j] = std::sqrt(rr / bb);
                itsOut[j] = h.its[j];
                if (relres[j] < tol || h.its[j] >= maxit || !(rr > 0)) {
                    done[j] = 1;
                    h.active[j] = 0;
                    zero |= 1u << j;
                    continue;
                }
                fct[j] = scale[j] / std::sqrt(rr);
                scale[j] = std::sqrt(rr);
                inv[j] = 1.0 / scale[j];
                itol[j] = std::max(delta, 0.5 * tol * normb[j] / scale[j]);
                h.rho[j] *= fct[j] * fct[j];
                h.restart[j] = (forced >> j) & 1u;
                itsAtUpdate[j] = h.its[j];
                keep |= 1u << j;
            }
            SGLD_HIP(hipMemcpyAsync(B->scal.p, inv.data(), sizeof(double) * NR, hipMemcpyHostToDevice, C->s));
            SGLD_HIP(hipMemcpyAsync(B->scal.p + NR, fct.data(), sizeof(double) * NR, hipMemcpyHostToDevice,
                                    C->s));
            hipLaunchKernelGGL(k_breplace, dim3(nbv), dim3(256), 0, C->s, B->r.p, p, r64, nd, B->scal.p,
                               B->scal.p + NR, keep, zero);
            SGLD_HIP(hipGetLastError());
            SGLD_TRY(upload_ctl(C));
            // the replaced residuals' share of rho (all cases: one pass)
            hipLaunchKernelGGL(k_bdinv_rr, dim3(nbv), dim3(256), 0, C->s, (const float *)B->r.p,
                               C->lv[0].dinv32.p, n16, B->part2.p);
            SGLD_HIP(hipGetLastError());
        }
        bool any = false;
        for (int j = 0; j < NR; ++j) any |= !done[j];
        if (!any) break;
        int nbX2 = 0;
        SGLD_TRY(enqueue_bvcycle_coarse(C, B->r.p, B->part3.p, &nbX2));
        SGLD_TRY(bfin(C, nbX2, BFIN_RHO, B->part3.p, B->part2.p, nbv));
    }
    for (int j = 0; j < nrhs; ++j) SGLD_TRY(from_lattice(C, x64 + j * nd, dU + j * nc));
    SGLD_HIP(hipStreamSynchronize(C->s));
    *updatesOut = updates;
    return SGLD_OK;
}

}  // namespace
}  // namespace sgld

using namespace sgld;

// ================================= C ABI ====================================
extern "C" int sgld_ctx_pcg_batched_device(sgld_ctx *C, int nrhs, const double *dF, double *dU, double tol,
                                           int maxIT, int precision, int *its, double *relres) {
    SGLD_TRY(check_ready(C));
    SGLD_ARG(nrhs >= 1 && nrhs <= NR, "sgld:arg: nrhs must be in 1..16");
    SGLD_ARG(dF && dU && its && relres, "sgld:arg: null pointer");
    SGLD_ARG(tol > 0 && maxIT > 0, "sgld:arg: need tol > 0 and maxIT > 0");
    const int64_t nc = 3 * C->lv[0].numNodes;
    const double t0 = now_s();
    int allConv = 1;
    if (precision == SGLD_PREC_FP64 || precision == SGLD_PREC_MIXED) {
        // sequential single-case solves: FP64 (the reference recurrence) or
        // mixed precision.  The mixed single-case solver beats the
        // interleaved batch on MI355X (1.90 vs 2.29 s for 16 cases of a 256^3
        // cuboid): per case the batch's vector passes cost the same and its
        // matrix-core K.u costs more than the FP32 mode-space kernel.
        int updates = 0;
        for (int j = 0; j < nrhs; ++j) {
            double rr = 0;
            const int rc = sgld_ctx_pcg_device(C, dF + j * nc, dU + j * nc, tol, maxIT, precision, its + j, &rr);
            if (rc != SGLD_OK && rc != SGLD_ERR_MAXIT) return rc;
            relres[j] = rr;
            updates += (int)C->stats[3];
            if (rc == SGLD_ERR_MAXIT) allConv = 0;
        }
        C->stats[0] = now_s() - t0;
        C->stats[3] = updates;
        if (!allConv) {
            set_error("sgld:maxit: some of %d cases did not reach tol %.3e within %d iterations", nrhs, tol, maxIT);
            return SGLD_ERR_MAXIT;
        }
        return SGLD_OK;
    }
    SGLD_ARG(precision == SGLD_PREC_INTERLEAVED,
             "sgld:arg: precision must be SGLD_PREC_FP64, SGLD_PREC_MIXED or SGLD_PREC_INTERLEAVED");
    int updates = 0;
    SGLD_TRY(solve_batched(C, nrhs, dF, dU, tol, maxIT, its, relres, &updates));
    C->stats[0] = now_s() - t0;
    C->stats[3] = updates;
    int maxits = 0;
    for (int j = 0; j < nrhs; ++j) {
        maxits = std::max(maxits, its[j]);
        if (!(relres[j] < tol)) allConv = 0;
    }
    C->stats[1] = maxits;
    if (!allConv) {
        set_error("sgld:maxit: %d of %d cases did not reach tol %.3e within %d iterations", nrhs, nrhs, tol, maxIT);
        return SGLD_ERR_MAXIT;
    }
    return SGLD_OK;
}

extern "C" int sgld_ctx_pcg_batched(sgld_ctx *C, int nrhs, const double *F, double *U, double tol, int maxIT,
                                    int precision, int *its, double *relres) {
    SGLD_TRY(check_ready(C));
    SGLD_ARG(nrhs >= 1 && nrhs <= NR, "sgld:arg: nrhs must be in 1..16");
    SGLD_ARG(F && U, "sgld:arg: null pointer");
    const int64_t nc = 3 * C->lv[0].numNodes;
    DevBuf<double> dF, dU;
    SGLD_TRY(dF.upload(F, (size_t)nc * nrhs));
    SGLD_TRY(dU.upload(U, (size_t)nc * nrhs));
    const int rc = sgld_ctx_pcg_batched_device(C, nrhs, dF.p, dU.p, tol, maxIT, precision, its, relres);
    if (rc != SGLD_OK && rc != SGLD_ERR_MAXIT) return rc;
    SGLD_HIP(hipMemcpy(U, dU.p, sizeof(double) * nc * nrhs, hipMemcpyDeviceToHost));
    return rc;
}

extern "C" int sgld_ctx_kbyu_batched_device(sgld_ctx *C, int nrhs, const double *dU, double *dY) {
    SGLD_ARG(C, "sgld:arg: null context");
    SGLD_ARG(C->haveModulus && C->haveFixed, "sgld:state: call sgld_ctx_set_fixed and sgld_ctx_set_modulus first");
    SGLD_ARG(nrhs >= 1 && nrhs <= NR, "sgld:arg: nrhs must be in 1..16");
    SGLD_ARG(dU && dY, "sgld:arg: null pointer");
    SGLD_TRY(batch_init(C));
    BatchState *B = C->batch;
    const int64_t nd = B->nd, nc = 3 * C->lv[0].numNodes;
    SGLD_HIP(hipMemsetAsync(B->r64.p, 0, sizeof(double) * nd * NR, C->s));
    for (int j = 0; j < nrhs; ++j) SGLD_TRY(to_lattice(C, dU + j * nc, B->r64.p + j * nd));
    const int nbv = bgrid(nd * NR);
    hipLaunchKernelGGL(k_bpack, dim3(nbv), dim3(256), 0, C->s, B->r64.p, nd, nrhs, B->p0.p);
    SGLD_HIP(hipGetLastError());
    int nb = 0;
    SGLD_TRY(enqueue_bkbyu(C, B->p0.p, B->q.p, &nb));
    hipLaunchKernelGGL(k_bunpack, dim3(nbv), dim3(256), 0, C->s, B->q.p, nd, nrhs, B->r64.p);
    SGLD_HIP(hipGetLastError());
    for (int j = 0; j < nrhs; ++j) SGLD_TRY(from_lattice(C, B->r64.p + j * nd, dY + j * nc));
    SGLD_HIP(hipStreamSynchronize(C->s));
    return SGLD_OK;
}

extern "C" int sgld_ctx_bench_kbyu_batched(sgld_ctx *C, int reps, double *msPerLaunch) {
    SGLD_ARG(C && msPerLaunch && reps > 0, "sgld:arg: bad arguments");
    SGLD_ARG(C->haveModulus && C->haveFixed, "sgld:state: call sgld_ctx_set_fixed and sgld_ctx_set_modulus first");
    SGLD_TRY(batch_init(C));
    BatchState *B = C->batch;
    int nb = 0;
    SGLD_TRY(enqueue_bkbyu(C, B->p0.p, B->q.p, &nb));  // warm-up
    SGLD_HIP(hipEventRecord(C->e0, C->s));
    for (int i = 0; i < reps; ++i) SGLD_TRY(enqueue_bkbyu(C, B->p0.p, B->q.p, &nb));
    SGLD_HIP(hipEventRecord(C->e1, C->s));
    SGLD_HIP(hipEventSynchronize(C->e1));
    float ms = 0;
    SGLD_HIP(hipEventElapsedTime(&ms, C->e0, C->e1));
    *msPerLaunch = ms / reps;
    return SGLD_OK;
}
