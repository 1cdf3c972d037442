// Device kernels of the structured (lattice) solver path.
//
// Layout in HBM.  Every multigrid level k is a full node lattice of
// (ny+1) x (nx+1) x (nz+1) nodes, id g = row + (ny+1) col + (ny+1)(nx+1) layer
// (row 0 at the top, as src/FEA_VoxelBasedDiscretization.m:71-85 numbers
// nodes); DOF vectors are interleaved, 3g + {0,1,2}.  Void nodes hold zeros,
// so the compressed numbering of the reference (solid nodes in ascending g) is
// a gather of this layout and, for a full cuboid, identical to it.  Element
// data (modulus E, Galerkin Ks) is indexed by the voxel lattice id
// v = row + ny col + ny nx layer.  Per-node flags: bits 0-2 fixed x/y/z, bit
// 3 solid node.
//
// Because the grid is implicit, the transfers are fixed trilinear stencils:
// R = P^T with P(n, N) = prod_d (1 - |n_d - s N_d| / s), which is exactly the
// reference's ghost-node operator (multiplicity-averaged duplicates of the same
// trilinear value, Solving_Interpolation_MatrixFree_mex.cpp:75-129).
#pragma once
#include "common.hpp"

namespace sgld {

constexpr uint8_t kSolid = 8;

struct Grid {
    int nx, ny, nz;  // elements per axis
    __host__ __device__ int64_t NX() const { return nx + 1; }
    __host__ __device__ int64_t NY() const { return ny + 1; }
    __host__ __device__ int64_t NZ() const { return nz + 1; }
    __host__ __device__ int64_t nodes() const { return NX() * NY() * NZ(); }
    __host__ __device__ int64_t elems() const { return (int64_t)nx * ny * nz; }
};

// Division by a fixed d for 0 <= n < 2^31 without a divide instruction:
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d) and
// m = floor(2^32 (2^s - d) / d) + 1 (Granlund & Montgomery 1994, unsigned
// round-up variant; the sum cannot wrap because umulhi(n, m) < n < 2^31).
struct FastDiv {
    uint32_t d = 1, m = 1, s = 0;
    static FastDiv make(uint32_t dv) {
        FastDiv f;
        f.d = dv;
        f.s = 0;
        while ((1ull << f.s) < dv) ++f.s;
        f.m = (uint32_t)((((1ull << 32) * ((1ull << f.s) - dv)) / dv) + 1);
        return f;
    }
    __host__ __device__ static constexpr FastDiv make3() { return FastDiv{3u, 1431655766u, 2u}; }
    // (64-bit product >> 32 of zero-extended operands compiles to v_mul_hi_u32)
    __host__ __device__ uint32_t div(uint32_t n) const {
        return ((uint32_t)(((uint64_t)n * m) >> 32) + n) >> s;
    }
};

// (row, col, layer) of lattice node n < 2^31 of a grid with NY x NX nodes per layer.
struct NodeDiv {
    FastDiv NY, NX;
    static NodeDiv make(const Grid &g) { return NodeDiv{FastDiv::make(g.ny + 1), FastDiv::make(g.nx + 1)}; }
    __host__ __device__ __forceinline__ void rcl(uint32_t n, int &r, int &c, int &l) const {
        const uint32_t q = NY.div(n);
        r = (int)(n - q * NY.d);
        const uint32_t lq = NX.div(q);
        c = (int)(q - lq * NX.d);
        l = (int)lq;
    }
};

// The nine numbers of the isotropic cube's mode-space element stiffness
// K' = H Ke H^T / 64 (see mode_forces in lattice.hip), a kernel argument.
template <class T>
struct ModeK {
    T v[9];
};

// Device-side PCG state (one per solve); all solver kernels read it so that a
// batch of iterations can be queued (or graph-captured) without host syncs.
struct Ctl {
    double rho, pq, rr, normb, lambda, beta, res, tol;
    int its, maxit, done, flags;
};
// Ctl::flags.  kCtlSoft: reaching tol does not stop the recursion but sets
// kCtlWant (the mixed solve's FP32 stretch keeps iterating until its reliable
// update runs, instead of idling until then).
constexpr int kCtlSoft = 1, kCtlWant = 2;

// Entries visited by a grid-stride pass over a vector of length n: all of
// them when `chunks` is null, else the listed chunks of 2^log entries (chunk
// c is [c 2^log, (c + 1) 2^log) clipped to n).  Irregular domains list the
// chunks that hold a solid node; entries outside them are void, stay zero and
// are skipped.  Position t of the pass maps to entry at(t).  The positions
// of one wavefront share t >> 6 (grid strides are multiples of 64), so its
// chunk ids are scalar loads: one for 64-entry chunks, two for 32-entry and
// four for 16-entry ones (which visit fewer void entries; a list carries
// three padding entries so that these loads stay inside it).
struct Span {
    const int32_t *chunks;
    int64_t n, count;  // count = n, or 2^log x listed chunks
    int log;
    __host__ __device__ static Span all(int64_t len) { return Span{nullptr, len, len, 6}; }
    __host__ __device__ static Span listed(const int32_t *ch, int64_t len, int64_t nChunks, int lg) {
        return Span{ch, len, nChunks << lg, lg};
    }
    __device__ __forceinline__ int64_t at(int64_t t) const {
        if (!chunks) return t;
        if (log == 6) {
            const int ci = __builtin_amdgcn_readfirstlane((int)(t >> 6));
            return ((int64_t)chunks[ci] << 6) + (t & 63);
        }
        const int cw = __builtin_amdgcn_readfirstlane((int)(t >> 6));
        if (log == 5) {
            const int32_t c0 = chunks[2 * cw], c1 = chunks[2 * cw + 1];
            return ((int64_t)((t & 32) ? c1 : c0) << 5) + (t & 31);
        }
        const int32_t c0 = chunks[4 * cw], c1 = chunks[4 * cw + 1], c2 = chunks[4 * cw + 2],
                      c3 = chunks[4 * cw + 3];
        const int q = (int)(t >> 4) & 3;
        return ((int64_t)(q == 0 ? c0 : q == 1 ? c1 : q == 2 ? c2 : c3) << 4) + (t & 15);
    }
};

// Packed storage of the coarsest level's inverse Cholesky factor W = L^-1 over
// its n free DOFs: row i of W (entries 0..i) at tri_row_off(i), column j
// (entries j..n-1, a row of W^T) at tri_col_off(j, n).
__host__ __device__ inline int64_t tri_row_off(int64_t i) { return i * (i + 1) / 2; }
__host__ __device__ inline int64_t tri_col_off(int64_t j, int64_t n) { return j * n - j * (j - 1) / 2; }
// Largest coarsest free-DOF count of the dense factor: W is stored packed
// twice (rows and columns, 8 n^2 bytes: 8.6 GB at the limit) and both copies
// are read once per V-cycle; set-up holds the dense n x n matrix on the host
// and two on the device and takes about 2 n^3 / 3 flops.
constexpr int64_t kCoarseMaxFree = 32768;

constexpr int kRedBlocks = 8192;  // partial-sum slots of one reduction (max grid of a reducing kernel)
constexpr int kVecBlocks = 2048;  // grid of the grid-stride vector kernels
constexpr int kRedSlots = 4;      // independent reductions per solve

}  // namespace sgld
