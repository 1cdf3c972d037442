// Host-side state of one solver context (the device-resident mesh hierarchy,
// the meshHierarchy_ of src/Data_CartesianMeshStruct.m) shared by the
// single-RHS solver (lattice.hip) and the batched multi-RHS solver
// (batched.hip).
#pragma once
#include <algorithm>
#include <vector>

#include "lattice_kernels.hpp"

namespace sgld {

constexpr int kTE = 16;        // element columns per K.u patch edge
constexpr int kTN = kTE + 1;   // staged nodes per patch edge
constexpr int kTO = kTE - 1;   // output nodes per patch edge
constexpr int kXCD = 8;        // accelerator complex dies (one L2 each)
constexpr int kStageSlots = (3 * kTN * kTN + 255) / 256;  // staged values per thread and node layer

// Work decomposition of the patch-tiled K.u: tilesR x tilesC patches of
// kTO x kTO output node columns, each cut into z-chunks of cz node layers.
struct TileMap {
    int tilesR, tilesC, tilesZ, cz;
    __host__ __device__ int64_t work() const { return (int64_t)tilesR * tilesC * tilesZ; }
};

// Schedule of the patch-tiled K.u for one precision: work decomposition and
// persistent grid = min(work items, cap), cap = resident blocks on the device.
struct TiledSchedule {
    TileMap tm{1, 1, 1, 8};
    int grid = 1, cap = 1;
    void set_chunk(int NZ, int cz) {
        tm.cz = cz;
        tm.tilesZ = (NZ + cz - 1) / cz;
        const int64_t w = tm.work();
        grid = (int)(w < cap ? w : cap);
        if (grid < 1) grid = 1;
    }
};

struct Level {
    Grid g{};
    int span = 1;
    int64_t numElements = 0, numNodes = 0;
    std::vector<uint8_t> esolid;  // host, lattice elements
    std::vector<uint8_t> hflag;   // host, lattice nodes
    DevBuf<uint8_t> nflag;
    DevBuf<double> dinv, r, x;
    DevBuf<float> dinv32, r32, x32;
    DevBuf<uint16_t> dinv16;  // finest level: bfloat16 Jacobi weights of the FP32 inner solve
    DevBuf<int32_t> rtiles;   // irregular domains: tiles of the tiled restriction into this level
    int64_t nRtiles = 0;
    int rtilesCz = 0;         // z-chunk the tile list was built for (0: none)
    DevBuf<int32_t> dofChunks;  // irregular domains, levels >= 2: DOF chunks with a solid node
    int64_t nDofChunks = 0;
};

struct BatchState;             // batched.hip
void batch_free(BatchState *b);

// Launches recorded instead of issued, for the multi-instance iteration:
// while a context's `rec` is set, the kernels of one PCG iteration are
// appended here (step: the launch's position in the iteration, a sort key
// shared by all instances; multi: the launcher of the same kernel over many
// instances; args: the kernel's argument block).
struct RecOp {
    int step;
    void (*multi)(const void *args, const int *start, int n, int grid, size_t shm, hipStream_t s);
    int grid;
    size_t shm;
    std::vector<char> args;
};
struct OpRec {
    std::vector<RecOp> ops;
    int base = 0;  // added to the steps of the launches recorded next
};

// HIP-event timing of the finest-level passes of a solve, per kernel and
// precision (ids in the order of SGLD_KT_* in sgldfea.h).  Each recorded
// launch carries its iteration's position in the queued batch (-1 outside
// the iterations); once the batch has completed, launches of iterations after
// the stopping one (their kernels see the device-side stop flag and return
// at once) are told apart from the ones that did the work.
struct KernelTimer {
    static constexpr int kN = 8;
    bool on = false, inSolve = false;
    int curIter = -1;
    struct Rec { int id, iter; hipEvent_t a, b; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[kN] = {}, msWork[kN] = {};
    int64_t n[kN] = {}, nWork[kN] = {};
    const char *name[kN] = {};  // kernel of the last timed launch per id
    void reset() {
        for (int k = 0; k < kN; ++k) { ms[k] = msWork[k] = 0.0; n[k] = nWork[k] = 0; }
    }
};

}  // namespace sgld

struct sgld_ctx {
    int L = 0, nonDyadic = 0;
    double nu = 0.3, h = 1.0, omega = 0.6;
    double Ke[576];
    sgld::DevBuf<double> KeDev;   // Ke on the device (FP64, column-major)
    sgld::DevBuf<float> KeDev32;  // and in FP32
    sgld::ModeK<double> km64{};   // mode-space coefficients of Ke (tiled K.u)
    int haveModes = 0;            // Ke has the isotropic-cube mode structure (km64 / km32 valid)
    sgld::ModeK<float> km32{};
    std::vector<sgld::Level> lv;
    bool full = true;
    std::vector<int32_t> hNodMap, hEleMap;  // compressed -> lattice (0-based)
    sgld::DevBuf<int32_t> nodMap, eleMap;
    sgld::DevBuf<int32_t> efwd;  // voxel -> 1-based element id (0 void), built on first use
    sgld::DevBuf<double> stressT;  // element nodal-stress operator (48 x 24), built on first use
    sgld::DevBuf<double> E;
    sgld::DevBuf<float> E32;
    // coarsest level: packed rows / columns of W = L^-1 (K_L(free, free) = L L^T),
    // its free lattice DOFs (ascending) and the intermediate W r (16 cases wide)
    sgld::DevBuf<double> cWrow, cWcol, cY;
    sgld::DevBuf<float> cWrow32, cWcol32;  // FP32 copies for the FP32 inner solve (n > 4096 only)
    sgld::DevBuf<int32_t> cFree;
    int64_t nCoarseFree = 0;
    sgld::DevBuf<double> b, x, r, z, p, p2, q;
    sgld::DevBuf<float> b32, r32, z32, p32, p32b, q32, x32;
    // Residual update fused with the first restriction (fuseRes: FP64,
    // fuseRes32: FP32 inner solve): the iterations alternate r / rB (FP64)
    // and r32 / r32B (FP32), an even batch ending in r / r32; allocated on the
    // first solve that fuses.  dinv16L: the FP32 inner solve's bfloat16
    // finest weights in lattice order (irregular domains; lv[0].dinv16 holds
    // them in span order).  Both are set by size in sgld_ctx_create.
    bool fuseRes = false, fuseRes32 = false;
    sgld::DevBuf<double> rB;
    sgld::DevBuf<float> r32B;
    sgld::DevBuf<uint16_t> dinv16L;
    sgld::DevBuf<double> partial;
    sgld::DevBuf<double> pdeWork;  // PDE filter work vectors, allocated on first use and kept
    sgld::DevBuf<double> filterWork;  // density filter: lattice copy of x and the weights, kept
    sgld::DevBuf<sgld::Ctl> ctl;
    sgld::Ctl *hctl = nullptr;  // pinned host mirror
    double *hpart = nullptr;    // pinned: ||r||^2 and ||f||^2 of a reliable update
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    bool haveFixed = false, haveModulus = false, assembled = false;
    int kbyuVariant = 0;  // 0 node-centric dense Ke, 1 tiled mode-space
    bool kbyuOneBar = true;  // tiled K.u with one block barrier per element layer (else two)
    bool kbyuOwner = true;   // tiled K.u with owner-computes node sums (else LDS atomics)
    // owner K.u, [0] FP64 / [1] FP32: outputs through LDS for contiguous stores
    // (set when the output vector exceeds the Infinity Cache share, see sgld_ctx_create)
    bool kbyuStageOut[2] = {true, true};
    int restrictTiled = -1;  // span-2 restriction: -1 by size, 0 node-centric, 1 LDS-tiled, 2 tiled into level 1 only
    int64_t restrictTiledMin = (int64_t)1 << 18;    // "by size": FP64 tiled from this many coarse nodes
    // and FP32 (125 M-DOF ellipsoids, 17 M coarse nodes: tiled 218 us, node-centric 222 us)
    int64_t restrictTiledMin32 = (int64_t)1 << 22;
    bool tailOn = true;  // small levels of the V-cycle as the fused two-launch tail
    bool restrictSkipVoid = true;  // irregular domains: the tiled restriction stages no void fine node
    // coarse layers per block of the tiled span-2 restriction and the fused
    // residual update: 3 (125 M-DOF ellipsoids, mixed / FP64 solve 760 / 1444
    // ms at 2, 755 / 1441 ms at 3; part_R256 FP64 55.6 -> 54.7 ms; the
    // 256x128x128 cuboid within 1 %)
    int restrictCz = 3;
    int vecBlocks = sgld::kVecBlocks;  // grid of the finest-level direction and residual updates
    bool dirXcd = true;  // direction update with one contiguous DOF range per XCD
    bool streamNT[2] = {false, false};  // [0] FP64 / [1] FP32 streaming passes with nontemporal loads
    bool dirCols = true;  // FP64 direction update with LDS-staged coarse columns (span 2, full cuboid)
    bool bRestrictTiled = true;  // batched span-2 restriction: LDS-tiled layer march (else node-centric)
    sgld::TiledSchedule ksched[2];  // tiled K.u schedule: [0] FP64, [1] FP32
    // Irregular domains (full == false): the finest-level passes visit only
    // what holds solid material.  dofChunks / nodeChunks list the chunks of
    // 2^chunkLog entries of the DOF and node vectors with a solid node; kwork[q] lists the
    // tiled K.u work items of schedule q whose output nodes touch a solid
    // element; hPatchLayer[l * patches + patch] says whether element layer l of
    // a patch's 16 x 16 element window holds a solid element.
    sgld::DevBuf<int32_t> dofChunks, nodeChunks, kwork[2];
    int64_t nDofChunks = 0, nNodeChunks = 0, nKwork[2] = {0, 0};
    int chunkLog = 5;  // chunks of 2^chunkLog entries: 16, 32 or 64
    std::vector<uint8_t> hPatchLayer;
    int kOcc[2] = {1, 1};           // resident tiled K.u blocks per compute unit
    int numCU = 0;
    double stats[8] = {0};
    sgld::KernelTimer kt;
    sgld::BatchState *batch = nullptr;  // multi-RHS buffers, allocated on first use
    sgld::OpRec *rec = nullptr;          // set while the multi-instance iteration is recorded
    int64_t nDof() const { return 3 * lv[0].g.nodes(); }
};

namespace sgld {

inline int grid_blocks(int64_t n) {
    return (int)std::min<int64_t>(kVecBlocks, std::max<int64_t>(1, (n + 255) / 256));
}
// Finest-level DOF / node passes: every entry of a full cuboid, the listed
// chunks of an irregular domain.
inline Span dof_span(const sgld_ctx *C) {
    return C->full ? Span::all(C->nDof()) : Span::listed(C->dofChunks.p, C->nDof(), C->nDofChunks, C->chunkLog);
}
inline Span node_span(const sgld_ctx *C) {
    const int64_t n = C->lv[0].g.nodes();
    return C->full ? Span::all(n) : Span::listed(C->nodeChunks.p, n, C->nNodeChunks, C->chunkLog);
}
// Sets the z-chunk of tiled K.u schedule q (0 FP64, 1 FP32) and, on an
// irregular domain, rebuilds its list of work items.
int set_kchunk(sgld_ctx *C, int q, int cz);
double now_s();
int to_lattice(sgld_ctx *C, const double *dsrc, double *ddst);    // compressed -> lattice DOFs
int from_lattice(sgld_ctx *C, const double *dsrc, double *ddst);  // lattice -> compressed DOFs
int check_ready(sgld_ctx *C);
// r = b - K x on lattice FP64 vectors (fixed and void DOFs 0), with
// rr = r.r and bb = b.b; host-synchronous.
int residual64(sgld_ctx *C, const double *b, const double *x, double *r, double *rr, double *bb);
int residual64_enqueue(sgld_ctx *C, const double *b, const double *x, double *r);
void residual64_sums(const sgld_ctx *C, double *rr, double *bb);
int ensure_efwd(sgld_ctx *C);  // voxel -> 1-based element id map of the finest level
// Clears the batched coarsest-level x (written at free DOFs only) if allocated.
int batch_clear_coarse(sgld_ctx *C);

}  // namespace sgld
