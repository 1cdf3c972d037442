/*
 * sgldfea.h -- C ABI of libsgldfea, the MI355X (gfx950) HIP implementation of
 * SGLDBench's voxel-FEA solve: matrix-free K.u, additive geometric-multigrid
 * V-cycle, PCG driver and unit compliance.
 *
 * Conventions (those of the reference's MEX interface):
 *   - int32 ids are 1-based; matrices are column-major;
 *   - DOF vectors are FP64, interleaved [ux uy uz] per compressed node;
 *   - n x 3 node arrays (transfers) are column-major (component-major);
 *   - voxel arrays are resY x resX x resZ column-major, row 0 at the top.
 * Every entry point returns an int status (SGLD_OK = 0); on failure
 * sgld_last_error() returns a thread-local message that starts with the
 * reference's error identifier where one exists.
 *
 * The library is thread-compatible, not thread-safe: one context must not be
 * used from two threads at once.  Stateless entry points take host pointers,
 * allocate device memory per call and copy results back, as the MEX calls
 * they replace hand arrays back to MATLAB.
 */
#ifndef SGLDFEA_H
#define SGLDFEA_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum {
    SGLD_OK = 0,
    SGLD_ERR_ARG = 1,       /* invalid argument (shape, range, null) */
    SGLD_ERR_HIP = 2,       /* HIP runtime failure */
    SGLD_ERR_NOT_SPD = 3,   /* coarsest matrix not positive definite */
    SGLD_ERR_MAXIT = 4,     /* PCG hit maxIT without reaching tol */
    SGLD_ERR_STATE = 5,     /* call out of order (e.g. solve before assemble) */
    SGLD_ERR_NOMEM = 6      /* device or host allocation failed */
};

enum { SGLD_PREC_FP64 = 0, SGLD_PREC_MIXED = 1,
       /* sgld_ctx_pcg_batched only: the 16 cases interleaved per DOF */
       SGLD_PREC_INTERLEAVED = 2 };

const char *sgld_last_error(void);
/* Library version string and the gfx target it was built for. */
const char *sgld_version(void);

/* ===================== stateless MEX replacements ===================== */

/* Replaces Solving_KbyU_MatrixFree_mex (src/MEXfuncs/Solving_KbyU_MatrixFree_mex.cpp:6-103),
 * called at src/Solving_KbyU_MatrixFree.m:9.  Y = sum_e E_e scatter(Ke gather_e(U)),
 * accumulated colour by colour (colors: concatenated 1-based element ids of
 * numColors colours, colorOffsets[numColors+1] prefix offsets). */
int sgld_kbyu(const double *U, int64_t numDOFs, const int32_t *eNodMat, int64_t numElements,
              const double *Ke, int64_t keRows, int64_t keCols, const double *E, int64_t numE,
              const int32_t *colors, const int64_t *colorOffsets, int numColors, double *Y);

/* Replaces Solving_KbyU_MatrixFree8x8_mex
 * (src/MEXfuncs/Solving_KbyU_MatrixFree8x8_mex.cpp:6-121), called at
 * src/TopOpti_PDEFiltering_matrixFree.m:46 (the PDE filter's CG).  One scalar
 * DOF per node: U, Y have numNodes entries, Ke is the 8x8 col-major element
 * matrix (no per-element factor); colours as for sgld_kbyu, any count.
 * Elements in no colour contribute nothing.  Two elements of one colour that
 * share a node race on it, as they do in the reference. */
int sgld_kbyu8x8(const double *U, int64_t numNodes, const int32_t *eNodMat, int64_t numElements,
                 const double *Ke, int64_t keRows, int64_t keCols, const int32_t *colors,
                 const int64_t *colorOffsets, int numColors, double *Y);

/* Replaces Solving_Restriction_MatrixFree_mex
 * (src/MEXfuncs/Solving_Restriction_MatrixFree_mex.cpp:6-191), called at
 * src/Solving_RestrictResidual.m:6-8.  rF: nF x 3, RIT: 8 x m (m = 27 or 125),
 * transferMat: m x numElesC, Y: nC x 3. */
int sgld_restrict(const double *rF, int64_t nF, const int32_t *transferMat, int64_t m,
                  const double *RIT, int64_t ritRows, int64_t ritCols,
                  const int32_t *eNodMatC, int64_t numElesC, int64_t nC, int64_t nGhost,
                  const int32_t *solidNodeMapCoarser2Finer, const double *transferMatCoeffi,
                  const int32_t *colors, const int64_t *colorOffsets, int numColors, double *Y);

/* sgld_restrict with every array in device memory except colorOffsets (host),
 * and no allocation: no counterpart in the reference, for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The ghost-node array the reference
 * mxMallocs per call lives in a fixed 32 MiB device array that each GPU holds
 * until the process ends, so nGhost <= 1,398,090; larger calls return
 * SGLD_ERR_ARG.  Ids are checked on the device before use.  Synchronous on the
 * null stream; calls that overlap on one GPU share the scratch and race. */
int sgld_restrict_noalloc(const double *rF, int64_t nF, const int32_t *transferMat, int64_t m,
                          const double *RIT, int64_t ritRows, int64_t ritCols,
                          const int32_t *eNodMatC, int64_t numElesC, int64_t nC, int64_t nGhost,
                          const int32_t *solidNodeMapCoarser2Finer, const double *transferMatCoeffi,
                          const int32_t *colors, const int64_t *colorOffsets, int numColors, double *Y);

/* Replaces Solving_Interpolation_MatrixFree_mex
 * (src/MEXfuncs/Solving_Interpolation_MatrixFree_mex.cpp:6-190), called at
 * src/Solving_InterpolationDeviation.m:7-9.  xC: nC x 3, RI: m x 8, Y: nF x 3. */
int sgld_interpolate(const double *xC, int64_t nC, const int32_t *transferMat, int64_t m,
                     const double *RI, int64_t riRows, int64_t riCols,
                     const int32_t *eNodMatC, int64_t numElesC, int64_t nF, int64_t nGhost,
                     const int32_t *solidNodeMapCoarser2Finer, const double *transferMatCoeffi,
                     const int32_t *colors, const int64_t *colorOffsets, int numColors, double *Y);

/* Replaces Solving_AssembleCmptStencilFromFinestLevel
 * (src/MEXfuncs/Solving_AssembleCmptStencilFromFinestLevel.cpp:12-169), called at
 * src/Solving_AssembleFEAstencil.m:33.  The trilinear interpolatingKe / localMapping
 * pair of the MEX is implied by span (2: 8 children, 4: 64).  Ks: 24 x 24 x nC. */
int sgld_assemble_finest(const double *Ke, int64_t keRows, int64_t keCols, const double *E,
                         int64_t numE, const int32_t *elementUpwardMap, int64_t nC,
                         int64_t numChildren, double *Ks);

/* Replaces Solving_AssembleCmptStencilFromNonFinestLevel
 * (src/MEXfuncs/Solving_AssembleCmptStencilFromNonFinestLevel.cpp:12-143), called at
 * src/Solving_AssembleFEAstencil.m:56.  KsPrev: 24 x 24 x numPrev. */
int sgld_assemble_nonfinest(const double *KsPrev, int64_t numPrev,
                            const int32_t *elementUpwardMap, int64_t nC, int64_t numChildren,
                            double *Ks);

/* Replaces TopOpti_CmptUnitCompliance_mex (src/MEXfuncs/TopOpti_CmptUnitCompliance_mex.cpp:6-59),
 * called at src/TopOpti_ComputeUnitCompliance.m:7.  ce_e = Ue^T Ke Ue. */
int sgld_unit_compliance(const double *U, int64_t numDOFs, const int32_t *eNodMat,
                         int64_t numElements, const double *Ke, int64_t keRows, int64_t keCols,
                         double *ce);

/* Replaces FEA_VoxelBasedElementStiffnessMatrix (src/FEA_VoxelBasedElementStiffnessMatrix.m:1-24). */
int sgld_element_stiffness(double poissonRatio, double cellSize, double *Ke);

/* ================ stateful solver (device-resident hierarchy) ================ */

typedef struct sgld_ctx sgld_ctx;

/* Builds meshHierarchy_ on the device: FEA_VoxelBasedDiscretization.m:56-89,
 * Solving_BuildingMeshHierarchy.m:1-165, FEA_SetupVoxelBased.m:6-10.
 * voxels: resY x resX x resZ uint8 col-major (nonzero = solid) or NULL for a full
 * cuboid.  numLevels and nonDyadic have the meaning of numLevels_ / nonDyadic_
 * (Solving_BuildingMeshHierarchy.m:12-23). */
int sgld_ctx_create(sgld_ctx **ctx, int resX, int resY, int resZ, const uint8_t *voxels,
                    int numLevels, int nonDyadic, double poissonRatio, double cellSize);
int sgld_ctx_destroy(sgld_ctx *ctx);
/* out[8]: resX resY resZ spanWidth numElements numNodes numDOFs latticeNodes
 * of level (1-based); returns the level count when level == 0 (out unused). */
int sgld_ctx_info(const sgld_ctx *ctx, int level, int64_t *out);

/* Fixed-DOF mask of level 1 (FEA_ApplyBoundaryCondition.m:37-41); coarse masks
 * follow FEA_SetupVoxelBased.m:13-26. */
int sgld_ctx_set_fixed(sgld_ctx *ctx, const uint8_t *fixedDOFs);
/* meshHierarchy_(1).eleModulus (numElements, compressed order). */
int sgld_ctx_set_modulus(sgld_ctx *ctx, const double *E);
/* Solving_AssembleFEAstencil.m:1-137 plus the Jacobi weight omega
 * (weightFactorJacobi_): Galerkin coarse operators, diagonals, coarsest factor. */
int sgld_ctx_assemble(sgld_ctx *ctx, double omega);

/* productMV = Solving_KbyU_MatrixFree(u) (src/Solving_KbyU_MatrixFree.m:1-41). */
int sgld_ctx_kbyu(sgld_ctx *ctx, const double *u, double *y);
/* rTilde = Solving_Vcycle(r) (src/Solving_Vcycle.m:1-52). */
int sgld_ctx_vcycle(sgld_ctx *ctx, const double *r, double *z);
/* [U, its] = Solving_PreconditionedConjugateGradientSolver(@KbyU, @Vcycle, F, tol,
 * maxIT, printP, U) (src/Solving_PreconditionedConjugateGradientSolver.m:1-37),
 * as Solving_CG_GMGS.m:7 calls it.  U: in x0, out solution.  precision:
 * SGLD_PREC_FP64 (the reference algorithm) or SGLD_PREC_MIXED (PCG in FP32 -
 * vectors, K.u, V-cycle - with FP64 dot products and coarsest inverse, and
 * reliable updates that fold the correction into the FP64 solution and
 * replace the residual by the FP64 true residual; stops on the latter).
 * relres receives the final FP64 true relative residual ||F - K U|| / ||F||.
 * Returns SGLD_ERR_MAXIT (with U filled) when maxIT is reached. */
int sgld_ctx_pcg(sgld_ctx *ctx, const double *F, double *U, double tol, int maxIT,
                 int precision, int *its, double *relres);
/* Same with F and U in device memory (compressed layout). */
int sgld_ctx_pcg_device(sgld_ctx *ctx, const double *dF, double *dU, double tol, int maxIT,
                        int precision, int *its, double *relres);
/* ceList = TopOpti_ComputeUnitCompliance() (src/TopOpti_ComputeUnitCompliance.m:1-21). */
int sgld_ctx_unit_compliance(sgld_ctx *ctx, const double *U, double *ce);
/* K.u with u and y in device memory (compressed layout); times K.u alone. */
int sgld_ctx_kbyu_device(sgld_ctx *ctx, const double *du, double *dy, int precision);

/* Times `reps` back-to-back K.u launches on the context's internal vectors
 * (hipEvents on the context stream); msPerLaunch = average kernel time. */
int sgld_ctx_bench_kbyu(sgld_ctx *ctx, int precision, int reps, double *msPerLaunch);

/* Batched solve of nrhs <= 16 load cases on one mesh hierarchy.  No
 * counterpart in the reference, which solves load cases one call of
 * Solving_CG_GMGS.m at a time; each case follows sgld_ctx_pcg's contract.
 * F, U: numDOF x nrhs, case-major (case j at F + j numDOF); U in: initial
 * guesses, out: solutions.  its[j], relres[j]: per-case iterations and FP64
 * true relative residual.  precision SGLD_PREC_FP64: nrhs sequential FP64
 * solves; SGLD_PREC_MIXED: nrhs sequential mixed-precision solves (the
 * fastest mode on one MI355X: 1.90 s for 16 cases of a 256^3 cuboid, against
 * 2.29 s interleaved and 3.13 s FP64); SGLD_PREC_INTERLEAVED: all cases
 * iterate together (16 cases interleaved per DOF, K.u on the BF16 matrix
 * cores as hi/lo split products with FP32 accumulation, FP32 V-cycle, FP64
 * per-case scalars and reliable updates).
 * Returns SGLD_ERR_MAXIT (solutions filled) if any case misses tol. */
int sgld_ctx_pcg_batched(sgld_ctx *ctx, int nrhs, const double *F, double *U, double tol,
                         int maxIT, int precision, int *its, double *relres);
/* Same with F and U in device memory. */
int sgld_ctx_pcg_batched_device(sgld_ctx *ctx, int nrhs, const double *dF, double *dU,
                                double tol, int maxIT, int precision, int *its,
                                double *relres);
/* Y = K U for nrhs <= 16 device vectors (compressed, case-major, FP64 in and
 * out) through the batched matrix-core K.u (FP32 vectors; per element the
 * displacements relative to its first node, scaled by E, enter the BF16
 * matrix cores as hi/lo pairs with FP32 accumulation; fixed DOFs 0).  No
 * counterpart in the reference. */
int sgld_ctx_kbyu_batched_device(sgld_ctx *ctx, int nrhs, const double *dU, double *dY);
/* Times `reps` launches of the batched (16-case) K.u kernel; msPerLaunch =
 * average kernel time (hipEvents on the context stream). */
int sgld_ctx_bench_kbyu_batched(sgld_ctx *ctx, int reps, double *msPerLaunch);

/* Compliance sensitivities after a solve (TopOpti_GlobalVolumeConstraint.m:
 * 113-118 with TopOpti_DeMaterialInterpolation.m:5), fused with the unit
 * compliance of TopOpti_ComputeUnitCompliance.m: per element
 *   ce_e = u_e^T Ke u_e,  dc_e = -penal (E0 - Emin) x_e^(penal-1) ce_e scale,
 * and compliance = scale sum_e E_e ce_e (E_e as set by sgld_ctx_set_modulus).
 * The reference normalises with scale = 1 / complianceSolid_.  U: numDOF,
 * x, dc: numElements (compressed order). */
int sgld_ctx_sensitivity(sgld_ctx *ctx, const double *U, const double *x, double penal, double E0,
                         double Emin, double scale, double *dc, double *compliance);
/* Same with U, x and dc in device memory. */
int sgld_ctx_sensitivity_device(sgld_ctx *ctx, const double *dU, const double *dx, double penal,
                                double E0, double Emin, double scale, double *ddc,
                                double *compliance);
/* Autotuner (no counterpart in the reference): times the K.u variants of this
 * mesh (patch-tiled mode-space with z-chunks of 8, 12, 16 or 32 node layers,
 * where a new context starts at 12, and the node-centric dense kernel; `reps` launches each, FP64 and FP32) and the
 * two span-2 restrictions (on FP32 V-cycles, once assembled), and keeps the
 * fastest: the z-chunk per precision, then tiled against dense for the
 * mixed solver (cost = FP32 K.u + FP64 K.u / 8).  report[8]: [0] K.u variant
 * (1 tiled, 0 dense), [1] FP64 z-chunk, [2] FP64 K.u ms, [3] FP32 K.u ms,
 * [4] restriction (1 tiled, 0 node-centric), [5], [6] V-cycle ms of the two
 * restrictions, [7] FP32 z-chunk. */
int sgld_ctx_autotune(sgld_ctx *ctx, int reps, double *report);
/* Pins the schedule the autotuner chooses: kbyuVariant 1 = patch-tiled
 * mode-space K.u with z-chunks of zChunk node layers (0: each precision keeps
 * its current z-chunk) and a persistent grid of
 * blocksPerCU blocks per compute unit (0 = as many as fit), both precisions;
 * 0 = node-centric
 * dense K.u; restrictTiled -1 = by level size, 0 = node-centric, 1 =
 * LDS-tiled span-2 restriction, 2 = LDS-tiled into level 1 and node-centric
 * below (one restriction kernel per level when many instances share the
 * launches of sgld_multi_pcg_device).  No counterpart in the reference. */
int sgld_ctx_set_schedule(sgld_ctx *ctx, int kbyuVariant, int zChunk, int restrictTiled,
                          int blocksPerCU);
/* n independent problems (one context each, distinct) solved together (no
 * counterpart in the reference, whose topology-optimisation loop solves one
 * design at a time).  Every kernel of the PCG iteration is launched once over
 * all instances (each instance runs its own block range with its own
 * arguments, so its iterates equal those of its own solve bit for bit); for
 * SGLD_PREC_MIXED the reliable updates run on the device as well, in gated
 * launches ahead of every 4th iteration (SGLD_MULTI_UPDATE_EVERY).  `batch`
 * (even) iterations are replayed as one graph between host checks of the
 * control blocks.  The V-cycle runs level by level during the call (the
 * contexts' fused small-level tail is off, restored on return).  A context on
 * the LDS-atomic K.u (SGLD_KU_OWNER=0), or SGLD_MULTI_GRAPHS=1, falls back to
 * one graph per instance replayed on the instance's own stream.  dF[i],
 * dU[i]: device vectors of instance i (U in: initial guess).  its, relres:
 * per instance (FP64 true relative residual); latency[i] (optional): seconds
 * from the call until instance i was seen converged.  Returns SGLD_ERR_MAXIT
 * if any instance missed tol. */
int sgld_multi_pcg_device(sgld_ctx *const *ctxs, int n, const double *const *dF, double *const *dU,
                          double tol, int maxIT, int precision, int batch, int *its,
                          double *relres, double *latency);
/* sgld_ctx_sensitivity_device for n instances (distinct contexts) in one
 * launch over all their elements and one host synchronisation (no
 * counterpart in the reference).  penal, E0, Emin and scale are shared;
 * dU[i], dx[i], ddc[i]: device vectors of instance i; compliance[i]: host
 * output.  Each instance's dc and compliance equal its own
 * sgld_ctx_sensitivity_device call bit for bit. */
int sgld_multi_sensitivity_device(sgld_ctx *const *ctxs, int n, const double *const *dU,
                                  const double *const *dx, double penal, double E0, double Emin,
                                  double scale, double *const *ddc, double *compliance);

/* Hs = TopOpti_SetupDensityFilter_mex(rMin, numElements, eleMapForward, resX,
 * resY, resZ) (src/MEXfuncs/TopOpti_SetupDensityFilter_mex.cpp:10-51):
 * Hs[e1] = sum over solid e2 within the box of half-width ceil(rMin)-1 of
 * max(0, rMin - |c_e1 - c_e2|).  eleMapForward: resX resY resZ voxels (row
 * fastest), 1-based element id or 0; Hs: numElements.  Host pointers. */
int sgld_density_filter_setup(double rMin, int64_t numElements, const int32_t *eleMapForward,
                              int resX, int resY, int resZ, double *Hs);
/* Y = TopOpti_PerformDensityFiltering_mex(x, rMin, numElements, eleMapForward,
 * resX, resY, resZ) (src/MEXfuncs/TopOpti_PerformDensityFiltering_mex.cpp:
 * 10-52): Y[e1] = sum_e2 max(0, rMin - |c_e1 - c_e2|) x[e2] (the caller
 * divides by Hs).  Bit-identical to the reference loop order in FP64. */
int sgld_density_filter(const double *x, double rMin, int64_t numElements,
                        const int32_t *eleMapForward, int resX, int resY, int resZ, double *Y);
/* The same filter on a context's finest mesh with device vectors
 * (numElements, compressed order); dx == NULL computes Hs.  The filter
 * kernel's time goes to sgld_ctx_stats out[5]. */
int sgld_ctx_density_filter_device(sgld_ctx *ctx, const double *dx, double rMin, double *dY);

/* [cartesianStressField, vonMisesStressField] = FEA_StressAnalysis()
 * (src/FEA_StressAnalysis.m:1-66): nodal stresses extrapolated from the 8
 * Gauss points of each element (E_e D B u_e), averaged over the solid
 * elements at each node; stress: numNodes x 6 column-major (xx yy zz yz zx
 * xy), vonMises: numNodes.  Host pointers; _device: device pointers. */
int sgld_ctx_stress(sgld_ctx *ctx, const double *U, double *stress, double *vonMises);
int sgld_ctx_stress_device(sgld_ctx *ctx, const double *dU, double *dstress, double *dvonMises);

/* tar = TopOpti_PDEFiltering_matrixFree(src) after
 * TopOpti_BuildPDEfilter_matrixFree() (src/TopOpti_PDEFiltering_matrixFree.m:
 * 1-56, src/TopOpti_BuildPDEfilter_matrixFree.m:7-37): element values to
 * nodes, Jacobi-preconditioned CG on (iRmin^2 K + M) with the 8 x 8 element
 * matrix of rHatMin (Solving_KbyU_MatrixFree8x8_mex.cpp product), tolerance
 * tol and maxIT as tol_ / maxIT_ (defaults 1e-3 / 500), nodes back to
 * elements.  src, tar: numElements; its: CG iterations.  Host pointers;
 * _device: device pointers. */
int sgld_ctx_pde_filter(sgld_ctx *ctx, const double *src, double rHatMin, double tol, int maxIT,
                        double *tar, int *its);
int sgld_ctx_pde_filter_device(sgld_ctx *ctx, const double *dsrc, double rHatMin, double tol,
                               int maxIT, double *dtar, int *its);

/* Statistics of the last solve: out[0] solve seconds, [1] iterations (inner
 * total), [2] true relres, [3] outer refinements, [4] setup seconds of the
 * last assemble, [5] milliseconds of the last density-filter kernel
 * (sgld_ctx_density_filter_device), [6] average K.u kernel milliseconds
 * (hipEvent, FP64), [7] average K.u kernel milliseconds (FP32).  For
 * SGLD_PREC_MIXED [3] counts reliable updates. */
int sgld_ctx_stats(const sgld_ctx *ctx, double *out);

/* Per-kernel timing of the finest-level passes of the following
 * sgld_ctx_pcg[_device] solves (no counterpart in the reference, whose
 * zTmp_Benchmarking_SpMV.m:73-100 times the SpMV and preconditioner calls
 * from MATLAB): with on != 0 every launch of the kernels below is bracketed
 * by HIP events on the context stream.  Resets the counters.  Kernel ids: */
enum { SGLD_KT_DIR64 = 0, /* direction update p = z + beta p (finest V-cycle level fused), FP64 */
       SGLD_KT_DIR32,     /* the same, FP32 inner solve of SGLD_PREC_MIXED */
       SGLD_KT_KU64,      /* K.u, FP64 (incl. the residuals of reliable updates) */
       SGLD_KT_KU32,      /* K.u, FP32 */
       SGLD_KT_UPD64,     /* residual update r -= lambda q, FP64 */
       SGLD_KT_UPD32,     /* the same, FP32 */
       SGLD_KT_RES64,     /* first restriction of the V-cycle, FP64 */
       SGLD_KT_RES32,     /* the same, FP32 */
       SGLD_KT_COUNT = 8 };
int sgld_ctx_kernel_timing(sgld_ctx *ctx, int on);
/* out[4 k + 0..3] for kernel id k < SGLD_KT_COUNT: launches, their total
 * milliseconds, launches that did work (not in an iteration queued after the
 * solve's stop), their total milliseconds. */
int sgld_ctx_kernel_times(const sgld_ctx *ctx, double *out);
/* Name of the kernel the last timed launch of id ran ("" if none). */
const char *sgld_ctx_kernel_name(const sgld_ctx *ctx, int id);

#ifdef __cplusplus
}
#endif
#endif
