/*
 * sgld_oracle.h -- CPU restatement of SGLDBench's voxel-FEA MG-PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libsgldfea, HIP) never links or calls it.
 *
 * Every function follows the reference file:line named in its comment.  All
 * arrays use the reference's conventions: int32 ids are 1-based, matrices are
 * column-major, DOF vectors are interleaved [ux uy uz] per compressed node.
 */
#ifndef SGLD_ORACLE_H
#define SGLD_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- free functions mirroring the six MEX kernels and the Ke builder ---- */

/* FEA_VoxelBasedElementStiffnessMatrix.m:1-24 (E = 1 per FEA_HookeLaw.m:4). */
void or_element_stiffness(double nu, double cellSize, double *Ke /*576*/);

/* Solving_KbyU_MatrixFree_mex.cpp:6-138.  colors: concatenated 1-based element
 * ids, colorOff[9] prefix offsets. */
void or_kbyu_mex(const double *U, int64_t numDOFs, const int32_t *eNodMat,
                 int64_t numElements, const double *Ke, const double *E,
                 const int32_t *colors, const int64_t *colorOff, double *Y);

/* Solving_KbyU_MatrixFree8x8_mex.cpp:6-121 (called at
 * TopOpti_PDEFiltering_matrixFree.m:46).  U, Y: numNodes; Ke: 8x8 col-major;
 * colors: concatenated 1-based element ids of numColors colours. */
void or_kbyu8x8_mex(const double *U, int64_t numNodes, const int32_t *eNodMat,
                    int64_t numElements, const double *Ke, const int32_t *colors,
                    const int64_t *colorOff, int numColors, double *Y);

/* Solving_Restriction_MatrixFree_mex.cpp:6-191. rF: nF x 3 col-major, Y: nC x 3. */
void or_restrict_mex(const double *rF, int64_t nF, const int32_t *transferMat,
                     int m, const double *RIT, const int32_t *eNodMatC,
                     int64_t numElesC, int64_t nC, int64_t nGhost,
                     const int32_t *map, const double *coeffi,
                     const int32_t *colors, const int64_t *colorOff, double *Y);

/* Solving_Interpolation_MatrixFree_mex.cpp:55-190. xC: nC x 3, Y: nF x 3. */
void or_interpolate_mex(const double *xC, int64_t nC, const int32_t *transferMat,
                        int m, const double *RI, const int32_t *eNodMatC,
                        int64_t numElesC, int64_t nF, int64_t nGhost,
                        const int32_t *map, const double *coeffi,
                        const int32_t *colors, const int64_t *colorOff, double *Y);

/* Solving_AssembleCmptStencilFromFinestLevel.cpp:12-169 (span 2: 8 children,
 * span 4: 64).  Ks: 24x24xnC col-major. */
void or_assemble_finest(const double *Ke, const double *E, const int32_t *upMap,
                        int64_t nC, int span, double *Ks);

/* Solving_AssembleCmptStencilFromNonFinestLevel.cpp:12-143. */
void or_assemble_nonfinest(const double *KsPrev, const int32_t *upMap,
                           int64_t nC, int span, double *Ks);

/* TopOpti_CmptUnitCompliance_mex.cpp:6-107. */
void or_unit_compliance_mex(const double *U, const int32_t *eNodMat,
                            int64_t numElements, const double *Ke, double *ce);

/* Solving_Operator4MultiGridRestrictionAndInterpolation.m 'inNODE':
 * (span+1)^3 x 8 col-major trilinear table. */
void or_transfer_table(int span, double *RI);

/* Solving_SubEleNodMat.m: span^3 x 8 col-major, 1-based block-node ids. */
void or_sub_ele_nod_mat(int span, int32_t *out);

/* ---- stateful mesh hierarchy (meshHierarchy_, Data_CartesianMeshStruct.m) ---- */
typedef struct or_mesh or_mesh;

/* FEA_VoxelBasedDiscretization.m:56-89 + Solving_BuildingMeshHierarchy.m:1-165.
 * voxels: resY x resX x resZ col-major (row fastest, row 0 = top), nonzero =
 * solid.  numLevels as configured; nonDyadic per Data_GlobalVariables.m:15. */
or_mesh *or_mesh_create(int resX, int resY, int resZ, const uint8_t *voxels,
                        int numLevels, int nonDyadic, double nu, double cellSize);
void or_mesh_destroy(or_mesh *m);
int or_mesh_num_levels(const or_mesh *m);
/* out: resX resY resZ span numElements numNodes numDOFs intermediateNumNodes */
void or_mesh_level_info(const or_mesh *m, int level, int64_t *out);
/* Copy a named per-level array into out; returns element count or -1.
 * Names: eleMapBack eNodMat nodMapBack elementUpwardMap transferMat
 *        solidNodeMapCoarser2Finer colors colorOff (int64) fixedDOFs (uint8)
 *        transferMatCoeffi RI diagK Ks Ke (double). */
int64_t or_mesh_get(const or_mesh *m, int level, const char *name, void *out);

/* FEA_SetupVoxelBased.m:13-26: level-1 fixed mask -> coarse masks. */
int or_mesh_set_fixed(or_mesh *m, const uint8_t *fixedDOFs);
/* eleModulus of level 1 (numElements). */
int or_mesh_set_modulus(or_mesh *m, const double *E);
/* lean != 0: the following or_mesh_assemble keeps no level-2 Galerkin
 * matrices (diagonals only; level 3 rebuilds them per element), with the
 * same results; for meshes whose level-2 matrices do not fit host memory. */
void or_mesh_set_lean(or_mesh *m, int lean);
/* Solving_AssembleFEAstencil.m:1-137: Galerkin Ks, diagK, coarsest Cholesky.
 * Returns 0, or 1 if the coarsest matrix is not SPD. */
int or_mesh_assemble(or_mesh *m);

/* Solving_KbyU_MatrixFree.m:1-41 (incl. fixed-DOF zeroing). */
void or_mesh_kbyu(const or_mesh *m, const double *u, double *y);
/* Solving_RestrictResidual.m / Solving_InterpolationDeviation.m between
 * level ii-1 and ii (1-based ii >= 2), interleaved vectors. */
void or_mesh_restrict(const or_mesh *m, int ii, const double *rF, double *rC);
void or_mesh_interpolate(const or_mesh *m, int ii, const double *xC, double *xF);
/* Solving_Vcycle.m:1-52. */
void or_mesh_vcycle(const or_mesh *m, double omega, const double *r, double *z);
/* Solving_PreconditionedConjugateGradientSolver.m:1-37.  x: in x0, out y.
 * resHist (maxIT entries, may be NULL) receives res per iteration. */
int or_mesh_pcg(const or_mesh *m, double omega, const double *b, double *x,
                double tol, int maxIT, double *resHist, int *itsOut);
/* TopOpti_ComputeUnitCompliance.m:1-21. */
void or_mesh_unit_compliance(const or_mesh *m, const double *U, double *ce);

/* Density filter of the topology optimisation (cone weights
 * w = max(0, rMin - |c_e1 - c_e2|) over the solid voxels within the
 * ceil(rMin)-1 box):
 *   or_density_filter_setup: Hs[e1] = sum_e2 w (TopOpti_SetupDensityFilter_mex.cpp:10-51)
 *   or_density_filter:       Y[e1]  = sum_e2 w x[e2] (TopOpti_PerformDensityFiltering_mex.cpp:10-52)
 * eleMapForward: resX resY resZ voxels (row fastest), 1-based element id or 0. */
void or_density_filter_setup(double rMin, int64_t numElements, const int32_t *eleMapForward,
                             int resX, int resY, int resZ, double *Hs);
void or_density_filter(const double *x, double rMin, int64_t numElements,
                       const int32_t *eleMapForward, int resX, int resY, int resZ, double *Y);

/* FEA_StressAnalysis.m:1-93: nodal Cartesian stresses (numNodes x 6,
 * column-major, xx yy zz yz zx xy) and von Mises stress (numNodes). */
void or_mesh_stress(const or_mesh *m, const double *U, double *stress, double *vonMises);

/* PDE (Helmholtz) density filter, TopOpti_BuildPDEfilter_matrixFree.m and
 * TopOpti_PDEFiltering_matrixFree.m: element matrix (8 x 8, column-major) and
 * the filter itself (element values in and out); returns the PCG iterations. */
void or_pde_element(double rHatMin, double *Ke8);
int or_mesh_pde_filter(const or_mesh *m, double rHatMin, const double *src, double tol, int maxIT,
                       double *tar);

/* Timing helper for the CPU baseline: runs `its` PCG iterations exactly
 * (no early exit) and returns the seconds they took; setup and the initial
 * z = M r excluded. */
double or_mesh_time_pcg_iterations(const or_mesh *m, double omega, const double *b,
                                   int its, double *kbyuSeconds, double *vcycleSeconds);

/* OpenMP team size of the following calls (timing at one thread and at all
 * cores); returns the previous maximum. */
int or_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
