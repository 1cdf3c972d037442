/*
 * sgld_oracle.c -- CPU restatement (FP64, OpenMP) of SGLDBench's voxel FEA
 * solve: element stiffness, mesh hierarchy, colour-scheduled matrix-free K.u,
 * ghost-node restriction/interpolation, Galerkin coarse stencils, additive
 * V-cycle, PCG and unit compliance.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker for libsgldfea (the HIP
 * product path) and the "port" CPU baseline in bench.py.  It is written from
 * the behaviour of the reference files cited per function; nothing of the
 * reference is compiled or linked here.  The reference MEX sources need
 * MathWorks' mex.h, which this image lacks, so they cannot be built.
 *
 * Pinning (tests/test_oracle_*.py): Ke against an independent exact
 * integration and the closed-form H8 matrix; K.u against a scipy-assembled
 * global K; transfers against dense trilinear P and R = P^T; Galerkin against
 * P^T K P; coarsest solve by residual; PCG against scipy spsolve; compliance
 * by the energy identity u^T K u = sum_e E_e ce_e.
 */
#include "sgld_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

#define OR_MAX_LEVELS 16

/* ------------------------------------------------------------------------- */
/* Element stiffness: FEA_VoxelBasedElementStiffnessMatrix.m:10-23,          */
/* FEA_GaussianIntegral.m:19-25 (points +-0.577350269189626, weights 1),     */
/* FEA_DeShapeFunction.m:3-22, FEA_ElementStrainMatrix.m:11-92 (strain order */
/* xx yy zz yz zx xy), FEA_HookeLaw.m:4-13 (E = 1), detJ = h^3/8, invJ = 2/h. */
/* ------------------------------------------------------------------------- */
static const int OR_CS[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
static const int OR_CT[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const int OR_CP[8] = {-1, -1, -1, -1, 1, 1, 1, 1};

void or_element_stiffness(double nu, double h, double *Ke) {
    const double g = 0.577350269189626;
    double S[6][6];
    memset(S, 0, sizeof(S));
    const double c1 = (1 + nu) * (1 - 2 * nu), c2 = 2 * (1 + nu);
    S[0][0] = S[1][1] = S[2][2] = (1 - nu) / c1;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) S[i][j] = nu / c1;
    S[3][3] = S[4][4] = S[5][5] = 1.0 / c2;
    const double detJ = h * h * h / 8.0, invJ = 2.0 / h;
    memset(Ke, 0, sizeof(double) * 576);
    for (int q = 0; q < 8; ++q) {
        double s = OR_CS[q] * g, t = OR_CT[q] * g, p = OR_CP[q] * g;
        double B[6][24];
        memset(B, 0, sizeof(B));
        for (int j = 0; j < 8; ++j) {
            double dNs = 0.125 * OR_CS[j] * (1 + OR_CT[j] * t) * (1 + OR_CP[j] * p);
            double dNt = 0.125 * OR_CT[j] * (1 + OR_CS[j] * s) * (1 + OR_CP[j] * p);
            double dNp = 0.125 * OR_CP[j] * (1 + OR_CS[j] * s) * (1 + OR_CT[j] * t);
            double dx = invJ * dNs, dy = invJ * dNt, dz = invJ * dNp;
            B[0][3 * j] = dx;
            B[1][3 * j + 1] = dy;
            B[2][3 * j + 2] = dz;
            B[3][3 * j + 1] = dz; B[3][3 * j + 2] = dy;
            B[4][3 * j] = dz;     B[4][3 * j + 2] = dx;
            B[5][3 * j] = dy;     B[5][3 * j + 1] = dx;
        }
        double DB[6][24];
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 24; ++k) {
                double acc = 0;
                for (int l = 0; l < 6; ++l) acc += S[i][l] * B[l][k];
                DB[i][k] = acc * detJ; /* weight w = 1 */
            }
        for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b) {
                double acc = 0;
                for (int i = 0; i < 6; ++i) acc += B[i][a] * DB[i][b];
                Ke[a + 24 * b] += acc;
            }
    }
}

/* Trilinear transfer table, Solving_Operator4MultiGridRestrictionAndInterpolation.m
 * 'inNODE' (span 2: :4-31, span 4: :118-243).  Row n = rr + (s+1)cc + (s+1)^2 ll
 * (block-local row rr counted from the top, y = (s-rr)/s), column J = coarse
 * local node J (corner order of FEA_GaussianIntegral.m:2-9). */
void or_transfer_table(int s, double *RI) {
    const int m = (s + 1) * (s + 1) * (s + 1);
    for (int ll = 0; ll <= s; ++ll)
        for (int cc = 0; cc <= s; ++cc)
            for (int rr = 0; rr <= s; ++rr) {
                int n = rr + (s + 1) * cc + (s + 1) * (s + 1) * ll;
                double tx = (double)cc / s, ty = (double)(s - rr) / s, tz = (double)ll / s;
                for (int J = 0; J < 8; ++J) {
                    int xJ = (OR_CS[J] + 1) / 2, yJ = (OR_CT[J] + 1) / 2, zJ = (OR_CP[J] + 1) / 2;
                    double wx = xJ ? tx : 1 - tx, wy = yJ ? ty : 1 - ty, wz = zJ ? tz : 1 - tz;
                    RI[n + (int64_t)m * J] = wx * wy * wz;
                }
            }
}

/* Solving_SubEleNodMat.m:1-80: child q = qr + s qc + s^2 ql, local node j ->
 * block node (qr + 1 - y_j, qc + x_j, ql + z_j), 1-based. */
void or_sub_ele_nod_mat(int s, int32_t *out) {
    const int nq = s * s * s;
    for (int ql = 0; ql < s; ++ql)
        for (int qc = 0; qc < s; ++qc)
            for (int qr = 0; qr < s; ++qr) {
                int q = qr + s * qc + s * s * ql;
                for (int j = 0; j < 8; ++j) {
                    int xj = (OR_CS[j] + 1) / 2, yj = (OR_CT[j] + 1) / 2, zj = (OR_CP[j] + 1) / 2;
                    int rr = qr + 1 - yj, cc = qc + xj, ll = ql + zj;
                    out[q + nq * j] = 1 + rr + (s + 1) * cc + (s + 1) * (s + 1) * ll;
                }
            }
}

/* ------------------------------------------------------------------------- */
/* MEX kernels                                                               */
/* ------------------------------------------------------------------------- */

/* Solving_KbyU_MatrixFree_mex.cpp:53-137: zero Y, then per colour (in order)
 * per element gather Ue, Ye = E * (Ke * Ue) column-sweep, scatter-add. */
void or_kbyu_mex(const double *U, int64_t numDOFs, const int32_t *eNodMat,
                 int64_t numElements, const double *Ke, const double *E,
                 const int32_t *colors, const int64_t *colorOff, double *Y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < numDOFs; ++i) Y[i] = 0.0;
    for (int c = 0; c < 8; ++c) {
        const int64_t n = colorOff[c + 1] - colorOff[c];
        const int32_t *ids = colors + colorOff[c];
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            int64_t e = ids[k] - 1;
            int64_t nod[8];
            double Ue[24], Ye[24];
            for (int j = 0; j < 8; ++j) nod[j] = eNodMat[e + j * numElements] - 1;
            for (int j = 0; j < 8; ++j)
                for (int a = 0; a < 3; ++a) Ue[3 * j + a] = U[3 * nod[j] + a];
            for (int j = 0; j < 24; ++j) Ye[j] = 0.0;
            for (int kk = 0; kk < 24; ++kk) {
                const double uk = Ue[kk];
                const double *col = Ke + kk * 24;
                for (int j = 0; j < 24; ++j) Ye[j] += col[j] * uk;
            }
            const double Ei = E[e];
            for (int j = 0; j < 24; ++j) Ye[j] *= Ei;
            for (int j = 0; j < 8; ++j)
                for (int a = 0; a < 3; ++a) Y[3 * nod[j] + a] += Ye[3 * j + a];
        }
    }
}

/* Solving_KbyU_MatrixFree8x8_mex.cpp:50-89 with apply_Ke_Ue_8 (:92-121, the
 * column-sweep branch): one scalar DOF per node, one uniform 8x8 Ke, no
 * per-element factor; elements in no colour contribute nothing. */
void or_kbyu8x8_mex(const double *U, int64_t numNodes, const int32_t *eNodMat,
                    int64_t numElements, const double *Ke, const int32_t *colors,
                    const int64_t *colorOff, int numColors, double *Y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < numNodes; ++i) Y[i] = 0.0;
    for (int c = 0; c < numColors; ++c) {
        const int64_t n = colorOff[c + 1] - colorOff[c];
        const int32_t *ids = colors + colorOff[c];
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            const int64_t e = ids[k] - 1;
            int64_t nod[8];
            double Ue[8], Ye[8];
            for (int j = 0; j < 8; ++j) nod[j] = eNodMat[e + j * numElements] - 1;
            for (int j = 0; j < 8; ++j) Ue[j] = U[nod[j]];
            for (int j = 0; j < 8; ++j) Ye[j] = 0.0;
            for (int kk = 0; kk < 8; ++kk) {
                const double uk = Ue[kk];
                for (int j = 0; j < 8; ++j) Ye[j] += Ke[kk * 8 + j] * uk;
            }
            for (int j = 0; j < 8; ++j) Y[nod[j]] += Ye[j];
        }
    }
}

/* Solving_Restriction_MatrixFree_mex.cpp:63-132 (efficiency-prior branch). */
void or_restrict_mex(const double *rF, int64_t nF, const int32_t *transferMat,
                     int m, const double *RIT, const int32_t *eNodMatC,
                     int64_t numElesC, int64_t nC, int64_t nGhost,
                     const int32_t *map, const double *coeffi,
                     const int32_t *colors, const int64_t *colorOff, double *Y) {
#pragma omp parallel for
    for (int64_t i = 0; i < 3 * nC; ++i) Y[i] = 0.0;
    double *ghost = (double *)calloc((size_t)(3 * nGhost), sizeof(double));
    for (int a = 0; a < 3; ++a) {
#pragma omp parallel for
        for (int64_t j = 0; j < nF; ++j) {
            int64_t g = map[j] - 1;
            ghost[g + a * nGhost] = rF[a * nF + j] / coeffi[g];
        }
    }
    for (int c = 0; c < 8; ++c) {
        const int64_t n = colorOff[c + 1] - colorOff[c];
        const int32_t *ids = colors + colorOff[c];
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            int64_t e = ids[k] - 1;
            double Ye[3][8];
            for (int a = 0; a < 3; ++a)
                for (int j = 0; j < 8; ++j) Ye[a][j] = 0.0;
            for (int kk = 0; kk < m; ++kk) {
                int64_t g = transferMat[(int64_t)m * e + kk] - 1;
                const double *col = RIT + kk * 8;
                for (int a = 0; a < 3; ++a) {
                    double uk = ghost[g + a * nGhost];
                    for (int j = 0; j < 8; ++j) Ye[a][j] += col[j] * uk;
                }
            }
            for (int j = 0; j < 8; ++j) {
                int64_t nd = eNodMatC[e + j * numElesC] - 1;
                for (int a = 0; a < 3; ++a) Y[nd + a * nC] += Ye[a][j];
            }
        }
    }
    free(ghost);
}

/* Solving_Interpolation_MatrixFree_mex.cpp:66-130 (efficiency-prior branch). */
void or_interpolate_mex(const double *xC, int64_t nC, const int32_t *transferMat,
                        int m, const double *RI, const int32_t *eNodMatC,
                        int64_t numElesC, int64_t nF, int64_t nGhost,
                        const int32_t *map, const double *coeffi,
                        const int32_t *colors, const int64_t *colorOff, double *Y) {
    double *ghost = (double *)calloc((size_t)(3 * nGhost), sizeof(double));
    for (int c = 0; c < 8; ++c) {
        const int64_t n = colorOff[c + 1] - colorOff[c];
        const int32_t *ids = colors + colorOff[c];
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < n; ++k) {
            double loc[3][125];
            int64_t e = ids[k] - 1;
            double Ue[3][8];
            for (int j = 0; j < 8; ++j) {
                int64_t nd = eNodMatC[e + j * numElesC] - 1;
                for (int a = 0; a < 3; ++a) Ue[a][j] = xC[nd + a * nC];
            }
            for (int a = 0; a < 3; ++a) {
                for (int j = 0; j < m; ++j) loc[a][j] = 0.0;
                for (int kk = 0; kk < 8; ++kk) {
                    double uk = Ue[a][kk];
                    const double *col = RI + (int64_t)kk * m;
                    for (int j = 0; j < m; ++j) loc[a][j] += col[j] * uk;
                }
            }
            for (int j = 0; j < m; ++j) {
                int64_t g = transferMat[(int64_t)m * e + j] - 1;
                for (int a = 0; a < 3; ++a) ghost[g + a * nGhost] += loc[a][j];
            }
        }
    }
    for (int a = 0; a < 3; ++a) {
#pragma omp parallel for
        for (int64_t j = 0; j < nF; ++j) {
            int64_t g = map[j] - 1;
            Y[j + a * nF] = ghost[g + a * nGhost] / coeffi[g];
        }
    }
    free(ghost);
}

/* Child-to-parent node prolongation p_q (8x8, row-major p[j][J]) for child q
 * of a span-s block: row j of the 'inNODE' table at block node SubEleNodMat(q,j).
 * P_q (24x24) = p_q (x) I3 in interleaved DOF order, which equals
 * interpolatingKe restricted to child q's DOFs (Solving_AssembleFEAstencil.m:11-20). */
static void child_prolongation(int s, int q, double p[8][8]) {
    const int m = (s + 1) * (s + 1) * (s + 1), nq = s * s * s;
    double *RI = (double *)malloc(sizeof(double) * m * 8);
    int32_t *sub = (int32_t *)malloc(sizeof(int32_t) * nq * 8);
    or_transfer_table(s, RI);
    or_sub_ele_nod_mat(s, sub);
    for (int j = 0; j < 8; ++j) {
        int n = sub[q + nq * j] - 1;
        for (int J = 0; J < 8; ++J) p[j][J] = RI[n + (int64_t)m * J];
    }
    free(RI);
    free(sub);
}

/* out += P_q^T K P_q, K 24x24 col-major, P_q = p (x) I3. */
static void galerkin_child_add(const double p[8][8], const double *K, double scale,
                               double *out) {
    double T[24 * 24]; /* T = K P_q, col-major */
    for (int J = 0; J < 8; ++J)
        for (int a = 0; a < 3; ++a) {
            double *Tc = T + 24 * (3 * J + a);
            for (int r = 0; r < 24; ++r) Tc[r] = 0.0;
            for (int j = 0; j < 8; ++j) {
                double w = p[j][J];
                if (w == 0.0) continue;
                const double *Kc = K + 24 * (3 * j + a);
                for (int r = 0; r < 24; ++r) Tc[r] += w * Kc[r];
            }
        }
    for (int c = 0; c < 24; ++c)
        for (int I = 0; I < 8; ++I)
            for (int a = 0; a < 3; ++a) {
                double acc = 0.0;
                for (int i = 0; i < 8; ++i) {
                    double w = p[i][I];
                    if (w != 0.0) acc += w * T[3 * i + a + 24 * c];
                }
                out[3 * I + a + 24 * c] += scale * acc;
            }
}

/* Solving_AssembleCmptStencilFromFinestLevel.cpp:71-169: templates
 * H_q = P_q^T Ke P_q, then Ks_e = sum_q E_{child q} H_q. */
static double *finest_templates(const double *Ke, int s) {
    const int nq = s * s * s;
    double *H = (double *)calloc((size_t)nq * 576, sizeof(double));
    for (int q = 0; q < nq; ++q) {
        double p[8][8];
        child_prolongation(s, q, p);
        galerkin_child_add(p, Ke, 1.0, H + 576 * q);
    }
    return H;
}

/* Level-2 matrix of coarse element i (0-based) from the templates, in the
 * order of or_assemble_finest. */
static void finest_element(const double *H, const double *E, const int32_t *upMap, int64_t nC,
                           int nq, int64_t i, double *acc) {
    for (int t = 0; t < 576; ++t) acc[t] = 0.0;
    for (int q = 0; q < nq; ++q) {
        int32_t child = upMap[i + q * nC];
        if (child == 0) continue;
        double Ev = E[child - 1];
        const double *Hq = H + 576 * q;
        for (int t = 0; t < 576; ++t) acc[t] += Ev * Hq[t];
    }
}

void or_assemble_finest(const double *Ke, const double *E, const int32_t *upMap,
                        int64_t nC, int s, double *Ks) {
    const int nq = s * s * s;
    double *H = finest_templates(Ke, s);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nC; ++i) {
        double acc[576];
        for (int t = 0; t < 576; ++t) acc[t] = 0.0;
        for (int q = 0; q < nq; ++q) {
            int32_t child = upMap[i + q * nC];
            if (child == 0) continue;
            double Ev = E[child - 1];
            const double *Hq = H + 576 * q;
            for (int t = 0; t < 576; ++t) acc[t] += Ev * Hq[t];
        }
        memcpy(Ks + 576 * i, acc, sizeof(acc));
    }
    free(H);
}

/* Solving_AssembleCmptStencilFromNonFinestLevel.cpp:78-143:
 * Ks_e = sum_q P_q^T Ks_prev(child q) P_q. */
void or_assemble_nonfinest(const double *KsPrev, const int32_t *upMap, int64_t nC,
                           int s, double *Ks) {
    const int nq = s * s * s;
    double (*P)[8][8] = malloc(sizeof(double) * 64 * nq);
    for (int q = 0; q < nq; ++q) child_prolongation(s, q, P[q]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nC; ++i) {
        double acc[576];
        for (int t = 0; t < 576; ++t) acc[t] = 0.0;
        for (int q = 0; q < nq; ++q) {
            int32_t child = upMap[i + q * nC];
            if (child <= 0) continue;
            galerkin_child_add(P[q], KsPrev + 576 * (int64_t)(child - 1), 1.0, acc);
        }
        memcpy(Ks + 576 * i, acc, sizeof(acc));
    }
    free(P);
}

/* TopOpti_CmptUnitCompliance_mex.cpp:40-107: ce_e = Ue^T Ke Ue. */
void or_unit_compliance_mex(const double *U, const int32_t *eNodMat,
                            int64_t numElements, const double *Ke, double *ce) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < numElements; ++i) {
        double Ue[24], y[24];
        for (int j = 0; j < 8; ++j) {
            int64_t nd = eNodMat[i + j * numElements] - 1;
            for (int a = 0; a < 3; ++a) Ue[3 * j + a] = U[3 * nd + a];
        }
        for (int r = 0; r < 24; ++r) y[r] = 0.0;
        for (int c = 0; c < 24; ++c) {
            double uj = Ue[c];
            const double *col = Ke + 24 * c;
            for (int r = 0; r < 24; ++r) y[r] += col[r] * uj;
        }
        double q = 0.0;
        for (int r = 0; r < 24; ++r) q += Ue[r] * y[r];
        ce[i] = q;
    }
}

/* ------------------------------------------------------------------------- */
/* Mesh hierarchy                                                            */
/* ------------------------------------------------------------------------- */
typedef struct {
    int resX, resY, resZ, span;
    int64_t numElements, numNodes, numDOFs, intermediateNumNodes;
    int32_t *eleMapBack, *eleMapForward, *nodMapBack, *nodMapForward, *eNodMat;
    int32_t *colors;
    int64_t colorOff[9];
    int32_t *elementUpwardMap, *transferMat, *solidNodeMapCoarser2Finer;
    double *transferMatCoeffi, *RI;
    uint8_t *fixedDOFs;
    double *diagK, *Ks;
} or_level;

struct or_mesh {
    int numLevels, nonDyadic, lean;
    double nu, cellSize;
    double Ke[576];
    double *eleModulus;
    or_level L[OR_MAX_LEVELS];
    int64_t nFreeC;
    int64_t *freeIdxC;
    double *cholL; /* nFreeC x nFreeC col-major, lower */
};

/* Solving_Coloring.m:1-13: colour = (row%2) + 2(col%2) + 4(layer%2), each
 * colour list in ascending lattice order. */
static void build_colors(or_level *L) {
    const int nx = L->resX, ny = L->resY;
    int64_t cnt[8] = {0};
    for (int64_t k = 0; k < L->numElements; ++k) {
        int64_t v = L->eleMapBack[k] - 1;
        int r = (int)(v % ny), c = (int)((v / ny) % nx), l = (int)(v / ((int64_t)ny * nx));
        cnt[(r & 1) + 2 * (c & 1) + 4 * (l & 1)]++;
    }
    L->colorOff[0] = 0;
    for (int i = 0; i < 8; ++i) L->colorOff[i + 1] = L->colorOff[i] + cnt[i];
    L->colors = (int32_t *)malloc(sizeof(int32_t) * (L->numElements + 1));
    int64_t pos[8];
    for (int i = 0; i < 8; ++i) pos[i] = L->colorOff[i];
    for (int64_t k = 0; k < L->numElements; ++k) {
        int64_t v = L->eleMapBack[k] - 1;
        int r = (int)(v % ny), c = (int)((v / ny) % nx), l = (int)(v / ((int64_t)ny * nx));
        L->colors[pos[(r & 1) + 2 * (c & 1) + 4 * (l & 1)]++] = (int32_t)(k + 1);
    }
}

/* FEA_VoxelBasedDiscretization.m:71-89 / Solving_BuildingMeshHierarchy.m:112-129:
 * eNodMat by offsets [0, ny+1, ny, -1] (+ (ny+1)(nx+1)) from node (row+1, col,
 * layer); nodes compressed in ascending lattice id. */
static void build_nodes(or_level *L) {
    const int64_t nx = L->resX, ny = L->resY, nz = L->resZ;
    const int64_t nNodLat = (nx + 1) * (ny + 1) * (nz + 1);
    const int64_t pl = (ny + 1) * (nx + 1);
    const int64_t off[8] = {0, ny + 1, ny, -1, pl, pl + ny + 1, pl + ny, pl - 1};
    L->nodMapForward = (int32_t *)calloc((size_t)nNodLat, sizeof(int32_t));
    const int64_t ne = L->numElements;
    L->eNodMat = (int32_t *)malloc(sizeof(int32_t) * 8 * (ne + 1));
    for (int64_t k = 0; k < ne; ++k) {
        int64_t v = L->eleMapBack[k] - 1;
        int64_t r = v % ny, c = (v / ny) % nx, l = v / (ny * nx);
        int64_t base = (r + 1) + (ny + 1) * c + pl * l;
        for (int j = 0; j < 8; ++j) {
            int64_t g = base + off[j];
            L->eNodMat[k + j * ne] = (int32_t)g; /* 0-based lattice id for now */
            L->nodMapForward[g] = 1;
        }
    }
    int64_t nn = 0;
    for (int64_t g = 0; g < nNodLat; ++g)
        if (L->nodMapForward[g]) L->nodMapForward[g] = (int32_t)(++nn);
    L->numNodes = nn;
    L->numDOFs = 3 * nn;
    L->nodMapBack = (int32_t *)malloc(sizeof(int32_t) * (nn + 1));
    for (int64_t g = 0; g < nNodLat; ++g)
        if (L->nodMapForward[g]) L->nodMapBack[L->nodMapForward[g] - 1] = (int32_t)(g + 1);
    for (int64_t i = 0; i < 8 * ne; ++i) L->eNodMat[i] = L->nodMapForward[L->eNodMat[i]];
}

/* Solving_BuildingMeshHierarchy.m:24-137 for one coarsening step. */
static void build_coarse(const or_level *F, or_level *C, int s) {
    const int64_t nx = F->resX / s, ny = F->resY / s, nz = F->resZ / s;
    C->resX = (int)nx; C->resY = (int)ny; C->resZ = (int)nz; C->span = s;
    const int nq = s * s * s, m = (s + 1) * (s + 1) * (s + 1);
    const int64_t nyf = F->resY, nxf = F->resX;
    const int64_t nCLat = nx * ny * nz;
    /* elementUpwardMap over all lattice elements, then keep non-empty ones */
    int32_t *upAll = (int32_t *)malloc(sizeof(int32_t) * nq * (nCLat + 1));
    C->eleMapForward = (int32_t *)calloc((size_t)nCLat, sizeof(int32_t));
    int64_t ne = 0;
    for (int64_t e = 0; e < nCLat; ++e) {
        int64_t r = e % ny, c = (e / ny) % nx, l = e / (ny * nx);
        int any = 0;
        for (int q = 0; q < nq; ++q) {
            int qr = q % s, qc = (q / s) % s, ql = q / (s * s);
            int64_t vf = (s * r + qr) + nyf * (s * c + qc) + nyf * nxf * (s * l + ql);
            int32_t id = F->eleMapForward[vf];
            upAll[e * nq + q] = id;
            any |= (id > 0);
        }
        if (any) C->eleMapForward[e] = (int32_t)(++ne);
    }
    C->numElements = ne;
    C->eleMapBack = (int32_t *)malloc(sizeof(int32_t) * (ne + 1));
    C->elementUpwardMap = (int32_t *)malloc(sizeof(int32_t) * nq * (ne + 1));
    for (int64_t e = 0; e < nCLat; ++e) {
        int32_t k = C->eleMapForward[e];
        if (!k) continue;
        C->eleMapBack[k - 1] = (int32_t)(e + 1);
        for (int q = 0; q < nq; ++q) C->elementUpwardMap[(k - 1) + q * ne] = upAll[e * nq + q];
    }
    free(upAll);
    /* ghost fine nodes: closure of non-empty blocks, ascending fine lattice id */
    const int64_t nFLat = (nxf + 1) * (nyf + 1) * (F->resZ + 1);
    int32_t *ghostId = (int32_t *)calloc((size_t)nFLat, sizeof(int32_t));
    const int64_t plf = (nyf + 1) * (nxf + 1);
    for (int64_t k = 0; k < ne; ++k) {
        int64_t e = C->eleMapBack[k] - 1;
        int64_t r = e % ny, c = (e / ny) % nx, l = e / (ny * nx);
        for (int ll = 0; ll <= s; ++ll)
            for (int cc = 0; cc <= s; ++cc)
                for (int rr = 0; rr <= s; ++rr)
                    ghostId[(s * r + rr) + (nyf + 1) * (s * c + cc) + plf * (s * l + ll)] = 1;
    }
    int64_t ng = 0;
    for (int64_t g = 0; g < nFLat; ++g)
        if (ghostId[g]) ghostId[g] = (int32_t)(++ng);
    C->intermediateNumNodes = ng;
    C->transferMat = (int32_t *)malloc(sizeof(int32_t) * m * (ne + 1));
    C->transferMatCoeffi = (double *)calloc((size_t)ng + 1, sizeof(double));
    for (int64_t k = 0; k < ne; ++k) {
        int64_t e = C->eleMapBack[k] - 1;
        int64_t r = e % ny, c = (e / ny) % nx, l = e / (ny * nx);
        for (int ll = 0; ll <= s; ++ll)
            for (int cc = 0; cc <= s; ++cc)
                for (int rr = 0; rr <= s; ++rr) {
                    int n = rr + (s + 1) * cc + (s + 1) * (s + 1) * ll;
                    int32_t gid = ghostId[(s * r + rr) + (nyf + 1) * (s * c + cc) + plf * (s * l + ll)];
                    C->transferMat[(int64_t)m * k + n] = gid;
                    C->transferMatCoeffi[gid - 1] += 1.0;
                }
    }
    C->solidNodeMapCoarser2Finer = (int32_t *)malloc(sizeof(int32_t) * (F->numNodes + 1));
    for (int64_t j = 0; j < F->numNodes; ++j)
        C->solidNodeMapCoarser2Finer[j] = ghostId[F->nodMapBack[j] - 1];
    free(ghostId);
    build_colors(C);
    build_nodes(C);
    C->RI = (double *)malloc(sizeof(double) * m * 8);
    or_transfer_table(s, C->RI);
}

or_mesh *or_mesh_create(int resX, int resY, int resZ, const uint8_t *voxels,
                        int numLevels, int nonDyadic, double nu, double cellSize) {
    if (numLevels < 2 || numLevels > OR_MAX_LEVELS) return NULL;
    or_mesh *M = (or_mesh *)calloc(1, sizeof(or_mesh));
    /* Solving_BuildingMeshHierarchy.m:12-16 */
    if (nonDyadic == 1 && numLevels >= 4) numLevels -= 1; else nonDyadic = 0;
    M->numLevels = numLevels;
    M->nonDyadic = nonDyadic;
    or_element_stiffness(nu, cellSize, M->Ke);
    M->nu = nu;
    M->cellSize = cellSize;
    or_level *L1 = &M->L[0];
    L1->resX = resX; L1->resY = resY; L1->resZ = resZ; L1->span = 1;
    const int64_t nLat = (int64_t)resX * resY * resZ;
    L1->eleMapForward = (int32_t *)calloc((size_t)nLat, sizeof(int32_t));
    int64_t ne = 0;
    for (int64_t v = 0; v < nLat; ++v)
        if (voxels[v]) L1->eleMapForward[v] = (int32_t)(++ne);
    L1->numElements = ne;
    L1->eleMapBack = (int32_t *)malloc(sizeof(int32_t) * (ne + 1));
    for (int64_t v = 0; v < nLat; ++v)
        if (L1->eleMapForward[v]) L1->eleMapBack[L1->eleMapForward[v] - 1] = (int32_t)(v + 1);
    build_colors(L1);
    build_nodes(L1);
    for (int ii = 1; ii < numLevels; ++ii) {
        int s = (ii == 1 && nonDyadic == 1) ? 4 : 2;
        const or_level *F = &M->L[ii - 1];
        if (F->resX % s || F->resY % s || F->resZ % s) { or_mesh_destroy(M); return NULL; }
        build_coarse(F, &M->L[ii], s);
    }
    for (int ii = 0; ii < numLevels; ++ii)
        M->L[ii].fixedDOFs = (uint8_t *)calloc((size_t)M->L[ii].numDOFs + 1, 1);
    M->eleModulus = (double *)malloc(sizeof(double) * (ne + 1));
    for (int64_t i = 0; i < ne; ++i) M->eleModulus[i] = 1.0;
    return M;
}

void or_mesh_destroy(or_mesh *M) {
    if (!M) return;
    for (int i = 0; i < OR_MAX_LEVELS; ++i) {
        or_level *L = &M->L[i];
        free(L->eleMapBack); free(L->eleMapForward); free(L->nodMapBack);
        free(L->nodMapForward); free(L->eNodMat); free(L->colors);
        free(L->elementUpwardMap); free(L->transferMat); free(L->solidNodeMapCoarser2Finer);
        free(L->transferMatCoeffi); free(L->RI); free(L->fixedDOFs); free(L->diagK); free(L->Ks);
    }
    free(M->eleModulus); free(M->freeIdxC); free(M->cholL);
    free(M);
}

int or_mesh_num_levels(const or_mesh *M) { return M->numLevels; }

void or_mesh_level_info(const or_mesh *M, int level, int64_t *out) {
    const or_level *L = &M->L[level - 1];
    out[0] = L->resX; out[1] = L->resY; out[2] = L->resZ; out[3] = L->span;
    out[4] = L->numElements; out[5] = L->numNodes; out[6] = L->numDOFs;
    out[7] = L->intermediateNumNodes;
}

int64_t or_mesh_get(const or_mesh *M, int level, const char *name, void *out) {
    if (level < 1 || level > M->numLevels) return -1;
    const or_level *L = &M->L[level - 1];
    const int s = L->span, nq = s * s * s, m = (s + 1) * (s + 1) * (s + 1);
    const void *src = NULL;
    size_t esz = 4;
    int64_t n = -1;
#define OR_GET(nm, ptr, cnt, sz) if (!strcmp(name, nm)) { src = (ptr); n = (cnt); esz = (sz); }
    OR_GET("eleMapBack", L->eleMapBack, L->numElements, 4)
    OR_GET("eNodMat", L->eNodMat, 8 * L->numElements, 4)
    OR_GET("nodMapBack", L->nodMapBack, L->numNodes, 4)
    OR_GET("colors", L->colors, L->numElements, 4)
    OR_GET("colorOff", L->colorOff, 9, 8)
    OR_GET("fixedDOFs", L->fixedDOFs, L->numDOFs, 1)
    OR_GET("diagK", L->diagK, L->diagK ? L->numDOFs : -1, 8)
    if (level == 1) { OR_GET("Ke", M->Ke, 576, 8) OR_GET("eleModulus", M->eleModulus, L->numElements, 8) }
    if (level >= 2) {
        OR_GET("elementUpwardMap", L->elementUpwardMap, (int64_t)nq * L->numElements, 4)
        OR_GET("transferMat", L->transferMat, (int64_t)m * L->numElements, 4)
        OR_GET("solidNodeMapCoarser2Finer", L->solidNodeMapCoarser2Finer, M->L[level - 2].numNodes, 4)
        OR_GET("transferMatCoeffi", L->transferMatCoeffi, L->intermediateNumNodes, 8)
        OR_GET("RI", L->RI, (int64_t)m * 8, 8)
        OR_GET("Ks", L->Ks, L->Ks ? 576 * L->numElements : -1, 8)
    }
#undef OR_GET
    if (n < 0 || !src) return -1;
    if (out) memcpy(out, src, esz * (size_t)n);
    return n;
}

/* interleaved (3n) <-> SoA n x 3 (Solving_RestrictResidual.m:4,22). */
static void to_soa(const double *v, int64_t n, double *o) {
    for (int64_t j = 0; j < n; ++j)
        for (int a = 0; a < 3; ++a) o[j + a * n] = v[3 * j + a];
}
static void from_soa(const double *o, int64_t n, double *v) {
    for (int64_t j = 0; j < n; ++j)
        for (int a = 0; a < 3; ++a) v[3 * j + a] = o[j + a * n];
}

void or_mesh_restrict(const or_mesh *M, int ii, const double *rF, double *rC) {
    const or_level *F = &M->L[ii - 2], *C = &M->L[ii - 1];
    const int s = C->span, m = (s + 1) * (s + 1) * (s + 1);
    double *a = malloc(sizeof(double) * F->numDOFs), *b = malloc(sizeof(double) * C->numDOFs);
    double *RIT = malloc(sizeof(double) * 8 * m);
    for (int n = 0; n < m; ++n)
        for (int J = 0; J < 8; ++J) RIT[J + 8 * n] = C->RI[n + (int64_t)m * J];
    to_soa(rF, F->numNodes, a);
    or_restrict_mex(a, F->numNodes, C->transferMat, m, RIT, C->eNodMat, C->numElements,
                    C->numNodes, C->intermediateNumNodes, C->solidNodeMapCoarser2Finer,
                    C->transferMatCoeffi, C->colors, C->colorOff, b);
    from_soa(b, C->numNodes, rC);
    free(a); free(b); free(RIT);
}

void or_mesh_interpolate(const or_mesh *M, int ii, const double *xC, double *xF) {
    const or_level *F = &M->L[ii - 2], *C = &M->L[ii - 1];
    const int s = C->span, m = (s + 1) * (s + 1) * (s + 1);
    double *a = malloc(sizeof(double) * C->numDOFs), *b = malloc(sizeof(double) * F->numDOFs);
    to_soa(xC, C->numNodes, a);
    or_interpolate_mex(a, C->numNodes, C->transferMat, m, C->RI, C->eNodMat, C->numElements,
                       F->numNodes, C->intermediateNumNodes, C->solidNodeMapCoarser2Finer,
                       C->transferMatCoeffi, C->colors, C->colorOff, b);
    from_soa(b, F->numNodes, xF);
    free(a); free(b);
}

/* FEA_SetupVoxelBased.m:13-26 */
int or_mesh_set_fixed(or_mesh *M, const uint8_t *fixed) {
    or_level *L1 = &M->L[0];
    memcpy(L1->fixedDOFs, fixed, (size_t)L1->numDOFs);
    double *ind = malloc(sizeof(double) * L1->numDOFs);
    for (int64_t i = 0; i < L1->numDOFs; ++i) ind[i] = fixed[i] ? 1.0 : 0.0;
    for (int ii = 2; ii <= M->numLevels; ++ii) {
        or_level *C = &M->L[ii - 1];
        double *nxt = malloc(sizeof(double) * C->numDOFs);
        or_mesh_restrict(M, ii, ind, nxt);
        for (int64_t i = 0; i < C->numDOFs; ++i) C->fixedDOFs[i] = (nxt[i] != 0.0);
        free(ind);
        ind = nxt;
    }
    free(ind);
    return 0;
}

int or_mesh_set_modulus(or_mesh *M, const double *E) {
    memcpy(M->eleModulus, E, sizeof(double) * M->L[0].numElements);
    return 0;
}

/* Dense Cholesky (lower, col-major), in place.  Returns 0 or 1 if not SPD. */
static int dense_cholesky(double *A, int64_t n) {
    for (int64_t j = 0; j < n; ++j) {
        double d = A[j + n * j];
        for (int64_t k = 0; k < j; ++k) d -= A[j + n * k] * A[j + n * k];
        if (!(d > 0.0)) return 1;
        d = sqrt(d);
        A[j + n * j] = d;
#pragma omp parallel for schedule(static) if (n - j > 256)
        for (int64_t i = j + 1; i < n; ++i) {
            double v = A[i + n * j];
            for (int64_t k = 0; k < j; ++k) v -= A[i + n * k] * A[j + n * k];
            A[i + n * j] = v / d;
        }
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < j; ++i) A[i + n * j] = 0.0;
    return 0;
}

/* Lean storage (or_mesh_set_lean): with three levels or more the level-2
 * Galerkin matrices (24 x 24 per coarse element, 23 GB for 5 M coarse
 * elements) are not kept.  Their diagonals are summed from the templates'
 * diagonal entries, and each level-3 element rebuilds its children's matrices
 * on the fly; every value is formed by the same operations in the same order
 * as in the stored path, so the results are identical. */
void or_mesh_set_lean(or_mesh *M, int lean) { M->lean = lean; }

static void assemble_level3_lean(const or_mesh *M, const double *H2, or_level *C) {
    const or_level *C2 = &M->L[1];
    const int s = C->span, nq = s * s * s, nq2 = C2->span * C2->span * C2->span;
    double (*P)[8][8] = malloc(sizeof(double) * 64 * nq);
    for (int q = 0; q < nq; ++q) child_prolongation(s, q, P[q]);
    const int64_t nC = C->numElements;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < nC; ++i) {
        double acc[576], K2[576];
        for (int t = 0; t < 576; ++t) acc[t] = 0.0;
        for (int q = 0; q < nq; ++q) {
            int32_t child = C->elementUpwardMap[i + q * nC];
            if (child <= 0) continue;
            finest_element(H2, M->eleModulus, C2->elementUpwardMap, C2->numElements, nq2, child - 1, K2);
            galerkin_child_add(P[q], K2, 1.0, acc);
        }
        memcpy(C->Ks + 576 * i, acc, sizeof(acc));
    }
    free(P);
}

/* Solving_AssembleFEAstencil.m:9-135 */
int or_mesh_assemble(or_mesh *M) {
    const int Lc = M->numLevels;
    const int lean2 = M->lean && Lc >= 3;
    double *H2 = NULL;
    for (int ii = 2; ii <= Lc; ++ii) {
        or_level *C = &M->L[ii - 1];
        free(C->Ks);
        C->Ks = NULL;
        if (ii == 2 && lean2) { /* diagonal only, :82-99 order */
            const int nq = C->span * C->span * C->span;
            H2 = finest_templates(M->Ke, C->span);
            free(C->diagK);
            C->diagK = (double *)calloc((size_t)C->numDOFs, sizeof(double));
            for (int64_t e = 0; e < C->numElements; ++e) {
                double d[24];
                for (int t = 0; t < 24; ++t) d[t] = 0.0;
                for (int q = 0; q < nq; ++q) {
                    int32_t child = C->elementUpwardMap[e + q * C->numElements];
                    if (child == 0) continue;
                    double Ev = M->eleModulus[child - 1];
                    for (int t = 0; t < 24; ++t) d[t] += Ev * H2[576 * q + 25 * t];
                }
                for (int j = 0; j < 8; ++j) {
                    int64_t nd = C->eNodMat[e + j * C->numElements] - 1;
                    for (int a = 0; a < 3; ++a) C->diagK[3 * nd + a] += d[3 * j + a];
                }
            }
            continue;
        }
        C->Ks = (double *)malloc(sizeof(double) * 576 * (C->numElements + 1));
        if (ii == 2)
            or_assemble_finest(M->Ke, M->eleModulus, C->elementUpwardMap, C->numElements, C->span, C->Ks);
        else if (ii == 3 && lean2)
            assemble_level3_lean(M, H2, C);
        else
            or_assemble_nonfinest(M->L[ii - 2].Ks, C->elementUpwardMap, C->numElements, C->span, C->Ks);
        if (ii < Lc) { /* :82-99 diag accumulated in ascending element order */
            free(C->diagK);
            C->diagK = (double *)calloc((size_t)C->numDOFs, sizeof(double));
            for (int64_t e = 0; e < C->numElements; ++e)
                for (int j = 0; j < 8; ++j) {
                    int64_t nd = C->eNodMat[e + j * C->numElements] - 1;
                    for (int a = 0; a < 3; ++a)
                        C->diagK[3 * nd + a] += C->Ks[576 * e + (3 * j + a) * 25];
                }
        }
    }
    /* :102-121 finest diagonal */
    or_level *L1 = &M->L[0];
    free(L1->diagK);
    L1->diagK = (double *)calloc((size_t)L1->numDOFs, sizeof(double));
    for (int64_t e = 0; e < L1->numElements; ++e)
        for (int j = 0; j < 8; ++j) {
            int64_t nd = L1->eNodMat[e + j * L1->numElements] - 1;
            for (int a = 0; a < 3; ++a)
                L1->diagK[3 * nd + a] += M->Ke[(3 * j + a) * 25] * M->eleModulus[e];
        }
    /* :124-135 coarsest: assemble, restrict to free DOFs, factor */
    or_level *C = &M->L[Lc - 1];
    const int64_t nd = C->numDOFs;
    free(M->freeIdxC);
    M->freeIdxC = (int64_t *)malloc(sizeof(int64_t) * (nd + 1));
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (nd + 1));
    int64_t nf = 0;
    for (int64_t i = 0; i < nd; ++i) {
        pos[i] = C->fixedDOFs[i] ? -1 : nf;
        if (!C->fixedDOFs[i]) M->freeIdxC[nf++] = i;
    }
    M->nFreeC = nf;
    free(M->cholL);
    M->cholL = (double *)calloc((size_t)(nf * nf + 1), sizeof(double));
    for (int64_t e = 0; e < C->numElements; ++e) {
        int64_t dofs[24];
        for (int j = 0; j < 8; ++j)
            for (int a = 0; a < 3; ++a)
                dofs[3 * j + a] = 3 * (int64_t)(C->eNodMat[e + j * C->numElements] - 1) + a;
        for (int c = 0; c < 24; ++c) {
            int64_t pc = pos[dofs[c]];
            if (pc < 0) continue;
            for (int r = 0; r < 24; ++r) {
                int64_t pr = pos[dofs[r]];
                if (pr < 0) continue;
                M->cholL[pr + nf * pc] += C->Ks[576 * e + r + 24 * c];
            }
        }
    }
    free(pos);
    free(H2);
    /* Intermediate-level Ks are no longer needed except for inspection. */
    return dense_cholesky(M->cholL, nf);
}

/* Solving_KbyU_MatrixFree.m:9,39 */
void or_mesh_kbyu(const or_mesh *M, const double *u, double *y) {
    const or_level *L = &M->L[0];
    or_kbyu_mex(u, L->numDOFs, L->eNodMat, L->numElements, M->Ke, M->eleModulus,
                L->colors, L->colorOff, y);
    for (int64_t i = 0; i < L->numDOFs; ++i)
        if (L->fixedDOFs[i]) y[i] = 0.0;
}

/* Solving_Vcycle.m:29-31 on free DOFs (dense factor in place of CHOLMOD). */
static void coarsest_solve(const or_mesh *M, const double *r, double *x) {
    const or_level *C = &M->L[M->numLevels - 1];
    const int64_t nf = M->nFreeC;
    double *w = malloc(sizeof(double) * (nf + 1));
    for (int64_t i = 0; i < nf; ++i) w[i] = r[M->freeIdxC[i]];
    const double *Lm = M->cholL;
    for (int64_t j = 0; j < nf; ++j) { /* L w = b */
        w[j] /= Lm[j + nf * j];
        const double wj = w[j];
        for (int64_t i = j + 1; i < nf; ++i) w[i] -= Lm[i + nf * j] * wj;
    }
    for (int64_t j = nf - 1; j >= 0; --j) { /* L^T x = w */
        double v = w[j];
        for (int64_t i = j + 1; i < nf; ++i) v -= Lm[i + nf * j] * w[i];
        w[j] = v / Lm[j + nf * j];
    }
    for (int64_t i = 0; i < C->numDOFs; ++i) x[i] = 0.0;
    for (int64_t i = 0; i < nf; ++i) x[M->freeIdxC[i]] = w[i];
    free(w);
}

/* Solving_Vcycle.m:1-52 */
void or_mesh_vcycle(const or_mesh *M, double omega, const double *r, double *z) {
    const int Lc = M->numLevels;
    double *rl[OR_MAX_LEVELS] = {0}, *xl[OR_MAX_LEVELS] = {0};
    const or_level *L1 = &M->L[0];
    for (int ii = 2; ii <= Lc; ++ii) {
        const or_level *C = &M->L[ii - 1];
        rl[ii - 1] = malloc(sizeof(double) * C->numDOFs);
        if (ii == 2) {
            for (int64_t i = 0; i < L1->numDOFs; ++i) z[i] = omega * r[i] / L1->diagK[i];
            or_mesh_restrict(M, 2, r, rl[1]);
        } else {
            const or_level *F = &M->L[ii - 2];
            xl[ii - 2] = malloc(sizeof(double) * F->numDOFs);
            for (int64_t i = 0; i < F->numDOFs; ++i) xl[ii - 2][i] = omega * rl[ii - 2][i] / F->diagK[i];
            or_mesh_restrict(M, ii, rl[ii - 2], rl[ii - 1]);
        }
    }
    xl[Lc - 1] = malloc(sizeof(double) * M->L[Lc - 1].numDOFs);
    coarsest_solve(M, rl[Lc - 1], xl[Lc - 1]);
    for (int ii = Lc; ii >= 2; --ii) {
        const or_level *F = &M->L[ii - 2];
        double *t = malloc(sizeof(double) * F->numDOFs);
        or_mesh_interpolate(M, ii, xl[ii - 1], t);
        if (ii == 2) {
            for (int64_t i = 0; i < F->numDOFs; ++i) z[i] = z[i] + t[i];
            for (int64_t i = 0; i < F->numDOFs; ++i) z[i] = z[i] + omega * r[i] / F->diagK[i];
        } else {
            double *x = xl[ii - 2];
            for (int64_t i = 0; i < F->numDOFs; ++i) x[i] = x[i] + t[i];
            for (int64_t i = 0; i < F->numDOFs; ++i) x[i] = x[i] + omega * rl[ii - 2][i] / F->diagK[i];
        }
        free(t);
    }
    for (int64_t i = 0; i < L1->numDOFs; ++i)
        if (L1->fixedDOFs[i]) z[i] = 0.0;
    for (int i = 0; i < OR_MAX_LEVELS; ++i) { free(rl[i]); free(xl[i]); }
}

static double dot(const double *a, const double *b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* Solving_PreconditionedConjugateGradientSolver.m:7-36 */
int or_mesh_pcg(const or_mesh *M, double omega, const double *b, double *y,
                double tol, int maxIT, double *resHist, int *itsOut) {
    const int64_t n = M->L[0].numDOFs;
    const double normB = sqrt(dot(b, b, n));
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n);
    or_mesh_kbyu(M, y, z);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - z[i];
    or_mesh_vcycle(M, omega, r, z);
    memcpy(p, z, sizeof(double) * n);
    double x1 = dot(z, r, n);
    int its = 0, conv = 0;
    for (its = 1; its <= maxIT; ++its) {
        or_mesh_kbyu(M, p, z);
        double lambda = x1 / dot(p, z, n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            y[i] = y[i] + lambda * p[i];
            r[i] = r[i] - lambda * z[i];
        }
        double res = sqrt(dot(r, r, n)) / normB;
        if (resHist) resHist[its - 1] = res;
        if (res < tol) { conv = 1; break; }
        or_mesh_vcycle(M, omega, r, z);
        double x2 = dot(z, r, n);
        const double beta = x2 / x1;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        x1 = x2;
    }
    if (!conv) its = maxIT;
    *itsOut = its;
    free(r); free(z); free(p);
    return conv ? 0 : 2;
}

void or_mesh_unit_compliance(const or_mesh *M, const double *U, double *ce) {
    const or_level *L = &M->L[0];
    or_unit_compliance_mex(U, L->eNodMat, L->numElements, M->Ke, ce);
}

int or_set_threads(int n) {
    const int prev = omp_get_max_threads();
    if (n > 0) omp_set_num_threads(n);
    return prev;
}

double or_mesh_time_pcg_iterations(const or_mesh *M, double omega, const double *b,
                                   int nits, double *tK, double *tV) {
    const int64_t n = M->L[0].numDOFs;
    double *y = calloc((size_t)n, sizeof(double));
    double *r = malloc(sizeof(double) * n), *z = malloc(sizeof(double) * n);
    double *p = malloc(sizeof(double) * n);
    double sK = 0, sV = 0;
    memcpy(r, b, sizeof(double) * n);
    or_mesh_vcycle(M, omega, r, z);
    memcpy(p, z, sizeof(double) * n);
    double x1 = dot(z, r, n);
    const double t0 = omp_get_wtime(); /* the iterations only (not the initial z = M r) */
    for (int it = 0; it < nits; ++it) {
        double a = omp_get_wtime();
        or_mesh_kbyu(M, p, z);
        sK += omp_get_wtime() - a;
        double lambda = x1 / dot(p, z, n);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            y[i] += lambda * p[i];
            r[i] -= lambda * z[i];
        }
        (void)dot(r, r, n);
        a = omp_get_wtime();
        or_mesh_vcycle(M, omega, r, z);
        sV += omp_get_wtime() - a;
        double x2 = dot(z, r, n), beta = x2 / x1;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        x1 = x2;
    }
    double t = omp_get_wtime() - t0;
    if (tK) *tK = sK;
    if (tV) *tV = sV;
    free(y); free(r); free(z); free(p);
    return t;
}


/* ------------------------------ density filter ------------------------------ */
/* Both MEX functions visit the neighbours of e1 with layer outermost, then
 * column, then row, and add max(0, rMin - sqrt(d2)) (times x) for solid ones;
 * the sums keep that order so results match to the last bit. */
static void density_filter_impl(const double *x, double rMin, int64_t numElements,
                                const int32_t *efwd, int resX, int resY, int resZ, double *Y) {
    const int R = (int)ceil(rMin) - 1;
    const int64_t resXY = (int64_t)resX * resY;
    (void)numElements;
#pragma omp parallel for collapse(3) schedule(static)
    for (int kk = 0; kk < resZ; ++kk)
        for (int ii = 0; ii < resX; ++ii)
            for (int jj = 0; jj < resY; ++jj) {
                const int32_t e1 = efwd[kk * resXY + (int64_t)ii * resY + jj];
                if (!e1) continue;
                double acc = 0.0;
                const int k0 = kk - R > 0 ? kk - R : 0, k1 = kk + R < resZ - 1 ? kk + R : resZ - 1;
                const int i0 = ii - R > 0 ? ii - R : 0, i1 = ii + R < resX - 1 ? ii + R : resX - 1;
                const int j0 = jj - R > 0 ? jj - R : 0, j1 = jj + R < resY - 1 ? jj + R : resY - 1;
                for (int k2 = k0; k2 <= k1; ++k2)
                    for (int i2 = i0; i2 <= i1; ++i2)
                        for (int j2 = j0; j2 <= j1; ++j2) {
                            const int32_t e2 = efwd[k2 * resXY + (int64_t)i2 * resY + j2];
                            if (!e2) continue;
                            const int d2 = (ii - i2) * (ii - i2) + (jj - j2) * (jj - j2) + (kk - k2) * (kk - k2);
                            const double w = rMin - sqrt((double)d2);
                            const double wp = w > 0.0 ? w : 0.0;
                            acc += x ? wp * x[e2 - 1] : wp;
                        }
                Y[e1 - 1] = acc;
            }
}

void or_density_filter_setup(double rMin, int64_t numElements, const int32_t *eleMapForward,
                             int resX, int resY, int resZ, double *Hs) {
    density_filter_impl(NULL, rMin, numElements, eleMapForward, resX, resY, resZ, Hs);
}

void or_density_filter(const double *x, double rMin, int64_t numElements,
                       const int32_t *eleMapForward, int resX, int resY, int resZ, double *Y) {
    density_filter_impl(x, rMin, numElements, eleMapForward, resX, resY, resZ, Y);
}


/* ---------------------------- stress recovery ------------------------------ */
/* FEA_StressAnalysis.m:1-66: per element the stresses at the 8 Gauss points
 * E_e D B_g u_e (FEA_HookeLaw.m, FEA_ElementStrainMatrix.m; Gauss points in
 * node order, FEA_GaussianIntegral.m), extrapolated to the element's nodes
 * with the inverse of the shape-function matrix at the Gauss points
 * (OuterInterpolationMat, :68-93), summed at the nodes and divided by the
 * number of solid elements at each node; von Mises per node (:55-59).
 * stress: numNodes x 6 column-major (xx yy zz yz zx xy). */
static void or_invert8(const double A[8][8], double Ainv[8][8]) {
    double M[8][16];
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 16; ++j) M[i][j] = j < 8 ? A[i][j] : (j - 8 == i ? 1.0 : 0.0);
    for (int c = 0; c < 8; ++c) {
        int piv = c;
        for (int r = c + 1; r < 8; ++r)
            if (fabs(M[r][c]) > fabs(M[piv][c])) piv = r;
        if (piv != c)
            for (int j = 0; j < 16; ++j) { double tmp = M[c][j]; M[c][j] = M[piv][j]; M[piv][j] = tmp; }
        const double d = M[c][c];
        for (int j = 0; j < 16; ++j) M[c][j] /= d;
        for (int r = 0; r < 8; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
            for (int j = 0; j < 16; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) Ainv[i][j] = M[i][j + 8];
}

void or_mesh_stress(const or_mesh *M, const double *U, double *stress, double *vonMises) {
    const or_level *L = &M->L[0];
    const double g = 0.577350269189626, nu = M->nu, h = M->cellSize;
    double S[6][6];
    memset(S, 0, sizeof(S));
    const double c1 = (1 + nu) * (1 - 2 * nu), c2 = 2 * (1 + nu);
    S[0][0] = S[1][1] = S[2][2] = (1 - nu) / c1;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) S[i][j] = nu / c1;
    S[3][3] = S[4][4] = S[5][5] = 1.0 / c2;
    const double invJ = 2.0 / h;
    double DB[8][6][24], N[8][8], Ninv[8][8];
    for (int q = 0; q < 8; ++q) {
        const double s = OR_CS[q] * g, t = OR_CT[q] * g, p = OR_CP[q] * g;
        double B[6][24];
        memset(B, 0, sizeof(B));
        for (int j = 0; j < 8; ++j) {
            N[q][j] = 0.125 * (1 + OR_CS[j] * s) * (1 + OR_CT[j] * t) * (1 + OR_CP[j] * p);
            const double dNs = 0.125 * OR_CS[j] * (1 + OR_CT[j] * t) * (1 + OR_CP[j] * p);
            const double dNt = 0.125 * OR_CT[j] * (1 + OR_CS[j] * s) * (1 + OR_CP[j] * p);
            const double dNp = 0.125 * OR_CP[j] * (1 + OR_CS[j] * s) * (1 + OR_CT[j] * t);
            const double dx = invJ * dNs, dy = invJ * dNt, dz = invJ * dNp;
            B[0][3 * j] = dx;
            B[1][3 * j + 1] = dy;
            B[2][3 * j + 2] = dz;
            B[3][3 * j + 1] = dz; B[3][3 * j + 2] = dy;
            B[4][3 * j] = dz;     B[4][3 * j + 2] = dx;
            B[5][3 * j] = dy;     B[5][3 * j + 1] = dx;
        }
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 24; ++k) {
                double acc = 0;
                for (int l = 0; l < 6; ++l) acc += S[i][l] * B[l][k];
                DB[q][i][k] = acc;
            }
    }
    or_invert8(N, Ninv);
    const int64_t nn = L->numNodes, ne = L->numElements;
    double *acc = (double *)calloc((size_t)nn * 6, sizeof(double));
    int *cnt = (int *)calloc((size_t)nn, sizeof(int));
    for (int64_t e = 0; e < ne; ++e) {
        double ue[24], sg[8][6];
        for (int k = 0; k < 8; ++k) {
            const int64_t node = L->eNodMat[e + ne * k] - 1;
            for (int a = 0; a < 3; ++a) ue[3 * k + a] = U[3 * node + a];
        }
        for (int q = 0; q < 8; ++q)
            for (int i = 0; i < 6; ++i) {
                double v = 0;
                for (int k = 0; k < 24; ++k) v += DB[q][i][k] * ue[k];
                sg[q][i] = v * M->eleModulus[e];
            }
        for (int n = 0; n < 8; ++n) {
            const int64_t node = L->eNodMat[e + ne * n] - 1;
            for (int i = 0; i < 6; ++i) {
                double v = 0;
                for (int q = 0; q < 8; ++q) v += Ninv[n][q] * sg[q][i];
                acc[node + nn * i] += v;
            }
            cnt[node] += 1;
        }
    }
    for (int64_t n = 0; n < nn; ++n) {
        double sv[6];
        for (int i = 0; i < 6; ++i) {
            sv[i] = cnt[n] ? acc[n + nn * i] / cnt[n] : 0.0;
            stress[n + nn * i] = sv[i];
        }
        vonMises[n] = sqrt(0.5 * ((sv[0] - sv[1]) * (sv[0] - sv[1]) + (sv[1] - sv[2]) * (sv[1] - sv[2]) +
                                  (sv[2] - sv[0]) * (sv[2] - sv[0])) +
                           3 * (sv[5] * sv[5] + sv[3] * sv[3] + sv[4] * sv[4]));
    }
    free(acc);
    free(cnt);
}


/* ------------------------------- PDE filter -------------------------------- */
/* TopOpti_BuildPDEfilter_matrixFree.m:7-37: KePDE = iRmin^2 dShape' dShape +
 * N' diag(w detJ) N with iRmin = rHatMin / 2 / sqrt(3) (unit cell, detJ 1/8),
 * dShape the parametric shape derivatives at the 8 Gauss points
 * (FEA_DeShapeFunction.m); Jacobi preconditioner 1 / sum of diag(KePDE). */
void or_pde_element(double rHatMin, double *Ke8) {
    const double g = 0.577350269189626;
    double KEF0[8][8], KEF1[8][8];
    memset(KEF0, 0, sizeof(KEF0));
    memset(KEF1, 0, sizeof(KEF1));
    for (int q = 0; q < 8; ++q) {
        const double s = OR_CS[q] * g, t = OR_CT[q] * g, p = OR_CP[q] * g;
        double N[8], dN[3][8];
        for (int j = 0; j < 8; ++j) {
            N[j] = 0.125 * (1 + OR_CS[j] * s) * (1 + OR_CT[j] * t) * (1 + OR_CP[j] * p);
            dN[0][j] = 0.125 * OR_CS[j] * (1 + OR_CT[j] * t) * (1 + OR_CP[j] * p);
            dN[1][j] = 0.125 * OR_CT[j] * (1 + OR_CS[j] * s) * (1 + OR_CP[j] * p);
            dN[2][j] = 0.125 * OR_CP[j] * (1 + OR_CS[j] * s) * (1 + OR_CT[j] * t);
        }
        for (int a = 0; a < 8; ++a)
            for (int b = 0; b < 8; ++b) {
                for (int d = 0; d < 3; ++d) KEF0[a][b] += dN[d][a] * dN[d][b];
                KEF1[a][b] += N[a] * (1.0 / 8.0) * N[b];
            }
    }
    const double iRmin = rHatMin / 2.0 / sqrt(3.0);
    for (int a = 0; a < 8; ++a)
        for (int b = 0; b < 8; ++b) Ke8[a + 8 * b] = iRmin * iRmin * KEF0[a][b] + KEF1[a][b];
}

/* TopOpti_PDEFiltering_matrixFree.m:1-33 with MatTimesVec_matrixFree (:35-55):
 * element values to nodes (sum of src/8 over the elements at a node), PCG
 * (Solving_PreconditionedConjugateGradientSolver.m, x0 = 0) with the Jacobi
 * preconditioner, nodes back to elements (mean of the 8 nodes).  Returns the
 * iteration count. */
int or_mesh_pde_filter(const or_mesh *M, double rHatMin, const double *src, double tol, int maxIT,
                       double *tar) {
    const or_level *L = &M->L[0];
    const int64_t nn = L->numNodes, ne = L->numElements;
    double Ke8[64];
    or_pde_element(rHatMin, Ke8);
    double *b = (double *)calloc(nn, sizeof(double)), *x = (double *)calloc(nn, sizeof(double));
    double *r = (double *)calloc(nn, sizeof(double)), *z = (double *)calloc(nn, sizeof(double));
    double *p = (double *)calloc(nn, sizeof(double)), *q = (double *)calloc(nn, sizeof(double));
    double *dinv = (double *)calloc(nn, sizeof(double));
    for (int jj = 0; jj < 8; ++jj)
        for (int64_t e = 0; e < ne; ++e) {
            const int64_t n = L->eNodMat[e + ne * jj] - 1;
            b[n] += src[e] * (1.0 / 8.0);
            dinv[n] += Ke8[jj + 8 * jj];
        }
    for (int64_t n = 0; n < nn; ++n) dinv[n] = 1.0 / dinv[n];
    double bb = 0;
    for (int64_t n = 0; n < nn; ++n) bb += b[n] * b[n];
    const double normB = sqrt(bb);
    for (int64_t n = 0; n < nn; ++n) { r[n] = b[n]; z[n] = dinv[n] * r[n]; p[n] = z[n]; }
    double rho = 0;
    for (int64_t n = 0; n < nn; ++n) rho += z[n] * r[n];
    int its = 0;
    for (its = 1; its <= maxIT; ++its) {
        memset(q, 0, sizeof(double) * nn);
        for (int64_t e = 0; e < ne; ++e) {
            double ue[8];
            int64_t nodes[8];
            for (int k = 0; k < 8; ++k) { nodes[k] = L->eNodMat[e + ne * k] - 1; ue[k] = p[nodes[k]]; }
            for (int j = 0; j < 8; ++j) {
                double y = 0;
                for (int k = 0; k < 8; ++k) y += Ke8[j + 8 * k] * ue[k];
                q[nodes[j]] += y;
            }
        }
        double pq = 0;
        for (int64_t n = 0; n < nn; ++n) pq += p[n] * q[n];
        const double lambda = rho / pq;
        double rr = 0;
        for (int64_t n = 0; n < nn; ++n) {
            x[n] += lambda * p[n];
            r[n] -= lambda * q[n];
            rr += r[n] * r[n];
        }
        if (normB == 0.0 || sqrt(rr) / normB < tol) break;
        double rho2 = 0;
        for (int64_t n = 0; n < nn; ++n) { z[n] = dinv[n] * r[n]; rho2 += z[n] * r[n]; }
        for (int64_t n = 0; n < nn; ++n) p[n] = z[n] + (rho2 / rho) * p[n];
        rho = rho2;
    }
    if (its > maxIT) its = maxIT;
    for (int64_t e = 0; e < ne; ++e) {
        double sum = 0;
        for (int k = 0; k < 8; ++k) sum += x[L->eNodMat[e + ne * k] - 1];
        tar[e] = sum * (1.0 / 8.0);
    }
    free(b); free(x); free(r); free(z); free(p); free(q); free(dinv);
    return its;
}
