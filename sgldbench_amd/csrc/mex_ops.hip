// Stateless replacements of the six MEX kernels on the reference's solve path
// and of the PDE filter's scalar K.u (Solving_KbyU_MatrixFree8x8_mex).
//
// They take the same arrays as the MEX calls (1-based int32 ids, column-major
// matrices) as host pointers, validate every index on the host before any
// launch, run on the GPU and copy the result back.  Element scatters follow
// the reference's colour schedule: one launch per colour, one thread per
// element, so every node accumulates its element contributions in colour
// order, with the per-element sums in the reference's loop order and FMA
// contraction disabled.  The results therefore equal the reference
// algorithm's bit for bit (tests compare them to the oracle with ==).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace {

using sgld::DevBuf;

// --- K.u (Solving_KbyU_MatrixFree_mex.cpp:69-136) ---------------------------
__global__ void __launch_bounds__(256) k_kbyu_colour(const double *__restrict__ U,
                                                     const int32_t *__restrict__ eNodMat,
                                                     int64_t numElements,
                                                     const double *__restrict__ Ke,
                                                     const double *__restrict__ E,
                                                     const int32_t *__restrict__ ids, int64_t n,
                                                     double *__restrict__ Y) {
#pragma clang fp contract(off)
    __shared__ double sK[576];
    for (int t = threadIdx.x; t < 576; t += blockDim.x) sK[t] = Ke[t];
    __syncthreads();
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t e = ids[k] - 1;
    int64_t nod[8];
    double Ue[24], Ye[24];
#pragma unroll
    for (int j = 0; j < 8; ++j) nod[j] = (int64_t)eNodMat[e + j * numElements] - 1;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int a = 0; a < 3; ++a) Ue[3 * j + a] = U[3 * nod[j] + a];
#pragma unroll
    for (int j = 0; j < 24; ++j) Ye[j] = 0.0;
    for (int kk = 0; kk < 24; ++kk) {
        const double uk = Ue[kk];
#pragma unroll
        for (int j = 0; j < 24; ++j) Ye[j] += sK[kk * 24 + j] * uk;
    }
    const double Ei = E[e];
#pragma unroll
    for (int j = 0; j < 24; ++j) Ye[j] *= Ei;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int a = 0; a < 3; ++a) Y[3 * nod[j] + a] += Ye[3 * j + a];
}

// --- scalar K.u of the PDE filter (Solving_KbyU_MatrixFree8x8_mex.cpp:50-121)
// One DOF per node, one uniform 8x8 Ke, no element factor.
__global__ void __launch_bounds__(256) k_kbyu8_colour(const double *__restrict__ U,
                                                      const int32_t *__restrict__ eNodMat,
                                                      int64_t numElements,
                                                      const double *__restrict__ Ke,
                                                      const int32_t *__restrict__ ids, int64_t n,
                                                      double *__restrict__ Y) {
#pragma clang fp contract(off)
    __shared__ double sK[64];
    if (threadIdx.x < 64) sK[threadIdx.x] = Ke[threadIdx.x];
    __syncthreads();
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t e = ids[k] - 1;
    int64_t nod[8];
    double Ue[8], Ye[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) nod[j] = (int64_t)eNodMat[e + j * numElements] - 1;
#pragma unroll
    for (int j = 0; j < 8; ++j) Ue[j] = U[nod[j]];
#pragma unroll
    for (int j = 0; j < 8; ++j) Ye[j] = 0.0;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
        const double uk = Ue[kk];
#pragma unroll
        for (int j = 0; j < 8; ++j) Ye[j] += sK[kk * 8 + j] * uk;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) Y[nod[j]] += Ye[j];
}

// --- restriction (Solving_Restriction_MatrixFree_mex.cpp:72-131) -----------
__global__ void k_ghost_fill(const double *__restrict__ rF, int64_t nF,
                             const int32_t *__restrict__ map, const double *__restrict__ coeffi,
                             int64_t nG, double *__restrict__ ghost) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nF) return;
    const int64_t g = map[j] - 1;
    const double c = coeffi[g];
    for (int a = 0; a < 3; ++a) ghost[g + a * nG] = rF[a * nF + j] / c;
}

__global__ void __launch_bounds__(256) k_restrict_colour(const double *__restrict__ ghost,
                                                         int64_t nG,
                                                         const int32_t *__restrict__ transferMat,
                                                         int m, const double *__restrict__ RIT,
                                                         const int32_t *__restrict__ eNodMatC,
                                                         int64_t numElesC, int64_t nC,
                                                         const int32_t *__restrict__ ids,
                                                         int64_t n, double *__restrict__ Y) {
#pragma clang fp contract(off)
    __shared__ double sR[8 * 125];
    for (int t = threadIdx.x; t < 8 * m; t += blockDim.x) sR[t] = RIT[t];
    __syncthreads();
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t e = ids[k] - 1;
    double Ye[3][8];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 8; ++j) Ye[a][j] = 0.0;
    for (int kk = 0; kk < m; ++kk) {
        const int64_t g = (int64_t)transferMat[(int64_t)m * e + kk] - 1;
        const double ux = ghost[g], uy = ghost[g + nG], uz = ghost[g + 2 * nG];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double w = sR[kk * 8 + j];
            Ye[0][j] += w * ux;
            Ye[1][j] += w * uy;
            Ye[2][j] += w * uz;
        }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t nd = (int64_t)eNodMatC[e + j * numElesC] - 1;
        Y[nd] += Ye[0][j];
        Y[nd + nC] += Ye[1][j];
        Y[nd + 2 * nC] += Ye[2][j];
    }
}

// --- interpolation (Solving_Interpolation_MatrixFree_mex.cpp:75-129) -------
__global__ void __launch_bounds__(256) k_interp_colour(const double *__restrict__ xC, int64_t nC,
                                                       const int32_t *__restrict__ transferMat,
                                                       int m, const double *__restrict__ RI,
                                                       const int32_t *__restrict__ eNodMatC,
                                                       int64_t numElesC,
                                                       const int32_t *__restrict__ ids, int64_t n,
                                                       double *__restrict__ ghost, int64_t nG) {
#pragma clang fp contract(off)
    __shared__ double sR[125 * 8];
    for (int t = threadIdx.x; t < 8 * m; t += blockDim.x) sR[t] = RI[t];
    __syncthreads();
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t e = ids[k] - 1;
    double Ue[3][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t nd = (int64_t)eNodMatC[e + j * numElesC] - 1;
        Ue[0][j] = xC[nd];
        Ue[1][j] = xC[nd + nC];
        Ue[2][j] = xC[nd + 2 * nC];
    }
    for (int j = 0; j < m; ++j) {
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
#pragma unroll
        for (int kk = 0; kk < 8; ++kk) {
            const double w = sR[kk * m + j];
            v0 += w * Ue[0][kk];
            v1 += w * Ue[1][kk];
            v2 += w * Ue[2][kk];
        }
        const int64_t g = (int64_t)transferMat[(int64_t)m * e + j] - 1;
        ghost[g] += v0;
        ghost[g + nG] += v1;
        ghost[g + 2 * nG] += v2;
    }
}

__global__ void k_ghost_gather(const double *__restrict__ ghost, int64_t nG,
                               const int32_t *__restrict__ map, const double *__restrict__ coeffi,
                               int64_t nF, double *__restrict__ Y) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nF) return;
    const int64_t g = map[j] - 1;
    const double c = coeffi[g];
    for (int a = 0; a < 3; ++a) Y[j + a * nF] = ghost[g + a * nG] / c;
}

// --- Galerkin (Solving_AssembleCmptStencilFromFinestLevel.cpp:134-169) ------
// One thread per (coarse element, entry): Ks_e[t] = sum_q E_{child q} H_q[t].
__global__ void k_assemble_finest(const double *__restrict__ H, const double *__restrict__ E,
                                  const int32_t *__restrict__ upMap, int64_t nC, int nq,
                                  double *__restrict__ Ks) {
#pragma clang fp contract(off)
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nC * 576;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx / 576;
        const int t = (int)(idx % 576);
        double acc = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int32_t child = upMap[i + q * nC];
            if (child == 0) continue;
            acc += E[child - 1] * H[576 * q + t];
        }
        Ks[idx] = acc;
    }
}

// Solving_AssembleCmptStencilFromNonFinestLevel.cpp:78-143 as
// Ks_e = sum_q P_q^T Ks_prev(child q) P_q, P_q = p_q (x) I3.  One block of 576
// threads per coarse element; thread t owns output entry (row, col) = (t%24,
// t/24) and column T entry of the intermediate K P_q.
__global__ void __launch_bounds__(576) k_assemble_nonfinest(const double *__restrict__ KsPrev,
                                                            const int32_t *__restrict__ upMap,
                                                            int64_t nC, int nq,
                                                            const double *__restrict__ P,
                                                            double *__restrict__ Ks) {
#pragma clang fp contract(off)
    __shared__ double sK[576], sT[576], sP[64];
    const int t = threadIdx.x;
    const int row = t % 24, col = t / 24;
    for (int64_t i = blockIdx.x; i < nC; i += gridDim.x) {
        double out = 0.0;
        for (int q = 0; q < nq; ++q) {
            const int32_t child = upMap[i + q * nC];
            if (child <= 0) continue;
            __syncthreads();
            sK[t] = KsPrev[576 * (int64_t)(child - 1) + t];
            if (t < 64) sP[t] = P[64 * q + t];  // p[j][J] at j*8 + J
            __syncthreads();
            {   // T[:, 3J+a] = sum_j p[j][J] K[:, 3j+a]
                const int J = col / 3, a = col % 3;
                double v = 0.0;
                for (int j = 0; j < 8; ++j) {
                    const double w = sP[j * 8 + J];
                    if (w == 0.0) continue;
                    v += w * sK[row + 24 * (3 * j + a)];
                }
                sT[t] = v;
            }
            __syncthreads();
            {   // out[3I+a, c] += sum_i p[i][I] T[3i+a, c]
                const int I = row / 3, a = row % 3;
                double acc = 0.0;
                for (int ii = 0; ii < 8; ++ii) {
                    const double w = sP[ii * 8 + I];
                    if (w != 0.0) acc += w * sT[3 * ii + a + 24 * col];
                }
                out += acc;
            }
        }
        Ks[576 * i + t] = out;
    }
}

// --- unit compliance (TopOpti_CmptUnitCompliance_mex.cpp:40-107) -----------
__global__ void __launch_bounds__(256) k_unit_compliance(const double *__restrict__ U,
                                                         const int32_t *__restrict__ eNodMat,
                                                         int64_t numElements,
                                                         const double *__restrict__ Ke,
                                                         double *__restrict__ ce) {
#pragma clang fp contract(off)
    __shared__ double sK[576];
    for (int t = threadIdx.x; t < 576; t += blockDim.x) sK[t] = Ke[t];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numElements) return;
    double Ue[24], y[24];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t nd = (int64_t)eNodMat[i + j * numElements] - 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) Ue[3 * j + a] = U[3 * nd + a];
    }
#pragma unroll
    for (int r = 0; r < 24; ++r) y[r] = 0.0;
    for (int c = 0; c < 24; ++c) {
        const double uj = Ue[c];
#pragma unroll
        for (int r = 0; r < 24; ++r) y[r] += sK[24 * c + r] * uj;
    }
    double q = 0.0;
#pragma unroll
    for (int r = 0; r < 24; ++r) q += Ue[r] * y[r];
    ce[i] = q;
}

// ---------------------------- host validation ------------------------------
int check_ids(const int32_t *a, int64_t n, int64_t hi, const char *what) {
    for (int64_t i = 0; i < n; ++i)
        if (a[i] < 1 || a[i] > hi) {
            sgld::set_error("sgld:arg: %s[%lld] = %d outside 1..%lld", what, (long long)i, a[i],
                            (long long)hi);
            return SGLD_ERR_ARG;
        }
    return SGLD_OK;
}

int check_ids0(const int32_t *a, int64_t n, int64_t hi, const char *what) {
    for (int64_t i = 0; i < n; ++i)
        if (a[i] < 0 || a[i] > hi) {
            sgld::set_error("sgld:arg: %s[%lld] = %d outside 0..%lld", what, (long long)i, a[i],
                            (long long)hi);
            return SGLD_ERR_ARG;
        }
    return SGLD_OK;
}

int check_colours(const int32_t *colors, const int64_t *off, int numColors, int64_t numElements) {
    SGLD_ARG(numColors >= 0, "Solving_KbyU:colorArray: colour count must be >= 0");
    if (numColors == 0) return SGLD_OK;
    SGLD_ARG(off, "Solving_KbyU:colorArray: colorOffsets is null");
    SGLD_ARG(off[0] == 0, "Solving_KbyU:colorCell: colorOffsets[0] must be 0");
    for (int c = 0; c < numColors; ++c)
        SGLD_ARG(off[c + 1] >= off[c], "Solving_KbyU:colorCell: colorOffsets must not decrease");
    if (off[numColors] > 0) {
        SGLD_ARG(colors, "Solving_KbyU:colorCell: colors is null");
        SGLD_TRY(check_ids(colors, off[numColors], numElements, "colors"));
    }
    return SGLD_OK;
}

}  // namespace

extern "C" int sgld_kbyu(const double *U, int64_t numDOFs, const int32_t *eNodMat,
                         int64_t numElements, const double *Ke, int64_t keRows, int64_t keCols,
                         const double *E, int64_t numE, const int32_t *colors,
                         const int64_t *colorOffsets, int numColors, double *Y) {
    SGLD_ARG(keRows == 24 && keCols == 24,
             "MATLAB:mexFunction:inputNot24x24: Input Ke must be a 24x24 double matrix.");
    SGLD_ARG(numDOFs >= 0 && numDOFs % 3 == 0 && (numDOFs == 0 || (U && Y)),
             "MATLAB:mexFunction:inputNotVector: Input U must be a column vector of 3 x nodes.");
    SGLD_ARG(numElements >= 0 && (numElements == 0 || eNodMat),
             "MATLAB:mexFunction:inputNotMatrix: Input eNodMat must be an Mx8 int32 matrix.");
    SGLD_ARG(numE == numElements && (numE == 0 || E),
             "MATLAB:mexFunction:inputNotVector: Input E must be a column vector (one per element).");
    SGLD_ARG(Ke, "MATLAB:mexFunction:inputNot24x24: Ke is null");
    SGLD_TRY(check_ids(eNodMat, 8 * numElements, numDOFs / 3, "eNodMat"));
    SGLD_TRY(check_colours(colors, colorOffsets, numColors, numElements));
    if (numDOFs == 0) return SGLD_OK;
    DevBuf<double> dU, dK, dE, dY;
    DevBuf<int32_t> dN, dC;
    SGLD_TRY(dU.upload(U, numDOFs));
    SGLD_TRY(dK.upload(Ke, 576));
    SGLD_TRY(dE.upload(E, numE));
    SGLD_TRY(dN.upload(eNodMat, 8 * numElements));
    const int64_t nIds = numColors ? colorOffsets[numColors] : 0;
    SGLD_TRY(dC.upload(colors, nIds));
    SGLD_TRY(dY.alloc(numDOFs));
    SGLD_TRY(dY.zero());
    for (int c = 0; c < numColors; ++c) {
        const int64_t n = colorOffsets[c + 1] - colorOffsets[c];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_kbyu_colour, dim3(sgld::blocks_for(n, 256)), dim3(256), 0, 0, dU.p,
                           dN.p, numElements, dK.p, dE.p, dC.p + colorOffsets[c], n, dY.p);
        SGLD_HIP(hipGetLastError());
    }
    SGLD_HIP(hipMemcpy(Y, dY.p, sizeof(double) * numDOFs, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_kbyu8x8(const double *U, int64_t numNodes, const int32_t *eNodMat,
                            int64_t numElements, const double *Ke, int64_t keRows,
                            int64_t keCols, const int32_t *colors, const int64_t *colorOffsets,
                            int numColors, double *Y) {
    SGLD_ARG(keRows == 8 && keCols == 8,
             "MATLAB:mexFunction:inputNot8x8: Input Ke must be a 8x8 double matrix.");
    SGLD_ARG(Ke, "MATLAB:mexFunction:inputNot8x8: Ke is null");
    SGLD_ARG(numNodes >= 0 && (numNodes == 0 || (U && Y)),
             "MATLAB:mexFunction:inputNotVector: Input U must be a column vector.");
    SGLD_ARG(numElements >= 0 && (numElements == 0 || eNodMat),
             "MATLAB:mexFunction:inputNotMatrix: Input eNodMat must be an Mx8 int32 matrix.");
    SGLD_TRY(check_ids(eNodMat, 8 * numElements, numNodes, "eNodMat"));
    SGLD_TRY(check_colours(colors, colorOffsets, numColors, numElements));
    if (numNodes == 0) return SGLD_OK;
    DevBuf<double> dU, dK, dY;
    DevBuf<int32_t> dN, dC;
    SGLD_TRY(dU.upload(U, numNodes));
    SGLD_TRY(dK.upload(Ke, 64));
    SGLD_TRY(dN.upload(eNodMat, 8 * numElements));
    const int64_t nIds = numColors ? colorOffsets[numColors] : 0;
    SGLD_TRY(dC.upload(colors, nIds));
    SGLD_TRY(dY.alloc(numNodes));
    SGLD_TRY(dY.zero());
    for (int c = 0; c < numColors; ++c) {
        const int64_t n = colorOffsets[c + 1] - colorOffsets[c];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_kbyu8_colour, dim3(sgld::blocks_for(n, 256)), dim3(256), 0, 0, dU.p,
                           dN.p, numElements, dK.p, dC.p + colorOffsets[c], n, dY.p);
        SGLD_HIP(hipGetLastError());
    }
    SGLD_HIP(hipMemcpy(Y, dY.p, sizeof(double) * numNodes, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

static int transfer_common_checks(int64_t m, int64_t nF, int64_t numElesC, int64_t nC,
                                  int64_t nGhost, const int32_t *transferMat,
                                  const int32_t *eNodMatC, const int32_t *map,
                                  const double *coeffi, const int32_t *colors,
                                  const int64_t *off, int numColors) {
    SGLD_ARG(m == 27 || m == 125,
             "MATLAB:mexFunction:inputNotMatrix: Input transferMat must be an 27xM or 125xM int32 matrix.");
    SGLD_ARG(nF >= 0 && nC >= 0 && nGhost >= 0 && numElesC >= 0, "sgld:arg: negative size");
    SGLD_ARG(numElesC == 0 || (transferMat && eNodMatC),
             "MATLAB:mexFunction:inputNotMatrix: Input eNodMat must be an Mx8 int32 matrix.");
    SGLD_ARG(nF == 0 || (map && coeffi),
             "MATLAB:mexFunction:inputNotVector: Input solidNodeMapCoarser2Finer must be a column vector.");
    SGLD_TRY(check_ids(transferMat, m * numElesC, nGhost, "transferMat"));
    SGLD_TRY(check_ids(eNodMatC, 8 * numElesC, nC, "eNodMat"));
    SGLD_TRY(check_ids(map, nF, nGhost, "solidNodeMapCoarser2Finer"));
    SGLD_TRY(check_colours(colors, off, numColors, numElesC));
    return SGLD_OK;
}

extern "C" int sgld_restrict(const double *rF, int64_t nF, const int32_t *transferMat, int64_t m,
                             const double *RIT, int64_t ritRows, int64_t ritCols,
                             const int32_t *eNodMatC, int64_t numElesC, int64_t nC,
                             int64_t nGhost, const int32_t *map, const double *coeffi,
                             const int32_t *colors, const int64_t *off, int numColors,
                             double *Y) {
    SGLD_ARG(ritRows == 8 && ritCols == m && RIT,
             "MATLAB:mexFunction:inputNotMx8: Input RIT must be a 8x 125/27 double matrix.");
    SGLD_TRY(transfer_common_checks(m, nF, numElesC, nC, nGhost, transferMat, eNodMatC, map,
                                    coeffi, colors, off, numColors));
    SGLD_ARG(nC == 0 || Y, "sgld:arg: Y is null");
    SGLD_ARG(nF == 0 || rF, "MATLAB:mexFunction:inputNotVector: Input rFiner must be a n-by-3 matrix.");
    if (nC == 0) return SGLD_OK;
    DevBuf<double> drF, dR, dCo, dG, dY;
    DevBuf<int32_t> dT, dN, dM, dC;
    SGLD_TRY(drF.upload(rF, 3 * nF));
    SGLD_TRY(dR.upload(RIT, 8 * m));
    SGLD_TRY(dCo.upload(coeffi, nGhost));
    SGLD_TRY(dT.upload(transferMat, m * numElesC));
    SGLD_TRY(dN.upload(eNodMatC, 8 * numElesC));
    SGLD_TRY(dM.upload(map, nF));
    const int64_t nIds = numColors ? off[numColors] : 0;
    SGLD_TRY(dC.upload(colors, nIds));
    SGLD_TRY(dG.alloc(3 * std::max<int64_t>(nGhost, 1)));
    SGLD_TRY(dG.zero());
    SGLD_TRY(dY.alloc(3 * nC));
    SGLD_TRY(dY.zero());
    if (nF) {
        hipLaunchKernelGGL(k_ghost_fill, dim3(sgld::blocks_for(nF, 256)), dim3(256), 0, 0, drF.p, nF,
                           dM.p, dCo.p, nGhost, dG.p);
        SGLD_HIP(hipGetLastError());
    }
    for (int c = 0; c < numColors; ++c) {
        const int64_t n = off[c + 1] - off[c];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_restrict_colour, dim3(sgld::blocks_for(n, 256)), dim3(256), 0, 0, dG.p,
                           nGhost, dT.p, (int)m, dR.p, dN.p, numElesC, nC, dC.p + off[c], n, dY.p);
        SGLD_HIP(hipGetLastError());
    }
    SGLD_HIP(hipMemcpy(Y, dY.p, sizeof(double) * 3 * nC, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_interpolate(const double *xC, int64_t nC, const int32_t *transferMat,
                                int64_t m, const double *RI, int64_t riRows, int64_t riCols,
                                const int32_t *eNodMatC, int64_t numElesC, int64_t nF,
                                int64_t nGhost, const int32_t *map, const double *coeffi,
                                const int32_t *colors, const int64_t *off, int numColors,
                                double *Y) {
    SGLD_ARG(riRows == m && riCols == 8 && RI,
             "MATLAB:mexFunction:inputNotMatrix: Input RI must be a 125/27 x 8 double matrix.");
    SGLD_TRY(transfer_common_checks(m, nF, numElesC, nC, nGhost, transferMat, eNodMatC, map,
                                    coeffi, colors, off, numColors));
    SGLD_ARG(nF == 0 || Y, "sgld:arg: Y is null");
    SGLD_ARG(nC == 0 || xC, "MATLAB:mexFunction:inputNotVector: Input xCoarser must be a n-by-3 matrix.");
    if (nF == 0) return SGLD_OK;
    DevBuf<double> dX, dR, dCo, dG, dY;
    DevBuf<int32_t> dT, dN, dM, dC;
    SGLD_TRY(dX.upload(xC, 3 * nC));
    SGLD_TRY(dR.upload(RI, 8 * m));
    SGLD_TRY(dCo.upload(coeffi, nGhost));
    SGLD_TRY(dT.upload(transferMat, m * numElesC));
    SGLD_TRY(dN.upload(eNodMatC, 8 * numElesC));
    SGLD_TRY(dM.upload(map, nF));
    const int64_t nIds = numColors ? off[numColors] : 0;
    SGLD_TRY(dC.upload(colors, nIds));
    SGLD_TRY(dG.alloc(3 * std::max<int64_t>(nGhost, 1)));
    SGLD_TRY(dG.zero());
    SGLD_TRY(dY.alloc(3 * nF));
    for (int c = 0; c < numColors; ++c) {
        const int64_t n = off[c + 1] - off[c];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_interp_colour, dim3(sgld::blocks_for(n, 256)), dim3(256), 0, 0, dX.p, nC,
                           dT.p, (int)m, dR.p, dN.p, numElesC, dC.p + off[c], n, dG.p, nGhost);
        SGLD_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_ghost_gather, dim3(sgld::blocks_for(nF, 256)), dim3(256), 0, 0, dG.p, nGhost,
                       dM.p, dCo.p, nF, dY.p);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpy(Y, dY.p, sizeof(double) * 3 * nF, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

static int span_of_children(int64_t nq) {
    if (nq == 8) return 2;
    if (nq == 64) return 4;
    return 0;
}

extern "C" int sgld_assemble_finest(const double *Ke, int64_t keRows, int64_t keCols,
                                    const double *E, int64_t numE, const int32_t *upMap,
                                    int64_t nC, int64_t nq, double *Ks) {
    SGLD_ARG(keRows == 24 && keCols == 24 && Ke,
             "MATLAB:mexFunction:inputNot24x24: Input Ke must be a 24x24 double matrix.");
    const int s = span_of_children(nq);
    SGLD_ARG(s, "MATLAB:mexFunction:inputNotMatrix: Input elementUpwardMap must be an M*64 or M*8 int32 matrix.");
    SGLD_ARG(nC >= 0 && (nC == 0 || (upMap && Ks)), "sgld:arg: null elementUpwardMap or Ks");
    SGLD_ARG(numE >= 0 && (numE == 0 || E), "sgld:arg: null E");
    SGLD_TRY(check_ids0(upMap, nq * nC, numE, "elementUpwardMap"));
    if (nC == 0) return SGLD_OK;
    std::vector<double> H(576 * nq, 0.0);
    for (int q = 0; q < nq; ++q) {
        double p[8][8];
        sgld::child_prolongation(s, q, p);
        sgld::galerkin_child_add(p, Ke, H.data() + 576 * q);
    }
    DevBuf<double> dH, dE, dKs;
    DevBuf<int32_t> dU;
    SGLD_TRY(dH.upload(H.data(), H.size()));
    SGLD_TRY(dE.upload(E, numE));
    SGLD_TRY(dU.upload(upMap, nq * nC));
    SGLD_TRY(dKs.alloc(576 * nC));
    hipLaunchKernelGGL(k_assemble_finest, dim3(sgld::bounded_blocks(576 * nC, 256)), dim3(256), 0, 0,
                       dH.p, dE.p, dU.p, nC, (int)nq, dKs.p);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpy(Ks, dKs.p, sizeof(double) * 576 * nC, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_assemble_nonfinest(const double *KsPrev, int64_t numPrev,
                                       const int32_t *upMap, int64_t nC, int64_t nq,
                                       double *Ks) {
    const int s = span_of_children(nq);
    SGLD_ARG(s, "MATLAB:mexFunction:inputNotMatrix: Input elementUpwardMap must be an M*64 or M*8 int32 matrix.");
    SGLD_ARG(numPrev >= 0 && (numPrev == 0 || KsPrev),
             "MATLAB:mexFunction:KsPreviousDim: KsPrevious must be 24x24xN.");
    SGLD_ARG(nC >= 0 && (nC == 0 || (upMap && Ks)), "sgld:arg: null elementUpwardMap or Ks");
    SGLD_TRY(check_ids0(upMap, nq * nC, numPrev, "elementUpwardMap"));
    if (nC == 0) return SGLD_OK;
    std::vector<double> P(64 * nq);
    for (int q = 0; q < nq; ++q) {
        double p[8][8];
        sgld::child_prolongation(s, q, p);
        std::memcpy(P.data() + 64 * q, p, sizeof(p));
    }
    DevBuf<double> dKp, dP, dKs;
    DevBuf<int32_t> dU;
    SGLD_TRY(dKp.upload(KsPrev, 576 * numPrev));
    SGLD_TRY(dP.upload(P.data(), P.size()));
    SGLD_TRY(dU.upload(upMap, nq * nC));
    SGLD_TRY(dKs.alloc(576 * nC));
    hipLaunchKernelGGL(k_assemble_nonfinest, dim3((unsigned)std::min<int64_t>(nC, 1 << 20)), dim3(576), 0, 0,
                       dKp.p, dU.p, nC,
                       (int)nq, dP.p, dKs.p);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpy(Ks, dKs.p, sizeof(double) * 576 * nC, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

extern "C" int sgld_unit_compliance(const double *U, int64_t numDOFs, const int32_t *eNodMat,
                                    int64_t numElements, const double *Ke, int64_t keRows,
                                    int64_t keCols, double *ce) {
    SGLD_ARG(keRows == 24 && keCols == 24 && Ke,
             "MATLAB:mexFunction:inputNot24x24: Input Ke must be a 24x24 double matrix.");
    SGLD_ARG(numDOFs >= 0 && numDOFs % 3 == 0 && (numDOFs == 0 || U),
             "MATLAB:mexFunction:inputNotVector: Input U must be a column vector.");
    SGLD_ARG(numElements >= 0 && (numElements == 0 || (eNodMat && ce)),
             "MATLAB:mexFunction:inputNotMatrix: Input eNodMat must be an Mx8 int32 matrix.");
    SGLD_TRY(check_ids(eNodMat, 8 * numElements, numDOFs / 3, "eNodMat"));
    if (numElements == 0) return SGLD_OK;
    DevBuf<double> dU, dK, dce;
    DevBuf<int32_t> dN;
    SGLD_TRY(dU.upload(U, numDOFs));
    SGLD_TRY(dK.upload(Ke, 576));
    SGLD_TRY(dN.upload(eNodMat, 8 * numElements));
    SGLD_TRY(dce.alloc(numElements));
    hipLaunchKernelGGL(k_unit_compliance, dim3(sgld::blocks_for(numElements, 256)), dim3(256), 0, 0,
                       dU.p, dN.p, numElements, dK.p, dce.p);
    SGLD_HIP(hipGetLastError());
    SGLD_HIP(hipMemcpy(ce, dce.p, sizeof(double) * numElements, hipMemcpyDeviceToHost));
    return SGLD_OK;
}

// ===================== allocation-free restriction =====================
// sgld_restrict_noalloc: the restriction of sgld_restrict on device arrays,
// for callers that cannot allocate device memory.  Its only scratch -- the
// ghost-node array the reference mxMallocs per call
// (Solving_Restriction_MatrixFree_mex.cpp:72,190), 3 x nGhost doubles, and a
// validation word -- lives in the fixed device array `scratch`, which every
// GPU the library runs on holds until the process ends.  Calls that overlap
// on one GPU share it and race.
namespace {
constexpr size_t kNoallocScratchBytes = (size_t)32 << 20;  // 32 MiB
constexpr size_t kNoallocHeader = 256;                     // validation word, then the ghost array
constexpr int64_t kNoallocMaxGhost = (int64_t)((kNoallocScratchBytes - kNoallocHeader) / (3 * sizeof(double)));
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
unsigned char *g_scratchPtr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy at its own address).
int scratch_address(unsigned char **out) {
    int dev = 0;
    SGLD_HIP(hipGetDevice(&dev));
    SGLD_ARG(dev >= 0 && dev < kMaxDevices, "sgld:arg: device %d beyond the %d this library tracks", dev,
             kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void *p = nullptr;
        SGLD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchPtr[dev] = (unsigned char *)p;
    }
    *out = g_scratchPtr[dev];
    return SGLD_OK;
}

__global__ void k_fill_f64(double *__restrict__ v, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        v[i] = 0.0;
}
__global__ void k_fill_i32(int *__restrict__ v, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = 0;
}
// flag = code if any id lies outside 1..hi (the first failing check wins).
__global__ void k_check_ids(const int32_t *__restrict__ a, int64_t n, int64_t hi, int code, int *flag) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (a[i] < 1 || a[i] > hi) atomicCAS(flag, 0, code);
}

}  // namespace

extern "C" int sgld_restrict_noalloc(const double *rF, int64_t nF, const int32_t *transferMat, int64_t m,
                                     const double *RIT, int64_t ritRows, int64_t ritCols,
                                     const int32_t *eNodMatC, int64_t numElesC, int64_t nC, int64_t nGhost,
                                     const int32_t *map, const double *coeffi, const int32_t *colors,
                                     const int64_t *off, int numColors, double *Y) {
    SGLD_ARG(ritRows == 8 && ritCols == m && RIT,
             "MATLAB:mexFunction:inputNotMx8: Input RIT must be a 8x 125/27 double matrix.");
    SGLD_ARG(m == 27 || m == 125,
             "MATLAB:mexFunction:inputNotMatrix: Input transferMat must be an 27xM or 125xM int32 matrix.");
    SGLD_ARG(nF >= 0 && nC >= 0 && nGhost >= 0 && numElesC >= 0, "sgld:arg: negative size");
    SGLD_ARG(nGhost <= kNoallocMaxGhost,
             "sgld:arg: nGhost = %lld exceeds the %lld ghost nodes of the allocation-free restriction's "
             "32 MiB scratch", (long long)nGhost, (long long)kNoallocMaxGhost);
    SGLD_ARG(numElesC == 0 || (transferMat && eNodMatC),
             "MATLAB:mexFunction:inputNotMatrix: Input eNodMat must be an Mx8 int32 matrix.");
    SGLD_ARG(nF == 0 || (map && coeffi && rF),
             "MATLAB:mexFunction:inputNotVector: Input solidNodeMapCoarser2Finer must be a column vector.");
    SGLD_ARG(nC == 0 || Y, "sgld:arg: Y is null");
    SGLD_ARG(numColors >= 0, "Solving_KbyU:colorArray: colour count must be >= 0");
    if (numColors) {
        SGLD_ARG(off && off[0] == 0, "Solving_KbyU:colorCell: colorOffsets[0] must be 0");
        for (int c = 0; c < numColors; ++c)
            SGLD_ARG(off[c + 1] >= off[c], "Solving_KbyU:colorCell: colorOffsets must not decrease");
        SGLD_ARG(off[numColors] == 0 || colors, "Solving_KbyU:colorCell: colors is null");
    }
    if (nC == 0) return SGLD_OK;
    unsigned char *base = nullptr;
    SGLD_TRY(scratch_address(&base));
    int *flag = (int *)base;
    double *ghost = (double *)(base + kNoallocHeader);
    // what must start at zero: the flag, the ghost array, the output
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, 0, flag, 64);
    if (nGhost)
        hipLaunchKernelGGL(k_fill_f64, dim3(sgld::bounded_blocks(3 * nGhost, 256)), dim3(256), 0, 0, ghost,
                           3 * nGhost);
    hipLaunchKernelGGL(k_fill_f64, dim3(sgld::bounded_blocks(3 * nC, 256)), dim3(256), 0, 0, Y, 3 * nC);
    // every id checked on the device before any kernel indexes with it
    const int64_t nIds = numColors ? off[numColors] : 0;
    struct Chk { const int32_t *a; int64_t n, hi; } chk[4] = {
        {transferMat, m * numElesC, nGhost}, {eNodMatC, 8 * numElesC, nC}, {map, nF, nGhost},
        {colors, nIds, numElesC}};
    for (int k = 0; k < 4; ++k)
        if (chk[k].n > 0)
            hipLaunchKernelGGL(k_check_ids, dim3(sgld::bounded_blocks(chk[k].n, 256)), dim3(256), 0, 0, chk[k].a,
                               chk[k].n, chk[k].hi, k + 1, flag);
    SGLD_HIP(hipGetLastError());
    int bad = 0;
    SGLD_HIP(hipMemcpy(&bad, flag, sizeof(int), hipMemcpyDeviceToHost));
    static const char *const names[4] = {"transferMat", "eNodMat", "solidNodeMapCoarser2Finer", "colors"};
    SGLD_ARG(bad == 0, "sgld:arg: %s holds an id outside 1..%lld", names[bad > 0 ? bad - 1 : 0],
             (long long)(bad > 0 ? chk[bad - 1].hi : 0));
    if (nF) {
        hipLaunchKernelGGL(k_ghost_fill, dim3(sgld::blocks_for(nF, 256)), dim3(256), 0, 0, rF, nF, map, coeffi,
                           nGhost, ghost);
        SGLD_HIP(hipGetLastError());
    }
    for (int c = 0; c < numColors; ++c) {
        const int64_t n = off[c + 1] - off[c];
        if (n == 0) continue;
        hipLaunchKernelGGL(k_restrict_colour, dim3(sgld::blocks_for(n, 256)), dim3(256), 0, 0, ghost, nGhost,
                           transferMat, (int)m, RIT, eNodMatC, numElesC, nC, colors + off[c], n, Y);
        SGLD_HIP(hipGetLastError());
    }
    SGLD_HIP(hipStreamSynchronize(nullptr));
    return SGLD_OK;
}
