// Host-side element mathematics of libsgldfea: element stiffness, trilinear
// transfer tables, per-child prolongations and Galerkin templates, and the
// thread-local error state.  These are setup-time computations (microseconds);
// the device kernels consume their results.
#include <cmath>
#include <cstring>
#include <string>

#include "common.hpp"

namespace sgld {

static thread_local std::string g_err;

void set_error(const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

// Ke = sum_q B_q^T S B_q detJ over the 2x2x2 Gauss points of
// src/FEA_GaussianIntegral.m:19-25 (+-0.577350269189626, weight 1), with the
// shape-function derivatives of src/FEA_DeShapeFunction.m:3-22 scaled by
// invJ = 2/h (src/FEA_VoxelBasedElementStiffnessMatrix.m:10-11), strain rows
// xx yy zz yz zx xy (src/FEA_ElementStrainMatrix.m:11-19) and the isotropic law of
// src/FEA_HookeLaw.m:4-13 with E = 1.
void element_stiffness(double nu, double h, double *Ke) {
    const double g = 0.577350269189626;
    const int sgn[2] = {-1, 1};
    double S[6][6] = {};
    const double c1 = (1 + nu) * (1 - 2 * nu), c2 = 2 * (1 + nu);
    S[0][0] = S[1][1] = S[2][2] = (1 - nu) / c1;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) S[i][j] = nu / c1;
    S[3][3] = S[4][4] = S[5][5] = 1.0 / c2;
    const double detJ = h * h * h / 8.0, invJ = 2.0 / h;
    std::memset(Ke, 0, sizeof(double) * 576);
    for (int q = 0; q < 8; ++q) {
        const double s = sgn[kCornerX[q]] * g, t = sgn[kCornerY[q]] * g, p = sgn[kCornerZ[q]] * g;
        double B[6][24] = {};
        for (int j = 0; j < 8; ++j) {
            const int cs = sgn[kCornerX[j]], ct = sgn[kCornerY[j]], cp = sgn[kCornerZ[j]];
            const double dNs = 0.125 * cs * (1 + ct * t) * (1 + cp * p);
            const double dNt = 0.125 * ct * (1 + cs * s) * (1 + cp * p);
            const double dNp = 0.125 * cp * (1 + cs * s) * (1 + ct * t);
            const double dx = invJ * dNs, dy = invJ * dNt, dz = invJ * dNp;
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
                DB[i][k] = acc * detJ;
            }
        for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b) {
                double acc = 0;
                for (int i = 0; i < 6; ++i) acc += B[i][a] * DB[i][b];
                Ke[a + 24 * b] += acc;
            }
    }
}

// Nodal-stress operator of one element for E = 1 (FEA_StressAnalysis.m:17-46):
// T (48 x 24, row 6 n + k: stress component k at element-local node n) =
// (N^-1 x I6) blockdiag(D) [B_1; ...; B_8], N the shape functions at the 8
// Gauss points (OuterInterpolationMat, :68-93; Gauss points in node order).
void stress_operator(double nu, double h, double *T) {
    const double g = 0.577350269189626;
    const int sgn[2] = {-1, 1};
    double D[6][6] = {};
    const double c1 = (1 + nu) * (1 - 2 * nu), c2 = 2 * (1 + nu);
    D[0][0] = D[1][1] = D[2][2] = (1 - nu) / c1;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            if (i != j) D[i][j] = nu / c1;
    D[3][3] = D[4][4] = D[5][5] = 1.0 / c2;
    const double invJ = 2.0 / h;
    double DB[8][6][24] = {};
    double M[8][16];  // [N | I] for Gauss-Jordan inversion
    for (int q = 0; q < 8; ++q) {
        const double s = sgn[kCornerX[q]] * g, t = sgn[kCornerY[q]] * g, p = sgn[kCornerZ[q]] * g;
        double B[6][24] = {};
        for (int j = 0; j < 8; ++j) {
            const int cs = sgn[kCornerX[j]], ct = sgn[kCornerY[j]], cp = sgn[kCornerZ[j]];
            M[q][j] = 0.125 * (1 + cs * s) * (1 + ct * t) * (1 + cp * p);
            M[q][8 + j] = q == j ? 1.0 : 0.0;
            const double dx = invJ * 0.125 * cs * (1 + ct * t) * (1 + cp * p);
            const double dy = invJ * 0.125 * ct * (1 + cs * s) * (1 + cp * p);
            const double dz = invJ * 0.125 * cp * (1 + cs * s) * (1 + ct * t);
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
                for (int l = 0; l < 6; ++l) acc += D[i][l] * B[l][k];
                DB[q][i][k] = acc;
            }
    }
    for (int c = 0; c < 8; ++c) {  // N is well conditioned (entries (1 +- g)^3 / 8): no pivoting needed
        const double d = M[c][c];
        for (int j = 0; j < 16; ++j) M[c][j] /= d;
        for (int r = 0; r < 8; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
            for (int j = 0; j < 16; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int n = 0; n < 8; ++n)
        for (int k = 0; k < 6; ++k)
            for (int m = 0; m < 24; ++m) {
                double acc = 0;
                for (int q = 0; q < 8; ++q) acc += M[n][8 + q] * DB[q][k][m];
                T[(6 * n + k) * 24 + m] = acc;
            }
}

// Trilinear weights of block node n = rr + (s+1)cc + (s+1)^2 ll (row rr from
// the top, so y = (s - rr)/s) on coarse corner J: the 'inNODE' tables of
// src/Solving_Operator4MultiGridRestrictionAndInterpolation.m:4-31 (span 2)
// and :118-243 (span 4).  All values are k/64, exact in every precision used.
void transfer_table(int s, double *RI) {
    const int m = (s + 1) * (s + 1) * (s + 1);
    for (int ll = 0; ll <= s; ++ll)
        for (int cc = 0; cc <= s; ++cc)
            for (int rr = 0; rr <= s; ++rr) {
                const int n = rr + (s + 1) * cc + (s + 1) * (s + 1) * ll;
                const double tx = (double)cc / s, ty = (double)(s - rr) / s, tz = (double)ll / s;
                for (int J = 0; J < 8; ++J) {
                    const double wx = kCornerX[J] ? tx : 1 - tx;
                    const double wy = kCornerY[J] ? ty : 1 - ty;
                    const double wz = kCornerZ[J] ? tz : 1 - tz;
                    RI[n + (int64_t)m * J] = wx * wy * wz;
                }
            }
}

// p[j][J]: weight of parent corner J at local node j of child q = qr + s qc +
// s^2 ql of a span-s block.  Child local node j sits at block node
// (qr + 1 - y_j, qc + x_j, ql + z_j) (src/Solving_SubEleNodMat.m:1-80).
void child_prolongation(int s, int q, double p[8][8]) {
    const int m = (s + 1) * (s + 1) * (s + 1);
    std::vector<double> RI((size_t)m * 8);
    transfer_table(s, RI.data());
    const int qr = q % s, qc = (q / s) % s, ql = q / (s * s);
    for (int j = 0; j < 8; ++j) {
        const int rr = qr + 1 - kCornerY[j], cc = qc + kCornerX[j], ll = ql + kCornerZ[j];
        const int n = rr + (s + 1) * cc + (s + 1) * (s + 1) * ll;
        for (int J = 0; J < 8; ++J) p[j][J] = RI[n + (int64_t)m * J];
    }
}

// out += P_q^T K P_q with P_q = p (x) I3 (interleaved DOFs), K 24x24
// col-major; the accumulation order matches the device Galerkin kernel.
void galerkin_child_add(const double p[8][8], const double *K, double *out) {
    double T[576];
    for (int J = 0; J < 8; ++J)
        for (int a = 0; a < 3; ++a) {
            double *Tc = T + 24 * (3 * J + a);
            for (int r = 0; r < 24; ++r) Tc[r] = 0.0;
            for (int j = 0; j < 8; ++j) {
                const double w = p[j][J];
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
                    const double w = p[i][I];
                    if (w != 0.0) acc += w * T[3 * i + a + 24 * c];
                }
                out[3 * I + a + 24 * c] += acc;
            }
}

}  // namespace sgld

extern "C" const char *sgld_last_error(void) { return sgld::g_err.c_str(); }

extern "C" const char *sgld_version(void) { return "libsgldfea 0.1 (gfx950)"; }

extern "C" int sgld_element_stiffness(double nu, double h, double *Ke) {
    SGLD_ARG(Ke, "sgld:arg: Ke is null");
    SGLD_ARG(h > 0 && nu > -1.0 && nu < 0.5, "sgld:arg: need cellSize > 0 and -1 < nu < 0.5");
    sgld::element_stiffness(nu, h, Ke);
    return SGLD_OK;
}
