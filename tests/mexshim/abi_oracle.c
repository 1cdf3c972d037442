/*
 * The C ABI of include/sgldfea.h that the MEX gateways call, implemented over
 * the CPU oracle (oracle/sgld_oracle.c).  Test infrastructure only: linking a
 * gateway against this file instead of libsgldfea.so exercises the gateway's
 * marshalling on a machine without a GPU, and its results can be compared
 * with the oracle called directly.  Validation mirrors the library's host
 * checks that the tests reach (shapes, id ranges) with the same messages.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/sgldfea.h"
#include "../../oracle/sgld_oracle.h"

static char g_err[512];
const char *sgld_last_error(void) { return g_err; }
const char *sgld_version(void) { return "abi_oracle (CPU test stub)"; }

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

static int ids_ok(const int32_t *a, int64_t n, int64_t hi, const char *what) {
    for (int64_t i = 0; i < n; ++i)
        if (a[i] < 1 || a[i] > hi)
            return fail(SGLD_ERR_ARG, "sgld:arg: %s[%lld] = %d outside 1..%lld", what, (long long)i, a[i],
                        (long long)hi);
    return SGLD_OK;
}

static int colours_ok(const int32_t *c, const int64_t *off, int n, int64_t ne) {
    if (n != 8) return fail(SGLD_ERR_ARG, "sgld:arg: the oracle stub needs 8 colours, got %d", n);
    return ids_ok(c, off[8], ne, "colors");
}

int sgld_kbyu(const double *U, int64_t numDOFs, const int32_t *eNodMat, int64_t numElements, const double *Ke,
              int64_t keRows, int64_t keCols, const double *E, int64_t numE, const int32_t *colors,
              const int64_t *off, int numColors, double *Y) {
    if (keRows != 24 || keCols != 24)
        return fail(SGLD_ERR_ARG, "MATLAB:mexFunction:inputNot24x24: Input Ke must be a 24x24 double matrix.");
    if (numE != numElements)
        return fail(SGLD_ERR_ARG, "MATLAB:mexFunction:inputNotVector: Input E must be a column vector (one per element).");
    int rc = ids_ok(eNodMat, 8 * numElements, numDOFs / 3, "eNodMat");
    if (rc) return rc;
    if ((rc = colours_ok(colors, off, numColors, numElements))) return rc;
    or_kbyu_mex(U, numDOFs, eNodMat, numElements, Ke, E, colors, off, Y);
    return SGLD_OK;
}

int sgld_restrict(const double *rF, int64_t nF, const int32_t *tm, int64_t m, const double *RIT, int64_t rr,
                  int64_t rc_, const int32_t *eN, int64_t ne, int64_t nC, int64_t nG, const int32_t *map,
                  const double *co, const int32_t *colors, const int64_t *off, int numColors, double *Y) {
    int rc;
    if (rr != 8 || rc_ != m) return fail(SGLD_ERR_ARG, "MATLAB:mexFunction:inputNotMx8: bad RIT");
    if ((rc = ids_ok(tm, m * ne, nG, "transferMat")) || (rc = ids_ok(eN, 8 * ne, nC, "eNodMat")) ||
        (rc = ids_ok(map, nF, nG, "solidNodeMapCoarser2Finer")) || (rc = colours_ok(colors, off, numColors, ne)))
        return rc;
    or_restrict_mex(rF, nF, tm, (int)m, RIT, eN, ne, nC, nG, map, co, colors, off, Y);
    return SGLD_OK;
}

int sgld_interpolate(const double *xC, int64_t nC, const int32_t *tm, int64_t m, const double *RI, int64_t rr,
                     int64_t rc_, const int32_t *eN, int64_t ne, int64_t nF, int64_t nG, const int32_t *map,
                     const double *co, const int32_t *colors, const int64_t *off, int numColors, double *Y) {
    int rc;
    if (rr != m || rc_ != 8) return fail(SGLD_ERR_ARG, "MATLAB:mexFunction:inputNotMx8: bad RI");
    if ((rc = ids_ok(tm, m * ne, nG, "transferMat")) || (rc = ids_ok(eN, 8 * ne, nC, "eNodMat")) ||
        (rc = ids_ok(map, nF, nG, "solidNodeMapCoarser2Finer")) || (rc = colours_ok(colors, off, numColors, ne)))
        return rc;
    or_interpolate_mex(xC, nC, tm, (int)m, RI, eN, ne, nF, nG, map, co, colors, off, Y);
    return SGLD_OK;
}

static int span_of(int64_t nq) { return nq == 8 ? 2 : nq == 64 ? 4 : 0; }

int sgld_assemble_finest(const double *Ke, int64_t kr, int64_t kc, const double *E, int64_t numE,
                         const int32_t *up, int64_t nC, int64_t nq, double *Ks) {
    if (kr != 24 || kc != 24 || !span_of(nq)) return fail(SGLD_ERR_ARG, "sgld:arg: bad Ke or children count");
    for (int64_t i = 0; i < nq * nC; ++i)
        if (up[i] < 0 || up[i] > numE) return fail(SGLD_ERR_ARG, "sgld:arg: elementUpwardMap id out of range");
    or_assemble_finest(Ke, E, up, nC, span_of(nq), Ks);
    return SGLD_OK;
}

int sgld_assemble_nonfinest(const double *KsPrev, int64_t numPrev, const int32_t *up, int64_t nC, int64_t nq,
                            double *Ks) {
    if (!span_of(nq)) return fail(SGLD_ERR_ARG, "sgld:arg: bad children count");
    for (int64_t i = 0; i < nq * nC; ++i)
        if (up[i] < 0 || up[i] > numPrev) return fail(SGLD_ERR_ARG, "sgld:arg: elementUpwardMap id out of range");
    or_assemble_nonfinest(KsPrev, up, nC, span_of(nq), Ks);
    return SGLD_OK;
}

int sgld_unit_compliance(const double *U, int64_t numDOFs, const int32_t *eN, int64_t ne, const double *Ke,
                         int64_t kr, int64_t kc, double *ce) {
    if (kr != 24 || kc != 24) return fail(SGLD_ERR_ARG, "MATLAB:mexFunction:inputNot24x24: bad Ke");
    const int rc = ids_ok(eN, 8 * ne, numDOFs / 3, "eNodMat");
    if (rc) return rc;
    or_unit_compliance_mex(U, eN, ne, Ke, ce);
    return SGLD_OK;
}

struct sgld_ctx {
    or_mesh *m;
    double omega;
};

int sgld_ctx_create(sgld_ctx **out, int rx, int ry, int rz, const uint8_t *vox, int L, int nd, double nu,
                    double h) {
    const size_t n = (size_t)rx * ry * rz;
    uint8_t *v = (uint8_t *)malloc(n);
    for (size_t i = 0; i < n; ++i) v[i] = vox ? (vox[i] != 0) : 1;
    or_mesh *m = or_mesh_create(rx, ry, rz, v, L, nd, nu, h);
    free(v);
    if (!m) return fail(SGLD_ERR_ARG, "sgld:arg: resolution not divisible by the level spans");
    *out = (sgld_ctx *)calloc(1, sizeof(sgld_ctx));
    (*out)->m = m;
    (*out)->omega = 0.6;
    return SGLD_OK;
}
int sgld_ctx_destroy(sgld_ctx *C) {
    if (C) {
        or_mesh_destroy(C->m);
        free(C);
    }
    return SGLD_OK;
}
int sgld_ctx_info(const sgld_ctx *C, int level, int64_t *out) {
    if (level == 0) return or_mesh_num_levels(C->m);
    or_mesh_level_info(C->m, level, out);
    return SGLD_OK;
}
int sgld_ctx_set_fixed(sgld_ctx *C, const uint8_t *f) { return or_mesh_set_fixed(C->m, f); }
int sgld_ctx_set_modulus(sgld_ctx *C, const double *E) { return or_mesh_set_modulus(C->m, E); }
int sgld_ctx_assemble(sgld_ctx *C, double omega) {
    C->omega = omega;
    return or_mesh_assemble(C->m) ? fail(SGLD_ERR_NOT_SPD, "sgld:notspd: coarsest matrix not SPD") : SGLD_OK;
}
static int64_t ndof(const sgld_ctx *C) {
    int64_t o[8];
    or_mesh_level_info(C->m, 1, o);
    return o[6];
}
int sgld_ctx_pcg(sgld_ctx *C, const double *F, double *U, double tol, int maxIT, int precision, int *its,
                 double *relres) {
    (void)precision;
    const int rc = or_mesh_pcg(C->m, C->omega, F, U, tol, maxIT, NULL, its);
    const int64_t n = ndof(C);
    double *r = (double *)malloc(sizeof(double) * n);
    or_mesh_kbyu(C->m, U, r);
    double rr = 0, bb = 0;
    for (int64_t i = 0; i < n; ++i) {
        rr += (F[i] - r[i]) * (F[i] - r[i]);
        bb += F[i] * F[i];
    }
    free(r);
    *relres = bb > 0 ? sqrt(rr / bb) : 0.0;
    return rc ? fail(SGLD_ERR_MAXIT, "sgld:maxit: PCG stopped at maxIT = %d", maxIT) : SGLD_OK;
}
int sgld_ctx_kbyu(sgld_ctx *C, const double *u, double *y) {
    or_mesh_kbyu(C->m, u, y);
    return SGLD_OK;
}
int sgld_ctx_vcycle(sgld_ctx *C, const double *r, double *z) {
    or_mesh_vcycle(C->m, C->omega, r, z);
    return SGLD_OK;
}
int sgld_ctx_unit_compliance(sgld_ctx *C, const double *U, double *ce) {
    or_mesh_unit_compliance(C->m, U, ce);
    return SGLD_OK;
}
