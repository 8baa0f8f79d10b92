/*
 * ORACLE -- test infrastructure, not product code.
 *
 * A CPU restatement, in C with OpenMP, of the reference's Nedelec-2 tetrahedral
 * curl-curl assembly and its COO->CSC map. Only the test suite, the smoke check and
 * the CPU-baseline leg of the benchmark load this library, as the checker; the
 * product path never calls it.
 *
 * Followed step by step (paths relative to src/emerge/_emerge/ in the reference):
 *   or_ref_tet_coefficients, or_ref_matinv, or_ref_coeff, or_ref_eval_point
 *                         (ned2_strict.c): the steps Numba compiles without fastmath
 *                         (tet_coefficients curlcurl.py:125-213, matinv
 *                         mth/optimized.py:315-357, the ccbf.py 3D basis functions and
 *                         curls of codes 64/66/128/129), in the reference's own
 *                         evaluation order, compiled without fast-math or contraction
 *   or_local_map          mth/optimized.py:262-296, physics/.../curlcurl.py:72-85
 *   or_parse_dofcode      compiled/ccbf.py:40-66
 *   or_tet_stiff_mass     physics/microwave/assembly/curlcurl.py:285-380 (11-point
 *                         Keast rule of :277-283, complex arithmetic, i<=j then mirror;
 *                         the material products, weights and sums are fastmath in the
 *                         reference and compiled here with -ffast-math)
 *   ned2_oracle_build     curlcurl.py:387-456 (_matrix_builder; COO block per tet,
 *                         row-major, p = iassy*ndof^2)
 *   ned2_oracle_csc_*     mth/csc_cast.py:65-177 (precompute_csc_pattern, scatter)
 *   ned2_oracle_interp    compiled/base/interp.py:286-387, 410-519 (field and curl;
 *                         basis written from the definitions in elements/dofsets.py:66-83)
 *   ned2_oracle_locate    compiled/base/index_inter.py:134-246 (brute force)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double _Complex cplx;

/* ned2_strict.c: the quadrature table and the steps the reference evaluates strictly */
extern const double OR_DPTS[5][11];
#define DPTS OR_DPTS
void or_ref_tet_coefficients(const double x[4], const double y[4], const double z[4],
                             double a[4], double b[4], double c[4], double d[4], double *V);
void or_ref_matinv(const cplx M[9], cplx out[9]);
void or_ref_coeff(const double a[4], const double b[4], const double c[4], const double d[4],
                  double V, double coeff[4][4]);
int or_ref_eval_point(const double coeff[4][4], const double vx[3][4], int q,
                      const int64_t *codes, const int (*ijk)[3], int ndof, double (*F)[3],
                      double (*CF)[3]);
#define NQ 11

enum { OR_OK = 0, OR_EBADCODE = -2, OR_EDEGEN = -5 };

/* Local vertex index of every global id of a small table column (optimized.py:262). */
static void or_local_map(const int64_t vid[4], const int64_t *gids, int nrow, int64_t *out) {
    for (int j = 0; j < nrow; ++j) {
        out[j] = 0;
        for (int k = 0; k < 4; ++k)
            if (vid[k] == gids[j]) { out[j] = k; break; }
    }
}

static inline void cross3(const double u[3], const double v[3], double w[3]) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

/* Basis function and curl at one point. l[4] barycentric values, g[4][3] gradients
 * (both carry the |V| sign of the coefficients). */
static int or_basis(int code, int i, int j, int k, const double l[4], const double g[4][3],
                    double f[3], double cf[3]) {
    double gij[3], w[3], t[3];
    switch (code) {
    case 64: /* L_i grad L_j - L_j grad L_i ; curl = 2 grad L_i x grad L_j */
        cross3(g[i], g[j], gij);
        for (int q = 0; q < 3; ++q) {
            f[q] = l[i] * g[j][q] - l[j] * g[i][q];
            cf[q] = 2.0 * gij[q];
        }
        return OR_OK;
    case 66: /* (L_i - L_j)(L_i grad L_j - L_j grad L_i) ; curl = 3 (L_i - L_j) gi x gj */
        cross3(g[i], g[j], gij);
        for (int q = 0; q < 3; ++q) {
            f[q] = (l[i] - l[j]) * (l[i] * g[j][q] - l[j] * g[i][q]);
            cf[q] = 3.0 * (l[i] - l[j]) * gij[q];
        }
        return OR_OK;
    case 128: /* L_i (L_j grad L_k - L_k grad L_j) */
        for (int q = 0; q < 3; ++q) w[q] = l[j] * g[k][q] - l[k] * g[j][q];
        cross3(g[i], w, t);
        cross3(g[j], g[k], gij);
        for (int q = 0; q < 3; ++q) {
            f[q] = l[i] * w[q];
            cf[q] = t[q] + 2.0 * l[i] * gij[q];
        }
        return OR_OK;
    case 129: /* L_j (L_i grad L_k - L_k grad L_i) */
        for (int q = 0; q < 3; ++q) w[q] = l[i] * g[k][q] - l[k] * g[i][q];
        cross3(g[j], w, t);
        cross3(g[i], g[k], gij);
        for (int q = 0; q < 3; ++q) {
            f[q] = l[j] * w[q];
            cf[q] = t[q] + 2.0 * l[j] * gij[q];
        }
        return OR_OK;
    default:
        return OR_EBADCODE;
    }
}

/* compiled/ccbf.py:40-66: type (0 edge, 1 face) and occurrence index per code. */
static void or_parse_dofcode(const int64_t *codes, int ndof, int64_t *typ, int64_t *idx) {
    int64_t ne[64] = {0}, nf[64] = {0};
    for (int i = 0; i < ndof; ++i) {
        int64_t c = codes[i] & 63;
        if ((codes[i] & 192) == 64) {
            typ[i] = 0;
            idx[i] = ne[c]++;
        } else {
            typ[i] = 1;
            idx[i] = nf[c]++;
        }
    }
}

#define MAXDOF 20

/* curlcurl.py:294-380 for one tet. vx[3][4] vertex coords in stored order. */
static int or_tet_stiff_mass(const double vx[3][4], const int64_t lem[2][6],
                             const int64_t ltm[3][4], const cplx Ms[9], const cplx Mm[9],
                             const int64_t *codes, const int64_t *typ, const int64_t *idx,
                             int ndof, cplx *Kout, cplx *Mout) {
    double a[4], b[4], c[4], d[4], V;
    or_ref_tet_coefficients(vx[0], vx[1], vx[2], a, b, c, d, &V);
    if (!(V > 0.0)) return OR_EDEGEN;
    double co[4][4]; /* co[0]=a/(6V), co[1]=b/(6V) ... (curlcurl.py:307-311) */
    or_ref_coeff(a, b, c, d, V, co);
    static __thread cplx F[MAXDOF][3][NQ], CF[MAXDOF][3][NQ], MsCF[MAXDOF][3][NQ],
        MmF[MAXDOF][3][NQ];
    int ijk[MAXDOF][3];
    for (int i = 0; i < ndof; ++i) {
        if (typ[i] == 0) {
            ijk[i][0] = (int)lem[0][idx[i]];
            ijk[i][1] = (int)lem[1][idx[i]];
            ijk[i][2] = 0;
        } else {
            ijk[i][0] = (int)ltm[0][idx[i]];
            ijk[i][1] = (int)ltm[1][idx[i]];
            ijk[i][2] = (int)ltm[2][idx[i]];
        }
    }
    for (int q = 0; q < NQ; ++q) {
        double f[MAXDOF][3], cf[MAXDOF][3];
        int st = or_ref_eval_point(co, vx, q, codes, (const int(*)[3])ijk, ndof, f, cf);
        if (st) return OR_EBADCODE;
        for (int i = 0; i < ndof; ++i)
            for (int r = 0; r < 3; ++r) {
                F[i][r][q] = f[i][r];
                CF[i][r][q] = cf[i][r];
            }
    }
    for (int i = 0; i < ndof; ++i) {
        for (int q = 0; q < NQ; ++q) {
            for (int r = 0; r < 3; ++r) {
                MsCF[i][r][q] = Ms[3 * r + 0] * CF[i][0][q] + Ms[3 * r + 1] * CF[i][1][q] +
                                Ms[3 * r + 2] * CF[i][2][q];
                MmF[i][r][q] = Mm[3 * r + 0] * F[i][0][q] + Mm[3 * r + 1] * F[i][1][q] +
                               Mm[3 * r + 2] * F[i][2][q];
            }
            for (int r = 0; r < 3; ++r) {
                MsCF[i][r][q] *= DPTS[0][q];
                MmF[i][r][q] *= DPTS[0][q];
            }
        }
    }
    for (int i = 0; i < ndof; ++i) {
        for (int j = i; j < ndof; ++j) {
            cplx aS = 0, aM = 0;
            for (int q = 0; q < NQ; ++q) {
                aS += MsCF[i][0][q] * CF[j][0][q] + MsCF[i][1][q] * CF[j][1][q] +
                      MsCF[i][2][q] * CF[j][2][q];
                aM += MmF[i][0][q] * F[j][0][q] + MmF[i][1][q] * F[j][1][q] +
                      MmF[i][2][q] * F[j][2][q];
            }
            Kout[i * ndof + j] = aS * V;
            Mout[i * ndof + j] = aM * V;
            Kout[j * ndof + i] = aS * V;
            Mout[j * ndof + i] = aM * V;
        }
    }
    return OR_OK;
}

/* _matrix_builder (curlcurl.py:387-456). Layouts follow the reference arrays:
 * nodes[c*Nn+n], tets[v*T+t], tris[v*F+f], edges[v*E+e], tet_to_field[d*T+t],
 * tet_to_edge[s*T+t], tet_to_tri[s*T+t], ur/er interleaved complex at [(p*3+q)*T+t].
 * Outputs dataE/dataB (interleaved complex) and rows/cols have Ta*ndof^2 entries.
 * local_frame = 0 evaluates the cofactors in physical coordinates exactly as the
 * reference does; local_frame = 1 first translates each tet so that its stored vertex
 * 0 is the origin. The element integrals are translation invariant, so both restate
 * the same algorithm; the translated form avoids the cancellation of the absolute-
 * coordinate cofactor sums on thin (sliver) tets.
 * Returns 0, or a negative status (bad DOF code, degenerate tet). */
int ned2_oracle_build(const double *nodes, int64_t Nn, const int64_t *tets, int64_t T,
                      const int64_t *tris, int64_t F, const int64_t *edges, int64_t E,
                      const int64_t *tet_to_field, const int64_t *tet_to_edge,
                      const int64_t *tet_to_tri, const double *ur, const double *er,
                      const int64_t *tetids, int64_t Ta, const int64_t *dofcodes, int ndof,
                      double *dataE, double *dataB, int64_t *rows, int64_t *cols,
                      int nthreads, int local_frame) {
    if (ndof > MAXDOF || ndof <= 0) return OR_EBADCODE;
    int64_t typ[MAXDOF], idx[MAXDOF];
    or_parse_dofcode(dofcodes, ndof, typ, idx);
    for (int i = 0; i < ndof; ++i)
        if (dofcodes[i] != 64 && dofcodes[i] != 66 && dofcodes[i] != 128 && dofcodes[i] != 129)
            return OR_EBADCODE;
    int status = OR_OK;
    const int64_t nd2 = (int64_t)ndof * ndof;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
    for (int64_t ia = 0; ia < Ta; ++ia) {
        const int64_t t = tetids ? tetids[ia] : ia;
        const int64_t p = ia * nd2;
        int64_t vid[4];
        double vx[3][4];
        for (int v = 0; v < 4; ++v) {
            vid[v] = tets[v * T + t];
            for (int cc = 0; cc < 3; ++cc) vx[cc][v] = nodes[cc * Nn + vid[v]];
        }
        if (local_frame)
            for (int cc = 0; cc < 3; ++cc)
                for (int v = 3; v >= 0; --v) vx[cc][v] -= vx[cc][0];
        /* local_tet_to_triid / local_tet_to_edgeid (curlcurl.py:72-85) */
        int64_t ltm[3][4], lem[2][6];
        for (int s = 0; s < 4; ++s) {
            int64_t fid = tet_to_tri[s * T + t], g3[3], o3[3];
            for (int r = 0; r < 3; ++r) g3[r] = tris[r * F + fid];
            or_local_map(vid, g3, 3, o3);
            for (int r = 0; r < 3; ++r) ltm[r][s] = o3[r];
        }
        for (int s = 0; s < 6; ++s) {
            int64_t eid = tet_to_edge[s * T + t], g2[2], o2[2];
            for (int r = 0; r < 2; ++r) g2[r] = edges[r * E + eid];
            or_local_map(vid, g2, 2, o2);
            for (int r = 0; r < 2; ++r) lem[r][s] = o2[r];
        }
        cplx urt[9], ert[9], Ms[9];
        for (int k = 0; k < 9; ++k) {
            urt[k] = ur[2 * (k * T + t)] + I * ur[2 * (k * T + t) + 1];
            ert[k] = er[2 * (k * T + t)] + I * er[2 * (k * T + t) + 1];
        }
        or_ref_matinv(urt, Ms);
        cplx Ke[MAXDOF * MAXDOF], Me[MAXDOF * MAXDOF];
        int st = or_tet_stiff_mass(vx, lem, ltm, Ms, ert, dofcodes, typ, idx, ndof, Ke, Me);
        if (st) {
#pragma omp atomic write
            status = st;
            continue;
        }
        for (int ii = 0; ii < ndof; ++ii)
            for (int jj = 0; jj < ndof; ++jj) {
                int64_t e = p + ii * ndof + jj;
                rows[e] = tet_to_field[ii * T + t];
                cols[e] = tet_to_field[jj * T + t];
            }
        memcpy(dataE + 2 * p, Ke, sizeof(cplx) * nd2);
        memcpy(dataB + 2 * p, Me, sizeof(cplx) * nd2);
    }
    return status;
}

/* ---------------- COO -> CSC (mth/csc_cast.py:65-177) ---------------- */

static const int64_t *g_sort_rows;
static int cmp_row_then_k(const void *pa, const void *pb) {
    int64_t ka = *(const int64_t *)pa, kb = *(const int64_t *)pb;
    int64_t ra = g_sort_rows[ka], rb = g_sort_rows[kb];
    if (ra != rb) return ra < rb ? -1 : 1;
    return ka < kb ? -1 : (ka > kb);
}

/* Phase 1 (sequential count/group) and phase 3 (indptr) of precompute_csc_pattern.
 * Returns nnz; fills col_offsets[N+1], perm[nnz_coo] (sorted by column, then row,
 * then COO index -- the reference's per-column sort is unstable, only the order of
 * equal rows differs), unique counts into indptr[N+1]. */
int64_t ned2_oracle_csc_pattern(const int64_t *rows, const int64_t *cols, int64_t nnz_coo,
                                int64_t N, int64_t *col_offsets, int64_t *perm,
                                int64_t *indptr) {
    int64_t *cnt = calloc((size_t)N, sizeof(int64_t));
    for (int64_t k = 0; k < nnz_coo; ++k) cnt[cols[k]]++;
    col_offsets[0] = 0;
    for (int64_t i = 0; i < N; ++i) col_offsets[i + 1] = col_offsets[i] + cnt[i];
    int64_t *cur = malloc(sizeof(int64_t) * (size_t)N);
    memcpy(cur, col_offsets, sizeof(int64_t) * (size_t)N);
    for (int64_t k = 0; k < nnz_coo; ++k) perm[cur[cols[k]]++] = k;
    g_sort_rows = rows;
    indptr[0] = 0;
    for (int64_t i = 0; i < N; ++i) {
        int64_t s = col_offsets[i], n = col_offsets[i + 1] - s;
        if (n > 1) qsort(perm + s, (size_t)n, sizeof(int64_t), cmp_row_then_k);
        int64_t u = n > 0 ? 1 : 0;
        for (int64_t j = 1; j < n; ++j) u += rows[perm[s + j]] != rows[perm[s + j - 1]];
        indptr[i + 1] = indptr[i] + u;
    }
    free(cnt);
    free(cur);
    return indptr[N];
}

/* Phase 4: indices[nnz] and target_csc[nnz_coo] (indexed by sorted position). */
void ned2_oracle_csc_fill(const int64_t *rows, int64_t N, const int64_t *col_offsets,
                          const int64_t *perm, const int64_t *indptr, int64_t *indices,
                          int64_t *target) {
    for (int64_t i = 0; i < N; ++i) {
        int64_t s = col_offsets[i], n = col_offsets[i + 1] - s;
        if (n == 0) continue;
        int64_t pos = indptr[i], prev = rows[perm[s]];
        indices[pos] = prev;
        target[s] = pos;
        for (int64_t j = 1; j < n; ++j) {
            int64_t r = rows[perm[s + j]];
            if (r != prev) {
                indices[++pos] = r;
                prev = r;
            }
            target[s + j] = pos;
        }
    }
}

/* scatter_to_csc_c16 (csc_cast.py:162-177): gather then per-column sum. */
void ned2_oracle_csc_scatter(const double *data_coo, const int64_t *perm, const int64_t *target,
                             const int64_t *col_offsets, int64_t N, double *data_csc,
                             int64_t nnz) {
    memset(data_csc, 0, sizeof(double) * 2 * (size_t)nnz);
    for (int64_t i = 0; i < N; ++i)
        for (int64_t k = col_offsets[i]; k < col_offsets[i + 1]; ++k) {
            data_csc[2 * target[k]] += data_coo[2 * perm[k]];
            data_csc[2 * target[k] + 1] += data_coo[2 * perm[k] + 1];
        }
}

/* Copy of the quadrature table (weights row, then 4 barycentric rows), for tests. */
void ned2_oracle_keast(double *out) { memcpy(out, DPTS, sizeof(DPTS)); }

/* compiled/base/interp.py:286-387 (ned2_tet_interp) and :410-519 (ned2_tet_interp_curl)
 * for points whose tet is known (tet_of[j], -1: NaN). Per point: tet_coefficients in
 * physical coordinates (or translated to stored vertex 0 when local_frame), coefficients
 * over 6V, local_mapping of the slot edges/faces, then sum_i sol[tet_to_field[i]] f_i
 * (curl: times c_tet[t] when given). coords [3][n]; out interleaved complex [3][n]. */
int ned2_oracle_interp(const double *nodes, int64_t Nn, const int64_t *tets, int64_t T,
                       const int64_t *tris, int64_t F, const int64_t *edges, int64_t E,
                       const int64_t *tet_to_field, const int64_t *tet_to_edge,
                       const int64_t *tet_to_tri, const int64_t *dofcodes, int ndof,
                       const double *coords, int64_t n, const int64_t *tet_of,
                       const cplx *sol, int curl, const cplx *c_tet, int local_frame,
                       cplx *out) {
    int64_t typ[MAXDOF], idx[MAXDOF];
    or_parse_dofcode(dofcodes, ndof, typ, idx);
    int status = OR_OK;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        const int64_t t = tet_of[j];
        if (t < 0) {
            for (int q = 0; q < 3; ++q) out[q * n + j] = NAN + NAN * I;
            continue;
        }
        int64_t vid[4];
        double vx[3][4], o[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k) vid[k] = tets[k * T + t];
        if (local_frame)
            for (int c = 0; c < 3; ++c) o[c] = nodes[c * Nn + vid[0]];
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) vx[c][k] = nodes[c * Nn + vid[k]] - o[c];
        int64_t lem[2][6], ltm[3][4];
        for (int e = 0; e < 6; ++e) {
            const int64_t ge = tet_to_edge[e * T + t];
            int64_t g2[2] = {edges[ge], edges[E + ge]};
            int64_t l2[2];
            or_local_map(vid, g2, 2, l2);
            lem[0][e] = l2[0];
            lem[1][e] = l2[1];
        }
        for (int f = 0; f < 4; ++f) {
            const int64_t gf = tet_to_tri[f * T + t];
            int64_t g3[3] = {tris[gf], tris[F + gf], tris[2 * F + gf]};
            int64_t l3[3];
            or_local_map(vid, g3, 3, l3);
            for (int r = 0; r < 3; ++r) ltm[r][f] = l3[r];
        }
        double a[4], b[4], c[4], d[4], V;
        or_ref_tet_coefficients(vx[0], vx[1], vx[2], a, b, c, d, &V);
        double g[4][3], l[4];
        const double px = coords[j] - o[0], py = coords[n + j] - o[1], pz = coords[2 * n + j] - o[2];
        for (int k = 0; k < 4; ++k) {
            g[k][0] = b[k] / (6.0 * V);
            g[k][1] = c[k] / (6.0 * V);
            g[k][2] = d[k] / (6.0 * V);
            l[k] = a[k] / (6.0 * V) + g[k][0] * px + g[k][1] * py + g[k][2] * pz;
        }
        cplx acc[3] = {0, 0, 0};
        for (int i = 0; i < ndof; ++i) {
            int i1, j1, k1;
            if (typ[i] == 0) {
                i1 = (int)lem[0][idx[i]];
                j1 = (int)lem[1][idx[i]];
                k1 = 0;
            } else {
                i1 = (int)ltm[0][idx[i]];
                j1 = (int)ltm[1][idx[i]];
                k1 = (int)ltm[2][idx[i]];
            }
            double f[3], cf[3];
            if (or_basis((int)dofcodes[i], i1, j1, k1, l, g, f, cf) != OR_OK) {
                status = OR_EBADCODE;
                continue;
            }
            const cplx w = sol[tet_to_field[i * T + t]];
            for (int q = 0; q < 3; ++q) acc[q] += (curl ? cf[q] : f[q]) * w;
        }
        const cplx k = (curl && c_tet) ? c_tet[t] : 1.0;
        for (int q = 0; q < 3; ++q) out[q * n + j] = acc[q] * k;
    }
    return status;
}

/* compiled/base/index_inter.py:134-246 by brute force: every candidate tet tests every
 * point with the same barycentric formula and tolerance; a point keeps the largest
 * containing tet id (the reference's answer for ascending candidate lists). */
void ned2_oracle_locate(const double *nodes, int64_t Nn, const int64_t *tets, int64_t T,
                        const int64_t *tetids, int64_t n_cand, const double *coords, int64_t n,
                        int64_t *out) {
    const double EPS = 1e-8;
    for (int64_t j = 0; j < n; ++j) out[j] = -1;
    for (int64_t it = 0; it < n_cand; ++it) {
        const int64_t t = tetids ? tetids[it] : it;
        const int64_t i1 = tets[t], i2 = tets[T + t], i3 = tets[2 * T + t], i4 = tets[3 * T + t];
        const double v1x = nodes[i1], v1y = nodes[Nn + i1], v1z = nodes[2 * Nn + i1];
        const double m00 = nodes[i2] - v1x, m10 = nodes[Nn + i2] - v1y, m20 = nodes[2 * Nn + i2] - v1z;
        const double m01 = nodes[i3] - v1x, m11 = nodes[Nn + i3] - v1y, m21 = nodes[2 * Nn + i3] - v1z;
        const double m02 = nodes[i4] - v1x, m12 = nodes[Nn + i4] - v1y, m22 = nodes[2 * Nn + i4] - v1z;
        const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
                           m02 * (m10 * m21 - m11 * m20);
        const double id = 1.0 / det;
        const double b00 = (m11 * m22 - m12 * m21) * id, b01 = (m02 * m21 - m01 * m22) * id,
                     b02 = (m01 * m12 - m02 * m11) * id;
        const double b10 = (m12 * m20 - m10 * m22) * id, b11 = (m00 * m22 - m02 * m20) * id,
                     b12 = (m10 * m02 - m00 * m12) * id;
        const double b20 = (m10 * m21 - m11 * m20) * id, b21 = (m20 * m01 - m00 * m21) * id,
                     b22 = (m00 * m11 - m10 * m01) * id;
        for (int64_t j = 0; j < n; ++j) {
            const double dx = coords[j] - v1x, dy = coords[n + j] - v1y, dz = coords[2 * n + j] - v1z;
            const double u = b00 * dx + b01 * dy + b02 * dz;
            const double v = b10 * dx + b11 * dy + b12 * dz;
            const double w = b20 * dx + b21 * dy + b22 * dz;
            if (u >= -EPS && v >= -EPS && w >= -EPS && u + v + w <= 1.0 + EPS)
                if (t > out[j]) out[j] = t;
        }
    }
}
