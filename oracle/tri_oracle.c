/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU restatement of the reference's Nedelec-2 surface (Robin / port) assembly on
 * boundary triangles: src/emerge/_emerge/physics/microwave/assembly/robinbc.py.
 * Only the test suite and the smoke check load it, as the checker.
 *
 * Followed step by step:
 *   or_local_vertices   robinbc.py:163-181 (construct_local_vertices: origin at vertex 0,
 *                       x along edge 0->1, z the face normal, y = z x x)
 *   or_tri_coefficients robinbc.py:197-217 (a, b, c over 2 sA; sA the signed area)
 *   DPTS2               robinbc.py:185-192 (6-point degree-4 triangle rule)
 *   or_f2d              compiled/ccbf.py 2D functions for codes 64 (_ne0_2d :80-91),
 *                       66 (_ne2_2d :106-120), 128 (_nf0_2d :268-281), 129 (_nf1_2d
 *                       :283-296), written here from their definitions
 *                       (elements/dofsets.py:66-83); edge index e -> local vertices
 *                       (0,1), (1,2), (0,2), face (0,1,2) (robinbc.py:256-272)
 *   or_tri_stiff        robinbc.py:316-381 (ned2_tri_stiff)
 *   or_tri_force        robinbc.py:220-275 (ned2_tri_force) and :453-514
 *                       (ned2_tri_force_scat: U + curl terms with the normal's sign)
 *   tri_oracle_bc_entries     robinbc.py:385-402 (compute_bc_entries)
 *   tri_oracle_force_entries  robinbc.py:278-313 / :517-569 (compute_force_entries[_scat])
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef double _Complex cplx;

/* weights, then L1, L2, L3 of the 6 points (robinbc.py:185-192) */
static const double DPTS2[4][6] = {
    {0.10995174365532200, 0.10995174365532200, 0.10995174365532200, 0.22338158967801100,
     0.22338158967801100, 0.22338158967801100},
    {0.81684757298045896, 0.09157621350977101, 0.09157621350977101, 0.10810301816807000,
     0.44594849091596500, 0.44594849091596500},
    {0.09157621350977101, 0.81684757298045896, 0.09157621350977101, 0.44594849091596500,
     0.10810301816807000, 0.44594849091596500},
    {0.09157621350977004, 0.09157621350977008, 0.81684757298045807, 0.44594849091596500,
     0.44594849091596500, 0.10810301816807000},
};

static const int64_t CODES2D[8] = {64, 64, 64, 66, 66, 66, 128, 129};
static const int EDGE_I[3] = {0, 1, 0}, EDGE_J[3] = {1, 2, 2};

static void cross(const double u[3], const double v[3], double w[3]) {
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

static void normalize(double a[3]) {
    const double n = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    a[0] /= n;
    a[1] /= n;
    a[2] /= n;
}

/* basis rows (xhat, yhat, zhat) and the local coordinates of the 3 vertices */
static void or_local_vertices(const double gv[3][3], double basis[3][3], double lx[3],
                              double ly[3]) {
    double e1[3], e2[3];
    for (int q = 0; q < 3; ++q) {
        e1[q] = gv[q][1] - gv[q][0];
        e2[q] = gv[q][2] - gv[q][0];
    }
    double z[3], x[3], y[3];
    cross(e1, e2, z);
    normalize(z);
    memcpy(x, e1, sizeof(x));
    normalize(x);
    cross(z, x, y);
    normalize(y);
    for (int q = 0; q < 3; ++q) {
        basis[0][q] = x[q];
        basis[1][q] = y[q];
        basis[2][q] = z[q];
    }
    for (int v = 0; v < 3; ++v) {
        double d[3];
        for (int q = 0; q < 3; ++q) d[q] = gv[q][v] - gv[q][0];
        lx[v] = basis[0][0] * d[0] + basis[0][1] * d[1] + basis[0][2] * d[2];
        ly[v] = basis[1][0] * d[0] + basis[1][1] * d[1] + basis[1][2] * d[2];
    }
}

/* coeff[0..2][v] = a, b, c of vertex v over 2 sA; returns |sA| */
static double or_tri_coefficients(const double x[3], const double y[3], double coeff[3][3]) {
    const double a[3] = {x[1] * y[2] - y[1] * x[2], x[2] * y[0] - y[2] * x[0],
                         x[0] * y[1] - y[0] * x[1]};
    const double b[3] = {y[1] - y[2], y[2] - y[0], y[0] - y[1]};
    const double c[3] = {x[2] - x[1], x[0] - x[2], x[1] - x[0]};
    const double sA = 0.5 * (b[0] * c[1] - b[1] * c[0]);
    for (int v = 0; v < 3; ++v) {
        coeff[0][v] = a[v] / (2 * sA);
        coeff[1][v] = b[v] / (2 * sA);
        coeff[2][v] = c[v] / (2 * sA);
    }
    return fabs(sA);
}

/* 2D basis function idof (0..7) at local point (px, py) */
static void or_f2d(int idof, const double coeff[3][3], double px, double py, double f[2]) {
    double L[3], gx[3], gy[3];
    for (int v = 0; v < 3; ++v) {
        L[v] = coeff[0][v] + coeff[1][v] * px + coeff[2][v] * py;
        gx[v] = coeff[1][v];
        gy[v] = coeff[2][v];
    }
    const int64_t code = CODES2D[idof];
    if (code == 64 || code == 66) {
        const int e = idof % 3, i = EDGE_I[e], j = EDGE_J[e];
        double s = 1.0;
        if (code == 66) s = L[i] - L[j];
        f[0] = s * (L[i] * gx[j] - L[j] * gx[i]);
        f[1] = s * (L[i] * gy[j] - L[j] * gy[i]);
    } else if (code == 128) { /* L0 (L1 grad L2 - L2 grad L1) */
        f[0] = L[0] * (L[1] * gx[2] - L[2] * gx[1]);
        f[1] = L[0] * (L[1] * gy[2] - L[2] * gy[1]);
    } else { /* 129: L1 (L0 grad L2 - L2 grad L0) */
        f[0] = L[1] * (L[0] * gx[2] - L[2] * gx[0]);
        f[1] = L[1] * (L[0] * gy[2] - L[2] * gy[0]);
    }
}

static void or_tri_setup(const double gv[3][3], double basis[3][3], double coeff[3][3],
                         double *A, double px[6], double py[6]) {
    double lx[3], ly[3];
    or_local_vertices(gv, basis, lx, ly);
    *A = or_tri_coefficients(lx, ly, coeff);
    for (int w = 0; w < 6; ++w) {
        px[w] = lx[0] * DPTS2[1][w] + lx[1] * DPTS2[2][w] + lx[2] * DPTS2[3][w];
        py[w] = ly[0] * DPTS2[1][w] + ly[1] * DPTS2[2][w] + ly[2] * DPTS2[3][w];
    }
}

/* robinbc.py:316-381 */
static void or_tri_stiff(const double gv[3][3], cplx gamma, cplx B[8][8]) {
    double basis[3][3], coeff[3][3], A, px[6], py[6];
    or_tri_setup(gv, basis, coeff, &A, px, py);
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
            double s = 0.0;
            for (int w = 0; w < 6; ++w) {
                double fi[2], fj[2];
                or_f2d(i, coeff, px[w], py[w], fi);
                or_f2d(j, coeff, px[w], py[w], fj);
                s += (fi[0] * fj[0] + fi[1] * fj[1]) * DPTS2[0][w];
            }
            B[i][j] = gamma * s * A;
        }
}

/* robinbc.py:220-275 (ucurl == NULL) and :453-514. U, ucurl: [3][6] complex global
 * components at the 6 points; nhat: the port normal (scattered-field form). */
static void or_tri_force(const double gv[3][3], const cplx U[3][6], const cplx (*ucurl)[6],
                         const double *nhat, cplx b[8]) {
    double basis[3][3], coeff[3][3], A, px[6], py[6];
    or_tri_setup(gv, basis, coeff, &A, px, py);
    cplx Ux[6], Uy[6];
    double sgn = 0.0;
    if (ucurl) {
        const double nz = basis[2][0] * nhat[0] + basis[2][1] * nhat[1] + basis[2][2] * nhat[2];
        sgn = (nz > 0) - (nz < 0);
    }
    for (int w = 0; w < 6; ++w) {
        Ux[w] = basis[0][0] * U[0][w] + basis[0][1] * U[1][w] + basis[0][2] * U[2][w];
        Uy[w] = basis[1][0] * U[0][w] + basis[1][1] * U[1][w] + basis[1][2] * U[2][w];
        if (ucurl) {
            const cplx cx = basis[0][0] * ucurl[0][w] + basis[0][1] * ucurl[1][w] +
                            basis[0][2] * ucurl[2][w];
            const cplx cy = basis[1][0] * ucurl[0][w] + basis[1][1] * ucurl[1][w] +
                            basis[1][2] * ucurl[2][w];
            Ux[w] += cy * sgn;
            Uy[w] -= cx * sgn;
        }
    }
    for (int i = 0; i < 8; ++i) {
        cplx s = 0.0;
        for (int w = 0; w < 6; ++w) {
            double fi[2];
            or_f2d(i, coeff, px[w], py[w], fi);
            s += DPTS2[0][w] * (fi[0] * Ux[w] + fi[1] * Uy[w]);
        }
        b[i] = -A * s;
    }
}

static void tri_vertices(const double *nodes, int64_t Nn, const int64_t *tris, int64_t F,
                         int64_t itri, double gv[3][3]) {
    for (int v = 0; v < 3; ++v) {
        const int64_t n = tris[v * F + itri];
        for (int q = 0; q < 3; ++q) gv[q][v] = nodes[q * Nn + n];
    }
}

/* robinbc.py:385-402: Bmat[itri*64 : +64] += ned2_tri_stiff(...).ravel() */
void tri_oracle_bc_entries(const double *nodes, int64_t Nn, const int64_t *tris, int64_t F,
                           const int64_t *surf, int64_t nsurf, double gre, double gim,
                           cplx *Bmat) {
    const cplx gamma = gre + gim * I;
    for (int64_t s = 0; s < nsurf; ++s) {
        double gv[3][3];
        tri_vertices(nodes, Nn, tris, F, surf[s], gv);
        cplx B[8][8];
        or_tri_stiff(gv, gamma, B);
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) Bmat[surf[s] * 64 + 8 * i + j] += B[i][j];
    }
}

/* robinbc.py:278-313 / :517-569: Bvec[tri_to_field[:, itri]] += bvec.
 * U, ucurl: complex [3][6][nsurf] (U_global_all); normals [3][nsurf] or NULL. */
void tri_oracle_force_entries(const double *nodes, int64_t Nn, const int64_t *tris, int64_t F,
                              const int64_t *tri_to_field, const int64_t *surf, int64_t nsurf,
                              const cplx *U, const cplx *ucurl, const double *normals,
                              cplx *Bvec) {
    for (int64_t s = 0; s < nsurf; ++s) {
        double gv[3][3];
        tri_vertices(nodes, Nn, tris, F, surf[s], gv);
        cplx Us[3][6], Cs[3][6];
        for (int q = 0; q < 3; ++q)
            for (int w = 0; w < 6; ++w) {
                Us[q][w] = U[(q * 6 + w) * nsurf + s];
                if (ucurl) Cs[q][w] = ucurl[(q * 6 + w) * nsurf + s];
            }
        double nh[3];
        if (ucurl)
            for (int q = 0; q < 3; ++q) nh[q] = normals[q * nsurf + s];
        cplx b[8];
        or_tri_force(gv, (const cplx(*)[6])Us, ucurl ? (const cplx(*)[6])Cs : NULL,
                     ucurl ? nh : NULL, b);
        for (int i = 0; i < 8; ++i) Bvec[tri_to_field[i * F + surf[s]]] += b[i];
    }
}

void tri_oracle_dpts(double *out) { memcpy(out, DPTS2, sizeof(DPTS2)); }
