/*
 * ORACLE -- test infrastructure, not product code.
 *
 * CPU restatement of the reference's Lagrange-2 heat-conduction assembly,
 * src/emerge/_emerge/physics/heatconduction/assembly/grad.py. Only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it, as the checker.
 *
 * Followed step by step:
 *   heat_tet_coefficients  grad.py:42-135 (cofactors a, b, c, d of the vertex matrix and
 *                          V = |det| / 6, in physical coordinates)
 *   grad_n1 / grad_n2      grad.py:139-172 ((4 l - 1) grad l; 4 (l_b grad l_a + l_a grad l_b))
 *   heat_tet_stiff         grad.py:199-261 (leg2_tet_stiff: all 100 entries,
 *                          (C grad N_i) . grad N_j summed with the rule, times V)
 *   heat_tet_mass          grad.py:265-338 (leg2_tet_mass: i <= j then mirrored, times
 *                          rho_cp V)
 *   heat_oracle_build      grad.py:346-414 (_stiff_matrix_builder / _mass_matrix_builder:
 *                          COO block per tet, row-major, p = itet * 100)
 * The quadrature rule is an argument (weights then L1..L4 of 11 points): the reference's
 * 8-digit table (grad.py:179-190), or any other rule for the pins.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

static const int LOCAL_EDGE[2][6] = {{0, 0, 0, 1, 1, 2}, {1, 2, 3, 2, 3, 3}};

static double det3h(double a0, double a1, double a2, double b0, double b1, double b2, double c0,
                    double c1, double c2) {
    return a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0);
}

/* cofactors of [[1, x_n, y_n, z_n]] (row n): a_n, b_n, c_n, d_n; V = |det| / 6 */
static void heat_tet_coefficients(const double x[4], const double y[4], const double z[4],
                                  double a[4], double b[4], double c[4], double d[4],
                                  double *V) {
    for (int n = 0; n < 4; ++n) {
        int r[3], m = 0;
        for (int k = 0; k < 4; ++k)
            if (k != n) r[m++] = k;
        const double sg = (n % 2 == 0) ? 1.0 : -1.0;
        a[n] = sg * det3h(x[r[0]], y[r[0]], z[r[0]], x[r[1]], y[r[1]], z[r[1]], x[r[2]],
                          y[r[2]], z[r[2]]);
        b[n] = -sg * det3h(1.0, y[r[0]], z[r[0]], 1.0, y[r[1]], z[r[1]], 1.0, y[r[2]], z[r[2]]);
        c[n] = sg * det3h(1.0, x[r[0]], z[r[0]], 1.0, x[r[1]], z[r[1]], 1.0, x[r[2]], z[r[2]]);
        d[n] = -sg * det3h(1.0, x[r[0]], y[r[0]], 1.0, x[r[1]], y[r[1]], 1.0, x[r[2]], y[r[2]]);
    }
    *V = fabs(a[0] + a[1] + a[2] + a[3]) / 6.0;
}

/* barycentric-like values l[4] and gradients g[4][3] at the rule's points, as the
 * reference evaluates them: coefficients over 6V at the physical points */
static void heat_setup(const double vx[3][4], const double *rule, double l[11][4],
                       double g[4][3], double *V) {
    double a[4], b[4], c[4], d[4];
    heat_tet_coefficients(vx[0], vx[1], vx[2], a, b, c, d, V);
    for (int n = 0; n < 4; ++n) {
        g[n][0] = b[n] / (6.0 * *V);
        g[n][1] = c[n] / (6.0 * *V);
        g[n][2] = d[n] / (6.0 * *V);
    }
    for (int q = 0; q < 11; ++q) {
        double p[3];
        for (int k = 0; k < 3; ++k)
            p[k] = vx[k][0] * rule[11 + q] + vx[k][1] * rule[22 + q] + vx[k][2] * rule[33 + q] +
                   vx[k][3] * rule[44 + q];
        for (int n = 0; n < 4; ++n)
            l[q][n] = a[n] / (6.0 * *V) + g[n][0] * p[0] + g[n][1] * p[1] + g[n][2] * p[2];
    }
}

static void grad_n(int i, const double l[4], const double g[4][3], double out[3]) {
    if (i < 4) {
        const double qv = 4.0 * l[i] - 1.0;
        for (int k = 0; k < 3; ++k) out[k] = qv * g[i][k];
    } else {
        const int e1 = LOCAL_EDGE[0][i - 4], e2 = LOCAL_EDGE[1][i - 4];
        for (int k = 0; k < 3; ++k) out[k] = 4.0 * l[e2] * g[e1][k] + 4.0 * l[e1] * g[e2][k];
    }
}

static double shape_n(int i, const double l[4]) {
    if (i < 4) return l[i] * (2.0 * l[i] - 1.0);
    return 4.0 * l[LOCAL_EDGE[0][i - 4]] * l[LOCAL_EDGE[1][i - 4]];
}

/* cond[9] row-major (p, q) */
static void heat_tet_stiff(const double vx[3][4], const double cond[9], const double *rule,
                           double K[100]) {
    double l[11][4], g[4][3], V;
    heat_setup(vx, rule, l, g, &V);
    for (int i = 0; i < 10; ++i)
        for (int j = 0; j < 10; ++j) {
            double s = 0.0;
            for (int q = 0; q < 11; ++q) {
                double gi[3], gj[3], cg[3];
                grad_n(i, l[q], g, gi);
                grad_n(j, l[q], g, gj);
                for (int r = 0; r < 3; ++r)
                    cg[r] = cond[3 * r] * gi[0] + cond[3 * r + 1] * gi[1] + cond[3 * r + 2] * gi[2];
                s += (cg[0] * gj[0] + cg[1] * gj[1] + cg[2] * gj[2]) * rule[q];
            }
            K[10 * i + j] = s * V;
        }
}

static void heat_tet_mass(const double vx[3][4], double rho_cp, const double *rule,
                          double M[100]) {
    double l[11][4], g[4][3], V;
    heat_setup(vx, rule, l, g, &V);
    for (int i = 0; i < 10; ++i)
        for (int j = i; j < 10; ++j) {
            double s = 0.0;
            for (int q = 0; q < 11; ++q) s += rule[q] * shape_n(i, l[q]) * shape_n(j, l[q]);
            M[10 * i + j] = s * rho_cp * V;
            M[10 * j + i] = s * rho_cp * V;
        }
}

/* grad.py:346-414. nodes[c*Nn+n], tets[v*T+t], tet_to_field[d*T+t], cond[(p*3+q)*T+t]
 * (the reference's (3,3,T) array), rho_cp[T]; K/M (either may be NULL) [100 T];
 * rows/cols [100 T] (may be NULL). */
void heat_oracle_build(const double *nodes, int64_t Nn, const int64_t *tets, int64_t T,
                       const int64_t *tet_to_field, const double *cond, const double *rho_cp,
                       const double *rule, double *K, double *M, int64_t *rows,
                       int64_t *cols, int local_frame) {
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < T; ++t) {
        /* local_frame translates the tet to its stored vertex 0 (same algorithm; avoids
         * the cancellation of absolute-coordinate cofactors on slivers) */
        double vx[3][4], o[3] = {0, 0, 0};
        if (local_frame)
            for (int c = 0; c < 3; ++c) o[c] = nodes[c * Nn + tets[t]];
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) vx[c][k] = nodes[c * Nn + tets[k * T + t]] - o[c];
        if (K) {
            double ct[9];
            for (int p = 0; p < 9; ++p) ct[p] = cond[p * T + t];
            heat_tet_stiff(vx, ct, rule, K + 100 * t);
        }
        if (M) heat_tet_mass(vx, rho_cp[t], rule, M + 100 * t);
        if (rows && cols)
            for (int ii = 0; ii < 10; ++ii)
                for (int jj = 0; jj < 10; ++jj) {
                    rows[100 * t + 10 * ii + jj] = tet_to_field[ii * T + t];
                    cols[100 * t + 10 * ii + jj] = tet_to_field[jj * T + t];
                }
    }
}
