/*
 * ORACLE -- test infrastructure, not product code.
 *
 * The parts of the reference's element computation that Numba compiles WITHOUT fastmath,
 * restated in the reference's own evaluation order so that their rounding is the
 * reference's bit for bit. This file is compiled with -fno-fast-math -ffp-contract=off
 * (no reassociation, no fused multiply-adds); see the Makefile.
 *
 *   or_ref_tet_coefficients  physics/microwave/assembly/curlcurl.py:125-213
 *                            (tet_coefficients: njit without fastmath). V is the
 *                            absolute value of a 24-term sum of products x_i y_j z_k, and
 *                            the cofactors a, b, c, d are 6-term sums, all in absolute
 *                            coordinates, each evaluated left to right.
 *   or_ref_matinv            mth/optimized.py:315-357 (matinv: njit without fastmath);
 *                            complex products (ac - bd) + (ad + bc)i; the scalar division
 *                            1 / M[p,p] by CPython's algorithm (_Py_c_quot) and the array
 *                            division out / det by NumPy's, which is how Numba lowers
 *                            scalar and array expressions.
 *   or_ref_eval_point        one quadrature point of curlcurl.py:313-331 (coordinates of
 *                            the point from the stored vertices and the rule's barycentric
 *                            row) and the 3D basis functions and curls of
 *                            compiled/ccbf.py:536-550 (_ne0_3d), 568-586 (_ne2_3d),
 *                            643-652 (_curl_ne0_3d), 658-682 (_curl_ne2_3d), 789-805
 *                            (_nf0_3d), 807-823 (_nf1_3d), 927-946 (_curl_nf0_3d),
 *                            948-967 (_curl_nf1_3d): each njit without fastmath; the
 *                            expressions keep the reference's operand order.
 *
 * Two steps of ned2_tet_stiff_mass are fastmath in the reference (curlcurl.py:285-292):
 * the coefficient rows aas/(6V) (curlcurl.py:307-311) and the point coordinates
 * (curlcurl.py:319-321). Numba may turn the division into a product by a reciprocal and
 * the coordinate sums into fused multiply-adds, depending on the host CPU. Here both are
 * evaluated plainly (a division per coefficient; the coordinate sum left to right,
 * unfused).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>

typedef double _Complex cplx;

/* 11-point Keast degree-4 rule: weights then barycentric points
 * (curlcurl.py:277-283, identical to keast_rules.py KEAST4). */
const double OR_DPTS[5][11] = {
    {-0.07893333333333329982, 0.04573333333333329948, 0.04573333333333329948,
     0.04573333333333329948, 0.04573333333333329948, 0.14933333333333329018,
     0.14933333333333329018, 0.14933333333333329018, 0.14933333333333329018,
     0.14933333333333329018, 0.14933333333333329018},
    {0.25000000000000000000, 0.78571428571428569843, 0.07142857142857139685,
     0.07142857142857139685, 0.07142857142857139685, 0.10059642383320080428,
     0.39940357616679922348, 0.39940357616679922348, 0.39940357616679922348,
     0.10059642383320080428, 0.10059642383320080428},
    {0.25000000000000000000, 0.07142857142857139685, 0.07142857142857139685,
     0.07142857142857139685, 0.78571428571428569843, 0.39940357616679922348,
     0.10059642383320080428, 0.39940357616679922348, 0.10059642383320080428,
     0.39940357616679922348, 0.10059642383320080428},
    {0.25000000000000000000, 0.07142857142857139685, 0.07142857142857139685,
     0.78571428571428569843, 0.07142857142857139685, 0.39940357616679922348,
     0.39940357616679922348, 0.10059642383320080428, 0.10059642383320080428,
     0.10059642383320080428, 0.39940357616679922348},
    {0.25000000000000000000, 0.07142857142857150787, 0.78571428571428580945,
     0.07142857142857150787, 0.07142857142857150787, 0.10059642383320077652,
     0.10059642383320077652, 0.10059642383320074877, 0.39940357616679922348,
     0.39940357616679922348, 0.39940357616679922348},
};

/* The three vertices other than n, ascending. */
static void others(int n, int o[3]) {
    int m = 0;
    for (int k = 0; k < 4; ++k)
        if (k != n) o[m++] = k;
}

/* six3: p(A,B,C) - p(A,C,B) - p(B,A,C) + p(B,C,A) + p(C,A,B) - p(C,B,A), left to right,
 * with p(i,j,k) = u_i v_j w_k; six2: the same signs over the pairs (A,B), (A,C), (B,A),
 * (B,C), (C,A), (C,B) with p(i,j) = u_i v_j. Every cofactor of curlcurl.py:163-212 is one
 * of these sums or its negation; a negated sum of negated terms rounds the same (rounding
 * to nearest is symmetric), so the sign may be applied at the end. */
static double six3(const double *u, const double *v, const double *w, const int o[3]) {
    const int A = o[0], B = o[1], C = o[2];
    return u[A] * v[B] * w[C] - u[A] * v[C] * w[B] - u[B] * v[A] * w[C] +
           u[B] * v[C] * w[A] + u[C] * v[A] * w[B] - u[C] * v[B] * w[A];
}
static double six2(const double *u, const double *v, const int o[3]) {
    const int A = o[0], B = o[1], C = o[2];
    return u[A] * v[B] - u[A] * v[C] - u[B] * v[A] + u[B] * v[C] + u[C] * v[A] - u[C] * v[B];
}

/* Parity of the permutation (i, j, k, l) of (0, 1, 2, 3): 0 even, 1 odd. */
static int parity4(int i, int j, int k, int l) {
    const int p[4] = {i, j, k, l};
    int inv = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = a + 1; b < 4; ++b) inv += p[a] > p[b];
    return inv & 1;
}

/* tet_coefficients (curlcurl.py:132-213): vertices in stored order. V sums the 24 terms
 * x_i y_j z_k (i ascending, then j, then k among the remaining vertices) with the sign
 * -(-1)^parity(i,j,k,l), left to right: the order of curlcurl.py:140-164. aas[n] is
 * +/- six3 over the other vertices (x, y, z), bbs[n] -/+ six2 (y, z), ccs[n] +/- six2
 * (x, z), dds[n] -/+ six2 (x, y), with + for even n (curlcurl.py:166-212). */
void or_ref_tet_coefficients(const double x[4], const double y[4], const double z[4],
                             double a[4], double b[4], double c[4], double d[4], double *V) {
    double acc = 0.0;
    int first = 1;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            for (int k = 0; k < 4; ++k) {
                if (k == i || k == j) continue;
                const int l = 6 - i - j - k;
                const double t = x[i] * y[j] * z[k];
                const int neg = !parity4(i, j, k, l);
                if (first) {
                    acc = neg ? -t : t;
                    first = 0;
                } else {
                    acc = neg ? acc - t : acc + t;
                }
            }
        }
    *V = fabs(acc) / 6;
    for (int n = 0; n < 4; ++n) {
        int o[3];
        others(n, o);
        const double s = (n % 2 == 0) ? 1.0 : -1.0;
        a[n] = s * six3(x, y, z, o);
        b[n] = -s * six2(y, z, o);
        c[n] = s * six2(x, z, o);
        d[n] = -s * six2(x, y, o);
    }
}

/* Numba's complex division (CPython's _Py_c_quot). */
static cplx cdiv_py(cplx a, cplx b) {
    const double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    if (fabs(br) >= fabs(bi)) {
        const double ratio = bi / br;
        const double denom = br + bi * ratio;
        return CMPLX((ar + ai * ratio) / denom, (ai - ar * ratio) / denom);
    }
    const double ratio = br / bi;
    const double denom = br * ratio + bi;
    return CMPLX((ar * ratio + ai) / denom, (ai * ratio - ar) / denom);
}

/* NumPy's complex division (Smith's method with a reciprocal scale), which Numba uses for
 * array expressions such as matinv's out / det. */
static cplx cdiv_np(cplx a, cplx b) {
    const double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    if (fabs(br) >= fabs(bi)) {
        const double rat = bi / br;
        const double scl = 1.0 / (br + bi * rat);
        return CMPLX((ar + ai * rat) * scl, (ai - ar * rat) * scl);
    }
    const double rat = br / bi;
    const double scl = 1.0 / (bi + br * rat);
    return CMPLX((ar * rat + ai) * scl, (ai * rat - ar) * scl);
}

/* Numba's complex product (a.re b.re - a.im b.im) + (a.re b.im + a.im b.re) i. */
static cplx cmul_py(cplx a, cplx b) {
    const double ar = creal(a), ai = cimag(a), br = creal(b), bi = cimag(b);
    return CMPLX(ar * br - ai * bi, ar * bi + ai * br);
}
static cplx cadd_py(cplx a, cplx b) { return CMPLX(creal(a) + creal(b), cimag(a) + cimag(b)); }
static cplx csub_py(cplx a, cplx b) { return CMPLX(creal(a) - creal(b), cimag(a) - cimag(b)); }
static cplx cneg_py(cplx a) { return CMPLX(-creal(a), -cimag(a)); }

/* matinv (mth/optimized.py:315-357), row-major M[3p+q]. */
void or_ref_matinv(const cplx M[9], cplx out[9]) {
    for (int i = 0; i < 9; ++i) out[i] = 0;
    if (M[1] == 0 && M[2] == 0 && M[3] == 0 && M[5] == 0 && M[6] == 0 && M[7] == 0) {
        out[0] = cdiv_py(1.0, M[0]);
        out[4] = cdiv_py(1.0, M[4]);
        out[8] = cdiv_py(1.0, M[8]);
        return;
    }
#define P3(a, b, c) cmul_py(cmul_py(M[a], M[b]), M[c])
    cplx det = P3(0, 4, 8);
    det = csub_py(det, P3(0, 5, 7));
    det = csub_py(det, P3(1, 3, 8));
    det = cadd_py(det, P3(1, 5, 6));
    det = cadd_py(det, P3(2, 3, 7));
    det = csub_py(det, P3(2, 4, 6));
#undef P3
#define MM(a, b) cmul_py(M[a], M[b])
    cplx o[9];
    o[0] = csub_py(MM(4, 8), MM(5, 7));
    o[1] = cadd_py(cmul_py(cneg_py(M[1]), M[8]), MM(2, 7));
    o[2] = csub_py(MM(1, 5), MM(2, 4));
    o[3] = cadd_py(cmul_py(cneg_py(M[3]), M[8]), MM(5, 6));
    o[4] = csub_py(MM(0, 8), MM(2, 6));
    o[5] = cadd_py(cmul_py(cneg_py(M[0]), M[5]), MM(2, 3));
    o[6] = csub_py(MM(3, 7), MM(4, 6));
    o[7] = cadd_py(cmul_py(cneg_py(M[0]), M[7]), MM(1, 6));
    o[8] = csub_py(MM(0, 4), MM(1, 3));
#undef MM
    for (int i = 0; i < 9; ++i) out[i] = cdiv_np(o[i], det);
}

/* coeff[r][n]: row r (a, b, c, d) of vertex n, i.e. cof / (6 V) (curlcurl.py:307-311). */
void or_ref_coeff(const double a[4], const double b[4], const double c[4], const double d[4],
                  double V, double coeff[4][4]) {
    const double sixV = 6 * V;
    for (int n = 0; n < 4; ++n) {
        coeff[0][n] = a[n] / sixV;
        coeff[1][n] = b[n] / sixV;
        coeff[2][n] = c[n] / sixV;
        coeff[3][n] = d[n] / sixV;
    }
}

/* L_n at (x, y, z) as every ccbf.py function writes it: a + b*x + c*y + d*z. */
static double lam(const double co[4][4], int n, double x, double y, double z) {
    return co[0][n] + co[1][n] * x + co[2][n] * y + co[3][n] * z;
}

/* Code 64, _ne0_3d / _curl_ne0_3d. */
static void ne0(const double co[4][4], int i, int j, double x, double y, double z, double f[3],
                double cf[3]) {
    const double bi = co[1][i], ci = co[2][i], di = co[3][i];
    const double bj = co[1][j], cj = co[2][j], dj = co[3][j];
    const double t0 = lam(co, j, x, y, z), t1 = lam(co, i, x, y, z);
    f[0] = -t0 * bi + t1 * bj;
    f[1] = -t0 * ci + t1 * cj;
    f[2] = -t0 * di + t1 * dj;
    cf[0] = 2 * ci * dj - 2 * cj * di;
    cf[1] = -2 * bi * dj + 2 * bj * di;
    cf[2] = 2 * bi * cj - 2 * bj * ci;
}

/* Code 66, _ne2_3d / _curl_ne2_3d. */
static void ne2(const double co[4][4], int i, int j, double x, double y, double z, double f[3],
                double cf[3]) {
    const double ai = co[0][i], bi = co[1][i], ci = co[2][i], di = co[3][i];
    const double aj = co[0][j], bj = co[1][j], cj = co[2][j], dj = co[3][j];
    {
        const double t0 = bj * x, t1 = cj * y, t2 = dj * z;
        const double t3 = ai + bi * x + ci * y + di * z;
        const double t4 = -t0 - t1 - t2 + t3 - aj;
        const double t5 = t0 + t1 + t2 + aj;
        f[0] = -t4 * (-t3 * bj + t5 * bi);
        f[1] = -t4 * (-t3 * cj + t5 * ci);
        f[2] = -t4 * (-t3 * dj + t5 * di);
    }
    const double t0 = di - dj;
    const double t1 = bj * x, t2 = cj * y, t3 = dj * z;
    const double t4 = t1 + t2 + t3 + aj;
    const double t5 = ai + bi * x + ci * y + di * z;
    const double t6 = t4 * ci - t5 * cj;
    const double t7 = ci - cj;
    const double t8 = t4 * di - t5 * dj;
    const double t9 = -2 * t1 - 2 * t2 - 2 * t3 + 2 * t5 - 2 * aj;
    const double t10 = t4 * bi - t5 * bj;
    const double t11 = bi - bj;
    cf[0] = t0 * t6 - t7 * t8 + t9 * (ci * dj - cj * di);
    cf[1] = -t0 * t10 + t11 * t8 - t9 * (bi * dj - bj * di);
    cf[2] = t10 * t7 - t11 * t6 + t9 * (bi * cj - bj * ci);
}

/* Code 128, _nf0_3d / _curl_nf0_3d; code 129 (_nf1_3d / _curl_nf1_3d) is the same
 * expressions with the roles of i and j exchanged. */
static void nf0(const double co[4][4], int i, int j, int k, double x, double y, double z,
                double f[3], double cf[3]) {
    const double ai = co[0][i], bi = co[1][i], ci = co[2][i], di = co[3][i];
    const double bj = co[1][j], cj = co[2][j], dj = co[3][j];
    const double bk = co[1][k], ck = co[2][k], dk = co[3][k];
    {
        const double t0 = lam(co, i, x, y, z), t1 = lam(co, k, x, y, z), t2 = lam(co, j, x, y, z);
        f[0] = -t0 * (t1 * bj - t2 * bk);
        f[1] = -t0 * (t1 * cj - t2 * ck);
        f[2] = -t0 * (t1 * dj - t2 * dk);
    }
    const double t0 = 2 * ai + 2 * bi * x + 2 * ci * y + 2 * di * z;
    const double t1 = lam(co, k, x, y, z), t2 = lam(co, j, x, y, z);
    const double t3 = t1 * cj - t2 * ck;
    const double t4 = t1 * dj - t2 * dk;
    const double t5 = t1 * bj - t2 * bk;
    cf[0] = t0 * (cj * dk - ck * dj) + t3 * di - t4 * ci;
    cf[1] = -t0 * (bj * dk - bk * dj) + t4 * bi - t5 * di;
    cf[2] = t0 * (bj * ck - bk * cj) - t3 * bi + t5 * ci;
}

/* Coordinates of quadrature point q (curlcurl.py:319-321) from the stored vertices
 * vx[coord][vertex], then every basis function and curl at it. ijk[i] holds the local
 * vertices (i, j, k) of function i (the local maps, curlcurl.py:72-85). Returns -2 on an
 * unknown DOF code (ccbf.py:1206-1207). */
int or_ref_eval_point(const double coeff[4][4], const double vx[3][4], int q,
                      const int64_t *codes, const int (*ijk)[3], int ndof, double (*F)[3],
                      double (*CF)[3]) {
    double p[3];
    for (int r = 0; r < 3; ++r)
        p[r] = vx[r][0] * OR_DPTS[1][q] + vx[r][1] * OR_DPTS[2][q] + vx[r][2] * OR_DPTS[3][q] +
               vx[r][3] * OR_DPTS[4][q];
    for (int n = 0; n < ndof; ++n) {
        const int i = ijk[n][0], j = ijk[n][1], k = ijk[n][2];
        switch (codes[n]) {
        case 64: ne0(coeff, i, j, p[0], p[1], p[2], F[n], CF[n]); break;
        case 66: ne2(coeff, i, j, p[0], p[1], p[2], F[n], CF[n]); break;
        case 128: nf0(coeff, i, j, k, p[0], p[1], p[2], F[n], CF[n]); break;
        case 129: nf0(coeff, j, i, k, p[0], p[1], p[2], F[n], CF[n]); break;
        default: return -2;
        }
    }
    return 0;
}
