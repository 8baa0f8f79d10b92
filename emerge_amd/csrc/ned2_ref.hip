// Nedelec-2 element matrices in the reference's own arithmetic (EM_PREC_FP64_REF).
//
// The default FP64 kernel (ned2_assemble.hip) computes the exact element integrals from
// vertex differences. The reference instead evaluates them with an 11-point rule in
// absolute coordinates (physics/microwave/assembly/curlcurl.py:285-380): barycentric
// coefficients from 24-term cofactor sums (tet_coefficients, curlcurl.py:125-213),
// normalised by |V|, then the basis functions of compiled/ccbf.py at the physical points.
// On thin tets far from the origin those sums cancel, so its values differ from the exact
// integrals (by up to ~1e-7 relative on seeded Delaunay meshes). This kernel repeats that
// arithmetic step by step -- the same operand order, no fused multiply-adds -- so that
// every step the reference compiles without fastmath (tet_coefficients, matinv and the
// ccbf.py functions) rounds as the reference's does, and the remaining steps (the
// fastmath material products, weights and sums of ned2_tet_stiff_mass) differ from any
// one evaluation order of them only by rounding of well-conditioned sums.
//
// Work layout: one workgroup of 256 threads per tile of kRefTile consecutive assembly
// indices. Phase 1: one thread per tet computes the coefficients, the 11 physical points,
// the local vertex maps and the material tensors (matinv with the reference's semantics)
// into LDS. Phase 2, per quadrature point: one thread per (tet, function) evaluates the
// function and its curl, their material products and weights (LDS); then thread p < 210
// owns the upper-triangle pair p = (i <= j) of every tet of the tile and accumulates
// sum_c (Ms curl f_i w)_c curl f_j,c and sum_c (Mm f_i w)_c f_j,c in registers, point by
// point in ascending order (curlcurl.py:358-376). Phase 3: times V, mirrored, stored.
// Compute-bound FP64 (about 1e5 flop per tet): a parity mode, not the fast path.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "em_internal.h"

#pragma clang fp contract(off)

namespace em {

namespace {

constexpr int kRefTile = 16;
constexpr int kRefBlock = 256;
constexpr int kNq = 11;

// 11-point Keast degree-4 rule: weights, then the barycentric weight of stored vertex
// 0..3 at each point (curlcurl.py:277-283).
__constant__ double kRefDpts[5][kNq] = {
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

// Stored vertex positions of the edge and face slots (mesh3d.py:757-774).
__constant__ int kEdgeSlot[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
__constant__ int kFaceSlot[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};

struct Cx {
    double re, im;
};
// Numba's complex arithmetic: (ac - bd) + (ad + bc)i, component sums.
__device__ __forceinline__ Cx xmul(Cx a, Cx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cx xadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx xsub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx xneg(Cx a) { return {-a.re, -a.im}; }
// Scalar complex division as Numba lowers it (CPython's algorithm).
__device__ __forceinline__ Cx xdiv_py(Cx a, Cx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double ratio = b.im / b.re;
        const double denom = b.re + b.im * ratio;
        return {(a.re + a.im * ratio) / denom, (a.im - a.re * ratio) / denom};
    }
    const double ratio = b.re / b.im;
    const double denom = b.re * ratio + b.im;
    return {(a.re * ratio + a.im) / denom, (a.im * ratio - a.re) / denom};
}
// Array complex division as Numba lowers it (NumPy's: a reciprocal scale).
__device__ __forceinline__ Cx xdiv_np(Cx a, Cx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double rat = b.im / b.re;
        const double scl = 1.0 / (b.re + b.im * rat);
        return {(a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl};
    }
    const double rat = b.re / b.im;
    const double scl = 1.0 / (b.im + b.re * rat);
    return {(a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl};
}

// The reference's (3,3) tensor of one tet, row-major, from a material descriptor (the
// entries a scalar or diagonal descriptor leaves out are exact zeros, as in the (3,3,T)
// arrays the reference reads).
__device__ __forceinline__ void ref_tensor(const MatDesc& m, long long t, Cx E[9]) {
    const double2* p = m.p + t * m.ts;
#pragma unroll
    for (int i = 0; i < 9; ++i) E[i] = {0.0, 0.0};
    if (m.kind == EM_MAT_SCALAR) {
        const double2 v = p[0];
        E[0] = E[4] = E[8] = {v.x, v.y};
    } else if (m.kind == EM_MAT_DIAG) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double2 v = p[d * m.cs];
            E[4 * d] = {v.x, v.y};
        }
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            const double2 v = p[i * m.cs];
            E[i] = {v.x, v.y};
        }
    }
}

__device__ __forceinline__ bool xzero(Cx a) { return a.re == 0.0 && a.im == 0.0; }

// matinv (mth/optimized.py:315-357).
__device__ void ref_matinv(const Cx M[9], Cx out[9]) {
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = {0.0, 0.0};
    if (xzero(M[1]) && xzero(M[2]) && xzero(M[3]) && xzero(M[5]) && xzero(M[6]) &&
        xzero(M[7])) {
        const Cx one = {1.0, 0.0};
        out[0] = xdiv_py(one, M[0]);
        out[4] = xdiv_py(one, M[4]);
        out[8] = xdiv_py(one, M[8]);
        return;
    }
    Cx det = xmul(xmul(M[0], M[4]), M[8]);
    det = xsub(det, xmul(xmul(M[0], M[5]), M[7]));
    det = xsub(det, xmul(xmul(M[1], M[3]), M[8]));
    det = xadd(det, xmul(xmul(M[1], M[5]), M[6]));
    det = xadd(det, xmul(xmul(M[2], M[3]), M[7]));
    det = xsub(det, xmul(xmul(M[2], M[4]), M[6]));
    Cx o[9];
    o[0] = xsub(xmul(M[4], M[8]), xmul(M[5], M[7]));
    o[1] = xadd(xmul(xneg(M[1]), M[8]), xmul(M[2], M[7]));
    o[2] = xsub(xmul(M[1], M[5]), xmul(M[2], M[4]));
    o[3] = xadd(xmul(xneg(M[3]), M[8]), xmul(M[5], M[6]));
    o[4] = xsub(xmul(M[0], M[8]), xmul(M[2], M[6]));
    o[5] = xadd(xmul(xneg(M[0]), M[5]), xmul(M[2], M[3]));
    o[6] = xsub(xmul(M[3], M[7]), xmul(M[4], M[6]));
    o[7] = xadd(xmul(xneg(M[0]), M[7]), xmul(M[1], M[6]));
    o[8] = xsub(xmul(M[0], M[4]), xmul(M[1], M[3]));
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = xdiv_np(o[i], det);
}

__device__ __forceinline__ int parity4(int i, int j, int k, int l) {
    return ((i > j) + (i > k) + (i > l) + (j > k) + (j > l) + (k > l)) & 1;
}

// Sums of the cofactors of tet_coefficients (curlcurl.py:166-212), left to right; see
// six3/six2 of the oracle's restatement: signs + - - + + - over (A,B,C), (A,C,B),
// (B,A,C), (B,C,A), (C,A,B), (C,B,A).
__device__ __forceinline__ double six3(const double* u, const double* v, const double* w,
                                       int A, int B, int C) {
    return u[A] * v[B] * w[C] - u[A] * v[C] * w[B] - u[B] * v[A] * w[C] +
           u[B] * v[C] * w[A] + u[C] * v[A] * w[B] - u[C] * v[B] * w[A];
}
__device__ __forceinline__ double six2(const double* u, const double* v, int A, int B, int C) {
    return u[A] * v[B] - u[A] * v[C] - u[B] * v[A] + u[B] * v[C] + u[C] * v[A] - u[C] * v[B];
}

// tet_coefficients (curlcurl.py:132-213) and coeff = cof / (6 V) (curlcurl.py:307-311).
__device__ void ref_coefficients(const double x[4], const double y[4], const double z[4],
                                 double co[4][4], double& V) {
    double acc = 0.0;
    bool first = true;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k == i || k == j) continue;
                const int l = 6 - i - j - k;
                const double t = x[i] * y[j] * z[k];
                const bool neg = !parity4(i, j, k, l);
                if (first) {
                    acc = neg ? -t : t;
                    first = false;
                } else {
                    acc = neg ? acc - t : acc + t;
                }
            }
        }
    V = fabs(acc) / 6;
    const double sixV = 6 * V;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        const int A = n == 0 ? 1 : 0, B = n <= 1 ? 2 : 1, C = n <= 2 ? 3 : 2;
        const double s = (n % 2 == 0) ? 1.0 : -1.0;
        co[0][n] = (s * six3(x, y, z, A, B, C)) / sixV;
        co[1][n] = (-s * six2(y, z, A, B, C)) / sixV;
        co[2][n] = (s * six2(x, z, A, B, C)) / sixV;
        co[3][n] = (-s * six2(x, y, A, B, C)) / sixV;
    }
}

// L_n at a point as every ccbf.py function writes it: a + b*x + c*y + d*z.
__device__ __forceinline__ double lam(const double (*co)[4], int n, double x, double y,
                                      double z) {
    return co[0][n] + co[1][n] * x + co[2][n] * y + co[3][n] * z;
}

// Basis function `fn` (slot order: codes 64 x6, 66 x6, 128 x4, 129 x4) and its curl at
// (x, y, z): compiled/ccbf.py _ne0_3d / _curl_ne0_3d (:536-552, :643-652), _ne2_3d /
// _curl_ne2_3d (:568-586, :658-682), _nf0_3d / _curl_nf0_3d (:789-805, :927-946); code
// 129 (_nf1_3d / _curl_nf1_3d, :807-823, :948-967) is code 128 with i and j exchanged.
__device__ void ref_basis(const double (*co)[4], int fn, int i, int j, int k, double x,
                          double y, double z, double f[3], double cf[3]) {
    if (fn >= 16) {
        const int t = i;
        i = j;
        j = t;
    }
    const double ai = co[0][i], bi = co[1][i], ci = co[2][i], di = co[3][i];
    const double aj = co[0][j], bj = co[1][j], cj = co[2][j], dj = co[3][j];
    if (fn < 6) {
        const double t0 = lam(co, j, x, y, z), t1 = lam(co, i, x, y, z);
        f[0] = -t0 * bi + t1 * bj;
        f[1] = -t0 * ci + t1 * cj;
        f[2] = -t0 * di + t1 * dj;
        cf[0] = 2 * ci * dj - 2 * cj * di;
        cf[1] = -2 * bi * dj + 2 * bj * di;
        cf[2] = 2 * bi * cj - 2 * bj * ci;
    } else if (fn < 12) {
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
    } else {
        const double bk = co[1][k], ck = co[2][k], dk = co[3][k];
        {
            const double t0 = lam(co, i, x, y, z), t1 = lam(co, k, x, y, z),
                         t2 = lam(co, j, x, y, z);
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
}

struct RefTetRec {
    double co[4][4];
    double p[kNq][3];
    double V;
    double ms[9][2];  // matinv(ur)
    double mm[9][2];  // er
    signed char ijk[20][3];
    int ia;
};

struct RefPointRec {
    double f[kRefTile][20][3];
    double cf[kRefTile][20][3];
    double msw[kRefTile][20][3][2];  // (Ms curl f) * w
    double mmw[kRefTile][20][3][2];  // (Mm f) * w
};

// (Ms v) * w for one component row: the sum m0 v0 + m1 v1 + m2 v2 of complex times
// real-valued complex products, left to right (matmul_inplace, curlcurl.py:110-119),
// then scaled by the weight (curlcurl.py:345-348).
__device__ __forceinline__ void mat_row_w(const double (*m)[2], int r, const double v[3],
                                          double w, double out[2]) {
    double re = m[3 * r][0] * v[0], im = m[3 * r][1] * v[0];
    re = re + m[3 * r + 1][0] * v[1];
    im = im + m[3 * r + 1][1] * v[1];
    re = re + m[3 * r + 2][0] * v[2];
    im = im + m[3 * r + 2][1] * v[2];
    out[0] = re * w;
    out[1] = im * w;
}

template <int MODE>
__global__ __launch_bounds__(kRefBlock) void ned2_ref_kernel(MeshDev m, MatDesc eps, MatDesc mu,
                                                             K0Pack k0p, int nf,
                                                             double2* __restrict__ out0,
                                                             double2* __restrict__ out1,
                                                             long long freq_stride) {
    __shared__ RefTetRec tets[kRefTile];
    __shared__ RefPointRec pt;
    const long long start = (long long)blockIdx.x * kRefTile;
    const int cnt = (int)min((long long)kRefTile, m.n_assy - start);
    const int tid = threadIdx.x;

    if (tid < cnt) {
        RefTetRec& R = tets[tid];
        const long long ia = start + tid;
        const long long t = m.tetids ? (long long)m.tetids[ia] : ia;
        const int4 nd = m.tets[t];
        const int n[4] = {nd.x, nd.y, nd.z, nd.w};
        double x[4], y[4], z[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const double4 X = m.xyz[n[v]];
            x[v] = X.x;
            y[v] = X.y;
            z[v] = X.z;
        }
        ref_coefficients(x, y, z, R.co, R.V);
        // point coordinates (curlcurl.py:319-321), left to right, unfused
        for (int q = 0; q < kNq; ++q) {
            R.p[q][0] = x[0] * kRefDpts[1][q] + x[1] * kRefDpts[2][q] + x[2] * kRefDpts[3][q] +
                        x[3] * kRefDpts[4][q];
            R.p[q][1] = y[0] * kRefDpts[1][q] + y[1] * kRefDpts[2][q] + y[2] * kRefDpts[3][q] +
                        y[3] * kRefDpts[4][q];
            R.p[q][2] = z[0] * kRefDpts[1][q] + z[1] * kRefDpts[2][q] + z[2] * kRefDpts[3][q] +
                        z[3] * kRefDpts[4][q];
        }
        // local maps (curlcurl.py:72-85): slot vertices in ascending global-id order
        for (int s = 0; s < 6; ++s) {
            int a = kEdgeSlot[s][0], b = kEdgeSlot[s][1];
            if (n[b] < n[a]) {
                const int tmp = a;
                a = b;
                b = tmp;
            }
            R.ijk[s][0] = R.ijk[s + 6][0] = (signed char)a;
            R.ijk[s][1] = R.ijk[s + 6][1] = (signed char)b;
            R.ijk[s][2] = R.ijk[s + 6][2] = 0;
        }
        for (int s = 0; s < 4; ++s) {
            int a = kFaceSlot[s][0], b = kFaceSlot[s][1], c = kFaceSlot[s][2];
            if (n[b] < n[a]) { const int tmp = a; a = b; b = tmp; }
            if (n[c] < n[b]) { const int tmp = b; b = c; c = tmp; }
            if (n[b] < n[a]) { const int tmp = a; a = b; b = tmp; }
            R.ijk[12 + s][0] = R.ijk[16 + s][0] = (signed char)a;
            R.ijk[12 + s][1] = R.ijk[16 + s][1] = (signed char)b;
            R.ijk[12 + s][2] = R.ijk[16 + s][2] = (signed char)c;
        }
        Cx E[9], G[9];
        ref_tensor(eps, t, E);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            R.mm[i][0] = E[i].re;
            R.mm[i][1] = E[i].im;
        }
        ref_tensor(mu, t, E);
        if (mu.invert) {
            ref_matinv(E, G);
        } else {
#pragma unroll
            for (int i = 0; i < 9; ++i) G[i] = E[i];
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            R.ms[i][0] = G[i].re;
            R.ms[i][1] = G[i].im;
        }
        R.ia = (int)ia;
    }

    // pair p -> (pi, pj), pi <= pj, row by row
    int pi = 0, pj = 0;
    const bool owner = tid < 210;
    {
        int p = owner ? tid : 0, row = 0;
        while (p >= 20 - row) {
            p -= 20 - row;
            ++row;
        }
        pi = row;
        pj = row + p;
    }
    double kre[kRefTile], kim[kRefTile], mre[kRefTile], mim[kRefTile];
#pragma unroll
    for (int lt = 0; lt < kRefTile; ++lt) kre[lt] = kim[lt] = mre[lt] = mim[lt] = 0.0;
    __syncthreads();

    for (int q = 0; q < kNq; ++q) {
        const double w = kRefDpts[0][q];
        for (int item = tid; item < cnt * 20; item += kRefBlock) {
            const int lt = item / 20, fn = item - 20 * lt;
            const RefTetRec& R = tets[lt];
            double f[3], cf[3];
            ref_basis(R.co, fn, R.ijk[fn][0], R.ijk[fn][1], R.ijk[fn][2], R.p[q][0], R.p[q][1],
                      R.p[q][2], f, cf);
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                pt.f[lt][fn][r] = f[r];
                pt.cf[lt][fn][r] = cf[r];
                mat_row_w(R.ms, r, cf, w, pt.msw[lt][fn][r]);
                mat_row_w(R.mm, r, f, w, pt.mmw[lt][fn][r]);
            }
        }
        __syncthreads();
        if (owner) {
#pragma unroll
            for (int lt = 0; lt < kRefTile; ++lt) {
                if (lt < cnt) {
                    // sum over c of the complex products with a real-valued complex:
                    // (a.re b, a.im b), summed left to right (curlcurl.py:358-370)
                    const double (*ms)[2] = pt.msw[lt][pi];
                    const double* c = pt.cf[lt][pj];
                    double sr = ms[0][0] * c[0], si = ms[0][1] * c[0];
                    sr = sr + ms[1][0] * c[1];
                    si = si + ms[1][1] * c[1];
                    sr = sr + ms[2][0] * c[2];
                    si = si + ms[2][1] * c[2];
                    kre[lt] = kre[lt] + sr;
                    kim[lt] = kim[lt] + si;
                    const double (*mm)[2] = pt.mmw[lt][pi];
                    const double* fv = pt.f[lt][pj];
                    double tr = mm[0][0] * fv[0], ti = mm[0][1] * fv[0];
                    tr = tr + mm[1][0] * fv[1];
                    ti = ti + mm[1][1] * fv[1];
                    tr = tr + mm[2][0] * fv[2];
                    ti = ti + mm[2][1] * fv[2];
                    mre[lt] = mre[lt] + tr;
                    mim[lt] = mim[lt] + ti;
                }
            }
        }
        __syncthreads();
    }

    if (!owner) return;
    const int e0 = 20 * pi + pj, e1 = 20 * pj + pi;
#pragma unroll
    for (int lt = 0; lt < kRefTile; ++lt) {
        if (lt >= cnt) break;
        const double V = tets[lt].V;
        // MatStiff * V, MatMass * V (curlcurl.py:378-379)
        const double Kr = kre[lt] * V, Ki = kim[lt] * V, Mr = mre[lt] * V, Mi = mim[lt] * V;
        const long long base = (long long)tets[lt].ia * 400;
        if (MODE == MODE_KM) {
            out0[base + e0] = make_double2(Kr, Ki);
            out0[base + e1] = make_double2(Kr, Ki);
            out1[base + e0] = make_double2(Mr, Mi);
            out1[base + e1] = make_double2(Mr, Mi);
        } else {
            // Evec - Bvec * K0**2 (assembler.py:494)
            for (int fq = 0; fq < nf; ++fq) {
                const double k2 = k0p.v[fq];
                const double2 a = make_double2(Kr - Mr * k2, Ki - Mi * k2);
                double2* o = out0 + (long long)fq * freq_stride + base;
                o[e0] = a;
                o[e1] = a;
            }
        }
    }
}

}  // namespace

hipError_t launch_assemble_ref(const MeshDev& m, const MatDesc& eps, const MatDesc& mu, int mode,
                               const double* k0sq, int nf, void* out0, void* out1,
                               long long freq_stride, hipStream_t stream) {
    if (m.n_assy == 0) return hipSuccess;
    const unsigned grid = (unsigned)((m.n_assy + kRefTile - 1) / kRefTile);
    if (mode == MODE_KM) {
        K0Pack p{};
        hipLaunchKernelGGL(ned2_ref_kernel<MODE_KM>, dim3(grid), dim3(kRefBlock), 0, stream, m,
                           eps, mu, p, 0, (double2*)out0, (double2*)out1, freq_stride);
        return hipGetLastError();
    }
    const int nchunk = (mode == MODE_AF) ? nf : 1;
    for (int f0 = 0; f0 < nchunk; f0 += kMaxK0) {
        K0Pack p;
        const int n = (mode == MODE_AF) ? min(kMaxK0, nf - f0) : 1;
        for (int f = 0; f < kMaxK0; ++f) p.v[f] = f < n ? k0sq[f0 + f] : 0.0;
        double2* o = (double2*)out0 + (long long)f0 * freq_stride;
        hipLaunchKernelGGL(ned2_ref_kernel<MODE_A>, dim3(grid), dim3(kRefBlock), 0, stream, m, eps,
                           mu, p, n, o, nullptr, freq_stride);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace em
