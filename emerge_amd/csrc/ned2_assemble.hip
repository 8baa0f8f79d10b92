// Nedelec-2 tetrahedral curl-curl assembly kernels for gfx950.
//
// Reference: src/emerge/_emerge/physics/microwave/assembly/curlcurl.py:285-456
// (ned2_tet_stiff_mass + _matrix_builder). The reference integrates every element with
// an 11-point rule in physical coordinates (~1.2e5 flop/tet). Here the same exact
// integrals are a contraction of per-tet metric tensors with constant reference-element
// tensors (ned2_tables.inc), which makes the kernel bound by its output stream.
//
// Per tet, with the 4 vertices put in ascending global-id order (the "canonical" order,
// in which every edge/face function has its vertices in local order):
//   J = [w1-w0, w2-w0, w3-w0],  C = rows (J1 x J2, J2 x J0, J0 x J1) = det(J) J^-1
//   G_M = C eps C^T / |det J|          (mass:      f = J^-T f_hat)
//   G_K = J^T mu^-1 J / |det J|        (stiffness: curl f = J curl f_hat / det J)
//   X(a,b) = sum_pq Gs_pq TSYM[ab][pq] -/+ sum_pq Ga_pq TANTI[ab][pq]
// then canonical function c goes to reference slot SLOT_OF[code][c], and entries with
// exactly one odd-degree function (codes 66/128/129 vs 64) take the factor
// s = sign(det of the stored-order Jacobian): the reference normalises barycentric
// coefficients by |V| (mth/optimized.py:142-167, curlcurl.py:306-311).
//
// Work layout: the plan groups assembly indices by vertex-rank code inside windows of
// consecutive tets and cuts each group into tiles of 32, 64 or 128 tets (autotuned), all
// of one code, interleaved so that the tiles of a window run on one XCD. One workgroup
// of 256 threads per tile. Phase 1: threads < tile compute one tet's metric tensors
// (FP64; materials gathered through strides) into LDS. Phase 2: threads 0..199 own two
// output entries of every 20x20 block, with the reference-tensor coefficients of their
// two canonical pairs in registers; per tet of the tile they read the metric tensors
// from LDS (broadcast), contract, and store their 16 (complex64) or 32 (complex128)
// bytes straight to the COO arrays with non-temporal 16-byte stores, contiguous across
// the wavefront. One barrier per tile; the output is not staged in LDS.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "em_internal.h"
#include "ned2_tables.inc"

namespace em {

constexpr int kBlock = 256;
constexpr int kTile = 64;

struct Cd {
    double re, im;
};
__device__ __forceinline__ Cd cmul(Cd a, Cd b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cd cadd(Cd a, Cd b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cd csub(Cd a, Cd b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cd cscale(Cd a, double s) { return {a.re * s, a.im * s}; }
__device__ __forceinline__ Cd cdiv(Cd a, Cd b) {
    double d = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// Tensor of one tet as a complex 3x3 (row-major).
__device__ __forceinline__ void load_tensor(const MatDesc& m, long long t, Cd E[9]) {
    const double2* p = m.p + t * m.ts;
    if (m.kind == EM_MAT_SCALAR) {
        double2 v = p[0];
#pragma unroll
        for (int i = 0; i < 9; ++i) E[i] = {0.0, 0.0};
        E[0] = E[4] = E[8] = {v.x, v.y};
    } else if (m.kind == EM_MAT_DIAG) {
#pragma unroll
        for (int i = 0; i < 9; ++i) E[i] = {0.0, 0.0};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double2 v = p[d * m.cs];
            E[4 * d] = {v.x, v.y};
        }
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            double2 v = p[i * m.cs];
            E[i] = {v.x, v.y};
        }
    }
}

// Inverse with the reference's semantics (mth/optimized.py:315-357): a tensor with
// all off-diagonal entries exactly zero is inverted entry-wise, otherwise by cofactors.
__device__ __forceinline__ void invert_tensor(Cd E[9]) {
    bool diag = true;
#pragma unroll
    for (int i = 0; i < 9; ++i)
        if (i != 0 && i != 4 && i != 8) diag = diag && (E[i].re == 0.0 && E[i].im == 0.0);
    const Cd one = {1.0, 0.0};
    if (diag) {
        E[0] = cdiv(one, E[0]);
        E[4] = cdiv(one, E[4]);
        E[8] = cdiv(one, E[8]);
        return;
    }
    Cd o[9];
    o[0] = csub(cmul(E[4], E[8]), cmul(E[5], E[7]));
    o[1] = csub(cmul(E[2], E[7]), cmul(E[1], E[8]));
    o[2] = csub(cmul(E[1], E[5]), cmul(E[2], E[4]));
    o[3] = csub(cmul(E[5], E[6]), cmul(E[3], E[8]));
    o[4] = csub(cmul(E[0], E[8]), cmul(E[2], E[6]));
    o[5] = csub(cmul(E[2], E[3]), cmul(E[0], E[5]));
    o[6] = csub(cmul(E[3], E[7]), cmul(E[4], E[6]));
    o[7] = csub(cmul(E[1], E[6]), cmul(E[0], E[7]));
    o[8] = csub(cmul(E[0], E[4]), cmul(E[1], E[3]));
    Cd det = cadd(cadd(cmul(E[0], o[0]), cmul(E[1], o[3])), cmul(E[2], o[6]));
#pragma unroll
    for (int i = 0; i < 9; ++i) E[i] = cdiv(o[i], det);
}

// G = A E A^T * scale for real 3x3 A (rows a0,a1,a2) and complex E (row-major).
__device__ __forceinline__ void congruence(const double A[3][3], const Cd E[9], int kind,
                                           double scale, Cd G[9]) {
    if (kind == EM_MAT_SCALAR) {
        // E = e I: G = e (A A^T) scale
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double d = (A[p][0] * A[q][0] + A[p][1] * A[q][1] + A[p][2] * A[q][2]) * scale;
                G[3 * p + q] = cscale(E[0], d);
            }
        return;
    }
    Cd AE[9];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            Cd acc = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; ++k) acc = cadd(acc, cscale(E[3 * k + c], A[p][k]));
            AE[3 * p + c] = acc;
        }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            Cd acc = {0.0, 0.0};
#pragma unroll
            for (int c = 0; c < 3; ++c) acc = cadd(acc, cscale(AE[3 * p + c], A[q][c]));
            G[3 * p + q] = cscale(acc, scale);
        }
}

template <typename R>
__device__ __forceinline__ void split_store(const Cd G[9], R (*sym)[2], R (*anti)[2]) {
    const Cd s01 = cscale(cadd(G[1], G[3]), 0.5), s02 = cscale(cadd(G[2], G[6]), 0.5),
             s12 = cscale(cadd(G[5], G[7]), 0.5);
    const Cd a01 = cscale(csub(G[1], G[3]), 0.5), a02 = cscale(csub(G[2], G[6]), 0.5),
             a12 = cscale(csub(G[5], G[7]), 0.5);
    const Cd sv[6] = {G[0], G[4], G[8], s01, s02, s12};
    const Cd av[3] = {a01, a02, a12};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        sym[c][0] = (R)sv[c].re;
        sym[c][1] = (R)sv[c].im;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        anti[c][0] = (R)av[c].re;
        anti[c][1] = (R)av[c].im;
    }
}

// Diagonal (or scalar) tensor of one tet as 3 complex values.
__device__ __forceinline__ void load_diag(const MatDesc& m, long long t, Cd e[3]) {
    const double2* p = m.p + t * m.ts;
    if (m.kind == EM_MAT_SCALAR) {
        const double2 v = p[0];
        e[0] = e[1] = e[2] = {v.x, v.y};
    } else {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double2 v = p[d * m.cs];
            e[d] = {v.x, v.y};
        }
    }
}

// Symmetric components (00, 11, 22, 01, 02, 12) of A diag(e) A^T * scale.
template <typename R>
__device__ __forceinline__ void diag_congruence_store(const double A[3][3], const Cd e[3],
                                                      double scale, R (*sym)[2]) {
    constexpr int P[6] = {0, 1, 2, 0, 0, 1}, Q[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double w = A[P[c]][k] * A[Q[c]][k];
            re += w * e[k].re;
            im += w * e[k].im;
        }
        sym[c][0] = (R)(re * scale);
        sym[c][1] = (R)(im * scale);
    }
}

// Metric tensors of one tet (FP64 throughout): writes the symmetric (and, when FULL,
// antisymmetric) parts of G_K and of mscale * G_M, and returns the vertex-rank code and
// the orientation sign.
template <typename R, bool FULL>
__device__ __forceinline__ void tet_geometry(const MeshDev& m, const MatDesc& eps,
                                             const MatDesc& mu, long long t, double mscale,
                                             R (*gk)[2], R (*gm)[2], R (*ak)[2], R (*am)[2],
                                             int& code, R& sgn) {
    const int4 nd = m.tets[t];
    const int n[4] = {nd.x, nd.y, nd.z, nd.w};
    int r[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        r[a] = (n[0] < n[a]) + (n[1] < n[a]) + (n[2] < n[a]) + (n[3] < n[a]);
    // canonical vertices: the node ids in ascending order (5-comparator network), so
    // the coordinates are gathered already in canonical order
    const int s0 = min(n[0], n[1]), s1 = max(n[0], n[1]);
    const int s2 = min(n[2], n[3]), s3 = max(n[2], n[3]);
    const int w0 = min(s0, s2), t2 = max(s0, s2);
    const int t1 = min(s1, s3), w3 = max(s1, s3);
    const int w1 = min(t1, t2), w2 = max(t1, t2);
    const double4 X0 = m.xyz[w0], X1 = m.xyz[w1], X2 = m.xyz[w2], X3 = m.xyz[w3];
    // Jc[q] = column q of J = w_{q+1} - w_0
    const double Jc[3][3] = {{X1.x - X0.x, X1.y - X0.y, X1.z - X0.z},
                             {X2.x - X0.x, X2.y - X0.y, X2.z - X0.z},
                             {X3.x - X0.x, X3.y - X0.y, X3.z - X0.z}};
    // rows of C = det(J) J^-1: J1 x J2, J2 x J0, J0 x J1
    const double C[3][3] = {
        {Jc[1][1] * Jc[2][2] - Jc[1][2] * Jc[2][1], Jc[1][2] * Jc[2][0] - Jc[1][0] * Jc[2][2],
         Jc[1][0] * Jc[2][1] - Jc[1][1] * Jc[2][0]},
        {Jc[2][1] * Jc[0][2] - Jc[2][2] * Jc[0][1], Jc[2][2] * Jc[0][0] - Jc[2][0] * Jc[0][2],
         Jc[2][0] * Jc[0][1] - Jc[2][1] * Jc[0][0]},
        {Jc[0][1] * Jc[1][2] - Jc[0][2] * Jc[1][1], Jc[0][2] * Jc[1][0] - Jc[0][0] * Jc[1][2],
         Jc[0][0] * Jc[1][1] - Jc[0][1] * Jc[1][0]}};
    const double det = Jc[0][0] * C[0][0] + Jc[0][1] * C[0][1] + Jc[0][2] * C[0][2];
    const double inv_adet = 1.0 / fabs(det);
    code = r[0] * 16 + r[1] * 4 + r[2];
    const int inv = (r[0] > r[1]) + (r[0] > r[2]) + (r[0] > r[3]) + (r[1] > r[2]) +
                    (r[1] > r[3]) + (r[2] > r[3]);
    sgn = ((det < 0.0) != (inv & 1)) ? (R)-1 : (R)1;

    if constexpr (!FULL) {
        Cd e[3];
        load_diag(eps, t, e);
        diag_congruence_store<R>(C, e, inv_adet * mscale, gm);
        load_diag(mu, t, e);
        if (mu.invert) {
            const Cd one = {1.0, 0.0};
#pragma unroll
            for (int d = 0; d < 3; ++d) e[d] = cdiv(one, e[d]);
        }
        diag_congruence_store<R>(Jc, e, inv_adet, gk);
    } else {
        Cd E[9], G[9];
        load_tensor(eps, t, E);
        congruence(C, E, eps.kind, inv_adet * mscale, G);
        split_store<R>(G, gm, am);
        load_tensor(mu, t, E);
        if (mu.invert) invert_tensor(E);
        congruence(Jc, E, mu.kind, inv_adet, G);
        split_store<R>(G, gk, ak);
    }
}

template <typename R>
struct OutVec;
template <>
struct OutVec<double> {
    using T = double2;
};
template <>
struct OutVec<float> {
    using T = float2;
};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// LDS record of a tile's tets: metric tensors, orientation sign, assembly index.
template <typename R, bool FULL, int TILE>
struct StreamRec {
    R gk[TILE][6][2];
    R gm[TILE][6][2];
    R ak[FULL ? TILE : 1][3][2];
    R am[FULL ? TILE : 1][3][2];
    R sgn[TILE];
    int ia[TILE];
};

// Store a thread's two entries of one block with 16-byte streaming stores: complex64
// entries e0 and e1 = e0 + 1 together, complex128 entries e0 and e1 separately.
__device__ __forceinline__ void store2(float2* blk, int e0, int e1, float v0r, float v0i,
                                       float v1r, float v1i) {
    (void)e1;
    const f32x4 v = {v0r, v0i, v1r, v1i};
    __builtin_nontemporal_store(v, reinterpret_cast<f32x4*>(blk + e0));
}
__device__ __forceinline__ void store2(double2* blk, int e0, int e1, double v0r, double v0i,
                                       double v1r, double v1i) {
    const f64x2 a = {v0r, v0i}, b = {v1r, v1i};
    __builtin_nontemporal_store(a, reinterpret_cast<f64x2*>(blk + e0));
    __builtin_nontemporal_store(b, reinterpret_cast<f64x2*>(blk + e1));
}

// Streams a tile's blocks (the thread's two entries of every tet). KRE / MRE: the imaginary
// parts of every K / M metric tensor of the tile are zero (a real mu_r / eps_r), so their
// products (all +-0, which leave a +0 accumulator at +0) are skipped; the result is the
// same bit for bit.
// A's entry from the K and M parts of one slot (G_M carries -k0^2), added without
// contraction so that every kernel forming it rounds the same way.
template <typename R>
__device__ __forceinline__ R add_km(R k, R m) {
#pragma clang fp contract(off)
    return k + m;
}

// K and M entries of one output slot of one tet: the slot's canonical-pair coefficients
// (tk, tm; qk, qm for the antisymmetric parts) contracted with the tet's metric tensors,
// times sg (the orientation sign when the pair joins a code-64 and an odd-degree
// function, else 1). KRE / MRE: the imaginary parts are known to be zero and their
// products are skipped (they are +-0 and leave a +0 accumulator at +0: same bits).
template <typename R, bool FULL, bool KRE, bool MRE>
__device__ __forceinline__ void pair_entry(const R (&tk)[6], const R (&tm)[6], const R (&qk)[3],
                                           const R (&qm)[3], const R (&gkr)[6],
                                           const R (&gki)[6], const R (&gmr)[6],
                                           const R (&gmi)[6], const R (*ak)[2],
                                           const R (*am)[2], R sg, R& kr, R& ki, R& mr,
                                           R& mi) {
    R a = 0, b = 0, c2 = 0, d = 0;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        a += tk[c] * gkr[c];
        if constexpr (!KRE) b += tk[c] * gki[c];
        c2 += tm[c] * gmr[c];
        if constexpr (!MRE) d += tm[c] * gmi[c];
    }
    if constexpr (FULL) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a += qk[c] * ak[c][0];
            if constexpr (!KRE) b += qk[c] * ak[c][1];
            c2 += qm[c] * am[c][0];
            if constexpr (!MRE) d += qm[c] * am[c][1];
        }
    }
    kr = sg * a;
    ki = sg * b;
    mr = sg * c2;
    mi = sg * d;
}

template <typename R, int MODE, bool FULL, int TILE, bool KRE, bool MRE>
__device__ __forceinline__ void stream_tile(const StreamRec<R, FULL, TILE>& rec, int cnt,
                                            const int (&E)[2], const R (&tk)[2][6],
                                            const R (&tm)[2][6], const R (&qk)[2][3],
                                            const R (&qm)[2][3], const bool (&odd)[2],
                                            const K0Pack& k0p, int nf, void* __restrict__ out0,
                                            void* __restrict__ out1, long long freq_stride) {
    using O = typename OutVec<R>::T;
    for (int lt = 0; lt < cnt; ++lt) {
        R gkr[6], gki[6], gmr[6], gmi[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            gkr[c] = rec.gk[lt][c][0];
            gki[c] = KRE ? (R)0 : rec.gk[lt][c][1];
            gmr[c] = rec.gm[lt][c][0];
            gmi[c] = MRE ? (R)0 : rec.gm[lt][c][1];
        }
        const R s = rec.sgn[lt];
        const long long base = (long long)rec.ia[lt] * 400;
        R kr[2], ki[2], mr[2], mi[2];
#pragma unroll
        for (int q = 0; q < 2; ++q)
            pair_entry<R, FULL, KRE, MRE>(tk[q], tm[q], qk[q], qm[q], gkr, gki, gmr, gmi,
                                          rec.ak[FULL ? lt : 0], rec.am[FULL ? lt : 0],
                                          odd[q] ? s : (R)1, kr[q], ki[q], mr[q], mi[q]);
        if (MODE == MODE_A) {
            // G_M already carries -k0^2
            store2(reinterpret_cast<O*>(out0) + base, E[0], E[1], add_km(kr[0], mr[0]),
                   add_km(ki[0], mi[0]), add_km(kr[1], mr[1]), add_km(ki[1], mi[1]));
        } else if (MODE == MODE_KM) {
            store2(reinterpret_cast<O*>(out0) + base, E[0], E[1], kr[0], ki[0], kr[1], ki[1]);
            store2(reinterpret_cast<O*>(out1) + base, E[0], E[1], mr[0], mi[0], mr[1], mi[1]);
        } else {
            for (int f = 0; f < nf; ++f) {
                const R kf = (R)k0p.v[f];
                store2(reinterpret_cast<O*>(out0) + (long long)f * freq_stride + base, E[0], E[1],
                       kr[0] - kf * mr[0], ki[0] - kf * mi[0], kr[1] - kf * mr[1],
                       ki[1] - kf * mi[1]);
            }
        }
    }
}

// Coefficients of slot (ii, jj) alone (as slot_coefs for one entry).
template <typename R, bool FULL, typename Tab>
__device__ __forceinline__ void slot_coefs1(const Tab& tab, int code, int ii, int jj,
                                            R (&tk)[6], R (&tm)[6], R (&qk)[3], R (&qm)[3]) {
    const int ca = tab.canon(code, ii), cb = tab.canon(code, jj);
    const int p = tab.pair(ca, cb);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        tk[c] = (R)tab.sk(p, c);
        tm[c] = (R)tab.sm(p, c);
    }
    const int slot_lo = (ca <= cb) ? ii : jj, slot_hi = (ca <= cb) ? jj : ii;
    const R sa = (slot_lo < slot_hi) ? (R)-1 : (R)1;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        qk[c] = FULL ? sa * (R)tab.ak(p, c) : (R)0;
        qm[c] = FULL ? sa * (R)tab.am(p, c) : (R)0;
    }
}

// One workgroup per tile: up to TILE tets that share one vertex-rank code (so every
// output slot maps to the same canonical pair for all of them). Threads 0..199 own two
// output entries of every tet's 20x20 block, with the reference-tensor coefficients of
// their two canonical pairs in registers; they read each tet's metric tensors from LDS
// (broadcast) and stream their 16 bytes (complex64) or 32 bytes (complex128) straight to
// the COO arrays: no LDS staging, one barrier per tile.
template <typename R, int MODE, bool FULL, int TILE>
__global__ __launch_bounds__(kBlock) void ned2_assemble_kernel(
    MeshDev m, MatDesc eps, MatDesc mu, K0Pack k0p, int nf, const int* __restrict__ order,
    const int2* __restrict__ tiles, void* __restrict__ out0, void* __restrict__ out1,
    long long freq_stride) {
    __shared__ StreamRec<R, FULL, TILE> rec;
    const int2 td = tiles[blockIdx.x];
    const int start = td.x, cnt = td.y & 0xffff, code = td.y >> 16;
    if (cnt == 0) return;  // padding tile (uniform across the workgroup)
    const double mscale = (MODE == MODE_A) ? -k0p.v[0] : 1.0;

    int kre = 1, mre = 1;  // this thread's tets have real K / M metric tensors
    for (int lt = threadIdx.x; lt < cnt; lt += kBlock) {
        const int ia = order[start + lt];
        const long long t = m.tetids ? (long long)m.tetids[ia] : (long long)ia;
        int c;
        tet_geometry<R, FULL>(m, eps, mu, t, mscale, rec.gk[lt], rec.gm[lt],
                              rec.ak[FULL ? lt : 0], rec.am[FULL ? lt : 0], c, rec.sgn[lt]);
        rec.ia[lt] = ia;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            kre &= rec.gk[lt][k][1] == (R)0;
            mre &= rec.gm[lt][k][1] == (R)0;
        }
        if constexpr (FULL) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                kre &= rec.ak[lt][k][1] == (R)0;
                mre &= rec.am[lt][k][1] == (R)0;
            }
        }
    }

    // the two output slots of this thread and their coefficients: complex64 entries go
    // in adjacent pairs (one 16-byte store), complex128 entries i and i + 200 (two
    // 16-byte stores, each contiguous across the wavefront)
    const bool active = threadIdx.x < 200;
    const int ti = active ? (int)threadIdx.x : 0;
    constexpr bool kPaired = sizeof(R) == 4;
    const int E[2] = {kPaired ? 2 * ti : ti, kPaired ? 2 * ti + 1 : ti + 200};
    R tk[2][6], tm[2][6], qk[2][3], qm[2][3];
    bool odd[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int ii = E[q] / 20, jj = E[q] % 20;
        const int ca = CANON[code][ii], cb = CANON[code][jj];
        const int p = PAIR_INDEX[ca][cb];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            tk[q][c] = (R)TSYM_K[p][c];
            tm[q][c] = (R)TSYM_M[p][c];
        }
        // antisymmetric sign: -1 when the pair's lower canonical function sits in the
        // lower reference slot
        const int slot_lo = (ca <= cb) ? ii : jj, slot_hi = (ca <= cb) ? jj : ii;
        const R sa = (slot_lo < slot_hi) ? (R)-1 : (R)1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            qk[q][c] = FULL ? sa * (R)TANTI_K[p][c] : (R)0;
            qm[q][c] = FULL ? sa * (R)TANTI_M[p][c] : (R)0;
        }
        odd[q] = (ii >= 6) != (jj >= 6);
    }
    // the barrier before streaming, with the tile's reality flags (uniform afterwards)
    const bool kreal = __syncthreads_and(kre) != 0;
    const bool mreal = __syncthreads_and(mre) != 0;
    if (!active) return;
#define EM_STREAM(KR, MR)                                                                     \
    stream_tile<R, MODE, FULL, TILE, KR, MR>(rec, cnt, E, tk, tm, qk, qm, odd, k0p, nf, out0, \
                                            out1, freq_stride)
    if (kreal) {
        if (mreal)
            EM_STREAM(true, true);
        else
            EM_STREAM(true, false);
    } else {
        if (mreal)
            EM_STREAM(false, true);
        else
            EM_STREAM(false, false);
    }
#undef EM_STREAM
}

// Folds ld(k) for k in [lo, hi) (grid-stride from i0) into one word, four loads in
// flight per thread: indices past hi load element hi - 1 (in bounds, as k0 < hi) and are
// masked out, so the loads need no branch and no wait between them.
template <typename Load>
__device__ __forceinline__ unsigned long long fold4(long long lo, long long hi, long long i0,
                                                   long long stride, Load ld) {
    unsigned long long x = 0;
    for (long long k0 = lo + i0; k0 < hi; k0 += 4 * stride) {
        unsigned long long v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const long long k = k0 + u * stride;
            v[u] = ld(min(k, hi - 1)) & (k < hi ? ~0ull : 0ull);
        }
        x ^= v[0] ^ v[1] ^ v[2] ^ v[3];
    }
    return x;
}

// Material values of tets [t0, t1) as the assembly kernel reads them (load_tensor).
__device__ __forceinline__ unsigned long long touch_material(const MatDesc& d, long long t0,
                                                            long long t1, long long i0,
                                                            long long stride) {
    unsigned long long x = 0;
    const int nc = d.kind == EM_MAT_SCALAR ? 1 : d.kind == EM_MAT_DIAG ? 3 : 9;
    if (d.ts == 0) t1 = min(t1, t0 + 1);  // one tensor for every tet
    for (int c = 0; c < nc; ++c) {
        const double2* p = d.p + (long long)c * d.cs;
        const long long ts = d.ts;
        x ^= fold4(t0, t1, i0, stride, [=](long long t) {
            return (unsigned long long)__double_as_longlong(p[t * ts].x);
        });
    }
    return x;
}

// Reads the inputs of one chunk of tiles so that the assembly launch right after it
// finds them in the memory-side cache instead of DRAM: the node coordinates (first chunk
// only), order positions and tetids [ia_lo, ia_hi), tet node ids and material values of
// tets [t_lo, t_hi). Grid-stride; the values are folded into one word that is stored
// only if it equals `key`, which keeps the loads without writing anything in practice.
__global__ __launch_bounds__(256) void prefetch_chunk_kernel(MeshDev m, long long n_nodes,
                                                             const int* __restrict__ order,
                                                             PrefetchChunk ch, MatDesc eps,
                                                             MatDesc mu, unsigned long long key,
                                                             unsigned long long* sink) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const double4* xyz = m.xyz;
    const int4* tets = m.tets;
    const int* ids = m.tetids;
    unsigned long long x = fold4(0, n_nodes, i0, stride, [=](long long j) {
        return (unsigned long long)__double_as_longlong(xyz[j].x);
    });
    x ^= fold4(ch.ia_lo, ch.ia_hi, i0, stride,
               [=](long long j) { return (unsigned long long)(unsigned)order[j]; });
    if (ids)
        x ^= fold4(ch.ia_lo, ch.ia_hi, i0, stride,
                   [=](long long j) { return (unsigned long long)(unsigned)ids[j] << 32; });
    x ^= fold4(ch.t_lo, ch.t_hi, i0, stride,
               [=](long long j) { return (unsigned long long)(unsigned)tets[j].x; });
    x ^= touch_material(eps, ch.t_lo, ch.t_hi, i0, stride);
    x ^= touch_material(mu, ch.t_lo, ch.t_hi, i0, stride) << 1;
    if (x == key) *sink = x;
}

// Launch over all tiles; MODE_AF launches once per chunk of up to kMaxK0 wavenumbers
// (passed by value, so the launch needs no device buffer and can be graph-captured).
// With prefetch chunks (ts.chunks) each slice of the tile list is its own launch, preceded
// by prefetch_chunk_kernel over the inputs of its tiles; every tile computes exactly what
// it computes in the single launch, so the outputs are the same bit for bit.
template <typename R, int MODE, bool FULL, int TILE>
static hipError_t launch_t(const MeshDev& m, const TileSet& ts, const MatDesc& eps,
                           const MatDesc& mu, const double* k0sq, int nf, void* out0, void* out1,
                           long long freq_stride, hipStream_t stream) {
    using O = typename OutVec<R>::T;
    if (ts.ntiles == 0) return hipSuccess;
    const int nchunk = (MODE == MODE_AF) ? nf : 1;
    for (int f0 = 0; f0 < nchunk; f0 += kMaxK0) {
        K0Pack p;
        const int n = (MODE == MODE_AF) ? min(kMaxK0, nf - f0) : 1;
        for (int f = 0; f < kMaxK0; ++f) p.v[f] = (k0sq && f < n) ? k0sq[f0 + f] : 0.0;
        void* o0 = reinterpret_cast<O*>(out0) + (long long)f0 * freq_stride;
        if (!ts.chunks) {
            hipLaunchKernelGGL((ned2_assemble_kernel<R, MODE, FULL, TILE>),
                               dim3((unsigned)ts.ntiles), dim3(kBlock), 0, stream, m, eps, mu, p,
                               n, ts.order, ts.tiles, o0, out1, freq_stride);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            continue;
        }
        for (int c = 0; c < ts.nchunks; ++c) {
            const PrefetchChunk& ch = ts.chunks[c];
            if (ch.tile_count == 0) continue;
            const long long span = max(max(ch.t_hi - ch.t_lo, ch.ia_hi - ch.ia_lo),
                                       c == 0 ? ts.n_nodes : 0ll);
            // four elements per thread and array: every load of the chunk in flight at once
            const unsigned pgrid = (unsigned)min(8192ll, max(1ll, (span + 1023) / 1024));
            hipLaunchKernelGGL(prefetch_chunk_kernel, dim3(pgrid), dim3(256), 0, stream, m,
                               c == 0 ? ts.n_nodes : 0ll, ts.order, ch, eps, mu,
                               0x9e3779b97f4a7c15ull, ts.sink);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL((ned2_assemble_kernel<R, MODE, FULL, TILE>),
                               dim3((unsigned)ch.tile_count), dim3(kBlock), 0, stream, m, eps, mu,
                               p, n, ts.order, ts.tiles + ch.tile_begin, o0, out1, freq_stride);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    return hipSuccess;
}

template <typename R, int MODE, bool FULL>
static hipError_t launch_tile(const MeshDev& m, const TileSet* ts, int tile, const MatDesc& eps,
                              const MatDesc& mu, const double* k0sq, int nf, void* out0,
                              void* out1, long long freq_stride, hipStream_t stream) {
    if (tile == 32)
        return launch_t<R, MODE, FULL, 32>(m, ts[0], eps, mu, k0sq, nf, out0, out1, freq_stride,
                                           stream);
    if (tile == 128)
        return launch_t<R, MODE, FULL, 128>(m, ts[2], eps, mu, k0sq, nf, out0, out1, freq_stride,
                                            stream);
    return launch_t<R, MODE, FULL, 64>(m, ts[1], eps, mu, k0sq, nf, out0, out1, freq_stride,
                                       stream);
}

template <typename R, int MODE>
static hipError_t launch_full(const MeshDev& m, const TileSet* ts, int tile, const MatDesc& eps,
                              const MatDesc& mu, int full, const double* k0sq, int nf, void* out0,
                              void* out1, long long freq_stride, hipStream_t stream) {
    if (full)
        return launch_tile<R, MODE, true>(m, ts, tile, eps, mu, k0sq, nf, out0, out1, freq_stride,
                                          stream);
    return launch_tile<R, MODE, false>(m, ts, tile, eps, mu, k0sq, nf, out0, out1, freq_stride,
                                       stream);
}

template <typename R>
static hipError_t launch_mode(const MeshDev& m, const TileSet* ts, int tile, const MatDesc& eps,
                              const MatDesc& mu, int mode, int full, const double* k0sq, int nf,
                              void* out0, void* out1, long long freq_stride, hipStream_t stream) {
    if (mode == MODE_A)
        return launch_full<R, MODE_A>(m, ts, tile, eps, mu, full, k0sq, 1, out0, out1,
                                      freq_stride, stream);
    if (mode == MODE_KM)
        return launch_full<R, MODE_KM>(m, ts, tile, eps, mu, full, nullptr, 0, out0, out1,
                                       freq_stride, stream);
    return launch_full<R, MODE_AF>(m, ts, tile, eps, mu, full, k0sq, nf, out0, out1, freq_stride,
                                   stream);
}

hipError_t launch_assemble(const MeshDev& m, const TileSet* ts, int tile, const MatDesc& eps,
                           const MatDesc& mu, int mode, int precision, int full,
                           const double* k0sq, int nf, void* out0, void* out1,
                           long long freq_stride, hipStream_t stream) {
    if (precision == EM_PREC_FP64)
        return launch_mode<double>(m, ts, tile, eps, mu, mode, full, k0sq, nf, out0, out1,
                                   freq_stride, stream);
    return launch_mode<float>(m, ts, tile, eps, mu, mode, full, k0sq, nf, out0, out1, freq_stride,
                              stream);
}

// ---------------------------------------------------------------------------
// Direct CSC assembly: the CSC values of A (or of K, or of M) without the COO arrays.
//
// The reference converts the COO blocks with to_csc (csc_cast.py:55-63, assembler.py:494);
// the column-incidence gather of csc_map.hip reads 8-16 B per contributor that the
// assembly kernel wrote just before. Here the contributors are computed where they are
// summed: a first kernel writes each tet's metric tensors, orientation sign and
// vertex-rank code (kRecLen values of R, 128-320 B), a second walks the columns as the
// incidence gather does (three columns per wavefront, one group of 20 lanes each; lane ii
// takes row slot ii of every incident (tet, slot jj) in incidence order = ascending COO
// index) and computes entry (ii, jj) of the tet's block from its record and the
// reference tensors (an LDS copy), with the assembly kernel's arithmetic (pair_entry), so
// the values equal the COO blocks' bit for bit and their FP64 sums equal to_csc of those
// blocks bit for bit. The block entry is computed directly, so symmetric and
// non-symmetric blocks take the same path. Per incidence: 4 B of incidence, 40 B of row
// positions and one record (mostly from L2: the 20 columns of a tet's DOFs read it)
// instead of 20 contributors.
// ---------------------------------------------------------------------------
template <bool FULL>
struct RecLayout {
    static constexpr int kLen = FULL ? 40 : 32;  // values of R per tet (padded)
    static constexpr int kGk = 0, kGm = 12, kAk = 24, kAm = 30;
    static constexpr int kSgn = kLen - 2, kCode = kLen - 1;
};

template <typename R, bool FULL>
__global__ __launch_bounds__(256) void ned2_metric_kernel(MeshDev m, MatDesc eps, MatDesc mu,
                                                          double mscale, R* __restrict__ rec) {
    using L = RecLayout<FULL>;
    const long long ia = (long long)blockIdx.x * 256 + threadIdx.x;
    if (ia >= m.n_assy) return;
    const long long t = m.tetids ? (long long)m.tetids[ia] : ia;
    R v[L::kLen];
#pragma unroll
    for (int k = 0; k < L::kLen; ++k) v[k] = (R)0;
    int code;
    tet_geometry<R, FULL>(m, eps, mu, t, mscale, reinterpret_cast<R(*)[2]>(v + L::kGk),
                          reinterpret_cast<R(*)[2]>(v + L::kGm),
                          reinterpret_cast<R(*)[2]>(v + L::kAk),
                          reinterpret_cast<R(*)[2]>(v + L::kAm), code, v[L::kSgn]);
    v[L::kCode] = (R)code;
    typedef R vec4 __attribute__((ext_vector_type(16 / sizeof(R))));
    constexpr int kPer = 16 / sizeof(R);
    vec4* dst = reinterpret_cast<vec4*>(rec + ia * L::kLen);
#pragma unroll
    for (int k = 0; k < L::kLen / kPer; ++k) {
        vec4 w;
#pragma unroll
        for (int u = 0; u < kPer; ++u) w[u] = v[k * kPer + u];
        dst[k] = w;
    }
}

// LDS copy of the reference-tensor tables for the direct CSC kernel.
template <typename R, bool FULL>
struct CoefLds {
    R sk[NED2_NPAIR][6];
    R sm[NED2_NPAIR][6];
    R ak[FULL ? NED2_NPAIR : 1][3];
    R am[FULL ? NED2_NPAIR : 1][3];
    unsigned char canon[64][20];
    unsigned char pidx[20][20];
};

template <typename R, bool FULL>
struct LdsTables {
    const CoefLds<R, FULL>& t;
    __device__ int canon(int code, int slot) const { return t.canon[code][slot]; }
    __device__ int pair(int a, int b) const { return t.pidx[a][b]; }
    __device__ R sk(int p, int c) const { return t.sk[p][c]; }
    __device__ R sm(int p, int c) const { return t.sm[p][c]; }
    __device__ R ak(int p, int c) const { return t.ak[FULL ? p : 0][c]; }
    __device__ R am(int p, int c) const { return t.am[FULL ? p : 0][c]; }
};

template <typename R, bool FULL>
__device__ __forceinline__ void fill_coef_lds(CoefLds<R, FULL>& cf) {
    for (int i = threadIdx.x; i < NED2_NPAIR * 6; i += blockDim.x) {
        cf.sk[i / 6][i % 6] = (R)TSYM_K[i / 6][i % 6];
        cf.sm[i / 6][i % 6] = (R)TSYM_M[i / 6][i % 6];
    }
    if constexpr (FULL)
        for (int i = threadIdx.x; i < NED2_NPAIR * 3; i += blockDim.x) {
            cf.ak[i / 3][i % 3] = (R)TANTI_K[i / 3][i % 3];
            cf.am[i / 3][i % 3] = (R)TANTI_M[i / 3][i % 3];
        }
    for (int i = threadIdx.x; i < 64 * 20; i += blockDim.x)
        cf.canon[i / 20][i % 20] = CANON[i / 20][i % 20];
    for (int i = threadIdx.x; i < 20 * 20; i += blockDim.x)
        cf.pidx[i / 20][i % 20] = PAIR_INDEX[i / 20][i % 20];
}

constexpr int kCscWaves = 4;   // wavefronts per workgroup
constexpr int kCscCols = 3;    // columns per wavefront (groups of 20 lanes)
constexpr int kCscCap = 128;   // entries per column window in LDS
constexpr long long kCscGrid = 2048;  // workgroups (8 per CU)

// Incidences per batch: a group stages the records of kCscBatch incidences in LDS with
// all their loads in flight at once, then computes from LDS.
template <typename R>
struct CscBatch {
    static constexpr int kN = sizeof(R) == 4 ? 8 : 4;
};

// WHICH: 0 = A (records built with G_M scaled by -k0^2), 1 = K, 2 = M.
template <typename R, int WHICH, bool FULL>
__global__ __launch_bounds__(64 * kCscWaves) void ned2_csc_kernel(
    const long long* __restrict__ indptr, const long long* __restrict__ inc_ptr,
    const int* __restrict__ inc, const unsigned short* __restrict__ pos,
    const long long* __restrict__ cols, long long N, const R* __restrict__ rec,
    typename OutVec<R>::T* __restrict__ out) {
    using L = RecLayout<FULL>;
    constexpr int kB = CscBatch<R>::kN;
    typedef R vec4 __attribute__((ext_vector_type(16 / sizeof(R))));
    constexpr int kChunks = L::kLen * (int)sizeof(R) / 16;  // 16-byte pieces per record
    __shared__ double acc[kCscWaves * kCscCols][kCscCap][2];
    __shared__ vec4 stage[kCscWaves * kCscCols][kB * kChunks];
    __shared__ CoefLds<R, FULL> cf;
    fill_coef_lds<R, FULL>(cf);
    __syncthreads();  // the only workgroup barrier: tables ready
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane / 20, ii = lane - 20 * grp;
    if (grp >= kCscCols) return;  // lanes 60..63: nothing to do
    const LdsTables<R, FULL> tab{cf};
    double (*a)[2] = acc[wave * kCscCols + grp];
    vec4* st = stage[wave * kCscCols + grp];
    const R* sr = reinterpret_cast<const R*>(st);
    const vec4* recv = reinterpret_cast<const vec4*>(rec);
    const long long stride = (long long)gridDim.x * kCscWaves * kCscCols;
    for (long long c = ((long long)blockIdx.x * kCscWaves + wave) * kCscCols + grp; c < N;
         c += stride) {
        const long long cb = indptr[c], len = indptr[c + 1] - cb;
        const long long fc = cols ? cols[c] : c;  // column of the incidence lists
        const long long s = inc_ptr[fc], ninc = inc_ptr[fc + 1] - s;
        for (long long w0 = 0; w0 < len; w0 += kCscCap) {
            const int wl = (int)(len - w0 < kCscCap ? len - w0 : kCscCap);
            for (int p = ii; p < wl; p += 20) {
                a[p][0] = 0.0;
                a[p][1] = 0.0;
            }
            wave_lds_sync();  // zeroed accumulators before any lane adds
            for (long long i0 = 0; i0 < ninc; i0 += kB) {
                const int nb = (int)(ninc - i0 < kB ? ninc - i0 : kB);
                // the batch's incidences (lane ii < nb loads incidence ii), shared by the
                // group's lanes through the wavefront
                const int my_e = ii < nb ? inc[s + i0 + ii] : 0;
                int e[kB];
#pragma unroll
                for (int u = 0; u < kB; ++u) e[u] = __shfl(my_e, grp * 20 + u);
                // every record piece of the batch and the lane's row positions, all in flight
                vec4 piece[(kB * kChunks + 19) / 20];
#pragma unroll
                for (int q = 0; q < (kB * kChunks + 19) / 20; ++q) {
                    const int k = ii + 20 * q, u = k / kChunks;
                    if (k < kB * kChunks && u < nb)
                        piece[q] = recv[(long long)(e[u] / 20) * kChunks + (k - u * kChunks)];
                }
                int p[kB];
#pragma unroll
                for (int u = 0; u < kB; ++u)
                    p[u] = u < nb ? (int)pos[20 * (s + i0 + u) + ii] - (int)w0 : -1;
#pragma unroll
                for (int q = 0; q < (kB * kChunks + 19) / 20; ++q) {
                    const int k = ii + 20 * q, u = k / kChunks;
                    if (k < kB * kChunks && u < nb) st[k] = piece[q];
                }
                wave_lds_sync();  // staged records visible to the group
                for (int u = 0; u < nb; ++u) {
                    const int jj = e[u] % 20;
                    const R* r = sr + u * L::kLen;
                    R gkr[6], gki[6], gmr[6], gmi[6], ak[3][2], am[3][2];
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        gkr[k] = r[L::kGk + 2 * k];
                        gki[k] = r[L::kGk + 2 * k + 1];
                        gmr[k] = r[L::kGm + 2 * k];
                        gmi[k] = r[L::kGm + 2 * k + 1];
                    }
                    if constexpr (FULL) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            ak[k][0] = r[L::kAk + 2 * k];
                            ak[k][1] = r[L::kAk + 2 * k + 1];
                            am[k][0] = r[L::kAm + 2 * k];
                            am[k][1] = r[L::kAm + 2 * k + 1];
                        }
                    }
                    const R sgn = r[L::kSgn];
                    const int code = (int)r[L::kCode];
                    R tk[6], tm[6], qk[3], qm[3];
                    slot_coefs1<R, FULL>(tab, code, ii, jj, tk, tm, qk, qm);
                    const bool odd = (ii >= 6) != (jj >= 6);
                    R kr, ki, mr, mi;
                    pair_entry<R, FULL, false, false>(tk, tm, qk, qm, gkr, gki, gmr, gmi, ak, am,
                                                      odd ? sgn : (R)1, kr, ki, mr, mi);
                    R vr, vi;
                    if (WHICH == 0) {
                        vr = add_km(kr, mr);
                        vi = add_km(ki, mi);
                    } else if (WHICH == 1) {
                        vr = kr;
                        vi = ki;
                    } else {
                        vr = mr;
                        vi = mi;
                    }
                    if (p[u] >= 0 && p[u] < wl) {
                        atomicAdd(&a[p[u]][0], (double)vr);
                        atomicAdd(&a[p[u]][1], (double)vi);
                    }
                }
                wave_lds_sync();  // records read before the next batch is staged
            }
            wave_lds_sync();  // every lane's adds before the read-back
            using O = typename OutVec<R>::T;
            for (int q = ii; q < wl; q += 20) {
                O o;
                o.x = (decltype(o.x))a[q][0];
                o.y = (decltype(o.y))a[q][1];
                out[cb + w0 + q] = o;
            }
            wave_lds_sync();  // read-back before the next window or column zeroes
        }
    }
}

template <typename R, bool FULL>
static hipError_t launch_csc_t(const MeshDev& m, const CscBuild& b, long long N,
                               const MatDesc& eps, const MatDesc& mu, int which, double k0sq,
                               void* rec, void* out, hipStream_t s) {
    if (m.n_assy > 0) {
        const double mscale = which == 0 ? -k0sq : 1.0;
        hipLaunchKernelGGL((ned2_metric_kernel<R, FULL>), dim3((unsigned)((m.n_assy + 255) / 256)),
                           dim3(256), 0, s, m, eps, mu, mscale, (R*)rec);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (N == 0) return hipSuccess;
    // a grid of a few workgroups per CU walking the columns (grid-stride): each workgroup
    // copies the reference tables to LDS once, so it should carry many columns
    const long long per_block = kCscWaves * kCscCols;
    long long nb = (N + per_block - 1) / per_block;
    if (nb > kCscGrid) nb = kCscGrid;
    using O = typename OutVec<R>::T;
    const dim3 g((unsigned)nb), blk(64 * kCscWaves);
    if (which == 0)
        hipLaunchKernelGGL((ned2_csc_kernel<R, 0, FULL>), g, blk, 0, s, b.indptr, b.inc_ptr, b.inc,
                           b.pos, b.cols, N, (const R*)rec, (O*)out);
    else if (which == 1)
        hipLaunchKernelGGL((ned2_csc_kernel<R, 1, FULL>), g, blk, 0, s, b.indptr, b.inc_ptr, b.inc,
                           b.pos, b.cols, N, (const R*)rec, (O*)out);
    else
        hipLaunchKernelGGL((ned2_csc_kernel<R, 2, FULL>), g, blk, 0, s, b.indptr, b.inc_ptr, b.inc,
                           b.pos, b.cols, N, (const R*)rec, (O*)out);
    return hipGetLastError();
}

size_t assemble_csc_record_bytes(long long n_assy, int precision, int full) {
    const size_t len = full ? RecLayout<true>::kLen : RecLayout<false>::kLen;
    return (size_t)n_assy * len * (precision == EM_PREC_FP64 ? 8 : 4);
}

hipError_t launch_assemble_csc(const MeshDev& m, const CscBuild& b, long long N,
                               const MatDesc& eps, const MatDesc& mu, int which, int precision,
                               int full, double k0sq, void* rec, void* out, hipStream_t s) {
    if (precision == EM_PREC_FP64)
        return full ? launch_csc_t<double, true>(m, b, N, eps, mu, which, k0sq, rec, out, s)
                    : launch_csc_t<double, false>(m, b, N, eps, mu, which, k0sq, rec, out, s);
    return full ? launch_csc_t<float, true>(m, b, N, eps, mu, which, k0sq, rec, out, s)
                : launch_csc_t<float, false>(m, b, N, eps, mu, which, k0sq, rec, out, s);
}

// ---------------------------------------------------------------------------
// Static COO pattern: rows/cols of every entry (curlcurl.py:449-453).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void ned2_rowcol_kernel(const int* __restrict__ t2f,
                                                          const int* __restrict__ tetids,
                                                          long long n_assy,
                                                          long long* __restrict__ rows,
                                                          long long* __restrict__ cols) {
    const long long n = n_assy * 400;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long ia = e / 400;
        const int r = (int)(e - ia * 400);
        const long long t = tetids ? (long long)tetids[ia] : ia;
        const int* d = t2f + t * 20;
        rows[e] = d[r / 20];
        cols[e] = d[r % 20];
    }
}

hipError_t launch_rowcol(const MeshDev& m, long long* rows, long long* cols, hipStream_t s) {
    const long long n = m.n_assy * 400;
    if (n == 0) return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(ned2_rowcol_kernel, dim3((unsigned)blocks), dim3(256), 0, s, m.t2f,
                       m.tetids, m.n_assy, rows, cols);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Plan validation: edge/face tables equal the sorted tuples of each tet's slots
// (mesh3d.py:707-774), and no tet has zero volume.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void sort2(long long& a, long long& b) {
    if (b < a) {
        long long t = a;
        a = b;
        b = t;
    }
}

__global__ __launch_bounds__(256) void check_tables_kernel(
    const int4* __restrict__ tets, long long T, const long long* __restrict__ edges, long long E,
    const long long* __restrict__ tris, long long F, const long long* __restrict__ t2e,
    const long long* __restrict__ t2t, unsigned long long* bad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int4 nd = tets[t];
    const long long n[4] = {nd.x, nd.y, nd.z, nd.w};
    const int es[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
    const int fs[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};
    unsigned long long nbad = 0;
    for (int s = 0; s < 6; ++s) {
        long long a = n[es[s][0]], b = n[es[s][1]];
        sort2(a, b);
        const long long id = t2e[s * T + t];
        if (id < 0 || id >= E || edges[id] != a || edges[E + id] != b) ++nbad;
    }
    for (int s = 0; s < 4; ++s) {
        long long a = n[fs[s][0]], b = n[fs[s][1]], c = n[fs[s][2]];
        sort2(a, b);
        sort2(b, c);
        sort2(a, b);
        const long long id = t2t[s * T + t];
        if (id < 0 || id >= F || tris[id] != a || tris[F + id] != b || tris[2 * F + id] != c)
            ++nbad;
    }
    if (nbad) atomicAdd(bad, nbad);
}

hipError_t launch_check_tables(const int4* tets, long long T, const long long* edges, long long E,
                               const long long* tris, long long F, const long long* t2e,
                               const long long* t2t, unsigned long long* bad, hipStream_t s) {
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(check_tables_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s,
                       tets, T, edges, E, tris, F, t2e, t2t, bad);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void check_volumes_kernel(const double4* __restrict__ xyz,
                                                            const int4* __restrict__ tets,
                                                            long long T,
                                                            unsigned long long* bad) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    const int4 nd = tets[t];
    const double4 p0 = xyz[nd.x], p1 = xyz[nd.y], p2 = xyz[nd.z], p3 = xyz[nd.w];
    const double ax = p1.x - p0.x, ay = p1.y - p0.y, az = p1.z - p0.z;
    const double bx = p2.x - p0.x, by = p2.y - p0.y, bz = p2.z - p0.z;
    const double cx = p3.x - p0.x, cy = p3.y - p0.y, cz = p3.z - p0.z;
    const double det = ax * (by * cz - bz * cy) - ay * (bx * cz - bz * cx) + az * (bx * cy - by * cx);
    if (!(det != 0.0) || !isfinite(det)) atomicAdd(bad, 1ull);
}

hipError_t launch_check_volumes(const double4* xyz, const int4* tets, long long T,
                                unsigned long long* bad, hipStream_t s) {
    if (T == 0) return hipSuccess;
    hipLaunchKernelGGL(check_volumes_kernel, dim3((unsigned)((T + 255) / 256)), dim3(256), 0, s,
                       xyz, tets, T, bad);
    return hipGetLastError();
}

}  // namespace em
