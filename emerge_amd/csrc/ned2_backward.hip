// Backward of the fused assembly A = K(mu_r^-1) - k0^2 M(eps_r) with respect to the
// per-tet material tensors (FP64).
//
// The reference has no gradients (its optimiser is derivative-free); this kernel is the
// reverse of ned2_assemble_kernel. A is linear in eps_r and mu_r^-1, so the gradient
// depends only on the geometry and the upstream gradient dL/dA of the COO values
// (PyTorch's convention for complex tensors: dL/dRe + i dL/dIm). With the forward's
// notation (ned2_assemble.hip): each reference entry (ii, jj) of a tet equals
//   sg * (sum_c gk_c TSYM_K[p][c] + sa sum_c ak_c TANTI_K[p][c]
//         - k0^2 (sum_c gm_c TSYM_M[p][c] + sa sum_c am_c TANTI_M[p][c]))
// for its canonical pair p, so with w = sg * (g(ii,jj) + g(jj,ii)) (g(ii,ii) once)
//   dL/dgk_c = sum_p w TSYM_K[p][c],  dL/dgm_c = -k0^2 sum_p w TSYM_M[p][c]
// (and the antisymmetric parts with sa). Per round of 8 tets the 210 pair-threads write
// the signed weights w into LDS (reading the upstream block straight from HBM, each
// entry once), then 192 threads form the 8 x 12 dot products of length 210 with the
// reference tensors held in LDS (a small GEMM). Then, from G_M = C eps C^T / |det J| and
// G_K = Jc mu^-1 Jc^T / |det J| (Jc rows = columns of J):
//   dL/deps = C^T W_M C / |det J|,   dL/dmu^-1 = Jc^T W_K Jc / |det J|
// reduced to the leaf's kind (scalar: trace, diagonal: diagonal, full: all 9).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "em_internal.h"
#include "ned2_tables.inc"

namespace em {

namespace {

constexpr int kBlock = 256;
constexpr int kTile = 64;

// Work split of one round of NW tets: the dot-product phase runs NW x NC x NSPLIT threads.
template <bool FULL>
struct Shape {
    static constexpr int NW = 8;                  // tets per round
    static constexpr int NC = FULL ? 18 : 12;     // complex outputs per tet (K then M)
    static constexpr int NSPLIT = FULL ? 1 : 2;   // pair ranges per output
};

template <bool FULL>
struct BwdSmem {
    static constexpr int NW = Shape<FULL>::NW, NC = Shape<FULL>::NC, NS = Shape<FULL>::NSPLIT;
    double Jc[kTile][9];
    double C[kTile][9];
    double iad[kTile];
    double sgn[kTile];
    long long tet[kTile];
    int code[kTile];
    double T[NED2_NPAIR][NC];          // reference tensors, M part scaled by -k0^2
    double2 W[NW][NED2_NPAIR];         // signed pair weights
    double2 Wa[FULL ? NW : 1][NED2_NPAIR];  // times the antisymmetric sign
    double2 R[NS][NW][NC];
    unsigned char slot[64][20];
};

// g_out (tet-major, complex) for a 3x3 complex gradient Gr (row-major) reduced to kind.
__device__ __forceinline__ void write_leaf(double2* out, long long t, int kind,
                                          const double Gre[9], const double Gim[9]) {
    if (!out) return;
    if (kind == EM_MAT_SCALAR) {
        out[t] = make_double2(Gre[0] + Gre[4] + Gre[8], Gim[0] + Gim[4] + Gim[8]);
    } else if (kind == EM_MAT_DIAG) {
#pragma unroll
        for (int d = 0; d < 3; ++d) out[3 * t + d] = make_double2(Gre[4 * d], Gim[4 * d]);
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) out[9 * t + i] = make_double2(Gre[i], Gim[i]);
    }
}

// dL/dX_rs = sum_pq W_pq A[p][r] A[q][s] * scale   (X = tensor, W = dL/dG, G = A X A^T)
__device__ __forceinline__ void pullback(const double A[9], const double Wre[9],
                                         const double Wim[9], double scale, double Ore[9],
                                         double Oim[9]) {
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    const double w = A[3 * p + r] * A[3 * q + s];
                    re += w * Wre[3 * p + q];
                    im += w * Wim[3 * p + q];
                }
            Ore[3 * r + s] = re * scale;
            Oim[3 * r + s] = im * scale;
        }
}

// Full 3x3 dL/dG (row-major, re/im) from the gradients of the symmetric components
// (00,11,22,01,02,12) and antisymmetric components (01,02,12) of the split.
__device__ __forceinline__ void unsplit(const double* ws, const double* wa, bool full,
                                       double Wre[9], double Wim[9]) {
    // ws/wa: interleaved re, im
    Wre[0] = ws[0];
    Wim[0] = ws[1];
    Wre[4] = ws[2];
    Wim[4] = ws[3];
    Wre[8] = ws[4];
    Wim[8] = ws[5];
    const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double sr = 0.5 * ws[2 * (3 + c)], si = 0.5 * ws[2 * (3 + c) + 1];
        const double ar = full ? 0.5 * wa[2 * c] : 0.0, ai = full ? 0.5 * wa[2 * c + 1] : 0.0;
        Wre[3 * P[c] + Q[c]] = sr + ar;
        Wim[3 * P[c] + Q[c]] = si + ai;
        Wre[3 * Q[c] + P[c]] = sr - ar;
        Wim[3 * Q[c] + P[c]] = si - ai;
    }
}

template <bool FULL>
__global__ __launch_bounds__(kBlock) void ned2_backward_kernel(
    MeshDev m, const double2* __restrict__ grad, double k0sq, int eps_kind, int mu_kind,
    double2* __restrict__ g_eps, double2* __restrict__ g_mu) {
    using S = Shape<FULL>;
    constexpr int NW = S::NW, NC = S::NC, NS = S::NSPLIT;
    constexpr int NP = NC;  // doubles per matrix in the flattened (re, im) result
    __shared__ BwdSmem<FULL> sm;
    for (int i = threadIdx.x; i < 64 * 20; i += kBlock) sm.slot[i / 20][i % 20] = SLOT_OF[i / 20][i % 20];
    for (int i = threadIdx.x; i < NED2_NPAIR * NC; i += kBlock) {
        const int p = i / NC, c = i % NC;
        double v;
        if (FULL) {
            v = c < 6 ? TSYM_K[p][c] : c < 9 ? TANTI_K[p][c - 6]
                : c < 15 ? -k0sq * TSYM_M[p][c - 9] : -k0sq * TANTI_M[p][c - 15];
        } else {
            v = c < 6 ? TSYM_K[p][c] : -k0sq * TSYM_M[p][c - 6];
        }
        sm.T[p][c] = v;
    }

    const long long tile0 = (long long)blockIdx.x * kTile;
    const int ntile = (int)min((long long)kTile, m.n_assy - tile0);
    if (threadIdx.x < ntile) {
        const int lt = threadIdx.x;
        const long long ia = tile0 + lt;
        const long long t = m.tetids ? (long long)m.tetids[ia] : ia;
        const int4 nd = m.tets[t];
        const int n[4] = {nd.x, nd.y, nd.z, nd.w};
        int r[4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
            r[a] = (n[0] < n[a]) + (n[1] < n[a]) + (n[2] < n[a]) + (n[3] < n[a]);
        const int s0 = min(n[0], n[1]), s1 = max(n[0], n[1]);
        const int s2 = min(n[2], n[3]), s3 = max(n[2], n[3]);
        const int w0 = min(s0, s2), t2 = max(s0, s2);
        const int t1 = min(s1, s3), w3 = max(s1, s3);
        const int w1 = min(t1, t2), w2 = max(t1, t2);
        const double4 X0 = m.xyz[w0], X1 = m.xyz[w1], X2 = m.xyz[w2], X3 = m.xyz[w3];
        const double J[9] = {X1.x - X0.x, X1.y - X0.y, X1.z - X0.z, X2.x - X0.x, X2.y - X0.y,
                             X2.z - X0.z, X3.x - X0.x, X3.y - X0.y, X3.z - X0.z};
        const double C[9] = {J[4] * J[8] - J[5] * J[7], J[5] * J[6] - J[3] * J[8],
                             J[3] * J[7] - J[4] * J[6], J[7] * J[2] - J[8] * J[1],
                             J[8] * J[0] - J[6] * J[2], J[6] * J[1] - J[7] * J[0],
                             J[1] * J[5] - J[2] * J[4], J[2] * J[3] - J[0] * J[5],
                             J[0] * J[4] - J[1] * J[3]};
        const double det = J[0] * C[0] + J[1] * C[1] + J[2] * C[2];
        const int inv = (r[0] > r[1]) + (r[0] > r[2]) + (r[0] > r[3]) + (r[1] > r[2]) +
                        (r[1] > r[3]) + (r[2] > r[3]);
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            sm.Jc[lt][i] = J[i];
            sm.C[lt][i] = C[i];
        }
        sm.iad[lt] = 1.0 / fabs(det);
        sm.sgn[lt] = ((det < 0.0) != (inv & 1)) ? -1.0 : 1.0;
        sm.code[lt] = r[0] * 16 + r[1] * 4 + r[2];
        sm.tet[lt] = t;
    }
    const int j = threadIdx.x;
    const bool pair_thread = j < NED2_NPAIR;
    const int pj = pair_thread ? j : 0;
    const int a = PAIR_A[pj], b = PAIR_B[pj];
    const bool aodd = a >= 6, bodd = b >= 6;
    // dot-phase role: (u, c, split)
    const bool dot_thread = threadIdx.x < NW * NC * NS;
    const int du = threadIdx.x / (NC * NS);
    const int dc = (threadIdx.x / NS) % NC;
    const int dsp = threadIdx.x % NS;
    const bool anti_c = FULL && ((dc >= 6 && dc < 9) || dc >= 15);
    const int p_lo = dsp * ((NED2_NPAIR + NS - 1) / NS);
    const int p_hi = min(NED2_NPAIR, p_lo + (NED2_NPAIR + NS - 1) / NS);
    __syncthreads();

    for (int lt0 = 0; lt0 < ntile; lt0 += NW) {
        const int nvalid = min(NW, ntile - lt0);
        // pair phase: signed weight of every canonical pair of every tet of the round
        if (pair_thread) {
            for (int u = 0; u < nvalid; ++u) {
                const int lt = lt0 + u;
                const int code = sm.code[lt];
                const int ra = sm.slot[code][a], rb = sm.slot[code][b];
                const double2* blk = grad + (tile0 + lt) * 400;
                double2 g = blk[ra * 20 + rb];
                if (ra != rb) {
                    const double2 g2 = blk[rb * 20 + ra];
                    g.x += g2.x;
                    g.y += g2.y;
                }
                const double s = sm.sgn[lt];
                const double sg = (aodd ? s : 1.0) * (bodd ? s : 1.0);
                sm.W[u][j] = make_double2(sg * g.x, sg * g.y);
                if constexpr (FULL) {
                    const double sa = (ra < rb) ? -sg : sg;
                    sm.Wa[u][j] = make_double2(sa * g.x, sa * g.y);
                }
            }
        }
        __syncthreads();
        // dot phase: R[split][u][c] = sum over pairs of W[u][p] * T[p][c]
        if (dot_thread && du < nvalid) {
            const double2* w = (FULL && anti_c) ? sm.Wa[FULL ? du : 0] : sm.W[du];
            double re = 0.0, im = 0.0;
            for (int p = p_lo; p < p_hi; ++p) {
                const double2 v = w[p];
                const double tv = sm.T[p][dc];
                re += v.x * tv;
                im += v.y * tv;
            }
            sm.R[dsp][du][dc] = make_double2(re, im);
        }
        __syncthreads();
        if (threadIdx.x < nvalid) {
            const int u = threadIdx.x;
            const int lt = lt0 + u;
            double w[2 * NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double re = 0.0, im = 0.0;
#pragma unroll
                for (int q = 0; q < NS; ++q) {
                    re += sm.R[q][u][c].x;
                    im += sm.R[q][u][c].y;
                }
                w[2 * c] = re;
                w[2 * c + 1] = im;
            }
            double Wre[9], Wim[9], Ore[9], Oim[9], A[9];
            const long long t = sm.tet[lt];
            const double iad = sm.iad[lt];
            if (g_mu) {
                unsplit(w, w + 12, FULL, Wre, Wim);
#pragma unroll
                for (int i = 0; i < 9; ++i) A[i] = sm.Jc[lt][i];
                pullback(A, Wre, Wim, iad, Ore, Oim);
                write_leaf(g_mu, t, mu_kind, Ore, Oim);
            }
            if (g_eps) {
                unsplit(w + NP, w + NP + 12, FULL, Wre, Wim);
#pragma unroll
                for (int i = 0; i < 9; ++i) A[i] = sm.C[lt][i];
                pullback(A, Wre, Wim, iad, Ore, Oim);
                write_leaf(g_eps, t, eps_kind, Ore, Oim);
            }
        }
        __syncthreads();  // the next round's dot phase rewrites R
    }
}

// Scalar/diagonal leaves (no antisymmetric parts): one wavefront per 64-tet tile, no
// barriers across wavefronts. Lanes 0..59 own one (component c of the 12, range of 42
// pairs) each, with those 42 reference-tensor coefficients in registers (the M ones
// times -k0^2). The upstream 20x20 blocks are read coalesced into registers kWAhead tets
// ahead, staged in LDS a group at a time (the loads of the next group are in flight while
// the current one is processed). Per tet every lane forms the signed weights of pairs
// lane, lane + 64, ... into LDS, each owning lane sums its 42 weight x coefficient
// products and five lanes reduce them by shuffles; the 12 sums of 16 tets are kept in LDS
// and 16 lanes then apply the chain rule, one tet each (geometry recomputed there).
constexpr int kWTile = 64;     // tets per wavefront tile
constexpr int kWChunk = 16;    // tets per chain-rule batch
constexpr int kWWaves = 4;     // wavefronts per workgroup
constexpr int kWSplit = 5;     // pair ranges per component
constexpr int kWPairs = 42;    // pairs per range (5 x 42 = 210)

constexpr int kWAhead = 2;     // upstream blocks staged (and prefetched) per group

struct WaveSmem {
    double2 blk[kWAhead][400];
    double2 w[NED2_NPAIR];
    double2 res[kWChunk][12];
    int code[kWTile];
    double sgn[kWTile];
};

// Geometry of assembly index ia: Jc (rows = columns of J, canonical vertex order), the
// cofactor matrix C, 1/|det J|, tet id, vertex-rank code and orientation sign.
__device__ __forceinline__ void bwd_geometry(const MeshDev& m, long long ia, double Jc[9],
                                             double C[9], double& iad, long long& tet,
                                             int& code, double& sgn) {
    tet = m.tetids ? (long long)m.tetids[ia] : ia;
    const int4 nd = m.tets[tet];
    const int n[4] = {nd.x, nd.y, nd.z, nd.w};
    int r[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
        r[a] = (n[0] < n[a]) + (n[1] < n[a]) + (n[2] < n[a]) + (n[3] < n[a]);
    const int s0 = min(n[0], n[1]), s1 = max(n[0], n[1]);
    const int s2 = min(n[2], n[3]), s3 = max(n[2], n[3]);
    const int w0 = min(s0, s2), t2 = max(s0, s2);
    const int t1 = min(s1, s3), w3 = max(s1, s3);
    const int w1 = min(t1, t2), w2 = max(t1, t2);
    const double4 X0 = m.xyz[w0], X1 = m.xyz[w1], X2 = m.xyz[w2], X3 = m.xyz[w3];
    const double J[9] = {X1.x - X0.x, X1.y - X0.y, X1.z - X0.z, X2.x - X0.x, X2.y - X0.y,
                         X2.z - X0.z, X3.x - X0.x, X3.y - X0.y, X3.z - X0.z};
    C[0] = J[4] * J[8] - J[5] * J[7];
    C[1] = J[5] * J[6] - J[3] * J[8];
    C[2] = J[3] * J[7] - J[4] * J[6];
    C[3] = J[7] * J[2] - J[8] * J[1];
    C[4] = J[8] * J[0] - J[6] * J[2];
    C[5] = J[6] * J[1] - J[7] * J[0];
    C[6] = J[1] * J[5] - J[2] * J[4];
    C[7] = J[2] * J[3] - J[0] * J[5];
    C[8] = J[0] * J[4] - J[1] * J[3];
#pragma unroll
    for (int i = 0; i < 9; ++i) Jc[i] = J[i];
    const double det = J[0] * C[0] + J[1] * C[1] + J[2] * C[2];
    const int inv = (r[0] > r[1]) + (r[0] > r[2]) + (r[0] > r[3]) + (r[1] > r[2]) +
                    (r[1] > r[3]) + (r[2] > r[3]);
    iad = 1.0 / fabs(det);
    sgn = ((det < 0.0) != (inv & 1)) ? -1.0 : 1.0;
    code = r[0] * 16 + r[1] * 4 + r[2];
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One tet of a wavefront tile whose upstream block is staged in blk: the signed pair
// weights and the 12 sums. When a 16-tet chunk is complete its chain rule runs, one lane
// per tet.
__device__ __forceinline__ void wave_tet(const MeshDev& m, WaveSmem& sm, const double2* blk,
                                         const unsigned char (*slot)[20],
                                         const double (&coef)[kWPairs], const int (&pa)[4],
                                         const int (&pb)[4], int lane, bool owner, int comp,
                                         int split, long long tile0, int lt, int ntile,
                                         int eps_kind, int mu_kind, double2* __restrict__ g_eps,
                                         double2* __restrict__ g_mu) {
    const int code = sm.code[lt];
    const double s = sm.sgn[lt];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = lane + 64 * k;
        if (p < NED2_NPAIR) {
            const int a = pa[k], b = pb[k];
            const int ra = slot[code][a], rb = slot[code][b];
            double2 v = blk[ra * 20 + rb];
            if (ra != rb) {
                const double2 v2 = blk[rb * 20 + ra];
                v.x += v2.x;
                v.y += v2.y;
            }
            const double sgp = (a >= 6 ? s : 1.0) * (b >= 6 ? s : 1.0);
            sm.w[p] = make_double2(sgp * v.x, sgp * v.y);
        }
    }
    wave_sync();
    double re = 0.0, im = 0.0;
    if (owner) {
        const double2* wp = sm.w + split * kWPairs;
#pragma unroll
        for (int k = 0; k < kWPairs; ++k) {
            const double2 v = wp[k];
            re += v.x * coef[k];
            im += v.y * coef[k];
        }
    }
    // sum over the 5 ranges of a component: lanes 5c .. 5c+4
    double rs = re, is = im;
#pragma unroll
    for (int d = 1; d < kWSplit; ++d) {
        rs += __shfl_down(re, d);
        is += __shfl_down(im, d);
    }
    const int u = lt % kWChunk;
    if (owner && split == 0) sm.res[u][comp] = make_double2(rs, is);
    wave_sync();
    if (u != kWChunk - 1 && lt != ntile - 1) return;
    // chain rule for the chunk's tets, one lane each (geometry recomputed)
    const int c0 = lt - u;
    if (lane <= u) {
        double Jc[9], C[9], iad, sg;
        long long tet;
        int cd;
        bwd_geometry(m, tile0 + c0 + lane, Jc, C, iad, tet, cd, sg);
        double w[24];
#pragma unroll
        for (int c = 0; c < 12; ++c) {
            const double2 v = sm.res[lane][c];
            w[2 * c] = v.x;
            w[2 * c + 1] = v.y;
        }
        double Wre[9], Wim[9], Ore[9], Oim[9];
        if (g_mu) {
            unsplit(w, w, false, Wre, Wim);
            pullback(Jc, Wre, Wim, iad, Ore, Oim);
            write_leaf(g_mu, tet, mu_kind, Ore, Oim);
        }
        if (g_eps) {
            unsplit(w + 12, w + 12, false, Wre, Wim);
            pullback(C, Wre, Wim, iad, Ore, Oim);
            write_leaf(g_eps, tet, eps_kind, Ore, Oim);
        }
    }
    wave_sync();
}

__global__ __launch_bounds__(64 * kWWaves) void ned2_backward_wave_kernel(
    MeshDev m, const double2* __restrict__ grad, double k0sq, int eps_kind, int mu_kind,
    double2* __restrict__ g_eps, double2* __restrict__ g_mu) {
    __shared__ WaveSmem smem[kWWaves];
    __shared__ unsigned char slot[64][20];
    for (int i = threadIdx.x; i < 64 * 20; i += 64 * kWWaves)
        slot[i / 20][i % 20] = SLOT_OF[i / 20][i % 20];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveSmem& sm = smem[wave];
    const long long tile0 = ((long long)blockIdx.x * kWWaves + wave) * kWTile;
    if (tile0 >= m.n_assy) return;  // whole wavefront
    const int ntile = (int)min((long long)kWTile, m.n_assy - tile0);
    // the upstream blocks of the next kWAhead tets, in flight while the current ones run
    double bre[kWAhead][7], bim[kWAhead][7];
#pragma unroll
    for (int q = 0; q < kWAhead; ++q)
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int e = lane + 64 * k;
            double2 v = make_double2(0.0, 0.0);
            if (e < 400 && q < ntile) v = grad[(tile0 + q) * 400 + e];
            bre[q][k] = v.x;
            bim[q][k] = v.y;
        }
    // this lane's (component, pair range) and coefficients, and its pairs' functions
    const bool owner = lane < 12 * kWSplit;
    const int comp = owner ? lane / kWSplit : 0;
    const int split = owner ? lane % kWSplit : 0;
    double coef[kWPairs];
#pragma unroll
    for (int k = 0; k < kWPairs; ++k) {
        const int p = split * kWPairs + k;
        coef[k] = comp < 6 ? TSYM_K[p][comp] : -k0sq * TSYM_M[p][comp - 6];
    }
    int pa[4], pb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = min(lane + 64 * k, NED2_NPAIR - 1);
        pa[k] = PAIR_A[p];
        pb[k] = PAIR_B[p];
    }
    if (lane < ntile) {
        double Jc[9], C[9], iad, sg;
        long long tet;
        int code;
        bwd_geometry(m, tile0 + lane, Jc, C, iad, tet, code, sg);
        sm.sgn[lane] = sg;
        sm.code[lane] = code;
    }
    for (int lt0 = 0; lt0 < ntile; lt0 += kWAhead) {
        // stage the blocks, then start the loads of the next group into the same registers
#pragma unroll
        for (int q = 0; q < kWAhead; ++q)
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int e = lane + 64 * k;
                if (e < 400) sm.blk[q][e] = make_double2(bre[q][k], bim[q][k]);
            }
#pragma unroll
        for (int q = 0; q < kWAhead; ++q)
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int e = lane + 64 * k;
                if (e < 400 && lt0 + kWAhead + q < ntile) {
                    const double2 v = grad[(tile0 + lt0 + kWAhead + q) * 400 + e];
                    bre[q][k] = v.x;
                    bim[q][k] = v.y;
                }
            }
        wave_sync();
        const int nq = min(kWAhead, ntile - lt0);
        for (int q = 0; q < nq; ++q)
            wave_tet(m, sm, sm.blk[q], slot, coef, pa, pb, lane, owner, comp, split, tile0,
                     lt0 + q, ntile, eps_kind, mu_kind, g_eps, g_mu);
    }
}

}  // namespace

hipError_t launch_backward(const MeshDev& m, const double2* grad, double k0sq, int eps_kind,
                           int mu_kind, double2* g_eps, double2* g_mu, hipStream_t stream) {
    const long long ntiles = (m.n_assy + kTile - 1) / kTile;
    if (ntiles == 0) return hipSuccess;
    const bool full = (g_eps && eps_kind == EM_MAT_FULL) || (g_mu && mu_kind == EM_MAT_FULL);
    if (full) {
        hipLaunchKernelGGL((ned2_backward_kernel<true>), dim3((unsigned)ntiles), dim3(kBlock), 0,
                           stream, m, grad, k0sq, eps_kind, mu_kind, g_eps, g_mu);
    } else {
        const long long wtiles = (m.n_assy + kWTile - 1) / kWTile;
        const unsigned blocks = (unsigned)((wtiles + kWWaves - 1) / kWWaves);
        hipLaunchKernelGGL(ned2_backward_wave_kernel, dim3(blocks), dim3(64 * kWWaves), 0, stream,
                           m, grad, k0sq, eps_kind, mu_kind, g_eps, g_mu);
    }
    return hipGetLastError();
}

}  // namespace em
