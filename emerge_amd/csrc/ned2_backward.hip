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
// (and the antisymmetric parts with sa): 12 (18) dot products of length 210 per tet, a
// small GEMM against the reference tensors (ned2_backward_wave_kernel below). Then, from
// G_M = C eps C^T / |det J| and G_K = Jc mu^-1 Jc^T / |det J| (Jc rows = columns of J):
//   dL/deps = C^T W_M C / |det J|,   dL/dmu^-1 = Jc^T W_K Jc / |det J|
// reduced to the leaf's kind (scalar: trace, diagonal: diagonal, full: all 9).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "em_internal.h"
#include "ned2_tables.inc"

namespace em {

namespace {


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

// The same as pullback in two steps, O = A^T (W A) * scale, with 18 live temporaries.
__device__ __forceinline__ void pullback2(const double A[9], const double Wre[9],
                                          const double Wim[9], double scale, double Ore[9],
                                          double Oim[9]) {
    double Tre[9], Tim[9];  // T[p][s] = sum_q W[p][q] A[q][s]
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                re += Wre[3 * p + q] * A[3 * q + c];
                im += Wim[3 * p + q] * A[3 * q + c];
            }
            Tre[3 * p + c] = re;
            Tim[3 * p + c] = im;
        }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double re = 0.0, im = 0.0;
#pragma unroll
            for (int p = 0; p < 3; ++p) {
                re += A[3 * p + r] * Tre[3 * p + c];
                im += A[3 * p + r] * Tim[3 * p + c];
            }
            Ore[3 * r + c] = re * scale;
            Oim[3 * r + c] = im * scale;
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

// Wavefront kernel: one wavefront per 64-tet tile, no barriers across wavefronts. The
// wavefront's four 16-lane rows own three of the 12 symmetric components each (row r:
// components 3r .. 3r+2; K then M) and lane j of a row owns pairs 14j .. 14j+13 (lane 15
// none), with those 3 x 14 reference-tensor coefficients in registers (the M ones times
// -k0^2); so one LDS read of a pair weight feeds six FMAs and the 16 partial sums of a row
// are reduced by DPP lane moves (no LDS traffic). With full tensor leaves (FULL) a second
// pass over the antisymmetric weights gives the 3 + 3 antisymmetric components (K on rows
// 0 and 2, M on rows 1 and 3, each row over half of its lanes' pairs). The upstream 20x20 blocks are read coalesced into registers one group ahead and
// staged in LDS a group at a time (the loads of the next group are in flight while the
// current one is processed). Per tet every lane forms the signed weights of pairs lane,
// lane + 64, ... into LDS; the sums of 16 tets are kept in LDS and the lanes holding those
// tets' geometry then apply the chain rule.
constexpr int kWTile = 64;     // tets per wavefront tile (one lane per tet for geometry)
constexpr int kWChunk = 16;    // tets per chain-rule batch
constexpr int kWComps = 3;     // components per lane (per 16-lane row)
constexpr int kWPairs = 14;    // pairs per lane (15 x 14 = 210)
constexpr int kWStride = 15;   // LDS slots per lane range (one pad: no bank conflicts)
constexpr int kWSlots = 16 * kWStride;

template <bool FULL>
struct WaveCfg {
    static constexpr int AHEAD = 1;               // upstream blocks staged per group
    static constexpr int NRES = FULL ? 18 : 12;   // K sym, M sym (, K anti, M anti)
    static constexpr int WAVES = FULL ? 8 : 4;    // wavefronts per workgroup
};

// Scalar/diagonal: 13.3 KB per wavefront (two 4-wavefront workgroups per CU, two
// wavefronts per SIMD). Full: 18.7 KB per wavefront; eight wavefronts and one copy of the
// antisymmetric coefficient table per workgroup fill a CU's 160 KB of LDS (one workgroup
// per CU, two wavefronts per SIMD; with two-wavefront workgroups the table copies left
// room for three, six wavefronts per CU).
template <bool FULL>
struct WaveSmem {
    double2 blk[WaveCfg<FULL>::AHEAD][400];
    double2 w[WaveCfg<FULL>::AHEAD][kWSlots];  // pair p at (p / 14) * 15 + p % 14
    double2 wa[FULL ? WaveCfg<FULL>::AHEAD : 1][FULL ? kWSlots : 1];  // antisymmetric
    double2 res[kWChunk][WaveCfg<FULL>::NRES];
};

// Cofactor matrix of J (row-major; C[3p + r]).
__device__ __forceinline__ void cofactors(const double J[9], double C[9]) {
    C[0] = J[4] * J[8] - J[5] * J[7];
    C[1] = J[5] * J[6] - J[3] * J[8];
    C[2] = J[3] * J[7] - J[4] * J[6];
    C[3] = J[7] * J[2] - J[8] * J[1];
    C[4] = J[8] * J[0] - J[6] * J[2];
    C[5] = J[6] * J[1] - J[7] * J[0];
    C[6] = J[1] * J[5] - J[2] * J[4];
    C[7] = J[2] * J[3] - J[0] * J[5];
    C[8] = J[0] * J[4] - J[1] * J[3];
}

// Diagonal of the pullback for a symmetric dL/dG given by its six components
// (00, 11, 22, 01, 02, 12; complex, in LDS): with a_p = A[3p + r],
//   dL/dX_rr = iad (w00 a0^2 + w11 a1^2 + w22 a2^2 + w01 a0 a1 + w02 a0 a2 + w12 a1 a2)
// (the off-diagonal entries of dL/dG are half the split's components), reduced to a
// scalar (trace) or diagonal leaf.
__device__ __forceinline__ void write_diag_leaf(double2* out, long long t, int kind,
                                                const double A[9], const double2* wsym,
                                                double iad) {
    double ore[3], oim[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double a0 = A[r], a1 = A[3 + r], a2 = A[6 + r];
        const double sq[6] = {a0 * a0, a1 * a1, a2 * a2, a0 * a1, a0 * a2, a1 * a2};
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const double2 v = wsym[c];
            re += sq[c] * v.x;
            im += sq[c] * v.y;
        }
        ore[r] = re * iad;
        oim[r] = im * iad;
    }
    if (kind == EM_MAT_SCALAR) {
        out[t] = make_double2(ore[0] + ore[1] + ore[2], oim[0] + oim[1] + oim[2]);
    } else {
#pragma unroll
        for (int r = 0; r < 3; ++r) out[3 * t + r] = make_double2(ore[r], oim[r]);
    }
}

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
    cofactors(J, C);
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

// x moved across lanes by the DPP control CTRL (a bitwise move of both halves).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double x) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Sum over the 16 lanes of each row, returned in every lane of the row: lane ^ 1, lane ^ 2
// (quad_perm), then the 8-lane half mirror and the 16-lane row mirror.
__device__ __forceinline__ double row_sum(double x) {
    x += dpp_move<0xb1>(x);
    x += dpp_move<0x4e>(x);
    x += dpp_move<0x141>(x);
    x += dpp_move<0x140>(x);
    return x;
}

// Geometry of the lane's own tet of the tile (lane t holds tet tile0 + t).
struct LaneGeom {
    double Jc[9], iad;
    long long tet;
    int code, neg;
};

// One group of nq <= AHEAD tets of a wavefront tile (lt0 ..), their upstream blocks
// staged in sm.blk: the signed pair weights and the component sums of every tet of the
// group. When a 16-tet chunk is complete its chain rule runs on the lanes that hold the
// chunk's geometry.
template <bool FULL>
__device__ __forceinline__ void wave_group(WaveSmem<FULL>& sm, const unsigned char (*slot)[20],
                                           const double (&coef)[kWComps][kWPairs],
                                           const double* coefa,
                                           const int (&pab)[4], const LaneGeom& gm, int lane,
                                           int row, int j, int lt0, int nq, int ntile,
                                           int eps_kind, int mu_kind,
                                           double2* __restrict__ g_eps,
                                           double2* __restrict__ g_mu) {
    constexpr int AH = WaveCfg<FULL>::AHEAD;
#pragma unroll
    for (int q = 0; q < AH; ++q) {
        const int lt = min(lt0 + q, ntile - 1);
        const int code = __builtin_amdgcn_readlane(gm.code, lt);
        const double s = __builtin_amdgcn_readlane(gm.neg, lt) ? -1.0 : 1.0;
        const double2* blk = sm.blk[q];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int p = lane + 64 * k;
            if (p < NED2_NPAIR) {
                const int a = pab[k] & 0xff, b = pab[k] >> 8;
                const int pw = (p / kWPairs) * kWStride + p % kWPairs;  // LDS slot
                const int ra = slot[code][a], rb = slot[code][b];
                double2 v = blk[ra * 20 + rb];
                if (ra != rb) {
                    const double2 v2 = blk[rb * 20 + ra];
                    v.x += v2.x;
                    v.y += v2.y;
                }
                const double sgp = (a >= 6 ? s : 1.0) * (b >= 6 ? s : 1.0);
                sm.w[q][pw] = make_double2(sgp * v.x, sgp * v.y);
                if constexpr (FULL) {
                    // antisymmetric sign: negative when the pair's lower canonical function
                    // sits in the lower reference slot
                    const double sa = ra < rb ? -sgp : sgp;
                    sm.wa[q][pw] = make_double2(sa * v.x, sa * v.y);
                }
            }
        }
    }
    wave_sync();
    double acc[AH][kWComps][2], acca[AH][kWComps][2];
#pragma unroll
    for (int q = 0; q < AH; ++q)
#pragma unroll
        for (int c = 0; c < kWComps; ++c) {
            acc[q][c][0] = acc[q][c][1] = 0.0;
            acca[q][c][0] = acca[q][c][1] = 0.0;
        }
#pragma unroll
    for (int k = 0; k < kWPairs; ++k) {
#pragma unroll
        for (int q = 0; q < AH; ++q) {
            const double2 v = sm.w[q][j * kWStride + k];
#pragma unroll
            for (int c = 0; c < kWComps; ++c) {
                acc[q][c][0] += v.x * coef[c][k];
                acc[q][c][1] += v.y * coef[c][k];
            }
            if (FULL && k < kWPairs / 2) {
                // antisymmetric pass: rows 0 and 2 K, rows 1 and 3 M; rows 0 and 1 take the
                // first 7 pairs of the lane's range, rows 2 and 3 the last 7 (the halves are
                // added across rows below). Lane 15's weights are zero.
                const int kk = k + (row >> 1) * (kWPairs / 2);
                const double2 va = sm.wa[q][j * kWStride + kk];
                const double* ca =
                    coefa + 3 * ((row & 1) * NED2_NPAIR + min(kWPairs * j + kk, NED2_NPAIR - 1));
                const double cc[3] = {ca[0], ca[1], ca[2]};
#pragma unroll
                for (int c = 0; c < kWComps; ++c) {
                    acca[q][c][0] += va.x * cc[c];
                    acca[q][c][1] += va.y * cc[c];
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < AH; ++q) {
        double2 mine = make_double2(0.0, 0.0), minea = make_double2(0.0, 0.0);
#pragma unroll
        for (int c = 0; c < kWComps; ++c) {
            const double re = row_sum(acc[q][c][0]), im = row_sum(acc[q][c][1]);
            if (j == c) mine = make_double2(re, im);
            if constexpr (FULL) {
                // row sums, then row r + 2's half added to row r's (lane l + 32)
                double ra = row_sum(acca[q][c][0]), ia = row_sum(acca[q][c][1]);
                ra += __shfl_xor(ra, 32);
                ia += __shfl_xor(ia, 32);
                if (j == c) minea = make_double2(ra, ia);
            }
        }
        if (q < nq && j < kWComps) {
            double2* r = sm.res[(lt0 + q) % kWChunk];
            r[kWComps * row + j] = mine;
            // antisymmetric K (row 0) and M (row 1; both whole sums after the cross-row
            // add) after the 12 symmetric components
            if (FULL && row < 2) r[12 + kWComps * row + j] = minea;
        }
    }
    wave_sync();
    const int last = lt0 + nq - 1;
    const int u = last % kWChunk;
    if (u != kWChunk - 1 && last != ntile - 1) return;
    // chain rule for the chunk's tets on the lanes that hold them
    const int c0 = last - u;
    if (lane >= c0 && lane <= last) {
        const double2* r = sm.res[lane - c0];
        if constexpr (!FULL) {
            if (g_mu) write_diag_leaf(g_mu, gm.tet, mu_kind, gm.Jc, r, gm.iad);
            if (g_eps) {
                double C[9];
                cofactors(gm.Jc, C);
                write_diag_leaf(g_eps, gm.tet, eps_kind, C, r + 6, gm.iad);
            }
        } else {
            double ws[12], wa[6], Wre[9], Wim[9], Ore[9], Oim[9];
            if (g_mu) {
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    ws[2 * c] = r[c].x;
                    ws[2 * c + 1] = r[c].y;
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    wa[2 * c] = r[12 + c].x;
                    wa[2 * c + 1] = r[12 + c].y;
                }
                unsplit(ws, wa, true, Wre, Wim);
                pullback2(gm.Jc, Wre, Wim, gm.iad, Ore, Oim);
                write_leaf(g_mu, gm.tet, mu_kind, Ore, Oim);
            }
            if (g_eps) {
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    ws[2 * c] = r[6 + c].x;
                    ws[2 * c + 1] = r[6 + c].y;
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    wa[2 * c] = r[15 + c].x;
                    wa[2 * c + 1] = r[15 + c].y;
                }
                double C[9];
                cofactors(gm.Jc, C);
                unsplit(ws, wa, true, Wre, Wim);
                pullback2(C, Wre, Wim, gm.iad, Ore, Oim);
                write_leaf(g_eps, gm.tet, eps_kind, Ore, Oim);
            }
        }
    }
    wave_sync();
}

template <bool FULL>
__global__ __launch_bounds__(64 * WaveCfg<FULL>::WAVES)
__attribute__((amdgpu_waves_per_eu(2, 2)))
void ned2_backward_wave_kernel(MeshDev m, const double2* __restrict__ grad, double k0sq,
                               int eps_kind, int mu_kind, double2* __restrict__ g_eps,
                               double2* __restrict__ g_mu) {
    constexpr int AH = WaveCfg<FULL>::AHEAD;
    constexpr int NW = WaveCfg<FULL>::WAVES;
    __shared__ WaveSmem<FULL> smem[NW];
    __shared__ unsigned char slot[64][20];
    // antisymmetric coefficients (K; M times -k0^2), three per pair
    __shared__ double coefa[FULL ? 2 : 1][FULL ? NED2_NPAIR : 1][3];
    // table copies with every thread's loads issued together (unrolled, no wait per entry)
#pragma unroll
    for (int r = 0; r < (64 * 20 + 64 * NW - 1) / (64 * NW); ++r) {
        const int i = threadIdx.x + r * 64 * NW;
        if (i < 64 * 20) slot[i / 20][i % 20] = SLOT_OF[i / 20][i % 20];
    }
    if constexpr (FULL) {
#pragma unroll
        for (int r = 0; r < (2 * NED2_NPAIR + 64 * NW - 1) / (64 * NW); ++r) {
            const int i = threadIdx.x + r * 64 * NW;
            if (i < 2 * NED2_NPAIR) {
                const int h = i / NED2_NPAIR, p = i % NED2_NPAIR;
                const double* tab = h == 0 ? TANTI_K[p] : TANTI_M[p];
                const double sc = h == 0 ? 1.0 : -k0sq;
#pragma unroll
                for (int c = 0; c < 3; ++c) coefa[h][p][c] = sc * tab[c];
            }
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveSmem<FULL>& sm = smem[wave];
    const long long tile0 = ((long long)blockIdx.x * NW + wave) * kWTile;
    if (tile0 >= m.n_assy) return;  // whole wavefront
    const int ntile = (int)min((long long)kWTile, m.n_assy - tile0);
    // the upstream blocks of the next AH tets, in flight while the current ones run
    double bre[AH][7], bim[AH][7];
#pragma unroll
    for (int q = 0; q < AH; ++q)
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int e = lane + 64 * k;
            double2 v = make_double2(0.0, 0.0);
            if (e < 400 && q < ntile) v = grad[(tile0 + q) * 400 + e];
            bre[q][k] = v.x;
            bim[q][k] = v.y;
        }
    // this lane's components and pair range with their coefficients, and the functions
    // and LDS slots of the pairs whose weights it forms
    const int row = lane >> 4, j = lane & 15;
    // One unconditional load per coefficient: the table and the scale are selected per
    // lane, so the 42 loads are issued back to back and waited for once (a select between
    // the two table reads becomes a branch and a wait per coefficient).
    double coef[kWComps][kWPairs];
#pragma unroll
    for (int c = 0; c < kWComps; ++c) {
        const int comp = kWComps * row + c;
        const bool isk = comp < 6;
        const double* tab = isk ? &TSYM_K[0][comp] : &TSYM_M[0][comp - 6];
        const double sc = j < 15 ? (isk ? 1.0 : -k0sq) : 0.0;
#pragma unroll
        for (int k = 0; k < kWPairs; ++k) {
            const int p = min(kWPairs * j + k, NED2_NPAIR - 1);
            coef[c][k] = sc * tab[6 * p];
        }
    }
    int pab[4];  // the functions of the lane's pairs, a | b << 8
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int p = min(lane + 64 * k, NED2_NPAIR - 1);
        pab[k] = PAIR_A[p] | PAIR_B[p] << 8;
    }
    // lane 15's range (read with zero coefficients) is never written by the pair phase
#pragma unroll
    for (int q = 0; q < AH; ++q)
        if (lane < kWStride) {
            sm.w[q][15 * kWStride + lane] = make_double2(0.0, 0.0);
            if (FULL) sm.wa[FULL ? q : 0][FULL ? 15 * kWStride + lane : 0] = make_double2(0.0, 0.0);
        }
    // geometry of the lane's tet
    LaneGeom gm;
    {
        double C[9], sg;
        bwd_geometry(m, tile0 + min(lane, ntile - 1), gm.Jc, C, gm.iad, gm.tet, gm.code, sg);
        gm.neg = sg < 0.0;
    }
    for (int lt0 = 0; lt0 < ntile; lt0 += AH) {
        // stage the blocks, then start the loads of the next group into the same registers
#pragma unroll
        for (int q = 0; q < AH; ++q)
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int e = lane + 64 * k;
                if (e < 400) sm.blk[q][e] = make_double2(bre[q][k], bim[q][k]);
            }
#pragma unroll
        for (int q = 0; q < AH; ++q)
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int e = lane + 64 * k;
                if (e < 400 && lt0 + AH + q < ntile) {
                    const double2 v = grad[(tile0 + lt0 + AH + q) * 400 + e];
                    bre[q][k] = v.x;
                    bim[q][k] = v.y;
                }
            }
        wave_sync();
        wave_group<FULL>(sm, slot, coef, &coefa[0][0][0], pab, gm, lane, row, j, lt0,
                         min(AH, ntile - lt0), ntile, eps_kind, mu_kind, g_eps, g_mu);
    }
}

}  // namespace

hipError_t launch_backward(const MeshDev& m, const double2* grad, double k0sq, int eps_kind,
                           int mu_kind, double2* g_eps, double2* g_mu, hipStream_t stream) {
    if (m.n_assy <= 0) return hipSuccess;
    const bool full = (g_eps && eps_kind == EM_MAT_FULL) || (g_mu && mu_kind == EM_MAT_FULL);
    const long long wtiles = (m.n_assy + kWTile - 1) / kWTile;
    if (full) {
        constexpr int nw = WaveCfg<true>::WAVES;
        hipLaunchKernelGGL(ned2_backward_wave_kernel<true>, dim3((unsigned)((wtiles + nw - 1) / nw)),
                           dim3(64 * nw), 0, stream, m, grad, k0sq, eps_kind, mu_kind, g_eps,
                           g_mu);
    } else {
        constexpr int nw = WaveCfg<false>::WAVES;
        hipLaunchKernelGGL(ned2_backward_wave_kernel<false>,
                           dim3((unsigned)((wtiles + nw - 1) / nw)), dim3(64 * nw), 0, stream, m,
                           grad, k0sq, eps_kind, mu_kind, g_eps, g_mu);
    }
    return hipGetLastError();
}

}  // namespace em
