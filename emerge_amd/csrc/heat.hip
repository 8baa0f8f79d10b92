// Lagrange-2 heat-conduction assembly: element conductance (stiffness) and capacity (mass)
// blocks of every tet, 10x10, in the reference's COO layout.
//
// Reference: src/emerge/_emerge/physics/heatconduction/assembly/grad.py:199-414
// (leg2_tet_stiff, leg2_tet_mass, _stiff_matrix_builder, _mass_matrix_builder).
// The reference integrates with its 8-digit 11-point table; HEAT_Q / HEAT_R are that rule
// applied to the polynomials (gen_heat_tables.py), so per tet
//   K_ij = V sum_ab (grad L_b . C grad L_a) HEAT_Q[s][ij][ab],  M_ij = rho_cp V HEAT_R[s][ij]
// with V = |det J| / 6, grad L_a the barycentric gradients (from vertex differences) and s
// the sign of det J (the reference's |V| normalisation). For a symmetric conductivity
// (scalar, diagonal) H is symmetric and the sum runs over its 10 distinct terms with the
// off-diagonal coefficient pairs added. HBM-bound: 1,600 B written per tet against
// ~100 B read; the contraction is 10 (16) FMA per K entry.
//
// Work layout: the output is written in 4 KiB chunks (512 entries of K and of M). Workgroup
// b writes chunks b, b + G, ..., b + (steps - 1) G of both arrays, one 16-byte store per
// thread and array per chunk, with G a multiple of 200 chunks. Workgroups are dispatched
// round-robin over the 8 XCDs, so each XCD writes only the chunks of its residue mod 8;
// a stores-only replica of this order moves 4.5 GB in 0.69 ms against 0.93 ms for runs of
// 64 consecutive tets per workgroup (that layout with its compute: 1.03 ms). And since 512 G = 0 (mod 100), each thread owns the same entry pair of its tet in every
// chunk, so its coefficients stay in registers.
#include <hip/hip_runtime.h>

#include "em_internal.h"
#include "heat_tables.inc"

namespace em {

namespace {

constexpr int kHeatBlock = 256;

// Coefficients of entries e and e + 1 for orientation s: the 16 HEAT_Q terms, or for a
// symmetric conductivity the 10 terms of the symmetric H (off-diagonal pairs summed).
template <int NT>
__device__ __forceinline__ void heat_coefs(int s, int e, double (&q)[2][NT]) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const double* Q = HEAT_Q[s][e + u];
        if (NT == 16) {
#pragma unroll
            for (int k = 0; k < 16; ++k) q[u][k] = Q[k];
        } else {
            // H index 4 a + b; symmetric order 00 11 22 33 01 02 03 12 13 23
            q[u][0] = Q[0];
            q[u][1] = Q[5];
            q[u][2] = Q[10];
            q[u][3] = Q[15];
            q[u][4] = Q[1] + Q[4];
            q[u][5] = Q[2] + Q[8];
            q[u][6] = Q[3] + Q[12];
            q[u][7] = Q[6] + Q[9];
            q[u][8] = Q[7] + Q[13];
            q[u][9] = Q[11] + Q[14];
        }
    }
}

// Conductance terms (16, or 10 when C is symmetric), V, rho_cp V and the orientation s
// (0: det J > 0) of tet t into one 160-byte row r: terms [0, 16), V [16], rho_cp V [17],
// s [18]. Returns s.
template <int NC>
__device__ __forceinline__ int heat_row(const double4* __restrict__ xyz,
                                        const int4* __restrict__ tets, long long t,
                                        const double* __restrict__ cond, long long cts,
                                        long long ccs, const double* __restrict__ rho,
                                        long long rts, double* r) {
    constexpr int NT = NC == 9 ? 16 : 10;
    const int4 v = tets[t];
    const double4 p0 = xyz[v.x], p1 = xyz[v.y], p2 = xyz[v.z], p3 = xyz[v.w];
    const double J[3][3] = {{p1.x - p0.x, p1.y - p0.y, p1.z - p0.z},
                            {p2.x - p0.x, p2.y - p0.y, p2.z - p0.z},
                            {p3.x - p0.x, p3.y - p0.y, p3.z - p0.z}};
    // grad L_{k+1} = (J_{k+1} x J_{k+2}) / det (rows of the inverse)
    double g[4][3];
    for (int k = 0; k < 3; ++k) {
        const double* a = J[(k + 1) % 3];
        const double* b = J[(k + 2) % 3];
        g[k + 1][0] = a[1] * b[2] - a[2] * b[1];
        g[k + 1][1] = a[2] * b[0] - a[0] * b[2];
        g[k + 1][2] = a[0] * b[1] - a[1] * b[0];
    }
    const double det = J[0][0] * g[1][0] + J[0][1] * g[1][1] + J[0][2] * g[1][2];
    for (int k = 1; k < 4; ++k)
        for (int c = 0; c < 3; ++c) g[k][c] /= det;
    for (int c = 0; c < 3; ++c) g[0][c] = -(g[1][c] + g[2][c] + g[3][c]);
    if (NC > 0) {
        const double* cp = cond + t * cts;
        double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (NC == 1) {
            C[0] = C[4] = C[8] = cp[0];
        } else if (NC == 3) {
            C[0] = cp[0];
            C[4] = cp[ccs];
            C[8] = cp[2 * ccs];
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) C[k] = cp[k * ccs];
        }
        // H_ab = grad L_b . C grad L_a
        double H[16];
        for (int a = 0; a < 4; ++a) {
            const double cg[3] = {C[0] * g[a][0] + C[1] * g[a][1] + C[2] * g[a][2],
                                  C[3] * g[a][0] + C[4] * g[a][1] + C[5] * g[a][2],
                                  C[6] * g[a][0] + C[7] * g[a][1] + C[8] * g[a][2]};
            for (int b = 0; b < 4; ++b)
                H[4 * a + b] = g[b][0] * cg[0] + g[b][1] * cg[1] + g[b][2] * cg[2];
        }
        if (NT == 16) {
#pragma unroll
            for (int k = 0; k < 16; ++k) r[k] = H[k];
        } else {
            r[0] = H[0];
            r[1] = H[5];
            r[2] = H[10];
            r[3] = H[15];
            r[4] = H[1];
            r[5] = H[2];
            r[6] = H[3];
            r[7] = H[6];
            r[8] = H[7];
            r[9] = H[11];
        }
    }
    const double V = fabs(det) / 6.0;
    r[16] = V;
    r[17] = rho ? rho[t * rts] * V : 0.0;
    const int sg = det > 0 ? 0 : 1;
    r[18] = (double)sg;
    return sg;
}

// Entries o, o + 1 of K (and M when M is given) from a tet's row and the entry pair's
// coefficients; 16-byte non-temporal stores (o is even).
template <int NC>
__device__ __forceinline__ void heat_store(const double* r, const double (&q)[2][NC == 9 ? 16 : 10],
                                           const double (&R)[2], long long o,
                                           double* __restrict__ K, double* __restrict__ M) {
    constexpr int NT = NC == 9 ? 16 : 10;
    if (NC > 0) {
        double h[NT];
#pragma unroll
        for (int k = 0; k < NT; ++k) h[k] = r[k];
        double k0a = 0.0, k0b = 0.0, k1a = 0.0, k1b = 0.0;
#pragma unroll
        for (int k = 0; k < NT; k += 2) {
            k0a = fma(h[k], q[0][k], k0a);
            k1a = fma(h[k], q[1][k], k1a);
            if (k + 1 < NT) {
                k0b = fma(h[k + 1], q[0][k + 1], k0b);
                k1b = fma(h[k + 1], q[1][k + 1], k1b);
            }
        }
        const double V = r[16];
        __builtin_nontemporal_store((k0a + k0b) * V, K + o);
        __builtin_nontemporal_store((k1a + k1b) * V, K + o + 1);
    }
    if (M) {
        const double rv = r[17];
        __builtin_nontemporal_store(rv * R[0], M + o);
        __builtin_nontemporal_store(rv * R[1], M + o + 1);
    }
}

// NC conductivity values per tet (0: K not asked for, 1 scalar, 3 diagonal, 9 full); NT
// conductance terms per tet (10 when C is symmetric, else 16). Workgroup b writes chunks
// b + k G, k < steps (4 KiB chunks of 512 entries of K and of M; G a multiple of 200). A
// chunk touches at most 7 tets: threads 7 k + j compute tet j of chunk k into LDS before
// the workgroup's one barrier; then thread i writes entries 2i, 2i + 1 of every chunk.
constexpr int kChunk = 512;
constexpr int kChunkTets = 7;
constexpr int kMaxChunkSteps = 32;  // 7 x 32 tets fill 224 of 256 threads

template <int NC>
__global__ __launch_bounds__(kHeatBlock) void heat_chunk_kernel(
    const double4* __restrict__ xyz, const int4* __restrict__ tets, long long T,
    const double* __restrict__ cond, long long cts, long long ccs,
    const double* __restrict__ rho, long long rts, double* __restrict__ K,
    double* __restrict__ M, long long G, int steps) {
    constexpr int NT = NC == 9 ? 16 : 10;
    __shared__ double rows[kMaxChunkSteps * kChunkTets][20];
    const long long n = T * 100;
    const int tid = threadIdx.x;
    const long long b = blockIdx.x;
    if (tid < steps * kChunkTets) {
        const int k = tid / kChunkTets, j = tid % kChunkTets;
        const long long e0 = (b + k * G) * kChunk;
        if (e0 < n) {
            const long long last = min((e0 + kChunk - 1) / 100, T - 1);
            const long long t = e0 / 100 + j;
            if (t <= last) heat_row<NC>(xyz, tets, t, cond, cts, ccs, rho, rts, rows[tid]);
        }
    }
    const int e = (int)((b * kChunk + 2 * tid) % 100);
    int sq = 0;  // orientation whose coefficients are loaded
    double q[2][NT];
    if (NC > 0) heat_coefs<NT>(sq, e, q);
    double R[2] = {HEAT_R[sq][e], HEAT_R[sq][e + 1]};
    __syncthreads();
    for (int k = 0; k < steps; ++k) {
        const long long c0 = (b + k * G) * kChunk;
        const long long o = c0 + 2 * tid;
        if (o >= n) break;
        const double* r = rows[k * kChunkTets + (int)(o / 100 - c0 / 100)];
        const int sl = (int)r[18];
        if (sl != sq) {  // a negatively oriented tet: the other coefficient set
            sq = sl;
            if (NC > 0) heat_coefs<NT>(sq, e, q);
            R[0] = HEAT_R[sq][e];
            R[1] = HEAT_R[sq][e + 1];
        }
        heat_store<NC>(r, q, R, o, K, M);
    }
}

}  // namespace

hipError_t heat_assemble(const double4* xyz, const int4* tets, long long T, const double* cond,
                         int cond_kind, long long cts, long long ccs, const double* rho,
                         long long rts, double* K, double* M, hipStream_t s) {
    if (T <= 0) return hipSuccess;
    // chunks per workgroup: 32 on large meshes (2.8M tets: 0.81 ms against 0.85-0.87 at
    // 16-24 and 0.91 at 8), 8 below 65,536 chunks (335k tets), where 32 would leave few
    // workgroups (21k tets: 0.015 ms at 8, 0.035 at 32)
    const long long nchunk = (T * 100 + kChunk - 1) / kChunk;
    const int steps = nchunk >= 65536 ? kMaxChunkSteps : 8;
    long long G = (nchunk + steps - 1) / steps;
    G = (G + 199) / 200 * 200;
    if (G > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 g((unsigned)G);
    if (!K) {
        hipLaunchKernelGGL(heat_chunk_kernel<0>, g, dim3(kHeatBlock), 0, s, xyz, tets, T, nullptr,
                           0LL, 0LL, rho, rts, nullptr, M, G, steps);
    } else if (cond_kind == EM_MAT_SCALAR) {
        hipLaunchKernelGGL(heat_chunk_kernel<1>, g, dim3(kHeatBlock), 0, s, xyz, tets, T, cond,
                           cts, ccs, rho, rts, K, M, G, steps);
    } else if (cond_kind == EM_MAT_DIAG) {
        hipLaunchKernelGGL(heat_chunk_kernel<3>, g, dim3(kHeatBlock), 0, s, xyz, tets, T, cond,
                           cts, ccs, rho, rts, K, M, G, steps);
    } else {
        hipLaunchKernelGGL(heat_chunk_kernel<9>, g, dim3(kHeatBlock), 0, s, xyz, tets, T, cond,
                           cts, ccs, rho, rts, K, M, G, steps);
    }
    return hipGetLastError();
}

namespace {

__global__ __launch_bounds__(256) void heat_rowcol_kernel(const int* __restrict__ t2f, long long T,
                                                          long long* __restrict__ rows,
                                                          long long* __restrict__ cols) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < T * 100;
         k += (long long)gridDim.x * blockDim.x) {
        const long long t = k / 100;
        const int e = (int)(k - t * 100);
        rows[k] = t2f[t * 10 + e / 10];
        cols[k] = t2f[t * 10 + e % 10];
    }
}

}  // namespace

hipError_t heat_rowcol(const int* t2f, long long T, long long* rows, long long* cols,
                       hipStream_t s) {
    if (T <= 0) return hipSuccess;
    long long b = (T * 100 + 255) / 256;
    const unsigned g = (unsigned)(b < 65536 * 4 ? b : 65536 * 4);
    hipLaunchKernelGGL(heat_rowcol_kernel, dim3(g), dim3(256), 0, s, t2f, T, rows, cols);
    return hipGetLastError();
}

}  // namespace em
