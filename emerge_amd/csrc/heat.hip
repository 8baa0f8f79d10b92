// Lagrange-2 heat-conduction assembly: element conductance (stiffness) and capacity (mass)
// blocks of every tet, 10x10, in the reference's COO layout.
//
// Reference: src/emerge/_emerge/physics/heatconduction/assembly/grad.py:199-414
// (leg2_tet_stiff, leg2_tet_mass, _stiff_matrix_builder, _mass_matrix_builder).
// The reference integrates with its 8-digit 11-point table; HEAT_Q / HEAT_R are that rule
// applied to the polynomials (gen_heat_tables.py), so per tet
//   K_ij = V sum_ab (grad L_b . C grad L_a) HEAT_Q[s][ij][ab],  M_ij = rho_cp V HEAT_R[s][ij]
// with V = |det J| / 6, grad L_a the barycentric gradients (from vertex differences) and s
// the sign of det J (the reference's |V| normalisation). HBM-bound: 1,600 B written per tet
// against ~100 B read; the contraction is 16 FMA per K entry.
#include <hip/hip_runtime.h>

#include "em_internal.h"
#include "heat_tables.inc"

namespace em {

namespace {

constexpr int kHeatBlock = 128;
constexpr int kHeatTile = 64;

struct HeatRec {
    double H[kHeatTile][16];  // grad L_b . C grad L_a, index 4 a + b
    double V[kHeatTile];
    double rV[kHeatTile];
    int s[kHeatTile];
};

__global__ __launch_bounds__(kHeatBlock) void heat_assemble_kernel(
    const double4* __restrict__ xyz, const int4* __restrict__ tets, long long T,
    const double* __restrict__ cond, int cond_kind, long long cts, long long ccs,
    const double* __restrict__ rho, long long rts, double* __restrict__ K,
    double* __restrict__ M) {
    __shared__ HeatRec rec;
    const int tid = threadIdx.x;
    // coefficients of this thread's entry for both orientations
    double q0[16], q1[16], r0 = 0.0, r1 = 0.0;
    const int e = tid < 100 ? tid : 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        q0[k] = HEAT_Q[0][e][k];
        q1[k] = HEAT_Q[1][e][k];
    }
    r0 = HEAT_R[0][e];
    r1 = HEAT_R[1][e];
    const long long ntiles = (T + kHeatTile - 1) / kHeatTile;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long long t0 = tile * kHeatTile;
        const int cnt = (int)(T - t0 < kHeatTile ? T - t0 : kHeatTile);
        if (tid < cnt) {
            const long long t = t0 + tid;
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
            // the conductivity is read only when K is asked for (cond may be null otherwise)
            double C[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (K) {
                const double* cp = cond + t * cts;
                if (cond_kind == EM_MAT_SCALAR) {
                    C[0] = C[4] = C[8] = cp[0];
                } else if (cond_kind == EM_MAT_DIAG) {
                    C[0] = cp[0];
                    C[4] = cp[ccs];
                    C[8] = cp[2 * ccs];
                } else {
                    for (int k = 0; k < 9; ++k) C[k] = cp[k * ccs];
                }
            }
            for (int a = 0; a < 4; ++a) {
                const double cg[3] = {C[0] * g[a][0] + C[1] * g[a][1] + C[2] * g[a][2],
                                      C[3] * g[a][0] + C[4] * g[a][1] + C[5] * g[a][2],
                                      C[6] * g[a][0] + C[7] * g[a][1] + C[8] * g[a][2]};
                for (int b = 0; b < 4; ++b)
                    rec.H[tid][4 * a + b] = g[b][0] * cg[0] + g[b][1] * cg[1] + g[b][2] * cg[2];
            }
            const double V = fabs(det) / 6.0;
            rec.V[tid] = V;
            rec.rV[tid] = rho ? rho[t * rts] * V : 0.0;
            rec.s[tid] = det > 0 ? 0 : 1;
        }
        __syncthreads();
        if (tid < 100) {
            for (int lt = 0; lt < cnt; ++lt) {
                const long long o = (t0 + lt) * 100 + tid;
                const bool neg = rec.s[lt] != 0;
                if (K) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < 16; ++k) acc += rec.H[lt][k] * (neg ? q1[k] : q0[k]);
                    __builtin_nontemporal_store(acc * rec.V[lt], K + o);
                }
                if (M) __builtin_nontemporal_store(rec.rV[lt] * (neg ? r1 : r0), M + o);
            }
        }
        __syncthreads();
    }
}

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

hipError_t heat_assemble(const double4* xyz, const int4* tets, long long T, const double* cond,
                         int cond_kind, long long cts, long long ccs, const double* rho,
                         long long rts, double* K, double* M, hipStream_t s) {
    if (T <= 0) return hipSuccess;
    long long ntiles = (T + kHeatTile - 1) / kHeatTile;
    const unsigned g = (unsigned)(ntiles < 65536 * 4 ? ntiles : 65536 * 4);
    hipLaunchKernelGGL(heat_assemble_kernel, dim3(g), dim3(kHeatBlock), 0, s, xyz, tets, T, cond,
                       cond_kind, cts, ccs, rho, rts, K, M);
    return hipGetLastError();
}

hipError_t heat_rowcol(const int* t2f, long long T, long long* rows, long long* cols,
                       hipStream_t s) {
    if (T <= 0) return hipSuccess;
    long long b = (T * 100 + 255) / 256;
    const unsigned g = (unsigned)(b < 65536 * 4 ? b : 65536 * 4);
    hipLaunchKernelGGL(heat_rowcol_kernel, dim3(g), dim3(256), 0, s, t2f, T, rows, cols);
    return hipGetLastError();
}

}  // namespace em
