// Robin / port surface assembly on boundary triangles (Nedelec-2, 8 functions per face).
//
// Reference: src/emerge/_emerge/physics/microwave/assembly/robinbc.py
//   ned2_tri_stiff :316-381 + compute_bc_entries :385-402   -> robin_blocks_kernel
//   ned2_tri_force :220-275 + compute_force_entries :278-313,
//   ned2_tri_force_scat :453-514 + compute_force_entries_scat :517-569 -> robin_force_kernel
//   FemData.empty_tri_rowcol (elements/femdata.py:75-94)       -> tri_rowcol_kernel
//   K += field.generate_csc(B_matrix_robin) (assembler.py:635)  -> csc_find_kernel +
//                                                                 csc_add_kernel
//
// The face's vertices are its node ids ascending (the mesh's sorted face tuples), so
// every function has its vertices in local order and no orientation sign enters.
// Blocks: B_ij = gamma int f_i . f_j dA = gamma / (2A) sum_c (e_a . e_b) TRI_E[ij][c]
// (exact integrals; the reference's 6-point rule is exact for these degree-4
// integrands). Force: the reference's 6-point rule on the caller's mode field at the
// rule's points, b_i = -A sum_w w_w f_i(x_w) . U'(x_w), U' = U - s (n x curl U) in the
// scattered-field form (s the sign of the port normal along the face normal), added to
// the global vector through tri_to_field. Everything is FP64; HBM/latency bound, small.
#include <hip/hip_runtime.h>

#include "em_internal.h"
#include "tri_tables.inc"

namespace em {

namespace {

// 6-point degree-4 triangle rule: weights, then the barycentrics (L1, L2, L3) of each
// point, with the digits the reference uses (robinbc.py:185-192).
__constant__ double kTriW[6] = {0.10995174365532200, 0.10995174365532200, 0.10995174365532200,
                                0.22338158967801100, 0.22338158967801100, 0.22338158967801100};
__constant__ double kTriL[3][6] = {
    {0.81684757298045896, 0.09157621350977101, 0.09157621350977101, 0.10810301816807000,
     0.44594849091596500, 0.44594849091596500},
    {0.09157621350977101, 0.81684757298045896, 0.09157621350977101, 0.44594849091596500,
     0.10810301816807000, 0.44594849091596500},
    {0.09157621350977004, 0.09157621350977008, 0.81684757298045807, 0.44594849091596500,
     0.44594849091596500, 0.10810301816807000}};

unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536 * 4) b = 65536 * 4;
    return (unsigned)b;
}

struct TriGeo {
    double v[3][3];  // vertices
    double e[3][3];  // e_0 = v2 - v1, e_1 = v0 - v2, e_2 = v1 - v0
    double n[3];     // unit normal (v1 - v0) x (v2 - v0)
    double A;        // area
};

// Loads face f; false when an id is out of range.
__device__ bool load_tri(const double4* __restrict__ xyz, long long n_nodes,
                         const long long* __restrict__ tris, long long F, long long f,
                         TriGeo& g) {
    if (f < 0 || f >= F) return false;
    for (int k = 0; k < 3; ++k) {
        const long long id = tris[k * F + f];
        if (id < 0 || id >= n_nodes) return false;
        const double4 p = xyz[id];
        g.v[k][0] = p.x;
        g.v[k][1] = p.y;
        g.v[k][2] = p.z;
    }
    for (int q = 0; q < 3; ++q) {
        g.e[0][q] = g.v[2][q] - g.v[1][q];
        g.e[1][q] = g.v[0][q] - g.v[2][q];
        g.e[2][q] = g.v[1][q] - g.v[0][q];
    }
    // (v1 - v0) x (v2 - v0) = e_2 x (-e_1)
    const double c0 = -(g.e[2][1] * g.e[1][2] - g.e[2][2] * g.e[1][1]);
    const double c1 = -(g.e[2][2] * g.e[1][0] - g.e[2][0] * g.e[1][2]);
    const double c2 = -(g.e[2][0] * g.e[1][1] - g.e[2][1] * g.e[1][0]);
    const double len = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
    g.A = 0.5 * len;
    g.n[0] = c0 / len;
    g.n[1] = c1 / len;
    g.n[2] = c2 / len;
    return true;
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// One thread per (listed face, block entry). out: by face id (face * 64) or compact
// (listed position * 64); both accumulate.
__global__ __launch_bounds__(256) void robin_blocks_kernel(
    const double4* __restrict__ xyz, long long n_nodes, const long long* __restrict__ tris,
    long long F, const long long* __restrict__ surf, long long nsurf, double gre, double gim,
    double2* __restrict__ out, int by_face, unsigned long long* __restrict__ bad) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nsurf * 64;
         t += (long long)gridDim.x * blockDim.x) {
        const long long s = t / 64;
        const int e = (int)(t - s * 64);
        const long long f = surf[s];
        TriGeo g;
        if (!load_tri(xyz, n_nodes, tris, F, f, g)) {
            atomicOr(bad, 1ull);
            continue;
        }
        const double E[6] = {dot3(g.e[0], g.e[0]), dot3(g.e[1], g.e[1]), dot3(g.e[2], g.e[2]),
                             dot3(g.e[0], g.e[1]), dot3(g.e[0], g.e[2]), dot3(g.e[1], g.e[2])};
        double val = 0.0;
#pragma unroll
        for (int c = 0; c < 6; ++c) val += E[c] * TRI_E[e][c];
        val /= 2.0 * g.A;
        double2* o = out + (by_face ? f : s) * 64 + e;
        double2 cur = *o;
        cur.x += gre * val;
        cur.y += gim * val;
        *o = cur;
    }
}

// f_i at barycentrics L (surface gradients gr[a] = n x e_a / 2A), as a 3D vector.
__device__ __forceinline__ void tri_func(int i, const double L[3], const double gr[3][3],
                                         double f[3]) {
    double pa[3] = {0.0, 0.0, 0.0};  // f = sum_a pa[a] gr[a]
    if (i < 6) {
        const int ed = i % 3;
        const int a = ed == 1 ? 1 : 0, b = ed == 0 ? 1 : 2;  // (0,1), (1,2), (0,2)
        const double s = i < 3 ? 1.0 : L[a] - L[b];
        pa[b] = s * L[a];
        pa[a] = -s * L[b];
    } else if (i == 6) {  // l0 (l1 g2 - l2 g1)
        pa[2] = L[0] * L[1];
        pa[1] = -L[0] * L[2];
    } else {  // l1 (l0 g2 - l2 g0)
        pa[2] = L[1] * L[0];
        pa[0] = -L[1] * L[2];
    }
    for (int q = 0; q < 3; ++q) f[q] = pa[0] * gr[0][q] + pa[1] * gr[1][q] + pa[2] * gr[2][q];
}

// One thread per (listed face, function). U, Uc: complex [3][6][nsurf]; normals [3][nsurf].
__global__ __launch_bounds__(256) void robin_force_kernel(
    const double4* __restrict__ xyz, long long n_nodes, const long long* __restrict__ tris,
    long long F, const long long* __restrict__ tri_to_field, long long n_field,
    const long long* __restrict__ surf, long long nsurf, const double2* __restrict__ U,
    const double2* __restrict__ Uc, const double* __restrict__ normals,
    double2* __restrict__ bvec, unsigned long long* __restrict__ bad) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nsurf * 8;
         t += (long long)gridDim.x * blockDim.x) {
        const long long s = t / 8;
        const int i = (int)(t - s * 8);
        const long long f = surf[s];
        TriGeo g;
        if (!load_tri(xyz, n_nodes, tris, F, f, g)) {
            atomicOr(bad, 1ull);
            continue;
        }
        const long long d = tri_to_field[i * F + f];
        if (d < 0 || d >= n_field) {
            atomicOr(bad, 2ull);
            continue;
        }
        double gr[3][3];
        for (int a = 0; a < 3; ++a) {
            // n x e_a / (2A)
            gr[a][0] = (g.n[1] * g.e[a][2] - g.n[2] * g.e[a][1]) / (2.0 * g.A);
            gr[a][1] = (g.n[2] * g.e[a][0] - g.n[0] * g.e[a][2]) / (2.0 * g.A);
            gr[a][2] = (g.n[0] * g.e[a][1] - g.n[1] * g.e[a][0]) / (2.0 * g.A);
        }
        double sgn = 0.0;
        if (Uc) {
            const double nz = g.n[0] * normals[s] + g.n[1] * normals[nsurf + s] +
                              g.n[2] * normals[2 * nsurf + s];
            sgn = (nz > 0.0) - (nz < 0.0);
        }
        double sr = 0.0, si = 0.0;
        for (int w = 0; w < 6; ++w) {
            const double L[3] = {kTriL[0][w], kTriL[1][w], kTriL[2][w]};
            double fv[3];
            tri_func(i, L, gr, fv);
            double ur[3], ui[3];
            for (int q = 0; q < 3; ++q) {
                const double2 u = U[(q * 6 + w) * nsurf + s];
                ur[q] = u.x;
                ui[q] = u.y;
            }
            if (Uc) {  // U - s (n x C)
                double cr[3], ci[3];
                for (int q = 0; q < 3; ++q) {
                    const double2 c = Uc[(q * 6 + w) * nsurf + s];
                    cr[q] = c.x;
                    ci[q] = c.y;
                }
                const double xr[3] = {g.n[1] * cr[2] - g.n[2] * cr[1],
                                      g.n[2] * cr[0] - g.n[0] * cr[2],
                                      g.n[0] * cr[1] - g.n[1] * cr[0]};
                const double xi[3] = {g.n[1] * ci[2] - g.n[2] * ci[1],
                                      g.n[2] * ci[0] - g.n[0] * ci[2],
                                      g.n[0] * ci[1] - g.n[1] * ci[0]};
                for (int q = 0; q < 3; ++q) {
                    ur[q] -= sgn * xr[q];
                    ui[q] -= sgn * xi[q];
                }
            }
            sr += kTriW[w] * dot3(fv, ur);
            si += kTriW[w] * dot3(fv, ui);
        }
        atomicAdd(&bvec[d].x, -g.A * sr);
        atomicAdd(&bvec[d].y, -g.A * si);
    }
}

__global__ __launch_bounds__(256) void tri_rowcol_kernel(const long long* __restrict__ tri_to_field,
                                                         long long F,
                                                         const long long* __restrict__ surf,
                                                         long long nsurf, long long* __restrict__ rows,
                                                         long long* __restrict__ cols,
                                                         unsigned long long* __restrict__ bad) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nsurf * 64;
         t += (long long)gridDim.x * blockDim.x) {
        const long long s = t / 64;
        const int e = (int)(t - s * 64);
        const long long f = surf[s];
        if (f < 0 || f >= F) {
            atomicOr(bad, 1ull);
            continue;
        }
        rows[t] = tri_to_field[(e / 8) * F + f];
        cols[t] = tri_to_field[(e % 8) * F + f];
    }
}

// Position of (row, col) in a CSC pattern with sorted rows (binary search), -1 if absent.
__global__ __launch_bounds__(256) void csc_find_kernel(
    const long long* __restrict__ indptr, const long long* __restrict__ indices, long long N,
    const long long* __restrict__ rows, const long long* __restrict__ cols,
    const long long* __restrict__ new_index, long long n, long long* __restrict__ pos) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        long long r = rows[t], c = cols[t], q = -1;
        if (new_index) {
            r = (r >= 0) ? new_index[r] : -1;
            c = (c >= 0) ? new_index[c] : -1;
        }
        if (r >= 0 && r < N && c >= 0 && c < N) {
            long long lo = indptr[c], hi = indptr[c + 1];
            while (lo < hi) {
                const long long mid = (lo + hi) / 2;
                if (indices[mid] < r) lo = mid + 1; else hi = mid;
            }
            if (lo < indptr[c + 1] && indices[lo] == r) q = lo;
        }
        pos[t] = q;
    }
}

template <typename T>
__device__ __forceinline__ void atomic_add_c(T* p, double re, double im);
template <>
__device__ __forceinline__ void atomic_add_c<double2>(double2* p, double re, double im) {
    atomicAdd(&p->x, re);
    atomicAdd(&p->y, im);
}
template <>
__device__ __forceinline__ void atomic_add_c<float2>(float2* p, double re, double im) {
    atomicAdd(&p->x, (float)re);
    atomicAdd(&p->y, (float)im);
}

template <typename T>
__global__ __launch_bounds__(256) void csc_add_kernel(T* __restrict__ data,
                                                      const long long* __restrict__ pos,
                                                      const double2* __restrict__ vals,
                                                      long long n, double sre, double sim) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const long long q = pos[t];
        if (q < 0) continue;
        const double2 v = vals[t];
        atomic_add_c<T>(data + q, sre * v.x - sim * v.y, sre * v.y + sim * v.x);
    }
}

hipError_t sync_bad(unsigned long long* dbad, long long& bad, hipStream_t s) {
    unsigned long long h = 0;
    hipError_t e = hipMemcpyAsync(&h, dbad, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    bad = (long long)h;
    return e;
}

}  // namespace

hipError_t robin_blocks(const double4* xyz, long long n_nodes, const long long* tris, long long F,
                        const long long* surf, long long nsurf, double gre, double gim,
                        double2* out, int by_face, long long& bad, hipStream_t s) {
    bad = 0;
    if (nsurf <= 0) return hipSuccess;
    unsigned long long* dbad = nullptr;
    hipError_t e = hipMalloc((void**)&dbad, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(dbad, 0, sizeof(unsigned long long), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(robin_blocks_kernel, dim3(grid_for(nsurf * 64)), dim3(256), 0, s, xyz,
                           n_nodes, tris, F, surf, nsurf, gre, gim, out, by_face, dbad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = sync_bad(dbad, bad, s);
    (void)hipFree(dbad);
    return e;
}

hipError_t robin_force(const double4* xyz, long long n_nodes, const long long* tris, long long F,
                       const long long* tri_to_field, long long n_field, const long long* surf,
                       long long nsurf, const double2* U, const double2* Uc, const double* normals,
                       double2* bvec, long long& bad, hipStream_t s) {
    bad = 0;
    if (nsurf <= 0) return hipSuccess;
    unsigned long long* dbad = nullptr;
    hipError_t e = hipMalloc((void**)&dbad, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(dbad, 0, sizeof(unsigned long long), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(robin_force_kernel, dim3(grid_for(nsurf * 8)), dim3(256), 0, s, xyz,
                           n_nodes, tris, F, tri_to_field, n_field, surf, nsurf, U, Uc, normals,
                           bvec, dbad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = sync_bad(dbad, bad, s);
    (void)hipFree(dbad);
    return e;
}

hipError_t tri_rowcol(const long long* tri_to_field, long long F, const long long* surf,
                      long long nsurf, long long* rows, long long* cols, long long& bad,
                      hipStream_t s) {
    bad = 0;
    if (nsurf <= 0) return hipSuccess;
    unsigned long long* dbad = nullptr;
    hipError_t e = hipMalloc((void**)&dbad, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(dbad, 0, sizeof(unsigned long long), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(tri_rowcol_kernel, dim3(grid_for(nsurf * 64)), dim3(256), 0, s,
                           tri_to_field, F, surf, nsurf, rows, cols, dbad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = sync_bad(dbad, bad, s);
    (void)hipFree(dbad);
    return e;
}

hipError_t csc_find(const long long* indptr, const long long* indices, long long N,
                    const long long* rows, const long long* cols, const long long* new_index,
                    long long n, long long* pos, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(csc_find_kernel, dim3(grid_for(n)), dim3(256), 0, s, indptr, indices, N,
                       rows, cols, new_index, n, pos);
    return hipGetLastError();
}

hipError_t csc_add(void* data, int complex64, const long long* pos, const double2* vals,
                   long long n, double sre, double sim, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (complex64)
        hipLaunchKernelGGL(csc_add_kernel<float2>, dim3(grid_for(n)), dim3(256), 0, s,
                           (float2*)data, pos, vals, n, sre, sim);
    else
        hipLaunchKernelGGL(csc_add_kernel<double2>, dim3(grid_for(n)), dim3(256), 0, s,
                           (double2*)data, pos, vals, n, sre, sim);
    return hipGetLastError();
}

}  // namespace em
