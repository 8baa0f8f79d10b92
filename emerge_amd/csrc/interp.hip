// Field interpolation of a Nedelec-2 solution at arbitrary points, and the point location
// it needs.
//
// Reference:
//   compiled/base/index_inter.py:37-298 (index_interp: uniform grid over the query points,
//     every candidate tet tests the points of the grid cells its bounding box overlaps with
//     barycentric tolerance EPS = 1e-8; a point gets the largest tet id containing it,
//     -1 if none)                                                 -> locate_* kernels
//   compiled/base/interp.py:286-387 (ned2_tet_interp) and :410-519 (ned2_tet_interp_curl:
//     the curl times a per-tet constant)                           -> ned2_interp_kernel
// Each function's vertices are in ascending global-id order (local_mapping); the
// reference's barycentric coefficients are normalised by |V|, so the odd-degree functions
// (codes 66, 128, 129) and their curls carry s = sign(det J) of the stored vertex order.
// Points without a tet give NaN.
#include <hip/hip_runtime.h>

#include "em_internal.h"

namespace em {

namespace {

constexpr double kEps = 1e-8;

unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536 * 4) b = 65536 * 4;
    return (unsigned)b;
}

struct Grid {
    double x0, y0, z0, dx, dy, dz;
    int res;
};

__device__ __forceinline__ int cell_index(const Grid& g, double x, double y, double z) {
    int ix = (int)((x - g.x0) / g.dx), iy = (int)((y - g.y0) / g.dy), iz = (int)((z - g.z0) / g.dz);
    ix = ix < 0 ? 0 : (ix >= g.res ? g.res - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= g.res ? g.res - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= g.res ? g.res - 1 : iz);
    return ix + iy * g.res + iz * g.res * g.res;
}

__global__ __launch_bounds__(256) void cell_count_kernel(const double* __restrict__ c,
                                                         long long n, Grid g,
                                                         unsigned long long* __restrict__ cnt) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x)
        atomicAdd(cnt + cell_index(g, c[j], c[n + j], c[2 * n + j]), 1ull);
}

__global__ __launch_bounds__(256) void cell_fill_kernel(const double* __restrict__ c, long long n,
                                                        Grid g, const long long* __restrict__ off,
                                                        unsigned long long* __restrict__ cur,
                                                        long long* __restrict__ contents) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        const int cell = cell_index(g, c[j], c[n + j], c[2 * n + j]);
        contents[off[cell] + (long long)atomicAdd(cur + cell, 1ull)] = j;
    }
}

// Bounding box of the query points: per-block min/max (NaN coordinates ignored), then one
// block folds the partials; out6 = (xmin, xmax, ymin, ymax, zmin, zmax).
constexpr int kBoxBlocks = 1024;

__global__ __launch_bounds__(256) void bbox_partial_kernel(const double* __restrict__ c,
                                                           long long n,
                                                           double* __restrict__ part) {
    __shared__ double red[6][256];
    double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const double x = c[q * n + j];
            v[2 * q] = fmin(v[2 * q], x);
            v[2 * q + 1] = fmax(v[2 * q + 1], x);
        }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int k = 0; k < 6; ++k)
                red[k][threadIdx.x] = (k & 1) ? fmax(red[k][threadIdx.x], red[k][threadIdx.x + w])
                                              : fmin(red[k][threadIdx.x], red[k][threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void bbox_final_kernel(const double* __restrict__ part, int nb,
                                                        double* __restrict__ out6) {
    if (threadIdx.x >= 6) return;
    const int k = threadIdx.x;
    double v = (k & 1) ? -INFINITY : INFINITY;
    for (int b = 0; b < nb; ++b) v = (k & 1) ? fmax(v, part[b * 6 + k]) : fmin(v, part[b * 6 + k]);
    out6[k] = v;
}

__global__ __launch_bounds__(256) void fill_i64_kernel(long long* __restrict__ p, long long n,
                                                       long long v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// One thread per candidate tet (index_inter.py:154-246).
__global__ __launch_bounds__(256) void locate_kernel(
    const double4* __restrict__ xyz, const int4* __restrict__ tets, long long n_tets,
    const long long* __restrict__ tetids, long long n_cand, const double* __restrict__ c,
    long long n, Grid g, const long long* __restrict__ off,
    const long long* __restrict__ contents, unsigned long long* __restrict__ bad,
    long long* __restrict__ assigned) {
    for (long long it = (long long)blockIdx.x * blockDim.x + threadIdx.x; it < n_cand;
         it += (long long)gridDim.x * blockDim.x) {
        const long long t = tetids ? tetids[it] : it;
        if (t < 0 || t >= n_tets) {
            atomicOr(bad, 1ull);
            continue;
        }
        const int4 v = tets[t];
        const double4 p1 = xyz[v.x], p2 = xyz[v.y], p3 = xyz[v.z], p4 = xyz[v.w];
        const double m00 = p2.x - p1.x, m10 = p2.y - p1.y, m20 = p2.z - p1.z;
        const double m01 = p3.x - p1.x, m11 = p3.y - p1.y, m21 = p3.z - p1.z;
        const double m02 = p4.x - p1.x, m12 = p4.y - p1.y, m22 = p4.z - p1.z;
        const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
                           m02 * (m10 * m21 - m11 * m20);
        const double id = 1.0 / det;
        const double b00 = (m11 * m22 - m12 * m21) * id, b01 = (m02 * m21 - m01 * m22) * id,
                     b02 = (m01 * m12 - m02 * m11) * id;
        const double b10 = (m12 * m20 - m10 * m22) * id, b11 = (m00 * m22 - m02 * m20) * id,
                     b12 = (m10 * m02 - m00 * m12) * id;
        const double b20 = (m10 * m21 - m11 * m20) * id, b21 = (m20 * m01 - m00 * m21) * id,
                     b22 = (m00 * m11 - m10 * m01) * id;
        const double bx0 = fmin(fmin(p1.x, p2.x), fmin(p3.x, p4.x));
        const double bx1 = fmax(fmax(p1.x, p2.x), fmax(p3.x, p4.x));
        const double by0 = fmin(fmin(p1.y, p2.y), fmin(p3.y, p4.y));
        const double by1 = fmax(fmax(p1.y, p2.y), fmax(p3.y, p4.y));
        const double bz0 = fmin(fmin(p1.z, p2.z), fmin(p3.z, p4.z));
        const double bz1 = fmax(fmax(p1.z, p2.z), fmax(p3.z, p4.z));
        // cells overlapping the box, one cell of margin; clamped to the grid
        int ix0 = (int)floor((bx0 - g.x0) / g.dx) - 1, ix1 = (int)floor((bx1 - g.x0) / g.dx) + 1;
        int iy0 = (int)floor((by0 - g.y0) / g.dy) - 1, iy1 = (int)floor((by1 - g.y0) / g.dy) + 1;
        int iz0 = (int)floor((bz0 - g.z0) / g.dz) - 1, iz1 = (int)floor((bz1 - g.z0) / g.dz) + 1;
        ix0 = ix0 < 0 ? 0 : ix0;
        iy0 = iy0 < 0 ? 0 : iy0;
        iz0 = iz0 < 0 ? 0 : iz0;
        ix1 = ix1 >= g.res ? g.res - 1 : ix1;
        iy1 = iy1 >= g.res ? g.res - 1 : iy1;
        iz1 = iz1 >= g.res ? g.res - 1 : iz1;
        for (int iz = iz0; iz <= iz1; ++iz)
            for (int iy = iy0; iy <= iy1; ++iy)
                for (int ix = ix0; ix <= ix1; ++ix) {
                    const int cell = ix + iy * g.res + iz * g.res * g.res;
                    for (long long ci = off[cell]; ci < off[cell + 1]; ++ci) {
                        const long long j = contents[ci];
                        const double dx = c[j] - p1.x, dy = c[n + j] - p1.y, dz = c[2 * n + j] - p1.z;
                        const double u = b00 * dx + b01 * dy + b02 * dz;
                        const double w1 = b10 * dx + b11 * dy + b12 * dz;
                        const double w2 = b20 * dx + b21 * dy + b22 * dz;
                        if (u >= -kEps && w1 >= -kEps && w2 >= -kEps && u + w1 + w2 <= 1.0 + kEps)
                            atomicMax(assigned + j, t);  // starts at -1
                    }
                }
    }
}

// Stored-vertex positions of the reference's edge and face slots (mesh3d.py:760-761).
__constant__ int kEdgeSlot[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {3, 1}, {2, 3}};
__constant__ int kFaceSlot[4][3] = {{0, 1, 2}, {0, 2, 3}, {0, 3, 1}, {1, 2, 3}};

__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// One thread per point: E (CURL = false) or c[t] * curl E (CURL = true) at the point.
template <bool CURL>
__global__ __launch_bounds__(256) void ned2_interp_kernel(
    const double4* __restrict__ xyz, const int4* __restrict__ tets, const int* __restrict__ t2f,
    long long n_tets, const double* __restrict__ c, long long n,
    const long long* __restrict__ tet_of, const double2* __restrict__ sol,
    const double2* __restrict__ ctet, double2* __restrict__ out) {
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (long long)gridDim.x * blockDim.x) {
        const long long t = tet_of[j];
        if (t < 0 || t >= n_tets) {
            const double nan = __builtin_nan("");
            for (int q = 0; q < 3; ++q) out[q * n + j] = make_double2(nan, nan);
            continue;
        }
        const int4 v = tets[t];
        const int gid[4] = {v.x, v.y, v.z, v.w};
        double P[4][3];
        for (int k = 0; k < 4; ++k) {
            const double4 p = xyz[gid[k]];
            P[k][0] = p.x;
            P[k][1] = p.y;
            P[k][2] = p.z;
        }
        // J columns = P_k - P_0; gradients of lambda_1..3 = rows of J^-1 (cofactors / det)
        double Jc[3][3];
        for (int k = 0; k < 3; ++k)
            for (int q = 0; q < 3; ++q) Jc[k][q] = P[k + 1][q] - P[0][q];
        double r[3][3];  // r[k] = J_{k+1} x J_{k+2} (cyclic) / det = grad lambda_k+1
        cross3(Jc[1], Jc[2], r[0]);
        cross3(Jc[2], Jc[0], r[1]);
        cross3(Jc[0], Jc[1], r[2]);
        const double det = Jc[0][0] * r[0][0] + Jc[0][1] * r[0][1] + Jc[0][2] * r[0][2];
        const double s = det > 0 ? 1.0 : -1.0;
        double G[4][3], L[4];
        const double d[3] = {c[j] - P[0][0], c[n + j] - P[0][1], c[2 * n + j] - P[0][2]};
        for (int k = 0; k < 3; ++k) {
            for (int q = 0; q < 3; ++q) G[k + 1][q] = r[k][q] / det;
            L[k + 1] = G[k + 1][0] * d[0] + G[k + 1][1] * d[1] + G[k + 1][2] * d[2];
        }
        for (int q = 0; q < 3; ++q) G[0][q] = -(G[1][q] + G[2][q] + G[3][q]);
        L[0] = 1.0 - L[1] - L[2] - L[3];
        double er[3] = {0, 0, 0}, ei[3] = {0, 0, 0};
        const int* map = t2f + t * 20;
        auto add = [&](int slot, const double f[3]) {
            const double2 a = sol[map[slot]];
            for (int q = 0; q < 3; ++q) {
                er[q] += f[q] * a.x;
                ei[q] += f[q] * a.y;
            }
        };
        for (int e = 0; e < 6; ++e) {
            int i = kEdgeSlot[e][0], k2 = kEdgeSlot[e][1];
            if (gid[i] > gid[k2]) {
                const int tmp = i;
                i = k2;
                k2 = tmp;
            }
            const int jv = k2;
            double f64[3], f66[3];
            if (!CURL) {
                for (int q = 0; q < 3; ++q) {
                    f64[q] = L[i] * G[jv][q] - L[jv] * G[i][q];
                    f66[q] = s * (L[i] - L[jv]) * f64[q];
                }
            } else {
                double x[3];
                cross3(G[i], G[jv], x);
                for (int q = 0; q < 3; ++q) {
                    f64[q] = 2.0 * x[q];
                    f66[q] = s * 3.0 * (L[i] - L[jv]) * x[q];
                }
            }
            add(e, f64);
            add(6 + e, f66);
        }
        for (int fc = 0; fc < 4; ++fc) {
            int a0 = kFaceSlot[fc][0], a1 = kFaceSlot[fc][1], a2 = kFaceSlot[fc][2];
            // ascending global id
            if (gid[a0] > gid[a1]) { const int tmp = a0; a0 = a1; a1 = tmp; }
            if (gid[a1] > gid[a2]) { const int tmp = a1; a1 = a2; a2 = tmp; }
            if (gid[a0] > gid[a1]) { const int tmp = a0; a0 = a1; a1 = tmp; }
            const int i = a0, jv = a1, k = a2;
            double f128[3], f129[3];
            if (!CURL) {
                for (int q = 0; q < 3; ++q) {
                    f128[q] = s * L[i] * (L[jv] * G[k][q] - L[k] * G[jv][q]);
                    f129[q] = s * L[jv] * (L[i] * G[k][q] - L[k] * G[i][q]);
                }
            } else {
                // curl(l_a (l_b g_c - l_c g_b)) = g_a x (l_b g_c - l_c g_b) + 2 l_a g_b x g_c
                double w[3], x1[3], x2[3];
                for (int q = 0; q < 3; ++q) w[q] = L[jv] * G[k][q] - L[k] * G[jv][q];
                cross3(G[i], w, x1);
                cross3(G[jv], G[k], x2);
                for (int q = 0; q < 3; ++q) f128[q] = s * (x1[q] + 2.0 * L[i] * x2[q]);
                for (int q = 0; q < 3; ++q) w[q] = L[i] * G[k][q] - L[k] * G[i][q];
                cross3(G[jv], w, x1);
                cross3(G[i], G[k], x2);
                for (int q = 0; q < 3; ++q) f129[q] = s * (x1[q] + 2.0 * L[jv] * x2[q]);
            }
            add(12 + fc, f128);
            add(16 + fc, f129);
        }
        double2 cc = make_double2(1.0, 0.0);
        if (CURL && ctet) cc = ctet[t];
        for (int q = 0; q < 3; ++q)
            out[q * n + j] = make_double2(cc.x * er[q] - cc.y * ei[q], cc.x * ei[q] + cc.y * er[q]);
    }
}

}  // namespace

hipError_t locate_points(const double4* xyz, const int4* tets, long long n_tets, const double* c,
                         long long n, const long long* tetids, long long n_cand,
                         long long* out, long long& bad, hipStream_t s) {
    bad = 0;
    if (n <= 0) return hipSuccess;
    hipError_t e;
    unsigned long long *cnt = nullptr, *dbad = nullptr;
    long long *off = nullptr, *contents = nullptr;
    double hc[6];
    Grid g;
    // bounding box of the points on the host side of the copy (one pass, small)
    double* box = nullptr;  // kBoxBlocks x 6 partials, then the 6 extremes
#define TRY(x)                          \
    do {                                \
        e = (x);                        \
        if (e != hipSuccess) goto done; \
    } while (0)
    {
        // bounding box on the device; only the 6 extremes come back to the host
        const long long nbl = (n + 255) / 256;
        const int nb = (int)(nbl < kBoxBlocks ? nbl : kBoxBlocks);
        TRY(hipMalloc((void**)&box, sizeof(double) * (6 * (size_t)nb + 6)));
        hipLaunchKernelGGL(bbox_partial_kernel, dim3(nb), dim3(256), 0, s, c, n, box);
        TRY(hipGetLastError());
        hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(64), 0, s, box, nb, box + 6 * nb);
        TRY(hipGetLastError());
        TRY(hipMemcpyAsync(hc, box + 6 * nb, sizeof(double) * 6, hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        for (int q = 0; q < 3; ++q)
            if (!(hc[2 * q] <= hc[2 * q + 1])) {  // no finite coordinate on this axis
                hc[2 * q] = 0.0;
                hc[2 * q + 1] = 0.0;
            }
        int res = (int)cbrt((double)n / 8.0);
        res = res < 20 ? 20 : (res > 200 ? 200 : res);
        const double pad = 1e-3;
        g.res = res;
        g.x0 = hc[0] - pad;
        g.y0 = hc[2] - pad;
        g.z0 = hc[4] - pad;
        g.dx = (hc[1] + pad - g.x0) / res;
        g.dy = (hc[3] + pad - g.y0) / res;
        g.dz = (hc[5] + pad - g.z0) / res;
        const long long ncell = (long long)res * res * res;
        TRY(hipMalloc((void**)&cnt, sizeof(unsigned long long) * (ncell + 1)));
        TRY(hipMalloc((void**)&off, sizeof(long long) * (ncell + 1)));
        TRY(hipMalloc((void**)&contents, sizeof(long long) * n));
        TRY(hipMalloc((void**)&dbad, sizeof(unsigned long long)));
        TRY(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * (ncell + 1), s));
        TRY(hipMemsetAsync(dbad, 0, sizeof(unsigned long long), s));
        hipLaunchKernelGGL(cell_count_kernel, dim3(grid_for(n)), dim3(256), 0, s, c, n, g, cnt);
        TRY(hipGetLastError());
        TRY(exclusive_scan_i64((const long long*)cnt, off, ncell + 1, s));
        TRY(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * (ncell + 1), s));
        hipLaunchKernelGGL(cell_fill_kernel, dim3(grid_for(n)), dim3(256), 0, s, c, n, g, off, cnt,
                           contents);
        TRY(hipGetLastError());
        hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_for(n)), dim3(256), 0, s, out, n, -1ll);
        TRY(hipGetLastError());
        if (n_cand > 0) {
            hipLaunchKernelGGL(locate_kernel, dim3(grid_for(n_cand)), dim3(256), 0, s, xyz, tets,
                               n_tets, tetids, n_cand, c, n, g, off, contents, dbad, out);
            TRY(hipGetLastError());
        }
        unsigned long long hb = 0;
        TRY(hipMemcpyAsync(&hb, dbad, sizeof(hb), hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        bad = (long long)hb;
    }
done:
#undef TRY
    (void)hipFree(box);
    (void)hipFree(cnt);
    (void)hipFree(off);
    (void)hipFree(contents);
    (void)hipFree(dbad);
    return e;
}

hipError_t ned2_interp(const double4* xyz, const int4* tets, const int* t2f, long long n_tets,
                       const double* c, long long n, const long long* tet_of, const double2* sol,
                       int curl, const double2* ctet, double2* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (curl)
        hipLaunchKernelGGL(ned2_interp_kernel<true>, dim3(grid_for(n)), dim3(256), 0, s, xyz, tets,
                           t2f, n_tets, c, n, tet_of, sol, ctet, out);
    else
        hipLaunchKernelGGL(ned2_interp_kernel<false>, dim3(grid_for(n)), dim3(256), 0, s, xyz,
                           tets, t2f, n_tets, c, n, tet_of, sol, ctet, out);
    return hipGetLastError();
}

}  // namespace em
