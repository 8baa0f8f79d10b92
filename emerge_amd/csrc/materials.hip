// Material tensors of one frequency, as Assembler.assemble_freq_matrix prepares them before
// the volume assembly (src/emerge/_emerge/physics/microwave/assembly/assembler.py:439-469;
// the same lines in assemble_scattering_matrix, :819-868, and assemble_eig_matrix,
// :1119-1142):
//   er' = er * (1 - 1j*tand) - 1j*cond / (W0*EPS0)      every one of the 9 components
//   conductor tets: cond[0,0,t] > mw_3d_peclim or cond[0,0,t] > mw_3d_surfimplim
//   frequency dependent: some component has 0 < cond < mw_3d_peclim
// NumPy orders complex numbers lexicographically (real part, then imaginary part); the
// comparisons here do the same. The arithmetic is NumPy's complex loops on x86-64: the
// product (a+bi)(c+di) = fma(a, c, -(b d)) + fma(a, d, b c) i, and a complex divided by a
// real d is multiplied by the reciprocal 1/d. Tensors are complex128 in the reference's
// (3,3,T) C order, entry (p,q) of tet t at (3p+q)T + t.
//
// Two launches, no device-wide scan:
//   material_prep_kernel  one element per thread (48 B read, 16 B written); the (0,0)
//                         components are the first T elements, so each wavefront of them
//                         leaves a 64-bit ballot of its conductor tets (masks), and each
//                         workgroup one frequency-dependence flag (plain stores: nothing
//                         has to be cleared before the launch);
//   conductor_ids_kernel  one workgroup: popcounts of the masks, a scan over its threads'
//                         contiguous mask segments, then the ids of the set bits in
//                         ascending order, the count and the OR of the flags.
// HBM-bound: 576 B per tet (three (3,3) complex128 tensors read, one written); the scratch
// is T/64 * 8 + 9T/256 bytes.
#include <hip/hip_runtime.h>

#include <mutex>

#include "em_internal.h"

namespace em {

namespace {

constexpr double kEps0 = 8.854187818814e-12;  // const.py:23
constexpr int kIdThreads = 1024;

__device__ __forceinline__ bool cgt(double2 a, double b) {  // a > b + 0j
    return a.x > b || (a.x == b && a.y > 0.0);
}
__device__ __forceinline__ bool clt(double2 a, double b) {  // a < b + 0j
    return a.x < b || (a.x == b && a.y < 0.0);
}

// Scratch layout: masks (uint64, one per 64 tets), flags (uint8, one per workgroup of
// material_prep_kernel), results (int64 count, int64 frequency dependence); each part
// starts on a 256-byte boundary.
struct PrepScratch {
    unsigned long long* masks;
    unsigned char* flags;
    long long* results;
};

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

PrepScratch carve(void* base, long long n_tets) {
    const long long nw = (n_tets + 63) / 64, nb = (9 * n_tets + 255) / 256;
    char* p = static_cast<char*>(base);
    PrepScratch s;
    s.masks = reinterpret_cast<unsigned long long*>(p);
    p += align256(sizeof(unsigned long long) * (size_t)nw);
    s.flags = reinterpret_cast<unsigned char*>(p);
    p += align256((size_t)nb);
    s.results = reinterpret_cast<long long*>(p);
    return s;
}

// out may be er (in place): every thread reads its element before writing it.
__global__ __launch_bounds__(256) void material_prep_kernel(
    const double2* er, const double2* __restrict__ tand, const double2* __restrict__ cond,
    long long n_tets, double inv_d, double pec_limit, double surf_limit, double2* out,
    unsigned long long* __restrict__ masks, unsigned char* __restrict__ flags) {
#pragma clang fp contract(off)
    const long long n = 9 * n_tets;
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    bool fd = false, conductor = false;
    if (k < n) {
        const double2 e = er[k], t = tand[k], c = cond[k];
        // w = 1 - 1j*tand; 1j*tand = (-Im tand, Re tand)
        const double wr = 1.0 - (-t.y), wi = 0.0 - t.x;
        const double pr = fma(e.x, wr, -(e.y * wi));
        const double pi = fma(e.x, wi, e.y * wr);
        // 1j*cond / d = (-Im cond, Re cond) * (1/d)
        const double qr = (-c.y) * inv_d, qi = c.x * inv_d;
        out[k] = make_double2(pr - qr, pi - qi);
        fd = cgt(c, 0.0) && clt(c, pec_limit);
        if (k < n_tets) conductor = cgt(c, pec_limit) || cgt(c, surf_limit);
    }
    // the workgroup starts at a multiple of 256, so lane 0 of a wavefront holds tet 64 w
    const unsigned long long m = __ballot(conductor);
    if ((threadIdx.x & 63) == 0 && k < n_tets) masks[k >> 6] = m;
    const int any_fd = __syncthreads_or(fd);
    if (threadIdx.x == 0) flags[blockIdx.x] = any_fd ? 1 : 0;
}

__global__ __launch_bounds__(kIdThreads) void conductor_ids_kernel(
    const unsigned long long* __restrict__ masks, long long nw,
    const unsigned char* __restrict__ flags, long long nb, long long* __restrict__ ids,
    long long* __restrict__ results) {
    __shared__ long long part[kIdThreads];
    const int tid = threadIdx.x;
    int fd = 0;
    for (long long b = tid; b < nb; b += kIdThreads) fd |= flags[b];
    const long long per = (nw + kIdThreads - 1) / kIdThreads;
    const long long w0 = min(nw, (long long)tid * per), w1 = min(nw, w0 + per);
    long long cnt = 0;
    for (long long w = w0; w < w1; ++w) cnt += __popcll(masks[w]);
    part[tid] = cnt;
    __syncthreads();
    for (int off = 1; off < kIdThreads; off <<= 1) {  // inclusive scan of the counts
        const long long v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long pos = part[tid] - cnt;
    const long long total = part[kIdThreads - 1];
    for (long long w = w0; w < w1; ++w) {
        unsigned long long m = masks[w];
        while (m) {
            ids[pos++] = 64 * w + __builtin_ctzll(m);
            m &= m - 1;
        }
    }
    const int any_fd = __syncthreads_or(fd);
    if (tid == 0) {
        results[0] = total;
        results[1] = any_fd ? 1 : 0;
    }
}

}  // namespace

size_t material_prep_scratch_bytes(long long n_tets) {
    const long long nw = (n_tets + 63) / 64, nb = (9 * n_tets + 255) / 256;
    return align256(sizeof(unsigned long long) * (size_t)nw) + align256((size_t)nb) +
           2 * sizeof(long long);
}

hipError_t material_prep(const double2* er, const double2* tand, const double2* cond,
                         long long n_tets, double omega, double pec_limit, double surf_limit,
                         double2* out, long long* conductor_tets, void* scratch,
                         long long& n_conductor, bool& frequency_dependent, hipStream_t s) {
    n_conductor = 0;
    frequency_dependent = false;
    if (n_tets == 0) return hipSuccess;
    const PrepScratch sc = carve(scratch, n_tets);
    const long long nw = (n_tets + 63) / 64, nb = (9 * n_tets + 255) / 256;
    const double inv_d = 1.0 / (omega * kEps0);
    hipLaunchKernelGGL(material_prep_kernel, dim3((unsigned)nb), dim3(256), 0, s, er, tand, cond,
                       n_tets, inv_d, pec_limit, surf_limit, out, sc.masks, sc.flags);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(conductor_ids_kernel, dim3(1), dim3(kIdThreads), 0, s, sc.masks, nw,
                       sc.flags, nb, conductor_tets, sc.results);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    long long res[2] = {0, 0};
    if ((e = hipMemcpyAsync(res, sc.results, sizeof(res), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
        return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    n_conductor = res[0];
    frequency_dependent = res[1] != 0;
    return hipSuccess;
}

}  // namespace em
