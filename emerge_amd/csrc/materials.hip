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
// (3,3,T) C order, entry (p,q) of tet t at (3p+q)T + t. One element per thread: 48 B read,
// 16 B written.
#include <hip/hip_runtime.h>

#include "em_internal.h"

namespace em {

namespace {

constexpr double kEps0 = 8.854187818814e-12;  // const.py:23

unsigned grid_for(long long n) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > 65536 * 4) b = 65536 * 4;
    return (unsigned)b;
}

__device__ __forceinline__ bool cgt(double2 a, double b) {  // a > b + 0j
    return a.x > b || (a.x == b && a.y > 0.0);
}
__device__ __forceinline__ bool clt(double2 a, double b) {  // a < b + 0j
    return a.x < b || (a.x == b && a.y < 0.0);
}

// out may be er (in place): every thread reads its element before writing it.
__global__ __launch_bounds__(256) void material_prep_kernel(
    const double2* er, const double2* __restrict__ tand, const double2* __restrict__ cond,
    long long n_tets, double inv_d, double pec_limit, double surf_limit, double2* out,
    long long* __restrict__ conductor, unsigned int* __restrict__ freq_dep) {
#pragma clang fp contract(off)
    const long long n = 9 * n_tets;
    bool fd = false;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n;
         k += (long long)gridDim.x * blockDim.x) {
        const double2 e = er[k], t = tand[k], c = cond[k];
        // w = 1 - 1j*tand; 1j*tand = (-Im tand, Re tand)
        const double wr = 1.0 - (-t.y), wi = 0.0 - t.x;
        const double pr = fma(e.x, wr, -(e.y * wi));
        const double pi = fma(e.x, wi, e.y * wr);
        // 1j*cond / d = (-Im cond, Re cond) * (1/d)
        const double qr = (-c.y) * inv_d, qi = c.x * inv_d;
        out[k] = make_double2(pr - qr, pi - qi);
        if (k < n_tets) conductor[k] = (cgt(c, pec_limit) || cgt(c, surf_limit)) ? 1 : 0;
        if (k == 0) conductor[n_tets] = 0;  // the scan's total lands here
        fd = fd || (cgt(c, 0.0) && clt(c, pec_limit));
    }
    if (__ballot(fd) && (threadIdx.x & 63) == 0) atomicOr(freq_dep, 1u);
}

// scan = exclusive scan of the conductor flags, [T + 1]: tet t is a conductor when its
// entry steps up, and goes to position scan[t] (ascending ids).
__global__ __launch_bounds__(256) void conductor_ids_kernel(const long long* __restrict__ scan,
                                                            long long n_tets,
                                                            long long* __restrict__ ids) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n_tets;
         t += (long long)gridDim.x * blockDim.x) {
        const long long p = scan[t];
        if (scan[t + 1] != p) ids[p] = t;
    }
}

}  // namespace

hipError_t material_prep(const double2* er, const double2* tand, const double2* cond,
                         long long n_tets, double omega, double pec_limit, double surf_limit,
                         double2* out, long long* conductor_tets, long long& n_conductor,
                         bool& frequency_dependent, hipStream_t s) {
    long long* scan = nullptr;  // conductor flags, then their exclusive scan
    unsigned int* dfd = nullptr;
    unsigned int hfd = 0;
    long long nc = 0;
    hipError_t e;
    n_conductor = 0;
    frequency_dependent = false;
    if (n_tets == 0) return hipSuccess;
    if ((e = hipMalloc((void**)&scan, sizeof(long long) * (n_tets + 1))) != hipSuccess) goto done;
    if ((e = hipMalloc((void**)&dfd, sizeof(unsigned int))) != hipSuccess) goto done;
    if ((e = hipMemsetAsync(dfd, 0, sizeof(unsigned int), s)) != hipSuccess) goto done;
    {
        const double inv_d = 1.0 / (omega * kEps0);
        hipLaunchKernelGGL(material_prep_kernel, dim3(grid_for(9 * n_tets)), dim3(256), 0, s, er,
                           tand, cond, n_tets, inv_d, pec_limit, surf_limit, out, scan, dfd);
        if ((e = hipGetLastError()) != hipSuccess) goto done;
    }
    if ((e = exclusive_scan_i64(scan, scan, n_tets + 1, s)) != hipSuccess) goto done;
    hipLaunchKernelGGL(conductor_ids_kernel, dim3(grid_for(n_tets)), dim3(256), 0, s, scan, n_tets,
                       conductor_tets);
    if ((e = hipGetLastError()) != hipSuccess) goto done;
    if ((e = hipMemcpyAsync(&nc, scan + n_tets, sizeof(nc), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
        goto done;
    if ((e = hipMemcpyAsync(&hfd, dfd, sizeof(hfd), hipMemcpyDeviceToHost, s)) != hipSuccess)
        goto done;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) goto done;
    n_conductor = nc;
    frequency_dependent = hfd != 0;
done:
    (void)hipStreamSynchronize(s);
    (void)hipFree(scan);
    (void)hipFree(dfd);
    return e;
}

}  // namespace em
