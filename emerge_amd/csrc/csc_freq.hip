// Per-frequency CSC values A_f = K - k0_f^2 M from the CSC values of K and M, which share
// one pattern (the map of the plan).
//
// The reference combines the cached element matrices before its COO -> CSC step,
// to_csc(Evec - Bvec * K0**2) (physics/microwave/assembly/assembler.py:471-494), so every
// frequency of a sweep (microwave_3d.py:1464-1470) pays a full gather over the COO values.
// K and M do not change with the frequency: summed once per mesh by the column gather,
// each frequency is then one streaming pass over the CSC values. The matrix is the same;
// each entry is summed in another order (sum(E) - k0^2 sum(B) instead of sum(E - k0^2 B)),
// which differs by rounding only.
//
// HBM-bound, no reuse: per launch K and M are read once (2 c bytes per entry) and n_k0
// value arrays written (n_k0 c bytes per entry), c = 16 (complex128) or 8 (complex64). One
// thread per 16-byte vector: one complex128 entry, or two complex64 entries. The sums are
// FP64 for both (complex64 inputs are widened, the result rounded once). Loads and stores
// are non-temporal (nothing is read twice by a launch): on the 2.8M-tet mesh's 780.6M
// complex64 entries that took one pass from 3.09 ms to 2.94 ms (0.76 -> 0.80 of 8 TB/s);
// two or four vectors per thread (3.2-3.4 ms) and a persistent grid (4.1 ms) were slower.
//
// A lossy medium adds a third array: er' = er (1 - j tand) - j cond / (w eps0) is affine in
// 1/w, so with C = the CSC mass values of cond, A_f = K - k0_f^2 M_a + j g_f C with M_a
// the mass values of er (1 - j tand) and g_f = k0_f^2 / (w_f eps0): one more array read
// per launch.
#include <hip/hip_runtime.h>

#include "em_internal.h"

namespace em {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// With a third array C (the CSC values of the mass matrix of a conductivity): A_f = K
// - k0_f^2 M + j g_f C, the affine frequency dependence of a lossy medium (see
// csc_freq_combine below). HAS_C = false reads neither C nor g.
template <bool HAS_C>
__global__ __launch_bounds__(256) void csc_freq_kernel_c128(const double2* __restrict__ K,
                                                            const double2* __restrict__ M,
                                                            const double2* __restrict__ C,
                                                            long long n, K0Pack k0p, K0Pack gp,
                                                            int nf, double2* __restrict__ out,
                                                            long long stride) {
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const f64x2 kv = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(K) + i);
    const f64x2 mv = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(M) + i);
    f64x2 cv = {0.0, 0.0};
    if (HAS_C) cv = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(C) + i);
    const double2 k = make_double2(kv.x, kv.y), m = make_double2(mv.x, mv.y);
    for (int f = 0; f < nf; ++f) {
        const double k2 = k0p.v[f];
        f64x2 a = {fma(-k2, m.x, k.x), fma(-k2, m.y, k.y)};
        if (HAS_C) {
            const double g = gp.v[f];
            a.x = fma(-g, cv.y, a.x);
            a.y = fma(g, cv.x, a.y);
        }
        __builtin_nontemporal_store(a, reinterpret_cast<f64x2*>(out + (long long)f * stride + i));
    }
}

// complex64: thread i owns entries 2i and 2i + 1 (one 16-byte load of K, of M and store
// per output); an odd n leaves the last entry to the thread of pair n / 2 alone.
template <bool HAS_C>
__global__ __launch_bounds__(256) void csc_freq_kernel_c64(const float2* __restrict__ K,
                                                           const float2* __restrict__ M,
                                                           const float2* __restrict__ C,
                                                           long long n, K0Pack k0p, K0Pack gp,
                                                           int nf, float2* __restrict__ out,
                                                           long long stride) {
    const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long e = 2 * p;
    if (e >= n) return;
    if (e + 1 < n) {
        const f32x4 k = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(K + e));
        const f32x4 m = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(M + e));
        f32x4 c = {0.f, 0.f, 0.f, 0.f};
        if (HAS_C) c = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(C + e));
        for (int f = 0; f < nf; ++f) {
            const double k2 = k0p.v[f];
            double a0 = fma(-k2, (double)m.x, (double)k.x), a1 = fma(-k2, (double)m.y, (double)k.y);
            double a2 = fma(-k2, (double)m.z, (double)k.z), a3 = fma(-k2, (double)m.w, (double)k.w);
            if (HAS_C) {
                const double g = gp.v[f];
                a0 = fma(-g, (double)c.y, a0);
                a1 = fma(g, (double)c.x, a1);
                a2 = fma(-g, (double)c.w, a2);
                a3 = fma(g, (double)c.z, a3);
            }
            const f32x4 a = {(float)a0, (float)a1, (float)a2, (float)a3};
            __builtin_nontemporal_store(a, reinterpret_cast<f32x4*>(out + (long long)f * stride + e));
        }
    } else {
        const float2 k = K[e], m = M[e];
        const float2 c = HAS_C ? C[e] : make_float2(0.f, 0.f);
        for (int f = 0; f < nf; ++f) {
            const double k2 = k0p.v[f];
            double a0 = fma(-k2, (double)m.x, (double)k.x), a1 = fma(-k2, (double)m.y, (double)k.y);
            if (HAS_C) {
                const double g = gp.v[f];
                a0 = fma(-g, (double)c.y, a0);
                a1 = fma(g, (double)c.x, a1);
            }
            out[(long long)f * stride + e] = make_float2((float)a0, (float)a1);
        }
    }
}

}  // namespace

hipError_t csc_freq_combine(long long n, const void* K, const void* M, const void* C,
                            const double* k0sq, const double* g, int nf, void* out,
                            long long stride, int complex64, hipStream_t s) {
    if (n == 0 || nf == 0) return hipSuccess;
    const long long threads = complex64 ? (n + 1) / 2 : n;
    const unsigned grid = (unsigned)((threads + 255) / 256);
    for (int f0 = 0; f0 < nf; f0 += kMaxK0) {
        K0Pack p, q;
        const int cnt = min(kMaxK0, nf - f0);
        for (int f = 0; f < kMaxK0; ++f) {
            p.v[f] = f < cnt ? k0sq[f0 + f] : 0.0;
            q.v[f] = (C && f < cnt) ? g[f0 + f] : 0.0;
        }
        if (complex64) {
            float2* o = (float2*)out + (long long)f0 * stride;
            if (C)
                hipLaunchKernelGGL(csc_freq_kernel_c64<true>, dim3(grid), dim3(256), 0, s,
                                   (const float2*)K, (const float2*)M, (const float2*)C, n, p, q,
                                   cnt, o, stride);
            else
                hipLaunchKernelGGL(csc_freq_kernel_c64<false>, dim3(grid), dim3(256), 0, s,
                                   (const float2*)K, (const float2*)M, nullptr, n, p, q, cnt, o,
                                   stride);
        } else {
            double2* o = (double2*)out + (long long)f0 * stride;
            if (C)
                hipLaunchKernelGGL(csc_freq_kernel_c128<true>, dim3(grid), dim3(256), 0, s,
                                   (const double2*)K, (const double2*)M, (const double2*)C, n, p,
                                   q, cnt, o, stride);
            else
                hipLaunchKernelGGL(csc_freq_kernel_c128<false>, dim3(grid), dim3(256), 0, s,
                                   (const double2*)K, (const double2*)M, nullptr, n, p, q, cnt, o,
                                   stride);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace em
