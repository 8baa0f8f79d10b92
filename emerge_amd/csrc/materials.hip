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
// Three launches, nothing to clear beforehand (every scratch word a launch reads was
// written by an earlier launch of the same call):
//   material_prep_kernel  one element per thread (48 B read, 16 B written); the (0,0)
//                         components are the first T elements, so each wavefront of them
//                         leaves a 64-bit ballot of its conductor tets (masks) and each
//                         workgroup of them its conductor count; every workgroup leaves
//                         one frequency-dependence flag;
//   conductor_scan_kernel exclusive scan of the workgroup counts in chunks of 1,024, with
//                         each chunk's total and the OR of the flags it covers;
//   conductor_ids_kernel  one thread per tet: a set mask bit writes the tet id at its
//                         workgroup's offset plus its rank among the set bits before it;
//                         the last workgroup writes the count and the flag.
// HBM-bound: 576 B per tet (three (3,3) complex128 tensors read, one written); the scratch
// is about T/64 * 8 + 9T/256 + T/32 bytes.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

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

// Scratch layout (each part on a 256-byte boundary): masks (uint64 per 64 tets), flags
// (uint8 per workgroup of material_prep_kernel), counts and offsets (uint32 per workgroup
// covering tets), chunk totals (int64) and chunk flags (uint32) per 1,024 of those
// workgroups, results (int64 count, int64 frequency dependence).
struct PrepScratch {
    unsigned long long* masks;
    unsigned char* flags;
    unsigned* counts;
    unsigned* offsets;
    long long* chunk_total;
    unsigned* chunk_fd;
    long long* results;
};

struct PrepSizes {
    long long nw, nb, nbt, nchunk;
    explicit PrepSizes(long long n_tets)
        : nw((n_tets + 63) / 64), nb((9 * n_tets + 255) / 256), nbt((n_tets + 255) / 256),
          nchunk(((n_tets + 255) / 256 + 1023) / 1024) {}
};

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Byte offsets of the parts; the last entry is the total size.
void layout(long long n_tets, size_t off[8]) {
    const PrepSizes z(n_tets);
    const size_t sz[7] = {8 * (size_t)z.nw,     (size_t)z.nb,          4 * (size_t)z.nbt,
                          4 * (size_t)z.nbt,    8 * (size_t)z.nchunk,  4 * (size_t)z.nchunk,
                          2 * sizeof(long long)};
    off[0] = 0;
    for (int i = 0; i < 7; ++i) off[i + 1] = off[i] + align256(sz[i]);
}

PrepScratch carve(void* base, long long n_tets) {
    size_t off[8];
    layout(n_tets, off);
    char* p = static_cast<char*>(base);
    PrepScratch s;
    s.masks = reinterpret_cast<unsigned long long*>(p + off[0]);
    s.flags = reinterpret_cast<unsigned char*>(p + off[1]);
    s.counts = reinterpret_cast<unsigned*>(p + off[2]);
    s.offsets = reinterpret_cast<unsigned*>(p + off[3]);
    s.chunk_total = reinterpret_cast<long long*>(p + off[4]);
    s.chunk_fd = reinterpret_cast<unsigned*>(p + off[5]);
    s.results = reinterpret_cast<long long*>(p + off[6]);
    return s;
}

// out may be er (in place): every thread reads its element before writing it.
__global__ __launch_bounds__(256) void material_prep_kernel(
    const double2* er, const double2* __restrict__ tand, const double2* __restrict__ cond,
    long long n_tets, double inv_d, double pec_limit, double surf_limit, double2* out,
    unsigned long long* __restrict__ masks, unsigned char* __restrict__ flags,
    unsigned* __restrict__ counts) {
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
    const int n_cond = __syncthreads_count(conductor);
    const int any_fd = __syncthreads_or(fd);
    if (threadIdx.x == 0) {
        flags[blockIdx.x] = any_fd ? 1 : 0;
        if ((long long)blockIdx.x * 256 < n_tets) counts[blockIdx.x] = (unsigned)n_cond;
    }
}

// Exclusive scan of counts[nbt] in chunks of kIdThreads workgroup counts (one chunk per
// workgroup of this kernel): offsets within the chunk, the chunk's total, and the OR of
// the frequency-dependence flags[nb] that fall to this chunk (flags strided over chunks).
__global__ __launch_bounds__(kIdThreads) void conductor_scan_kernel(
    const unsigned* __restrict__ counts, long long nbt, const unsigned char* __restrict__ flags,
    long long nb, unsigned* __restrict__ offsets, long long* __restrict__ chunk_total,
    unsigned* __restrict__ chunk_fd) {
    __shared__ unsigned wsum[kIdThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long i = (long long)blockIdx.x * kIdThreads + tid;
    const unsigned v = i < nbt ? counts[i] : 0u;
    unsigned x = v;  // inclusive scan inside the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    int fd = 0;
    for (long long b = i; b < nb; b += (long long)gridDim.x * kIdThreads) fd |= flags[b];
    __syncthreads();
    if (wave == 0) {  // scan of the wavefront totals
        unsigned w = lane < kIdThreads / 64 ? wsum[lane] : 0u;
#pragma unroll
        for (int d = 1; d < kIdThreads / 64; d <<= 1) {
            const unsigned y = __shfl_up(w, d, 64);
            if (lane >= d) w += y;
        }
        if (lane < kIdThreads / 64) wsum[lane] = w;
    }
    __syncthreads();
    const unsigned before = wave > 0 ? wsum[wave - 1] : 0u;
    if (i < nbt) offsets[i] = before + x - v;
    const int any_fd = __syncthreads_or(fd);
    if (tid == 0) {
        chunk_total[blockIdx.x] = wsum[kIdThreads / 64 - 1];
        chunk_fd[blockIdx.x] = any_fd ? 1u : 0u;
    }
}

// One thread per tet, one workgroup per kernel-A workgroup of tets: the tet's id goes to
// (totals of the earlier chunks) + offsets[b] + (set bits of the earlier wavefronts of the
// workgroup) + (set bits of the earlier lanes). The last workgroup writes the results.
__global__ __launch_bounds__(256) void conductor_ids_kernel(
    const unsigned long long* __restrict__ masks, long long n_tets,
    const unsigned* __restrict__ offsets, const long long* __restrict__ chunk_total,
    const unsigned* __restrict__ chunk_fd, long long nchunk, long long* __restrict__ ids,
    long long* __restrict__ results) {
    const long long b = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long chunk = b / kIdThreads;
    long long base = offsets[b];
    for (long long c = 0; c < chunk; ++c) base += chunk_total[c];  // at most 63 chunks
    const long long w0 = 4 * b;
    const long long nw = (n_tets + 63) / 64;
    long long before = 0;
    for (int w = 0; w < wave; ++w)
        if (w0 + w < nw) before += __popcll(masks[w0 + w]);
    const long long t = 256 * b + threadIdx.x;
    if (t < n_tets) {
        const unsigned long long m = masks[w0 + wave];
        if ((m >> lane) & 1ull) {
            const unsigned long long lower = lane ? (m & (~0ull >> (64 - lane))) : 0ull;
            ids[base + before + __popcll(lower)] = t;
        }
    }
    if (b == gridDim.x - 1 && threadIdx.x == 0) {
        long long total = 0;
        unsigned fd = 0;
        for (long long c = 0; c < nchunk; ++c) {
            total += chunk_total[c];
            fd |= chunk_fd[c];
        }
        results[0] = total;
        results[1] = fd ? 1 : 0;
    }
}

}  // namespace

size_t material_prep_scratch_bytes(long long n_tets) {
    size_t off[8];
    layout(n_tets, off);
    return off[7];
}

hipError_t material_prep(const double2* er, const double2* tand, const double2* cond,
                         long long n_tets, double omega, double pec_limit, double surf_limit,
                         double2* out, long long* conductor_tets, void* scratch,
                         long long& n_conductor, bool& frequency_dependent, hipStream_t s) {
    n_conductor = 0;
    frequency_dependent = false;
    if (n_tets == 0) return hipSuccess;
    const PrepScratch sc = carve(scratch, n_tets);
    const PrepSizes z(n_tets);
    const double inv_d = 1.0 / (omega * kEps0);
    hipLaunchKernelGGL(material_prep_kernel, dim3((unsigned)z.nb), dim3(256), 0, s, er, tand,
                       cond, n_tets, inv_d, pec_limit, surf_limit, out, sc.masks, sc.flags,
                       sc.counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(conductor_scan_kernel, dim3((unsigned)z.nchunk), dim3(kIdThreads), 0, s,
                       sc.counts, z.nbt, sc.flags, z.nb, sc.offsets, sc.chunk_total,
                       sc.chunk_fd);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(conductor_ids_kernel, dim3((unsigned)z.nbt), dim3(256), 0, s, sc.masks,
                       n_tets, sc.offsets, sc.chunk_total, sc.chunk_fd, z.nchunk,
                       conductor_tets, sc.results);
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

// ---------------------------------------------------------------------------------------
// em_material_prep_noalloc: em_material_prep for callers that cannot allocate device
// memory. The scratch is `scratch`, a file-scope device array of kNoAllocScratchBytes
// (4 MiB: material_prep_scratch_bytes(2^24) = 3,212,288 bytes), one copy per GPU in each
// process, held until the process ends. On the host the array's name stands for no device
// address: its address on the current device is obtained once per device with
// hipGetSymbolAddress and only that pointer is handed to the kernels. Every byte a call
// reads from it, the call writes first, so nothing is cleared between calls. Calls from
// the threads of one process take a per-device lock; calls that overlap on one GPU
// without it (the array is shared) would race.
// ---------------------------------------------------------------------------------------

namespace {

constexpr long long kNoAllocMaxTets = 1ll << 24;
constexpr size_t kNoAllocScratchBytes = (size_t)4 << 20;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct NoAllocSlot {
    std::mutex call;       // one call at a time on this device
    void* ptr = nullptr;   // address of `scratch` on this device
    bool resolved = false;
};
NoAllocSlot g_noalloc[kNoAllocMaxDevices];
std::mutex g_noalloc_resolve;

// The device address of `scratch` on device `dev` (the current device), resolved on the
// first call on that device.
hipError_t noalloc_scratch(int dev, void** ptr) {
    std::lock_guard<std::mutex> lk(g_noalloc_resolve);
    NoAllocSlot& s = g_noalloc[dev];
    if (!s.resolved) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        s.ptr = p;
        s.resolved = true;
    }
    *ptr = s.ptr;
    return hipSuccess;
}

}  // namespace

extern "C" int em_material_prep_noalloc(const void* er, const void* tand, const void* cond,
                                        int64_t n_tets, double omega, double pec_limit,
                                        double surf_limit, void* er_out, int64_t* conductor_tets,
                                        int64_t* n_conductor, int32_t* frequency_dependent,
                                        void* stream) {
    if (int st = em::material_prep_check_args(er, tand, cond, n_tets, omega, er_out,
                                              conductor_tets, n_conductor, frequency_dependent))
        return st;
    if (n_tets > kNoAllocMaxTets) {
        em::set_error("material_prep_noalloc: " + std::to_string(n_tets) +
                      " tets exceed the fixed scratch (at most 16,777,216 tets); use "
                      "em_material_prep");
        return EM_EINVAL;
    }
    *n_conductor = 0;
    *frequency_dependent = 0;
    if (n_tets == 0) return EM_OK;
    if (em::material_prep_scratch_bytes(n_tets) > kNoAllocScratchBytes) {
        em::set_error("material_prep_noalloc: scratch too small");
        return EM_EINVAL;
    }
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        em::set_error(std::string("material_prep_noalloc: hipGetDevice: ") + hipGetErrorString(e));
        return EM_EHIP;
    }
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        em::set_error("material_prep_noalloc: device index beyond 63");
        return EM_EUNSUPPORTED;
    }
    void* ptr = nullptr;
    if ((e = noalloc_scratch(dev, &ptr)) != hipSuccess) {
        em::set_error(std::string("material_prep_noalloc: hipGetSymbolAddress: ") +
                      hipGetErrorString(e));
        return EM_EHIP;
    }
    std::lock_guard<std::mutex> lk(g_noalloc[dev].call);
    long long nc = 0;
    bool fd = false;
    e = em::material_prep((const double2*)er, (const double2*)tand, (const double2*)cond, n_tets,
                          omega, pec_limit, surf_limit, (double2*)er_out,
                          (long long*)conductor_tets, ptr, nc, fd, (hipStream_t)stream);
    if (e != hipSuccess) {
        em::set_error(std::string("material_prep_noalloc: ") + hipGetErrorString(e));
        return EM_EHIP;
    }
    *n_conductor = nc;
    *frequency_dependent = fd ? 1 : 0;
    return EM_OK;
}
