// CSC pattern of the assembled COO matrix, and the COO -> CSC value reduction.
//
// Reference: src/emerge/_emerge/mth/csc_cast.py:65-160 (precompute_csc_pattern) and
// :162-177 (scatter_to_csc_c16). The reference sorts all COO (row, col) pairs; here the
// pattern is built from the DOF map, since COO entry k = 400*ia + 20*ii + jj of
// assembly index ia has row t2f[ii], col t2f[jj]:
//   1. incidence lists: for every DOF d, the (ia, jj) with t2f[ia][jj] == d
//      (histogram + exclusive scan + fill, then each list sorted so the build is
//      deterministic);
//   2. column c's COO entries are the 20 rows of each incidence; one wavefront sorts
//      its (row, position) keys in LDS, flags the first entry of every distinct row and
//      scans the flags with ballots (columns of more than 2,048 keys: one workgroup ranks
//      them, csc_long_column_kernel). Pass 1 counts distinct rows (-> indptr by scan),
//      pass 2 writes indices, perm (COO index of every entry in (col, row, k) order,
//      the reference's fully_sorted_perm) and seg (first perm position of every CSC
//      entry; the reference's target_csc is its inverse).
// to_csc: each CSC entry sums its contributors data[perm[seg[q] .. seg[q+1])] in FP64
// (complex64 inputs are accumulated in FP64 and rounded once).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "em_internal.h"

namespace em {

constexpr int kScanThreads = 256;
constexpr int kScanTile = 1024;  // 4 items per thread

__global__ __launch_bounds__(kScanThreads) void scan_tiles_kernel(const long long* __restrict__ in,
                                                                  long long* __restrict__ out,
                                                                  long long n,
                                                                  long long* __restrict__ sums) {
    __shared__ long long s[kScanThreads];
    const long long base = (long long)blockIdx.x * kScanTile + threadIdx.x * 4;
    long long v[4], local = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        local += v[i];
    }
    s[threadIdx.x] = local;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        long long t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = s[threadIdx.x] - local;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == kScanThreads - 1) sums[blockIdx.x] = s[threadIdx.x];
}

__global__ __launch_bounds__(kScanThreads) void scan_add_kernel(long long* __restrict__ out,
                                                                long long n,
                                                                const long long* __restrict__ offs) {
    const long long base = (long long)blockIdx.x * kScanTile + threadIdx.x * 4;
    const long long o = offs[blockIdx.x];
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (base + i < n) out[base + i] += o;
}

// Exclusive scan of n int64 values (in may equal out). Temporaries are allocated here.
hipError_t exclusive_scan_i64(const long long* in, long long* out, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const long long ntiles = (n + kScanTile - 1) / kScanTile;
    long long *sums = nullptr, *offs = nullptr;
    hipError_t e = hipMalloc((void**)&sums, sizeof(long long) * ntiles);
    if (e != hipSuccess) return e;
    e = hipMalloc((void**)&offs, sizeof(long long) * ntiles);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in,
                           out, n, sums);
        e = hipGetLastError();
    }
    if (e == hipSuccess && ntiles > 1) {
        e = exclusive_scan_i64(sums, offs, ntiles, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(scan_add_kernel, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s,
                               out, n, offs);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // temporaries are freed below
    (void)hipFree(sums);
    (void)hipFree(offs);
    return e;
}

__device__ __forceinline__ long long tet_of(const int* tetids, long long ia) {
    return tetids ? (long long)tetids[ia] : ia;
}

template <int NDOF>
__global__ __launch_bounds__(256) void inc_count_kernel(const int* __restrict__ t2f,
                                                        const int* __restrict__ tetids,
                                                        long long n_assy,
                                                        unsigned long long* __restrict__ cnt) {
    const long long n = n_assy * NDOF;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long ia = e / NDOF;
        const int d = t2f[tet_of(tetids, ia) * NDOF + (int)(e - ia * NDOF)];
        atomicAdd(cnt + d, 1ull);
    }
}

template <int NDOF>
__global__ __launch_bounds__(256) void inc_fill_kernel(const int* __restrict__ t2f,
                                                       const int* __restrict__ tetids,
                                                       long long n_assy,
                                                       const long long* __restrict__ inc_ptr,
                                                       unsigned long long* __restrict__ cursor,
                                                       int* __restrict__ inc) {
    const long long n = n_assy * NDOF;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (long long)gridDim.x * blockDim.x) {
        const long long ia = e / NDOF;
        const int d = t2f[tet_of(tetids, ia) * NDOF + (int)(e - ia * NDOF)];
        const unsigned long long pos = atomicAdd(cursor + d, 1ull);
        inc[inc_ptr[d] + (long long)pos] = (int)e;
    }
}

// Sort every incidence list (short) ascending; also record the longest list.
__global__ __launch_bounds__(256) void inc_sort_kernel(const long long* __restrict__ inc_ptr,
                                                       int* __restrict__ inc, long long N,
                                                       unsigned long long* __restrict__ maxlen) {
    const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    const long long s = inc_ptr[d], n = inc_ptr[d + 1] - s;
    int* a = inc + s;
    for (long long i = 1; i < n; ++i) {
        const int v = a[i];
        long long j = i - 1;
        while (j >= 0 && a[j] > v) {
            a[j + 1] = a[j];
            --j;
        }
        a[j + 1] = v;
    }
    atomicMax(maxlen, (unsigned long long)n);
}

constexpr int kColCap = 2048;  // COO entries per column handled in LDS (102 incidences)

template <bool FILL, int NDOF>
__global__ __launch_bounds__(64) void csc_column_kernel(
    const int* __restrict__ t2f, const int* __restrict__ tetids,
    const long long* __restrict__ inc_ptr, const int* __restrict__ inc, long long N,
    long long* __restrict__ ucount, const long long* __restrict__ indptr,
    long long* __restrict__ indices, long long* __restrict__ perm, long long* __restrict__ seg,
    unsigned short* __restrict__ pos) {
    __shared__ unsigned long long srt[kColCap];
    const int lane = threadIdx.x;
    for (long long c = blockIdx.x; c < N; c += gridDim.x) {
        const long long s = inc_ptr[c];
        const int n = (int)(NDOF * (inc_ptr[c + 1] - s));
        if (n > kColCap) continue;  // csc_long_column_kernel builds it
        int npad = 1;
        while (npad < n) npad <<= 1;
        for (int i = lane; i < npad; i += 64) {
            unsigned long long k = ~0ull;  // padding sorts last
            if (i < n) {
                const int e = inc[s + i / NDOF];
                const long long ia = e / NDOF;
                const int row = t2f[tet_of(tetids, ia) * NDOF + (i % NDOF)];
                k = ((unsigned long long)(unsigned)row << 32) | (unsigned)i;
            }
            srt[i] = k;
        }
        __syncthreads();
        // bitonic sort of the (row, local index) keys, all distinct, in LDS
        for (int k = 2; k <= npad; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = lane; i < npad; i += 64) {
                    const int ixj = i ^ j;
                    if (ixj > i) {
                        const unsigned long long a = srt[i], b = srt[ixj];
                        if ((a > b) == ((i & k) == 0)) {
                            srt[i] = b;
                            srt[ixj] = a;
                        }
                    }
                }
                __syncthreads();
            }
        long long carry = 0;
        for (int base = 0; base < n; base += 64) {
            const int i = base + lane;
            const bool valid = i < n;
            const unsigned long long cur = valid ? srt[i] : 0ull;
            const bool flag = valid && (i == 0 || (cur >> 32) != (srt[i - 1] >> 32));
            const unsigned long long mask = __ballot(flag);
            if (FILL && valid) {
                const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
                const long long q = indptr[c] + carry + __popcll(mask & upto) - 1;
                const int li = (int)(cur & 0xffffffffu);
                const int e = inc[s + li / NDOF];
                const long long ia = e / NDOF;
                const int jj = e % NDOF, ii = li % NDOF;
                perm[NDOF * s + i] = (long long)NDOF * NDOF * ia + NDOF * ii + jj;
                if (pos) pos[NDOF * s + li] = (unsigned short)(q - indptr[c]);
                if (flag) {
                    indices[q] = (long long)(cur >> 32);
                    seg[q] = NDOF * s + i;
                }
            }
            carry += __popcll(mask);
        }
        if (!FILL && lane == 0) ucount[c] = carry;
        __syncthreads();  // LDS reuse by the next column
    }
}

// Columns of more than kColCap keys (a DOF in more than 102 Nedelec-2 or 204 Lagrange-2
// tets: an edge on the axis of a fine fan, a vertex at a sharp corner) are rare. One
// 256-thread workgroup takes each: the column's rows go to LDS, and every key is ranked
// against all others (O(n^2) LDS reads). The keys (row, local index) are distinct, so a
// key's rank is its position in the sorted order, and the outputs equal those of
// csc_column_kernel's sort. The first key of a row (smallest local index among equal
// rows) carries the row's entry: its unique position is the number of distinct rows
// below it. Columns of up to kLongLds keys keep their rows in LDS; longer ones (GLOBAL)
// in a scratch slice of g_stride keys per workgroup in global memory, up to
// kCscMaxColumnKeys (the uint16 row positions), beyond which the host refuses.
constexpr int kLongLds = 12288;
constexpr int kLongThreads = 256;
static_assert(kLongLds % 64 == 0, "first-key bitmask words");
static_assert(kCscMaxColumnKeys <= 65535, "row positions are uint16");

template <bool FILL, int NDOF, bool GLOBAL>
__global__ __launch_bounds__(kLongThreads) void csc_long_column_kernel(
    const int* __restrict__ t2f, const int* __restrict__ tetids,
    const long long* __restrict__ inc_ptr, const int* __restrict__ inc, long long N,
    long long* __restrict__ ucount, const long long* __restrict__ indptr,
    long long* __restrict__ indices, long long* __restrict__ perm, long long* __restrict__ seg,
    unsigned short* __restrict__ pos, int* __restrict__ g_rows,
    unsigned long long* __restrict__ g_first, int g_stride) {
    __shared__ int s_rows[GLOBAL ? 1 : kLongLds];
    __shared__ unsigned long long s_first[GLOBAL ? 1 : kLongLds / 64];
    int* rows = GLOBAL ? g_rows + (size_t)blockIdx.x * g_stride : s_rows;
    unsigned long long* first = GLOBAL ? g_first + (size_t)blockIdx.x * (g_stride / 64) : s_first;
    const long long lo = GLOBAL ? kLongLds : kColCap, hi = GLOBAL ? g_stride : kLongLds;
    const int tid = threadIdx.x;
    for (long long c = blockIdx.x; c < N; c += gridDim.x) {
        const long long s = inc_ptr[c];
        const long long nl = NDOF * (inc_ptr[c + 1] - s);
        if (nl <= lo || nl > hi) continue;  // same c for the whole workgroup
        const int n = (int)nl, nw = (n + 63) / 64;
        for (int i = tid; i < n; i += kLongThreads)
            rows[i] = t2f[tet_of(tetids, inc[s + i / NDOF] / NDOF) * NDOF + (i % NDOF)];
        __syncthreads();
        // i runs over whole 64-key words, so every lane of a wavefront takes the ballot
        for (int i = tid; i < 64 * nw; i += kLongThreads) {
            bool f = i < n;
            if (f) {
                const int r = rows[i];
                for (int j = 0; j < i; ++j)
                    if (rows[j] == r) {
                        f = false;
                        break;
                    }
            }
            const unsigned long long m = __ballot(f);
            if ((tid & 63) == 0) first[i >> 6] = m;
        }
        __syncthreads();
        if (!FILL) {
            if (tid == 0) {
                long long u = 0;
                for (int w = 0; w < nw; ++w) u += __popcll(first[w]);
                ucount[c] = u;
            }
        } else {
            const long long q0 = indptr[c];
            for (int i = tid; i < n; i += kLongThreads) {
                const int r = rows[i];
                int rank = 0, q = 0;
                for (int j = 0; j < n; ++j) {
                    const int rj = rows[j];
                    rank += (rj < r) || (rj == r && j < i);
                    q += (rj < r) && ((first[j >> 6] >> (j & 63)) & 1ull);
                }
                const int e = inc[s + i / NDOF];
                const long long ia = e / NDOF;
                const int jj = e % NDOF, ii = i % NDOF;
                perm[NDOF * s + rank] = (long long)NDOF * NDOF * ia + NDOF * ii + jj;
                if (pos) pos[NDOF * s + i] = (unsigned short)q;
                if ((first[i >> 6] >> (i & 63)) & 1ull) {
                    indices[q0 + q] = r;
                    seg[q0 + q] = NDOF * s + rank;
                }
            }
        }
        __syncthreads();  // row storage reuse by the next column
    }
}

template <typename In>
struct LoadC;
template <>
struct LoadC<double2> {
    __device__ static double2 get(const double2* p, long long k) { return p[k]; }
};
template <>
struct LoadC<float2> {
    __device__ static double2 get(const float2* p, long long k) {
        const float2 v = p[k];
        return make_double2((double)v.x, (double)v.y);
    }
};

template <typename T>
__device__ __forceinline__ void store_c(T* p, long long q, double2 v);
template <>
__device__ __forceinline__ void store_c<double2>(double2* p, long long q, double2 v) {
    p[q] = v;
}
template <>
__device__ __forceinline__ void store_c<float2>(float2* p, long long q, double2 v) {
    p[q] = make_float2((float)v.x, (float)v.y);
}

// One COO contribution E[k] - k0sq * B[k] in FP64 (one rounding order for every gather).
template <typename T>
__device__ __forceinline__ double2 contrib(const T* __restrict__ E, const T* __restrict__ B,
                                           double k0sq, long long k) {
    double2 v = LoadC<T>::get(E, k);
    if (B) {
        const double2 b = LoadC<T>::get(B, k);
        v.x = fma(-k0sq, b.x, v.x);
        v.y = fma(-k0sq, b.y, v.y);
    }
    return v;
}

// COO index of the mirrored entry (jj, ii) of entry k = 400 ia + 20 ii + jj.
__device__ __forceinline__ long long mirror_coo(long long k) {
    const long long blk = k / 400, e = k - blk * 400;
    const long long ii = e / 20, jj = e - ii * 20;
    return blk * 400 + jj * 20 + ii;
}

// data_csc[q] = sum over contributors of (E[k] - k0sq * B[k]) (B may be null: E[k]).
// MIRROR: the blocks are symmetric, so read entry k from its mirror. The contributors of
// one column then lie in contiguous 20-entry rows of the blocks (one per incident tet)
// instead of 20-entry strides, and a wavefront's reads share cache lines; the values and
// their summation order are unchanged.
template <typename T, bool MIRROR>
__global__ __launch_bounds__(256) void csc_gather_kernel(const long long* __restrict__ seg,
                                                         const long long* __restrict__ perm,
                                                         long long nnz, const T* __restrict__ E,
                                                         const T* __restrict__ B, double k0sq,
                                                         T* __restrict__ out) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
         q += (long long)gridDim.x * blockDim.x) {
        const long long j0 = seg[q], j1 = seg[q + 1];
        double2 acc = make_double2(0.0, 0.0);
        for (long long j = j0; j < j1; ++j) {
            const long long k = MIRROR ? mirror_coo(perm[j]) : perm[j];
            const double2 v = contrib(E, B, k0sq, k);
            acc.x += v.x;
            acc.y += v.y;
        }
        store_c<T>(out, q, acc);
    }
}

// to_csc by column incidences. A column's contributors are the 20 rows ii of every
// incident (tet ia, slot jj). Each wavefront carries three columns, one per group of 20
// lanes (lane ii of a group handles row slot ii); a group walks its column's incidence
// list in order, loads the incidence's 20 values (with MIRROR one contiguous 20-entry
// row of the block) and adds each into the column's FP64 accumulator in LDS at pos[]
// (the entry index of the row within the column). The 20 lanes of one incidence hit
// distinct entries and a group issues its incidences in list order, so every entry sums
// its contributors in incidence order = ascending COO index, as the perm/seg gather does:
// the result is the same bit for bit. The zeroing, the adds and the read-back of a column
// window are separated by wavefront LDS syncs (lanes hand entries to each other through
// LDS). Columns longer than kIncCap entries are done in windows.
constexpr int kIncCap = 128;
constexpr int kIncWaves = 4;  // wavefronts per workgroup
constexpr int kIncCols = 3;   // columns per wavefront

// Workgroup order of the gathers. Workgroups are dispatched round-robin over the 8 XCDs
// (each with its own L2). The mirrored gathers keep that order (workgroup b, columns in
// index order across the XCDs). The plain gathers read each incident block down a
// column, 20 strided entries, and profit from neighbouring columns (a tet's DOFs) sharing
// an L2: there workgroup b = 8 j + x takes ((j / kIncWindow) * 8 + x) * kIncWindow +
// j % kIncWindow, runs of kIncWindow consecutive workgroups per XCD (551k tets, complex64
// combine: 5.5 -> 4.4 ms; heat 5.4 -> 4.8 ms). For the mirrored gathers the same order
// was no faster (1.33 vs 1.36 ms), and one contiguous eighth of the columns per XCD was
// 1.4x slower: the heavy columns come first in DOF order. The launchers round the grid
// up to whole window rows.
constexpr long long kIncWindow = 128;
template <bool WINDOWED>
__device__ __forceinline__ long long inc_block() {
    if (!WINDOWED) return blockIdx.x;
    const long long j = blockIdx.x / 8;
    return ((j / kIncWindow) * 8 + blockIdx.x % 8) * kIncWindow + j % kIncWindow;
}

static unsigned inc_grid(long long nb, long long cap) {
    const long long g = nb < cap ? nb : cap, span = 8 * kIncWindow;
    return (unsigned)((g + span - 1) / span * span);
}

template <typename T, bool MIRROR, bool HAS_B, typename TO = T>
__global__ __launch_bounds__(64 * kIncWaves) void csc_gather_inc_kernel(
    const long long* __restrict__ indptr, const long long* __restrict__ inc_ptr,
    const int* __restrict__ inc, const unsigned short* __restrict__ pos,
    const long long* __restrict__ cols, long long N, const T* __restrict__ E,
    const T* __restrict__ B, double k0sq, TO* __restrict__ out) {
    __shared__ double acc[kIncWaves * kIncCols][kIncCap][2];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane / 20, ii = lane - 20 * grp;
    if (grp >= kIncCols) return;  // lanes 60..63: no barriers below, nothing to do
    double (*a)[2] = acc[wave * kIncCols + grp];
    const long long stride = (long long)gridDim.x * kIncWaves * kIncCols;
    for (long long c = (inc_block<!MIRROR>() * kIncWaves + wave) * kIncCols + grp; c < N;
         c += stride) {
        const long long cb = indptr[c], len = indptr[c + 1] - cb;
        const long long fc = cols ? cols[c] : c;  // column of the incidence lists
        const long long s = inc_ptr[fc], ninc = inc_ptr[fc + 1] - s;
        for (long long w0 = 0; w0 < len; w0 += kIncCap) {
            const int wl = (int)(len - w0 < kIncCap ? len - w0 : kIncCap);
            for (int p = ii; p < wl; p += 20) {
                a[p][0] = 0.0;
                a[p][1] = 0.0;
            }
            wave_lds_sync();  // zeroed accumulators before any lane adds
            // batches of kU incidences: all loads of a batch are issued before the adds
            constexpr int kU = 4;
            for (long long i0 = 0; i0 < ninc; i0 += kU) {
                int p[kU], e[kU];
                double2 v[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const long long i = i0 + u < ninc ? i0 + u : ninc - 1;  // in range
                    e[u] = inc[s + i];
                    p[u] = i0 + u < ninc ? (int)pos[20 * (s + i) + ii] - (int)w0 : -1;
                }
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const long long ia = e[u] / 20;
                    const int jj = e[u] - 20 * (int)ia;
                    const long long k = 400 * ia + (MIRROR ? 20 * jj + ii : 20 * ii + jj);
                    v[u] = LoadC<T>::get(E, k);
                    if (HAS_B) {
                        const double2 b = LoadC<T>::get(B, k);
                        v[u].x = fma(-k0sq, b.x, v[u].x);
                        v[u].y = fma(-k0sq, b.y, v[u].y);
                    }
                }
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (p[u] >= 0 && p[u] < wl) {
                        atomicAdd(&a[p[u]][0], v[u].x);
                        atomicAdd(&a[p[u]][1], v[u].y);
                    }
            }
            wave_lds_sync();  // every lane's adds before the read-back
            for (int p = ii; p < wl; p += 20) store_c<TO>(out, cb + w0 + p,
                                                         make_double2(a[p][0], a[p][1]));
            wave_lds_sync();  // read-back before the next window or column zeroes
        }
    }
}

__global__ __launch_bounds__(256) void seg_to_target_kernel(const long long* __restrict__ seg,
                                                            long long nnz,
                                                            long long* __restrict__ target) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
         q += (long long)gridDim.x * blockDim.x)
        for (long long j = seg[q]; j < seg[q + 1]; ++j) target[j] = q;
}

__global__ __launch_bounds__(256) void scale_kernel(const long long* __restrict__ in, long long n,
                                                    long long mul, long long* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = in[i] * mul;
}

static unsigned grid_for(long long n, long long cap = 65536 * 4) {
    long long b = (n + 255) / 256;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// Builds the pattern of a DOF map t2f [T][NDOF] over n_assy assembly indices (N DOFs).
// On success the caller owns b's buffers, including the incidence lists and the
// incidence-gather row positions (pos).
template <int NDOF>
static hipError_t csc_build_n(const int* t2f, const int* tetids, long long n_assy, long long N,
                              CscBuild& b, long long& max_inc, hipStream_t s) {
    struct {
        const int* t2f;
        const int* tetids;
        long long n_assy;
    } m = {t2f, tetids, n_assy};
    const long long nnz_coo = m.n_assy * NDOF * NDOF;
    hipError_t e;
    unsigned long long *cnt = nullptr, *maxlen = nullptr;
    int* inc = nullptr;
    long long* ucount = nullptr;
    bool long_cols = false, huge_cols = false;
    int* g_rows = nullptr;  // scratch rows of the columns longer than the LDS holds
    unsigned long long* g_first = nullptr;
    int g_stride = 0;
    unsigned gh = 0;
#define TRY(x)                          \
    do {                                \
        e = (x);                        \
        if (e != hipSuccess) goto fail; \
    } while (0)
    TRY(hipMalloc((void**)&cnt, sizeof(unsigned long long) * (N + 1)));
    TRY(hipMalloc((void**)&b.inc_ptr, sizeof(long long) * (N + 1)));
    TRY(hipMalloc((void**)&maxlen, sizeof(unsigned long long)));
    TRY(hipMalloc((void**)&inc,
                  sizeof(int) * (size_t)(m.n_assy * NDOF > 0 ? m.n_assy * NDOF : 1)));
    TRY(hipMalloc((void**)&b.pos, sizeof(unsigned short) * (size_t)(nnz_coo > 0 ? nnz_coo : 1)));
    TRY(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * (N + 1), s));
    TRY(hipMemsetAsync(maxlen, 0, sizeof(unsigned long long), s));
    if (m.n_assy > 0) {
        hipLaunchKernelGGL(inc_count_kernel<NDOF>, dim3(grid_for(m.n_assy * NDOF)), dim3(256), 0, s,
                           m.t2f,
                           m.tetids, m.n_assy, cnt);
        TRY(hipGetLastError());
    }
    TRY(exclusive_scan_i64((const long long*)cnt, b.inc_ptr, N + 1, s));
    TRY(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * (N + 1), s));
    if (m.n_assy > 0) {
        hipLaunchKernelGGL(inc_fill_kernel<NDOF>, dim3(grid_for(m.n_assy * NDOF)), dim3(256), 0, s,
                           m.t2f,
                           m.tetids, m.n_assy, b.inc_ptr, cnt, inc);
        TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(inc_sort_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s,
                       b.inc_ptr, inc, N, maxlen);
    TRY(hipGetLastError());
    {
        unsigned long long ml = 0;
        TRY(hipMemcpyAsync(&ml, maxlen, sizeof(ml), hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        max_inc = (long long)ml;
        if (NDOF * max_inc > kCscMaxColumnKeys) {
            e = hipErrorInvalidValue;
            goto fail;
        }
    }
    long_cols = NDOF * max_inc > kColCap;
    huge_cols = NDOF * max_inc > kLongLds;
    if (huge_cols) {
        g_stride = (int)((NDOF * max_inc + 63) / 64 * 64);
        gh = (unsigned)(N < 512 ? N : 512);
        TRY(hipMalloc((void**)&g_rows, sizeof(int) * (size_t)gh * g_stride));
        TRY(hipMalloc((void**)&g_first, sizeof(unsigned long long) * (size_t)gh * (g_stride / 64)));
    }
    // pass 1: distinct rows per column
    ucount = reinterpret_cast<long long*>(cnt);  // reuse: N+1 int64
    TRY(hipMemsetAsync(ucount, 0, sizeof(long long) * (N + 1), s));
    {
        const unsigned g = (unsigned)(N < (1 << 20) ? N : (1 << 20));
        const unsigned gl = (unsigned)(N < (1 << 16) ? N : (1 << 16));
        if (N > 0) {
            hipLaunchKernelGGL((csc_column_kernel<false, NDOF>), dim3(g), dim3(64), 0, s, m.t2f,
                               m.tetids,
                               b.inc_ptr, inc, N, ucount, nullptr, nullptr, nullptr, nullptr,
                               nullptr);
            TRY(hipGetLastError());
        }
        if (long_cols) {
            hipLaunchKernelGGL((csc_long_column_kernel<false, NDOF, false>), dim3(gl),
                               dim3(kLongThreads), 0, s, m.t2f, m.tetids, b.inc_ptr, inc, N,
                               ucount, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                               nullptr, 0);
            TRY(hipGetLastError());
        }
        if (huge_cols) {
            hipLaunchKernelGGL((csc_long_column_kernel<false, NDOF, true>), dim3(gh),
                               dim3(kLongThreads), 0, s, m.t2f, m.tetids, b.inc_ptr, inc, N,
                               ucount, nullptr, nullptr, nullptr, nullptr, nullptr, g_rows,
                               g_first, g_stride);
            TRY(hipGetLastError());
        }
        TRY(hipMalloc((void**)&b.indptr, sizeof(long long) * (N + 1)));
        TRY(exclusive_scan_i64(ucount, b.indptr, N + 1, s));
        TRY(hipMemcpyAsync(&b.nnz, b.indptr + N, sizeof(long long), hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        TRY(hipMalloc((void**)&b.indices, sizeof(long long) * (b.nnz > 0 ? b.nnz : 1)));
        TRY(hipMalloc((void**)&b.seg, sizeof(long long) * (b.nnz + 1)));
        TRY(hipMalloc((void**)&b.perm, sizeof(long long) * (nnz_coo > 0 ? nnz_coo : 1)));
        TRY(hipMemcpyAsync(b.seg + b.nnz, &nnz_coo, sizeof(long long), hipMemcpyHostToDevice, s));
        if (N > 0) {
            hipLaunchKernelGGL((csc_column_kernel<true, NDOF>), dim3(g), dim3(64), 0, s, m.t2f,
                               m.tetids,
                               b.inc_ptr, inc, N, nullptr, b.indptr, b.indices, b.perm, b.seg,
                               b.pos);
            TRY(hipGetLastError());
        }
        if (long_cols) {
            hipLaunchKernelGGL((csc_long_column_kernel<true, NDOF, false>), dim3(gl),
                               dim3(kLongThreads), 0, s, m.t2f, m.tetids, b.inc_ptr, inc, N,
                               nullptr, b.indptr, b.indices, b.perm, b.seg, b.pos, nullptr,
                               nullptr, 0);
            TRY(hipGetLastError());
        }
        if (huge_cols) {
            hipLaunchKernelGGL((csc_long_column_kernel<true, NDOF, true>), dim3(gh),
                               dim3(kLongThreads), 0, s, m.t2f, m.tetids, b.inc_ptr, inc, N,
                               nullptr, b.indptr, b.indices, b.perm, b.seg, b.pos, g_rows,
                               g_first, g_stride);
            TRY(hipGetLastError());
        }
        TRY(hipStreamSynchronize(s));
    }
    (void)hipFree(cnt);
    (void)hipFree(maxlen);
    (void)hipFree(g_rows);
    (void)hipFree(g_first);
    b.inc = inc;  // kept for the incidence gather
    return hipSuccess;
fail:
    (void)hipStreamSynchronize(s);
    (void)hipFree(cnt);
    (void)hipFree(maxlen);
    (void)hipFree(g_rows);
    (void)hipFree(g_first);
    (void)hipFree(inc);
    (void)hipFree(b.indptr);
    (void)hipFree(b.indices);
    (void)hipFree(b.perm);
    (void)hipFree(b.seg);
    (void)hipFree(b.inc_ptr);
    (void)hipFree(b.pos);
    b = CscBuild();
    return e;
#undef TRY
}

hipError_t csc_build(const MeshDev& m, long long N, CscBuild& b, long long& max_inc,
                     hipStream_t s) {
    b.ndof = 20;
    return csc_build_n<20>(m.t2f, m.tetids, m.n_assy, N, b, max_inc, s);
}

hipError_t csc_build_ndof10(const int* t2f, long long n_tets, long long N, CscBuild& b,
                            long long& max_inc, hipStream_t s) {
    b.ndof = 10;
    hipError_t e = csc_build_n<10>(t2f, nullptr, n_tets, N, b, max_inc, s);
    if (e == hipSuccess) b.ndof = 10;
    return e;
}

// Real-valued data, column-incidence form (as csc_gather_inc_kernel): one group of NDOF
// lanes per column, 64 / NDOF columns per wavefront; lane ii of a group reads entry
// (ii, jj) of every incident block and adds it at its row position in an LDS column, in
// incidence (= COO) order, so the result equals the segment gather's bit for bit.
// MIRROR (bit-symmetric blocks): entry (ii, jj) is read from (jj, ii), so a group's reads
// are one contiguous NDOF-entry row of each incident block.
template <int NDOF, bool MIRROR, bool HAS_B>
__global__ __launch_bounds__(64 * kIncWaves) void csc_gather_inc_real_kernel(
    const long long* __restrict__ indptr, const long long* __restrict__ inc_ptr,
    const int* __restrict__ inc, const unsigned short* __restrict__ pos, long long N,
    const double* __restrict__ E, const double* __restrict__ B, double scale,
    double* __restrict__ out) {
    constexpr int kCols = 64 / NDOF;
    __shared__ double acc[kIncWaves * kCols][kIncCap];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane / NDOF, ii = lane - NDOF * grp;
    if (grp >= kCols) return;  // idle lanes: no barriers below
    double* a = acc[wave * kCols + grp];
    const long long stride = (long long)gridDim.x * kIncWaves * kCols;
    for (long long c = (inc_block<!MIRROR>() * kIncWaves + wave) * kCols + grp; c < N;
         c += stride) {
        const long long cb = indptr[c], len = indptr[c + 1] - cb;
        const long long s = inc_ptr[c], ninc = inc_ptr[c + 1] - s;
        for (long long w0 = 0; w0 < len; w0 += kIncCap) {
            const int wl = (int)(len - w0 < kIncCap ? len - w0 : kIncCap);
            for (int p = ii; p < wl; p += NDOF) a[p] = 0.0;
            wave_lds_sync();  // zeroed accumulators before any lane adds
            constexpr int kU = 4;
            for (long long i0 = 0; i0 < ninc; i0 += kU) {
                int p[kU];
                double v[kU];
#pragma unroll
                for (int u = 0; u < kU; ++u) {
                    const long long i = i0 + u < ninc ? i0 + u : ninc - 1;  // in range
                    const int e = inc[s + i];
                    p[u] = i0 + u < ninc ? (int)pos[NDOF * (s + i) + ii] - (int)w0 : -1;
                    const long long ia = e / NDOF;
                    const int jj = e - NDOF * (int)ia;
                    const long long k =
                        (long long)NDOF * NDOF * ia + (MIRROR ? NDOF * jj + ii : NDOF * ii + jj);
                    v[u] = HAS_B ? fma(scale, B[k], E[k]) : E[k];
                }
#pragma unroll
                for (int u = 0; u < kU; ++u)
                    if (p[u] >= 0 && p[u] < wl) atomicAdd(&a[p[u]], v[u]);
            }
            wave_lds_sync();  // every lane's adds before the read-back
            for (int p = ii; p < wl; p += NDOF) out[cb + w0 + p] = a[p];
            wave_lds_sync();  // read-back before the next window or column zeroes
        }
    }
}

hipError_t csc_gather_inc_real(const CscBuild& b, long long N, const double* E, const double* B,
                               double scale, bool mirror, double* out, hipStream_t s) {
    if (b.nnz == 0 || N == 0) return hipSuccess;
    if (b.ndof != 10) return hipErrorInvalidValue;
    constexpr int kCols = 64 / 10;
    const long long groups = (N + kIncWaves * kCols - 1) / (kIncWaves * kCols);
    const unsigned g = inc_grid(groups, 65536 * 4);
#define EM_INC_REAL(MIR, HB)                                                                    \
    hipLaunchKernelGGL((csc_gather_inc_real_kernel<10, MIR, HB>), dim3(g), dim3(64 * kIncWaves), \
                       0, s, b.indptr, b.inc_ptr, b.inc, b.pos, N, E, B, scale, out)
    if (mirror) {
        if (B)
            EM_INC_REAL(true, true);
        else
            EM_INC_REAL(true, false);
    } else {
        if (B)
            EM_INC_REAL(false, true);
        else
            EM_INC_REAL(false, false);
    }
#undef EM_INC_REAL
    return hipGetLastError();
}

// Real-valued data (one double per COO entry): data_csc[q] = sum of its contributors.
__global__ __launch_bounds__(256) void csc_gather_real_kernel(const long long* __restrict__ seg,
                                                              const long long* __restrict__ perm,
                                                              long long nnz,
                                                              const double* __restrict__ E,
                                                              const double* __restrict__ B,
                                                              double scale,
                                                              double* __restrict__ out) {
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nnz;
         q += (long long)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (long long j = seg[q]; j < seg[q + 1]; ++j) {
            const long long k = perm[j];
            acc += B ? fma(scale, B[k], E[k]) : E[k];
        }
        out[q] = acc;
    }
}

hipError_t csc_gather_real(const long long* seg, const long long* perm, long long nnz,
                           const double* E, const double* B, double scale, double* out,
                           hipStream_t s) {
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(csc_gather_real_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, seg, perm,
                       nnz, E, B, scale, out);
    return hipGetLastError();
}

template <typename T>
static void launch_gather(bool mirror, const long long* seg, const long long* perm, long long nnz,
                          const void* E, const void* B, double k0sq, void* out, hipStream_t s) {
    if (mirror)
        hipLaunchKernelGGL((csc_gather_kernel<T, true>), dim3(grid_for(nnz)), dim3(256), 0, s, seg,
                           perm, nnz, (const T*)E, (const T*)B, k0sq, (T*)out);
    else
        hipLaunchKernelGGL((csc_gather_kernel<T, false>), dim3(grid_for(nnz)), dim3(256), 0, s,
                           seg, perm, nnz, (const T*)E, (const T*)B, k0sq, (T*)out);
}

hipError_t csc_gather(const long long* seg, const long long* perm, long long nnz, const void* E,
                      const void* B, double k0sq, void* out, int complex64, int mirror,
                      hipStream_t s) {
    if (nnz == 0) return hipSuccess;
    if (complex64)
        launch_gather<float2>(mirror != 0, seg, perm, nnz, E, B, k0sq, out, s);
    else
        launch_gather<double2>(mirror != 0, seg, perm, nnz, E, B, k0sq, out, s);
    return hipGetLastError();
}

template <typename T, bool MIRROR>
static void launch_inc(const CscBuild& b, long long N, const void* E, const void* B, double k0sq,
                       void* out, hipStream_t s) {
    const long long per_block = kIncWaves * kIncCols;
    const long long nb = (N + per_block - 1) / per_block;
    const dim3 g(inc_grid(nb, 1 << 20)), blk(64 * kIncWaves);
    if (B)
        hipLaunchKernelGGL((csc_gather_inc_kernel<T, MIRROR, true>), g, blk, 0, s, b.indptr,
                           b.inc_ptr, b.inc, b.pos, b.cols, N, (const T*)E, (const T*)B, k0sq,
                           (T*)out);
    else
        hipLaunchKernelGGL((csc_gather_inc_kernel<T, MIRROR, false>), g, blk, 0, s, b.indptr,
                           b.inc_ptr, b.inc, b.pos, b.cols, N, (const T*)E, (const T*)B, k0sq,
                           (T*)out);
}

// Staging for blocks that are not symmetric: work[400 ia + 20 jj + ii] = the FP64
// contribution E - k0sq B of entry (ii, jj) (the value the plain gather would form), one
// wavefront per block through LDS, so that the mirrored gather over work reads each
// incident block's row contiguously and sums the same values in the same order.
template <typename T, bool HAS_B>
__global__ __launch_bounds__(256) void coo_stage_transpose_kernel(const T* __restrict__ E,
                                                                  const T* __restrict__ B,
                                                                  double k0sq, long long n_blk,
                                                                  double2* __restrict__ work) {
    __shared__ double2 tile[4][400];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (long long blk = (long long)blockIdx.x * 4 + wave; blk < n_blk;
         blk += (long long)gridDim.x * 4) {
        const long long base = blk * 400;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int e = lane + 64 * k;
            if (e < 400) {
                const int ii = e / 20, jj = e - 20 * ii;
                tile[wave][20 * jj + ii] = contrib<T>(E, HAS_B ? B : nullptr, k0sq, base + e);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            const int e = lane + 64 * k;
            if (e < 400) work[base + e] = tile[wave][e];
        }
        wave_lds_sync();
    }
}

hipError_t csc_gather_inc_staged(const CscBuild& b, long long N, long long n_blk, const void* E,
                                 const void* B, double k0sq, double2* work, void* out,
                                 int complex64, hipStream_t s) {
    if (N == 0 || b.nnz == 0) return hipSuccess;
    const long long gb = (n_blk + 3) / 4;
    const dim3 gt((unsigned)(gb < 65536 * 4 ? gb : 65536 * 4));
    if (complex64) {
        if (B)
            hipLaunchKernelGGL((coo_stage_transpose_kernel<float2, true>), gt, dim3(256), 0, s,
                               (const float2*)E, (const float2*)B, k0sq, n_blk, work);
        else
            hipLaunchKernelGGL((coo_stage_transpose_kernel<float2, false>), gt, dim3(256), 0, s,
                               (const float2*)E, (const float2*)B, k0sq, n_blk, work);
    } else {
        if (B)
            hipLaunchKernelGGL((coo_stage_transpose_kernel<double2, true>), gt, dim3(256), 0, s,
                               (const double2*)E, (const double2*)B, k0sq, n_blk, work);
        else
            hipLaunchKernelGGL((coo_stage_transpose_kernel<double2, false>), gt, dim3(256), 0, s,
                               (const double2*)E, (const double2*)B, k0sq, n_blk, work);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long per_block = kIncWaves * kIncCols;
    const long long nb = (N + per_block - 1) / per_block;
    const dim3 g(inc_grid(nb, 1 << 20)), blk(64 * kIncWaves);
    if (complex64)
        hipLaunchKernelGGL((csc_gather_inc_kernel<double2, true, false, float2>), g, blk, 0, s,
                           b.indptr, b.inc_ptr, b.inc, b.pos, b.cols, N, (const double2*)work,
                           (const double2*)nullptr, 0.0, (float2*)out);
    else
        hipLaunchKernelGGL((csc_gather_inc_kernel<double2, true, false, double2>), g, blk, 0, s,
                           b.indptr, b.inc_ptr, b.inc, b.pos, b.cols, N, (const double2*)work,
                           (const double2*)nullptr, 0.0, (double2*)out);
    return hipGetLastError();
}

hipError_t csc_gather_inc(const CscBuild& b, long long N, const void* E, const void* B,
                          double k0sq, void* out, int complex64, int mirror, hipStream_t s) {
    if (N == 0 || b.nnz == 0) return hipSuccess;
    if (complex64) {
        if (mirror)
            launch_inc<float2, true>(b, N, E, B, k0sq, out, s);
        else
            launch_inc<float2, false>(b, N, E, B, k0sq, out, s);
    } else {
        if (mirror)
            launch_inc<double2, true>(b, N, E, B, k0sq, out, s);
        else
            launch_inc<double2, false>(b, N, E, B, k0sq, out, s);
    }
    return hipGetLastError();
}

hipError_t csc_target(const long long* seg, long long nnz, long long* target, hipStream_t s) {
    if (nnz == 0) return hipSuccess;
    hipLaunchKernelGGL(seg_to_target_kernel, dim3(grid_for(nnz)), dim3(256), 0, s, seg, nnz, target);
    return hipGetLastError();
}

hipError_t csc_col_offsets(const long long* inc_ptr, long long N, long long ndof, long long* out,
                           hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(N + 1)), dim3(256), 0, s, inc_ptr, N + 1, ndof,
                       out);
    return hipGetLastError();
}

}  // namespace em
