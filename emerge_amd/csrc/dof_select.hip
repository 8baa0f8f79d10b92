// PEC elimination: the prescribed-DOF mask, the solve-id compaction and the selection of
// the free-DOF block of a CSC matrix.
//
// Reference: the PEC DOF collection of Assembler.assemble_freq_matrix
// (src/emerge/_emerge/physics/microwave/assembly/assembler.py:533-564: every DOF of a
// conductor tet, and every edge and face DOF of a PEC boundary triangle), solve_ids as the
// ascending complement (:742), and the solver's selection A[:, solve_ids][solve_ids, :]
// (src/emerge/_emerge/solver.py:2071; :2229 A[np.ix_(solve_ids, solve_ids)] is the same
// matrix), with b[solve_ids, :] (:2072) and the scatter of the solution back to all DOFs.
//
// All of it is integer/byte work bound by HBM: mark kernels, one exclusive scan, and one
// wavefront per column that compacts the column's kept rows with ballots (the rows stay
// ascending, so the output is scipy's canonical CSC). The restricted CSC map keeps, for the
// kept entries only, the segments of the full map's COO permutation, so the per-frequency
// to_csc of the restricted map produces the free-DOF block straight from the COO values.
#include <hip/hip_runtime.h>

#include "em_internal.h"

namespace em {

namespace {

unsigned grid_for(long long n, long long per_block = 256, long long cap = 65536 * 4) {
    long long b = (n + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

constexpr int kWavesPerBlock = 4;  // one wavefront per column

__global__ __launch_bounds__(256) void fill_u8_kernel(unsigned char* __restrict__ p, long long n,
                                                      unsigned char v) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// keep[d] = 0 for the 20 DOFs of every conductor tet and the 8 DOFs (3 edges x 2, face x 2)
// of every PEC face. Ids out of range set *bad and are skipped.
__global__ __launch_bounds__(256) void pec_mark_kernel(
    const int* __restrict__ t2f, long long n_tets, const long long* __restrict__ cond,
    long long n_cond, const long long* __restrict__ tri_to_field, long long n_tris,
    const long long* __restrict__ pec_tris, long long n_pec, long long n_field,
    unsigned char* __restrict__ keep, unsigned long long* __restrict__ bad) {
    const long long n_items = n_cond * 20 + n_pec * 8;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_items;
         i += (long long)gridDim.x * blockDim.x) {
        long long d;
        if (i < n_cond * 20) {
            const long long t = cond[i / 20];
            if (t < 0 || t >= n_tets) {
                atomicOr(bad, 1ull);
                continue;
            }
            d = t2f[t * 20 + i % 20];
        } else {
            const long long j = i - n_cond * 20;
            const long long f = pec_tris[j / 8];
            if (f < 0 || f >= n_tris) {
                atomicOr(bad, 2ull);
                continue;
            }
            d = tri_to_field[(j % 8) * n_tris + f];
        }
        if (d < 0 || d >= n_field) {
            atomicOr(bad, 4ull);
            continue;
        }
        keep[d] = 0;
    }
}

__global__ __launch_bounds__(256) void keep_to_i64_kernel(const unsigned char* __restrict__ keep,
                                                          long long n, long long* __restrict__ out) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i <= n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (i < n && keep[i]) ? 1 : 0;
}

// scan holds the exclusive scan of keep; becomes new_index (-1 for removed DOFs)
__global__ __launch_bounds__(256) void compact_kernel(const unsigned char* __restrict__ keep,
                                                      long long n, long long* __restrict__ scan,
                                                      long long* __restrict__ solve_ids) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const long long s = scan[i];
        if (keep[i]) {
            solve_ids[s] = i;
        } else {
            scan[i] = -1;
        }
    }
}

__device__ __forceinline__ bool is_zero(const void* data, int complex64, long long q) {
    if (complex64) {
        const float2 v = ((const float2*)data)[q];
        return v.x == 0.0f && v.y == 0.0f;
    }
    const double2 v = ((const double2*)data)[q];
    return v.x == 0.0 && v.y == 0.0;
}

// One wavefront per old column: the number of kept entries of each kept column.
__global__ __launch_bounds__(64 * kWavesPerBlock) void select_count_kernel(
    long long n, const long long* __restrict__ indptr, const long long* __restrict__ indices,
    const void* __restrict__ data, int complex64, int drop_zeros,
    const long long* __restrict__ new_index, long long* __restrict__ cnt,
    unsigned long long* __restrict__ bad) {
    const int lane = threadIdx.x & 63;
    for (long long c = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); c < n;
         c += (long long)gridDim.x * kWavesPerBlock) {
        const long long nc = new_index[c];
        if (nc < 0) continue;
        long long total = 0;
        for (long long j0 = indptr[c]; j0 < indptr[c + 1]; j0 += 64) {
            const long long j = j0 + lane;
            bool k = false;
            if (j < indptr[c + 1]) {
                const long long r = indices[j];
                if (r < 0 || r >= n) {
                    atomicOr(bad, 1ull);  // the fill pass is not run
                } else {
                    k = new_index[r] >= 0;
                }
            }
            if (k && drop_zeros) k = !is_zero(data, complex64, j);
            total += __popcll(__ballot(k));
        }
        if (lane == 0) cnt[nc] = total;
    }
}

template <typename T>
__global__ __launch_bounds__(64 * kWavesPerBlock) void select_fill_kernel(
    long long n, const long long* __restrict__ indptr, const long long* __restrict__ indices,
    const T* __restrict__ data, int drop_zeros, const long long* __restrict__ new_index,
    const long long* __restrict__ sel_indptr, long long* __restrict__ sel_indices,
    T* __restrict__ sel_data) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (long long c = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); c < n;
         c += (long long)gridDim.x * kWavesPerBlock) {
        const long long nc = new_index[c];
        if (nc < 0) continue;
        long long out = sel_indptr[nc];
        for (long long j0 = indptr[c]; j0 < indptr[c + 1]; j0 += 64) {
            const long long j = j0 + lane;
            long long r = -1;
            T v{};
            if (j < indptr[c + 1]) {
                const long long ri = indices[j];  // out of range: reported by the count pass
                r = (ri >= 0 && ri < n) ? new_index[ri] : -1;
                v = data ? data[j] : T{};
            }
            bool k = r >= 0;
            if (k && drop_zeros) k = !(v.x == 0 && v.y == 0);
            const unsigned long long m = __ballot(k);
            if (k) {
                const long long p = out + __popcll(m & below);
                sel_indices[p] = r;
                if (sel_data) sel_data[p] = v;
            }
            out += __popcll(m);
        }
    }
}

// Restricted map, pass 1: kept entries and their COO contributors per kept column.
__global__ __launch_bounds__(64 * kWavesPerBlock) void restrict_count_kernel(
    long long n, const long long* __restrict__ indptr, const long long* __restrict__ indices,
    const long long* __restrict__ seg, const long long* __restrict__ new_index,
    long long* __restrict__ ent_cnt, long long* __restrict__ perm_cnt) {
    const int lane = threadIdx.x & 63;
    for (long long c = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); c < n;
         c += (long long)gridDim.x * kWavesPerBlock) {
        const long long nc = new_index[c];
        if (nc < 0) continue;
        long long ne = 0, np = 0;
        for (long long j0 = indptr[c]; j0 < indptr[c + 1]; j0 += 64) {
            const long long j = j0 + lane;
            const bool k = j < indptr[c + 1] && new_index[indices[j]] >= 0;
            long long len = k ? seg[j + 1] - seg[j] : 0;
            ne += __popcll(__ballot(k));
            for (int off = 32; off > 0; off >>= 1) len += __shfl_xor(len, off);
            np += len;
        }
        if (lane == 0) {
            ent_cnt[nc] = ne;
            perm_cnt[nc] = np;
        }
    }
}

// Restricted map, pass 2: indices, seg and perm of the kept entries.
__global__ __launch_bounds__(64 * kWavesPerBlock) void restrict_fill_kernel(
    long long n, const long long* __restrict__ indptr, const long long* __restrict__ indices,
    const long long* __restrict__ seg, const long long* __restrict__ perm,
    const long long* __restrict__ new_index, const long long* __restrict__ sub_indptr,
    const long long* __restrict__ sub_colperm, long long* __restrict__ sub_indices,
    long long* __restrict__ sub_seg, long long* __restrict__ sub_perm,
    long long* __restrict__ newq) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (long long c = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); c < n;
         c += (long long)gridDim.x * kWavesPerBlock) {
        const long long nc = new_index[c];
        if (nc < 0) continue;
        long long out = sub_indptr[nc], pout = sub_colperm[nc];
        for (long long j0 = indptr[c]; j0 < indptr[c + 1]; j0 += 64) {
            const long long j = j0 + lane;
            long long r = -1;
            if (j < indptr[c + 1]) r = new_index[indices[j]];
            const bool k = r >= 0;
            const long long len = k ? seg[j + 1] - seg[j] : 0;
            // inclusive scan of len across the wavefront
            long long inc = len;
            for (int off = 1; off < 64; off <<= 1) {
                const long long t = __shfl_up(inc, off);
                if (lane >= off) inc += t;
            }
            const unsigned long long m = __ballot(k);
            if (j < indptr[c + 1]) newq[j] = k ? out + __popcll(m & below) : -1;
            if (k) {
                const long long p = out + __popcll(m & below);
                const long long ps = pout + inc - len;
                sub_indices[p] = r;
                sub_seg[p] = ps;
                for (long long i = 0; i < len; ++i) sub_perm[ps + i] = perm[seg[j] + i];
            }
            out += __popcll(m);
            pout += __shfl(inc, 63);
        }
    }
}

// Restricted map, pass 3 (incidence gather tables): for every kept column, the entry
// index of each incidence row within the block's column (0xffff: row removed), and the
// full-matrix column of every block column.
__global__ __launch_bounds__(64 * kWavesPerBlock) void restrict_pos_kernel(
    long long n, const long long* __restrict__ indptr, const long long* __restrict__ inc_ptr,
    const unsigned short* __restrict__ pos, const long long* __restrict__ new_index,
    const long long* __restrict__ newq, const long long* __restrict__ sub_indptr,
    unsigned short* __restrict__ sub_pos, long long* __restrict__ cols) {
    const int lane = threadIdx.x & 63;
    for (long long c = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6); c < n;
         c += (long long)gridDim.x * kWavesPerBlock) {
        const long long nc = new_index[c];
        if (nc < 0) continue;
        if (lane == 0) cols[nc] = c;
        const long long x0 = 20 * inc_ptr[c], x1 = 20 * inc_ptr[c + 1];
        for (long long x = x0 + lane; x < x1; x += 64) {
            const long long nq = newq[indptr[c] + pos[x]];
            sub_pos[x] = nq >= 0 ? (unsigned short)(nq - sub_indptr[nc]) : (unsigned short)0xffff;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void dof_gather_kernel(const long long* __restrict__ ids,
                                                         long long n_ids, long long ncols,
                                                         const T* __restrict__ src,
                                                         T* __restrict__ dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_ids * ncols;
         i += (long long)gridDim.x * blockDim.x) {
        const long long r = i / ncols, c = i % ncols;
        dst[i] = src[ids[r] * ncols + c];
    }
}

template <typename T>
__global__ __launch_bounds__(256) void dof_scatter_kernel(const long long* __restrict__ ids,
                                                          long long n_ids, long long ncols,
                                                          const T* __restrict__ src,
                                                          T* __restrict__ dst) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_ids * ncols;
         i += (long long)gridDim.x * blockDim.x) {
        const long long r = i / ncols, c = i % ncols;
        dst[ids[r] * ncols + c] = src[i];
    }
}

unsigned col_grid(long long n) {
    return grid_for(n, kWavesPerBlock, 1 << 20);
}

}  // namespace

hipError_t pec_mark(const int* t2f, long long n_tets, const long long* cond, long long n_cond,
                    const long long* tri_to_field, long long n_tris, const long long* pec_tris,
                    long long n_pec, long long n_field, unsigned char* keep, long long& bad,
                    hipStream_t s) {
    bad = 0;
    unsigned long long* dbad = nullptr;
    hipError_t e = hipMalloc((void**)&dbad, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    unsigned long long hbad = 0;
    if ((e = hipMemsetAsync(dbad, 0, sizeof(unsigned long long), s)) != hipSuccess) goto done;
    if (n_field > 0) {
        hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(n_field)), dim3(256), 0, s, keep, n_field,
                           (unsigned char)1);
        if ((e = hipGetLastError()) != hipSuccess) goto done;
    }
    if (n_cond * 20 + n_pec * 8 > 0) {
        hipLaunchKernelGGL(pec_mark_kernel, dim3(grid_for(n_cond * 20 + n_pec * 8)), dim3(256), 0,
                           s, t2f, n_tets, cond, n_cond, tri_to_field, n_tris, pec_tris, n_pec,
                           n_field, keep, dbad);
        if ((e = hipGetLastError()) != hipSuccess) goto done;
    }
    if ((e = hipMemcpyAsync(&hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost, s)) != hipSuccess)
        goto done;
    e = hipStreamSynchronize(s);
    bad = (long long)hbad;
done:
    (void)hipFree(dbad);
    return e;
}

hipError_t dof_compact(long long n, const unsigned char* keep, long long* new_index,
                       long long* solve_ids, long long& n_solve, hipStream_t s) {
    // new_index needs n + 1 entries of room for the scan: use a temporary of n + 1
    long long* scan = nullptr;
    hipError_t e = hipMalloc((void**)&scan, sizeof(long long) * (n + 1));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(keep_to_i64_kernel, dim3(grid_for(n + 1)), dim3(256), 0, s, keep, n, scan);
    if ((e = hipGetLastError()) != hipSuccess) goto done;
    if ((e = exclusive_scan_i64(scan, scan, n + 1, s)) != hipSuccess) goto done;
    if (n > 0) {
        hipLaunchKernelGGL(compact_kernel, dim3(grid_for(n)), dim3(256), 0, s, keep, n, scan,
                           solve_ids);
        if ((e = hipGetLastError()) != hipSuccess) goto done;
    }
    if ((e = hipMemcpyAsync(&n_solve, scan + n, sizeof(long long), hipMemcpyDeviceToHost, s)) !=
        hipSuccess)
        goto done;
    if (n > 0 &&
        (e = hipMemcpyAsync(new_index, scan, sizeof(long long) * n, hipMemcpyDeviceToDevice, s)) !=
            hipSuccess)
        goto done;
    e = hipStreamSynchronize(s);
done:
    (void)hipFree(scan);
    return e;
}

hipError_t csc_select_count(long long n, const long long* indptr, const long long* indices,
                            const void* data, int complex64, int drop_zeros,
                            const long long* new_index, long long n_solve, long long* sel_indptr,
                            long long& nnz_sel, bool& bad_rows, hipStream_t s) {
    unsigned long long* dbad = nullptr;
    unsigned long long hbad = 0;
    hipError_t e = hipMalloc((void**)&dbad, sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    if ((e = hipMemsetAsync(dbad, 0, sizeof(unsigned long long), s)) != hipSuccess) goto done;
    if ((e = hipMemsetAsync(sel_indptr, 0, sizeof(long long) * (n_solve + 1), s)) != hipSuccess)
        goto done;
    if (n > 0) {
        hipLaunchKernelGGL(select_count_kernel, dim3(col_grid(n)), dim3(64 * kWavesPerBlock), 0, s,
                           n, indptr, indices, data, complex64, drop_zeros, new_index, sel_indptr,
                           dbad);
        if ((e = hipGetLastError()) != hipSuccess) goto done;
    }
    if ((e = exclusive_scan_i64(sel_indptr, sel_indptr, n_solve + 1, s)) != hipSuccess) goto done;
    if ((e = hipMemcpyAsync(&nnz_sel, sel_indptr + n_solve, sizeof(long long),
                            hipMemcpyDeviceToHost, s)) != hipSuccess)
        goto done;
    if ((e = hipMemcpyAsync(&hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost, s)) != hipSuccess)
        goto done;
    e = hipStreamSynchronize(s);
    bad_rows = hbad != 0;
done:
    (void)hipFree(dbad);
    return e;
}

hipError_t csc_select_fill(long long n, const long long* indptr, const long long* indices,
                           const void* data, int complex64, int drop_zeros,
                           const long long* new_index, const long long* sel_indptr,
                           long long* sel_indices, void* sel_data, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (complex64)
        hipLaunchKernelGGL(select_fill_kernel<float2>, dim3(col_grid(n)),
                           dim3(64 * kWavesPerBlock), 0, s, n, indptr, indices,
                           (const float2*)data, drop_zeros, new_index, sel_indptr, sel_indices,
                           (float2*)sel_data);
    else
        hipLaunchKernelGGL(select_fill_kernel<double2>, dim3(col_grid(n)),
                           dim3(64 * kWavesPerBlock), 0, s, n, indptr, indices,
                           (const double2*)data, drop_zeros, new_index, sel_indptr, sel_indices,
                           (double2*)sel_data);
    return hipGetLastError();
}

hipError_t csc_restrict(const CscBuild& full, long long n, const long long* new_index,
                        long long n_solve, CscBuild& sub, long long& n_perm, hipStream_t s) {
    long long* colperm = nullptr;
    long long* newq = nullptr;
    const bool inc_tables = full.inc && full.pos;
    long long n_inc = 0;
    sub.ndof = full.ndof;
    hipError_t e;
#define TRY(x)                          \
    do {                                \
        e = (x);                        \
        if (e != hipSuccess) goto fail; \
    } while (0)
    TRY(hipMalloc((void**)&sub.indptr, sizeof(long long) * (n_solve + 1)));
    TRY(hipMalloc((void**)&colperm, sizeof(long long) * (n_solve + 1)));
    TRY(hipMemsetAsync(sub.indptr, 0, sizeof(long long) * (n_solve + 1), s));
    TRY(hipMemsetAsync(colperm, 0, sizeof(long long) * (n_solve + 1), s));
    if (n > 0) {
        hipLaunchKernelGGL(restrict_count_kernel, dim3(col_grid(n)), dim3(64 * kWavesPerBlock), 0,
                           s, n, full.indptr, full.indices, full.seg, new_index, sub.indptr,
                           colperm);
        TRY(hipGetLastError());
    }
    TRY(exclusive_scan_i64(sub.indptr, sub.indptr, n_solve + 1, s));
    TRY(exclusive_scan_i64(colperm, colperm, n_solve + 1, s));
    TRY(hipMemcpyAsync(&sub.nnz, sub.indptr + n_solve, sizeof(long long), hipMemcpyDeviceToHost, s));
    TRY(hipMemcpyAsync(&n_perm, colperm + n_solve, sizeof(long long), hipMemcpyDeviceToHost, s));
    TRY(hipStreamSynchronize(s));
    TRY(hipMalloc((void**)&sub.indices, sizeof(long long) * (sub.nnz > 0 ? sub.nnz : 1)));
    TRY(hipMalloc((void**)&sub.seg, sizeof(long long) * (sub.nnz + 1)));
    TRY(hipMalloc((void**)&sub.perm, sizeof(long long) * (n_perm > 0 ? n_perm : 1)));
    TRY(hipMemcpyAsync(sub.seg + sub.nnz, &n_perm, sizeof(long long), hipMemcpyHostToDevice, s));
    TRY(hipMalloc((void**)&newq, sizeof(long long) * (full.nnz > 0 ? full.nnz : 1)));
    if (n > 0) {
        hipLaunchKernelGGL(restrict_fill_kernel, dim3(col_grid(n)), dim3(64 * kWavesPerBlock), 0,
                           s, n, full.indptr, full.indices, full.seg, full.perm, new_index,
                           sub.indptr, colperm, sub.indices, sub.seg, sub.perm, newq);
        TRY(hipGetLastError());
    }
    if (inc_tables) {
        // incidence-gather tables: the full map's incidence lists, row positions within the
        // block's columns, and the full column of each block column
        TRY(hipMemcpyAsync(&n_inc, full.inc_ptr + n, sizeof(long long), hipMemcpyDeviceToHost, s));
        TRY(hipStreamSynchronize(s));
        TRY(hipMalloc((void**)&sub.inc_ptr, sizeof(long long) * (n + 1)));
        TRY(hipMalloc((void**)&sub.inc, sizeof(int) * (n_inc > 0 ? n_inc : 1)));
        TRY(hipMalloc((void**)&sub.pos, sizeof(unsigned short) * (n_inc > 0 ? 20 * n_inc : 1)));
        TRY(hipMalloc((void**)&sub.cols, sizeof(long long) * (n_solve > 0 ? n_solve : 1)));
        TRY(hipMemcpyAsync(sub.inc_ptr, full.inc_ptr, sizeof(long long) * (n + 1),
                           hipMemcpyDeviceToDevice, s));
        if (n_inc > 0) {
            TRY(hipMemcpyAsync(sub.inc, full.inc, sizeof(int) * n_inc, hipMemcpyDeviceToDevice, s));
            TRY(hipMemsetAsync(sub.pos, 0xff, sizeof(unsigned short) * 20 * n_inc, s));
        }
        if (n > 0) {
            hipLaunchKernelGGL(restrict_pos_kernel, dim3(col_grid(n)), dim3(64 * kWavesPerBlock),
                               0, s, n, full.indptr, full.inc_ptr, full.pos, new_index, newq,
                               sub.indptr, sub.pos, sub.cols);
            TRY(hipGetLastError());
        }
    }
    TRY(hipStreamSynchronize(s));
    (void)hipFree(colperm);
    (void)hipFree(newq);
    return hipSuccess;
fail:
    (void)hipStreamSynchronize(s);
    (void)hipFree(colperm);
    (void)hipFree(newq);
    (void)hipFree(sub.indptr);
    (void)hipFree(sub.indices);
    (void)hipFree(sub.seg);
    (void)hipFree(sub.perm);
    (void)hipFree(sub.inc_ptr);
    (void)hipFree(sub.inc);
    (void)hipFree(sub.pos);
    (void)hipFree(sub.cols);
    sub = CscBuild();
    sub.ndof = full.ndof;
    return e;
#undef TRY
}

hipError_t dof_move(const long long* ids, long long n_ids, long long ncols, const void* src,
                    void* dst, int complex64, int scatter, hipStream_t s) {
    if (n_ids * ncols <= 0) return hipSuccess;
    const unsigned g = grid_for(n_ids * ncols);
    if (complex64) {
        if (scatter)
            hipLaunchKernelGGL(dof_scatter_kernel<float2>, dim3(g), dim3(256), 0, s, ids, n_ids,
                               ncols, (const float2*)src, (float2*)dst);
        else
            hipLaunchKernelGGL(dof_gather_kernel<float2>, dim3(g), dim3(256), 0, s, ids, n_ids,
                               ncols, (const float2*)src, (float2*)dst);
    } else {
        if (scatter)
            hipLaunchKernelGGL(dof_scatter_kernel<double2>, dim3(g), dim3(256), 0, s, ids, n_ids,
                               ncols, (const double2*)src, (double2*)dst);
        else
            hipLaunchKernelGGL(dof_gather_kernel<double2>, dim3(g), dim3(256), 0, s, ids, n_ids,
                               ncols, (const double2*)src, (double2*)dst);
    }
    return hipGetLastError();
}

}  // namespace em
