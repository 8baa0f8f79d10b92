// Internal declarations shared by the HIP sources of libemerge_amd (not part of the
// public C-ABI, which is include/emerge_amd.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/emerge_amd.h"

namespace em {

// Device view of a per-tet material tensor field, complex128 interleaved.
// value(t, component) = p[t * ts + component * cs]; ts = 0 makes it uniform.
struct MatDesc {
    const double2* p;
    int kind;    // EM_MAT_SCALAR / EM_MAT_DIAG / EM_MAT_FULL
    int invert;  // 1: the stored tensor is mu_r and must be inverted (reference matinv)
    long long ts;
    long long cs;
};

// Mesh data resident on the device for one plan.
struct MeshDev {
    const double4* xyz;     // [Nn] node coordinates, padded to 32 B
    const int4* tets;       // [T] node ids of each tet, stored vertex order
    const int* t2f;         // [T][20] DOF map (tet_to_field transposed, int32)
    const int* tetids;      // [Ta] assembled tet of each assembly index, or nullptr
    long long n_assy;
};

enum OutMode { MODE_A = 0, MODE_KM = 1, MODE_AF = 2 };

// Hand-off of LDS data between the lanes of one wavefront: every lane's earlier LDS
// accesses are complete and visible before any lane's later ones (wavefront-scope
// release and acquire fences around a wave barrier; no workgroup barrier needed).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavenumbers of one frequency-sweep launch, passed by value as a kernel argument.
constexpr int kMaxK0 = 128;
struct K0Pack {
    double v[kMaxK0];  // k0^2
};

// Tets grouped by vertex-rank code inside windows of consecutive tets: order[] lists
// assembly indices window by window, bucket by bucket (ascending within a bucket);
// tiles[] = (first position in order, count | code << 16) with at most `tile` tets each,
// interleaved so that all tiles of a window land on one XCD; count 0 marks a padding tile.
// A slice of a tile list that is launched right after a prefetch kernel has read the
// inputs its tiles use: tiles [tile_begin, tile_begin + tile_count), assembly indices (and
// order positions) [ia_lo, ia_hi), tet ids [t_lo, t_hi).
struct PrefetchChunk {
    long long tile_begin = 0, tile_count = 0;
    long long ia_lo = 0, ia_hi = 0;
    long long t_lo = 0, t_hi = 0;
};

struct TileSet {
    const int* order = nullptr;
    const int2* tiles = nullptr;
    long long ntiles = 0;
    // Host array of nchunks slices covering the list in order, each launched after a
    // prefetch of its inputs; nullptr: one launch over the whole list, no prefetch.
    const PrefetchChunk* chunks = nullptr;
    int nchunks = 0;
    long long n_nodes = 0;             // node coordinates, read by the first prefetch
    unsigned long long* sink = nullptr;  // device word for the prefetch's never-taken store
};
constexpr int kTileSizes[3] = {32, 64, 128};

// Launchers (defined in ned2_assemble.hip). All return hipError_t of the launch.
// ts: the three TileSets for tile sizes 32, 64, 128; tile selects one.
// k0sq: host array of nf values of k0^2 (MODE_A uses k0sq[0]; MODE_KM ignores it).
// full: material tensors may be non-diagonal (general 3x3 path with antisymmetric terms).
hipError_t launch_assemble(const MeshDev& m, const TileSet* ts, int tile, const MatDesc& eps,
                           const MatDesc& mu, int mode, int precision, int full,
                           const double* k0sq, int nf, void* out0, void* out1,
                           long long freq_stride, hipStream_t stream);
// The reference's own arithmetic (ned2_ref.hip, EM_PREC_FP64_REF): 11-point rule in
// absolute coordinates, complex128 out; modes as launch_assemble.
hipError_t launch_assemble_ref(const MeshDev& m, const MatDesc& eps, const MatDesc& mu, int mode,
                               const double* k0sq, int nf, void* out0, void* out1,
                               long long freq_stride, hipStream_t stream);
hipError_t launch_rowcol(const MeshDev& m, long long* rows, long long* cols, hipStream_t s);
hipError_t launch_check_tables(const int4* tets, long long T, const long long* edges,
                               long long E, const long long* tris, long long F,
                               const long long* t2e, const long long* t2t,
                               unsigned long long* bad, hipStream_t s);
hipError_t launch_check_volumes(const double4* xyz, const int4* tets, long long T,
                                unsigned long long* bad, hipStream_t s);

// Backward of the fused assembly w.r.t. eps_r and mu_r^-1 (ned2_backward.hip).
hipError_t launch_backward(const MeshDev& m, const double2* grad, double k0sq, int eps_kind,
                           int mu_kind, double2* g_eps, double2* g_mu, hipStream_t stream);

// CSC map (csc_map.hip)
struct CscBuild {
    long long* indptr = nullptr;   // [N+1]
    long long* indices = nullptr;  // [nnz]
    long long* perm = nullptr;     // [nnz_coo]
    long long* seg = nullptr;      // [nnz+1]
    long long* inc_ptr = nullptr;  // [N+1]
    int* inc = nullptr;             // [20 Ta] incidences (20 ia + slot) of each DOF, ascending
    unsigned short* pos = nullptr;  // [400 Ta] entry index of row ii within the column
    long long* cols = nullptr;      // restricted maps: full-matrix column of each column
    long long nnz = 0;
    int ndof = 20;                  // DOFs per tet of the COO blocks (20 Nedelec, 10 Lagrange)
};
// Largest CSC column the map build takes, in COO entries (DOFs per tet x tets of the
// column's DOF; the row positions inside a column are uint16): 3,276 Nedelec-2 or 6,553
// Lagrange-2 tets. Longer columns fail the build with hipErrorInvalidValue and max_inc set.
constexpr long long kCscMaxColumnKeys = 65535;
hipError_t csc_build(const MeshDev& m, long long N, CscBuild& b, long long& max_inc,
                     hipStream_t s);
// mirror: the COO blocks are symmetric; read each contributor from its mirrored entry.
// Pattern of a 10-DOF-per-tet map (Lagrange-2: t2f [T][10], every tet assembled).
hipError_t csc_build_ndof10(const int* t2f, long long n_tets, long long N, CscBuild& b,
                            long long& max_inc, hipStream_t s);
// out[q] = sum of E[k] (+ scale * B[k]) over the contributors of entry q (real data).
// Direct CSC assembly (ned2_assemble.hip): per-tet records into `rec` (at least
// assemble_csc_record_bytes), then the CSC values of A (which 0, G_M scaled by -k0sq),
// K (1) or M (2) for the columns of map b.
size_t assemble_csc_record_bytes(long long n_assy, int precision, int full);
hipError_t launch_assemble_csc(const MeshDev& m, const CscBuild& b, long long N,
                               const MatDesc& eps, const MatDesc& mu, int which, int precision,
                               int full, double k0sq, void* rec, void* out, hipStream_t s);
// Real-valued column-incidence gather (10-DOF maps; needs b.inc and b.pos).
hipError_t csc_gather_inc_real(const CscBuild& b, long long N, const double* E, const double* B,
                               double scale, bool mirror, double* out, hipStream_t s);
hipError_t csc_gather_real(const long long* seg, const long long* perm, long long nnz,
                           const double* E, const double* B, double scale, double* out,
                           hipStream_t s);
hipError_t csc_gather(const long long* seg, const long long* perm, long long nnz, const void* E,
                      const void* B, double k0sq, void* out, int complex64, int mirror,
                      hipStream_t s);
hipError_t csc_target(const long long* seg, long long nnz, long long* target, hipStream_t s);
hipError_t csc_col_offsets(const long long* inc_ptr, long long N, long long ndof, long long* out,
                           hipStream_t s);
// Heat (Lagrange-2) assembly (heat.hip): K and/or M [100 T] of every tet; cond indexed
// cond[t * cts + comp * ccs] with EM_MAT_* kind, rho[t * rts].
hipError_t heat_assemble(const double4* xyz, const int4* tets, long long T, const double* cond,
                         int cond_kind, long long cts, long long ccs, const double* rho,
                         long long rts, double* K, double* M, hipStream_t s);
hipError_t heat_rowcol(const int* t2f, long long T, long long* rows, long long* cols,
                       hipStream_t s);
// to_csc by column incidences (full maps: inc and pos present); same result as csc_gather.
// Plain (non-symmetric) incidence gather through a block-transposed FP64 staging copy
// (work: complex128 [400 n_blk]); the same result bit for bit.
hipError_t csc_gather_inc_staged(const CscBuild& b, long long N, long long n_blk, const void* E,
                                 const void* B, double k0sq, double2* work, void* out,
                                 int complex64, hipStream_t s);
hipError_t csc_gather_inc(const CscBuild& b, long long N, const void* E, const void* B,
                          double k0sq, void* out, int complex64, int mirror, hipStream_t s);
// out[f * stride + i] = K[i] - k0sq[f] M[i] (+ j g[f] C[i] when C is given) over n CSC
// values, nf frequencies (csc_freq.hip); complex128 or complex64 (FP64 arithmetic).
hipError_t csc_freq_combine(long long n, const void* K, const void* M, const void* C,
                            const double* k0sq, const double* g, int nf, void* out,
                            long long stride, int complex64, hipStream_t s);
// Exclusive scan of n int64 values (in may equal out); allocates temporaries, blocks.
hipError_t exclusive_scan_i64(const long long* in, long long* out, long long n, hipStream_t s);

// PEC elimination (dof_select.hip). keep[d] = 0 for prescribed DOFs; `bad` gets a bit set
// for out-of-range ids (1 conductor tet, 2 face, 4 DOF). Blocks.
hipError_t pec_mark(const int* t2f, long long n_tets, const long long* cond, long long n_cond,
                    const long long* tri_to_field, long long n_tris, const long long* pec_tris,
                    long long n_pec, long long n_field, unsigned char* keep, long long& bad,
                    hipStream_t s);
// er' = er(1 - j tand) - j cond/(omega eps0) over (3,3,T) complex128 tensors (out may be
// er), the ascending ids of the conductor tets (cond[0,0] above either limit) and whether
// some component lies strictly between 0 and pec_limit (materials.hip). Device scratch of
// material_prep_scratch_bytes(T) bytes (256-byte aligned; every byte the call reads it
// writes first, so it needs no clearing). Blocks until done.
size_t material_prep_scratch_bytes(long long n_tets);
// Argument checks shared by the C-ABI material preparation entry points (capi.hip):
// EM_OK or EM_EINVAL with the message set.
int material_prep_check_args(const void* er, const void* tand, const void* cond, int64_t n_tets,
                             double omega, const void* er_out, const int64_t* conductor_tets,
                             const int64_t* n_conductor, const int32_t* frequency_dependent);
hipError_t material_prep(const double2* er, const double2* tand, const double2* cond,
                         long long n_tets, double omega, double pec_limit, double surf_limit,
                         double2* out, long long* conductor_tets, void* scratch,
                         long long& n_conductor, bool& frequency_dependent, hipStream_t s);
// new_index[d] = position of d among the kept DOFs or -1; solve_ids = kept DOFs ascending.
hipError_t dof_compact(long long n, const unsigned char* keep, long long* new_index,
                       long long* solve_ids, long long& n_solve, hipStream_t s);
hipError_t csc_select_count(long long n, const long long* indptr, const long long* indices,
                            const void* data, int complex64, int drop_zeros,
                            const long long* new_index, long long n_solve, long long* sel_indptr,
                            long long& nnz_sel, bool& bad_rows, hipStream_t s);
hipError_t csc_select_fill(long long n, const long long* indptr, const long long* indices,
                           const void* data, int complex64, int drop_zeros,
                           const long long* new_index, const long long* sel_indptr,
                           long long* sel_indices, void* sel_data, hipStream_t s);
// Map of the kept rows/columns of `full` (N columns): perm/seg of the kept contributors
// (n_perm of them) and, when `full` has them, incidence-gather tables (inc_ptr/inc of
// the full columns, pos with 0xffff for removed rows, cols).
hipError_t csc_restrict(const CscBuild& full, long long n, const long long* new_index,
                        long long n_solve, CscBuild& sub, long long& n_perm, hipStream_t s);
hipError_t dof_move(const long long* ids, long long n_ids, long long ncols, const void* src,
                    void* dst, int complex64, int scatter, hipStream_t s);

// Robin / port surface assembly (robin.hip). `bad` bits: 1 face or node id out of range,
// 2 DOF out of range. The first three block until done.
hipError_t robin_blocks(const double4* xyz, long long n_nodes, const long long* tris, long long F,
                        const long long* surf, long long nsurf, double gre, double gim,
                        double2* out, int by_face, long long& bad, hipStream_t s);
hipError_t robin_force(const double4* xyz, long long n_nodes, const long long* tris, long long F,
                       const long long* tri_to_field, long long n_field, const long long* surf,
                       long long nsurf, const double2* U, const double2* Uc, const double* normals,
                       double2* bvec, long long& bad, hipStream_t s);
hipError_t tri_rowcol(const long long* tri_to_field, long long F, const long long* surf,
                      long long nsurf, long long* rows, long long* cols, long long& bad,
                      hipStream_t s);
hipError_t csc_find(const long long* indptr, const long long* indices, long long N,
                    const long long* rows, const long long* cols, const long long* new_index,
                    long long n, long long* pos, hipStream_t s);
hipError_t csc_add(void* data, int complex64, const long long* pos, const double2* vals,
                   long long n, double sre, double sim, hipStream_t s);

// Point location and field interpolation (interp.hip). locate_points: out[j] = the largest
// candidate tet id containing point j (coords [3][n]), -1 if none; bad bit 1: candidate id
// out of range. Blocks.
hipError_t locate_points(const double4* xyz, const int4* tets, long long n_tets, const double* c,
                         long long n, const long long* tetids, long long n_cand,
                         long long* out, long long& bad, hipStream_t s);
hipError_t ned2_interp(const double4* xyz, const int4* tets, const int* t2f, long long n_tets,
                       const double* c, long long n, const long long* tet_of, const double2* sol,
                       int curl, const double2* ctet, double2* out, hipStream_t s);

void set_error(const std::string& msg);

}  // namespace em
