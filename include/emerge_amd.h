/*
 * emerge_amd -- MI355X (gfx950) Nedelec-2 tetrahedral curl-curl system assembly.
 *
 * C-ABI of libemerge_amd.so. Plain pointers and sizes; no C++ types cross it. Every
 * function returns an int status (EM_OK = 0, negative on error); em_last_error()
 * returns a thread-local message for the last failure on the calling thread.
 *
 * Reference interfaces replaced (paths under src/emerge/_emerge/ of EMerge 2.8.0):
 *   physics/microwave/assembly/curlcurl.py:387-456  _matrix_builder(nodes, tets, tris,
 *       edges, tet_to_field, tet_to_edge, tet_to_tri, ur, er, tetids, dofcodes)
 *       -> (dataE, dataB, rows, cols)
 *   physics/microwave/assembly/curlcurl.py:222-270  tet_mass_stiffness_matrices
 *   mth/csc_cast.py:51-53, 65-160                   CSCMapping.from_rowcol /
 *                                                   precompute_csc_pattern
 *   mth/csc_cast.py:55-63, 162-177                  CSCMapping.to_csc / scatter_to_csc_c16
 *   physics/microwave/assembly/assembler.py:494      to_csc(Evec - Bvec * K0**2)
 *
 * Array layouts on input follow the reference's numpy arrays (C order, int64 ids):
 *   nodes (3, Nn) f8, tets (4, T), tris (3, F), edges (2, E), tet_to_field (20, T),
 *   tet_to_edge (6, T), tet_to_tri (4, T), tetids (Ta,), dofcodes (20,).
 * Host pointers are used for the mesh tables given to em_ned2_plan_create (copied to
 * the device once) and for the k0 array of em_ned2_assemble_a (its values travel as
 * kernel arguments). Material tensors and every output are DEVICE pointers; launches are
 * asynchronous on the given HIP stream (hipStream_t passed as void*, NULL = default
 * stream). Functions that say "Blocks" return after their device work is done.
 *
 * COO output layout (identical to curlcurl.py:449-455): the 20x20 block of assembly
 * index ia (tet tetids[ia]) occupies entries [400*ia, 400*ia + 400), row-major
 * (entry 20*ii + jj has row tet_to_field[ii], col tet_to_field[jj]).
 */
#ifndef EMERGE_AMD_H
#define EMERGE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define EM_OK 0
#define EM_EINVAL (-1)        /* bad shape, dtype, size or argument */
#define EM_EUNSUPPORTED (-2)  /* DOF set / DOF code not supported (ccbf.py:1206-1207) */
#define EM_EHIP (-3)          /* HIP runtime error */
#define EM_ENOMEM (-4)        /* device allocation failed */
#define EM_EDEGENERATE (-5)   /* zero-volume tet (reference divides by V = 0) */
#define EM_ETABLES (-6)       /* edge/face tables inconsistent with the tets */

/* ---- material tensor kinds (complex128, interleaved re/im) ---- */
#define EM_MAT_SCALAR 0  /* one value per tet: tensor = v * I */
#define EM_MAT_DIAG 1    /* 3 values per tet: diag(v0, v1, v2) */
#define EM_MAT_FULL 2    /* 9 values per tet, row-major (p, q) */

/* flags of em_material_t */
#define EM_MAT_IS_INVERSE 1 /* (mu only) data already holds mu_r^-1: skip matinv */

/*
 * A per-tet tensor field. value(t, c) = data[t * tet_stride + c * comp_stride]
 * (complex128 elements). The reference's (3, 3, T) arrays are EM_MAT_FULL with
 * tet_stride = 1, comp_stride = T. tet_stride = 0 gives one tensor for all tets.
 */
typedef struct {
    const void* data;      /* device pointer, complex128 */
    int32_t kind;          /* EM_MAT_* */
    int32_t flags;         /* EM_MAT_IS_INVERSE or 0 */
    int64_t tet_stride;    /* in complex elements, indexed by TET id (not assembly idx) */
    int64_t comp_stride;   /* in complex elements */
} em_material_t;

/* ---- precision / output type ---- */
#define EM_PREC_FP64 0   /* FP64 geometry and contraction, complex128 output */
#define EM_PREC_MIXED 1  /* FP64 geometry, FP32 contraction, complex64 output */
/*
 * The reference's own FP64 arithmetic, complex128 output (em_ned2_assemble_km and
 * em_ned2_assemble_a only): the 11-point rule of curlcurl.py:277-380 in absolute
 * coordinates, with tet_coefficients (curlcurl.py:125-213), matinv
 * (mth/optimized.py:315-357) and the compiled/ccbf.py basis functions evaluated in the
 * reference's operand order without fused multiply-adds. EM_PREC_FP64 computes the exact
 * element integrals from vertex differences instead; on thin tets far from the origin the
 * reference's cofactor sums cancel and the two differ (up to ~1e-7 relative on seeded
 * Delaunay meshes). Compute-bound and several times slower than EM_PREC_FP64.
 */
#define EM_PREC_FP64_REF 2

typedef struct em_ned2_plan_s* em_ned2_plan_t;
typedef struct em_csc_map_s* em_csc_map_t;
typedef struct em_graph_s* em_graph_t;
typedef struct em_heat_plan_s* em_heat_plan_t;

const char* em_last_error(void);
const char* em_version(void);

/*
 * Material tensors of one frequency as Assembler.assemble_freq_matrix prepares them
 * (physics/microwave/assembly/assembler.py:439-469; the same in assemble_scattering_matrix
 * :819-868 and assemble_eig_matrix :1119-1142). er, tand, cond and er_out are device complex128 (3,3,T) in C order (entry
 * (p,q) of tet t at (3p+q)T + t); er_out may be er. Computes
 *   er_out = er * (1 - 1j*tand) - 1j*cond / (omega*EPS0)      (:453, EPS0 of const.py:23)
 * with omega = 2*pi*f; conductor_tets (device, room for T) gets the ascending ids of the
 * tets with cond[0,0,t] > pec_limit or > surf_limit (:461-469; the reference passes
 * settings mw_3d_peclim = 1e8 and mw_3d_surfimplim = 1e4), *n_conductor their count, and
 * *frequency_dependent = 1 when some component of cond lies strictly between 0 and
 * pec_limit (:455-457). Comparisons order complex numbers as NumPy does (real part, then
 * imaginary part). omega must be positive (EM_EINVAL otherwise; the reference divides by
 * zero at f = 0). Blocks until done.
 */
int em_material_prep(const void* er, const void* tand, const void* cond, int64_t n_tets,
                     double omega, double pec_limit, double surf_limit, void* er_out,
                     int64_t* conductor_tets, int64_t* n_conductor,
                     int32_t* frequency_dependent, void* stream);
/*
 * em_material_prep without any device allocation, for callers that cannot allocate (code
 * between allocation-free sections of a serving loop, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same arguments, results and blocking behaviour as
 * em_material_prep; its scratch is a fixed 4 MiB device array per GPU, held until the
 * process ends, so n_tets is limited to 16,777,216 (EM_EINVAL beyond). Every call on a GPU
 * shares that array and leaves its contents behind; calls that overlapped on one GPU would
 * race on it, so the library serialises the calls of one process per device.
 */
int em_material_prep_noalloc(const void* er, const void* tand, const void* cond, int64_t n_tets,
                             double omega, double pec_limit, double surf_limit, void* er_out,
                             int64_t* conductor_tets, int64_t* n_conductor,
                             int32_t* frequency_dependent, void* stream);

/*
 * Create an assembly plan: copies the mesh tables to device `device`, validates
 * them (ids in range, edge/face tables equal to the sorted-tuple tables of the tets as
 * mesh3d.py:707-774 builds them, no zero-volume tet) and keeps them resident.
 * Replaces the per-call table arguments of _matrix_builder (curlcurl.py:387-417).
 * tetids may be NULL (all tets, in order); otherwise it lists the Ta assembled tets
 * (the conductor-free compaction of curlcurl.py:248-250). dofcodes must be the
 * SECOND_MIXED_VOLAKIS codes [64]*6 + [66]*6 + [128]*4 + [129]*4.
 */
int em_ned2_plan_create(em_ned2_plan_t* plan, int device, const double* nodes,
                        int64_t n_nodes, const int64_t* tets, int64_t n_tets,
                        const int64_t* tris, int64_t n_tris, const int64_t* edges,
                        int64_t n_edges, const int64_t* tet_to_field,
                        const int64_t* tet_to_edge, const int64_t* tet_to_tri,
                        int64_t n_field, const int64_t* tetids, int64_t n_assy,
                        const int64_t* dofcodes, int32_t ndof);
int em_ned2_plan_destroy(em_ned2_plan_t plan);
/* n_assy (Ta), n_field (N), entries per COO array (400 * Ta) */
int em_ned2_plan_sizes(em_ned2_plan_t plan, int64_t* n_assy, int64_t* n_field,
                       int64_t* nnz_coo);
/*
 * Launch shape of the assembly kernels of this plan: tets per workgroup tile (32, 64 or
 * 128; default 64). No counterpart in the reference (Numba picks its own schedule); set
 * by the project's autotuner (emerge_amd/autotune.py).
 */
int em_ned2_plan_set_tuning(em_ned2_plan_t plan, int32_t tile);
int em_ned2_plan_get_tuning(em_ned2_plan_t plan, int32_t* tile);
/*
 * The other launch-shape knob (no counterpart in the reference either): tets per window of
 * the tile order (32 .. 16,777,216; default 1,024). The assembly kernels group each
 * window's tets by vertex order into tiles and run a window's tiles on one XCD; smaller
 * windows keep the blocks written together close in memory, too small ones leave tiles
 * nearly empty. Rebuilds the tile lists on the host, uploads them and synchronises the
 * device (call it between launches, not in a captured stream). Set by the autotuner.
 */
int em_ned2_plan_set_window(em_ned2_plan_t plan, int64_t window);
int em_ned2_plan_get_window(em_ned2_plan_t plan, int64_t* window);
/*
 * A third launch-shape knob (no counterpart in the reference): prefetched launches. With
 * chunk_tets > 0 (4,096 .. 2^31) em_ned2_assemble_km / em_ned2_assemble_a split the tile
 * list into chunks of about chunk_tets tets; each chunk is one launch, right after a small
 * kernel that reads the node coordinates (first chunk), tet ids and material values its
 * tiles use, so that the assembly finds them in the memory-side cache instead of DRAM.
 * Worth it when the inputs are not already cached: meshes whose inputs exceed the cache,
 * or launches after other work. Same outputs bit for bit; 0 is one launch. Default: 2^20
 * for plans of at least 2^21 assembled tets (measured neutral warm and faster cold there),
 * 0 below. Host-side only (no device work, no synchronisation); set by the autotuner.
 */
int em_ned2_plan_set_prefetch(em_ned2_plan_t plan, int64_t chunk_tets);
int em_ned2_plan_get_prefetch(em_ned2_plan_t plan, int64_t* chunk_tets);

/*
 * Element stiffness and mass blocks, reference-style (curlcurl.py:387-456):
 * dataE = K (with mu_r^-1), dataB = M (with eps_r), complex128 [400 * Ta] each
 * (precision EM_PREC_FP64) or complex64 (EM_PREC_MIXED). Outputs must be 16-byte aligned
 * (EM_EINVAL otherwise); at 128-byte alignment every store fills whole cache lines, and a
 * view that is only 16-byte aligned writes about 1.5x slower.
 */
int em_ned2_assemble_km(em_ned2_plan_t plan, const em_material_t* er,
                        const em_material_t* ur, void* dataE, void* dataB,
                        int32_t precision, void* stream);

/*
 * Fused A_f = K - k0[f]^2 M for n_k0 wavenumbers (assembler.py:485-494 combine step,
 * batched over a frequency sweep). out holds n_k0 COO value arrays, array f starting
 * at element f * freq_stride (freq_stride >= 400 * Ta). k0 is a HOST array; its values
 * travel as kernel arguments (one launch per 128 wavenumbers), so the call allocates
 * nothing, never synchronises and can be captured into a HIP graph.
 * complex128 for EM_PREC_FP64, complex64 for EM_PREC_MIXED. Alignment as for
 * em_ned2_assemble_km.
 */
int em_ned2_assemble_a(em_ned2_plan_t plan, const em_material_t* er,
                       const em_material_t* ur, const double* k0, int32_t n_k0, void* out,
                       int64_t freq_stride, int32_t precision, void* stream);

/*
 * Backward of em_ned2_assemble_a (one k0) with respect to per-tet eps_r and mu_r^-1
 * (FP64). No counterpart in the reference, which has no gradients (its optimiser,
 * optim.py:36-41, is derivative-free); this is the reverse of the assembly it replaces.
 * grad_coo: dL/dA for the COO values, complex128 [400 * Ta], in PyTorch's convention
 * dL/dRe + i dL/dIm. Outputs (either may be NULL) are complex128, tet-major, indexed by
 * tet id: grad_eps [T * w(eps_kind)], grad_mu_inv [T * w(mu_kind)] with w = 1 (scalar:
 * derivative w.r.t. the scalar multiplying I), 3 (diagonal) or 9 (full, row-major).
 * Entries of tets outside the assembly are not written.
 */
int em_ned2_backward(em_ned2_plan_t plan, const void* grad_coo, double k0, int32_t eps_kind,
                     void* grad_eps, int32_t mu_kind, void* grad_mu_inv, void* stream);

/* The static COO pattern: rows/cols int64 [400 * Ta] (curlcurl.py:449-453). */
int em_ned2_rowcol(em_ned2_plan_t plan, int64_t* rows, int64_t* cols, void* stream);

/*
 * CSC map of a plan's COO pattern (CSCMapping.from_rowcol -> precompute_csc_pattern,
 * csc_cast.py:51-53, 65-160), built on the device from the DOF map. Rows inside a
 * column are sorted ascending with duplicates merged; duplicates are summed in COO
 * index order. Blocks until the build is done (a one-time cost per mesh).
 * Fails with EM_EUNSUPPORTED if some DOF lies in more than 3,276 tets.
 */
int em_csc_map_create(em_csc_map_t* map, em_ned2_plan_t plan, void* stream);
int em_csc_map_destroy(em_csc_map_t map);
/* N (columns), nnz (CSC entries), nnz_coo (COO entries) */
int em_csc_map_sizes(em_csc_map_t map, int64_t* n, int64_t* nnz, int64_t* nnz_coo);
/*
 * Copy the map's arrays to caller DEVICE buffers (any may be NULL):
 * indptr [N+1], indices [nnz], perm [nnz_coo] (COO index of each entry in sorted
 * order, csc_cast.py fully_sorted_perm), target_csc [nnz_coo] (CSC position of each
 * sorted entry), col_offsets [N+1] (first sorted entry of each column).
 */
int em_csc_map_export(em_csc_map_t map, int64_t* indptr, int64_t* indices, int64_t* perm,
                      int64_t* target_csc, int64_t* col_offsets, void* stream);
/*
 * CSCMapping.to_csc (csc_cast.py:55-63, 162-177): data_csc[q] = sum of the COO values
 * of entry q. complex128 in/out (EM_PREC_FP64) or complex64 in/out (EM_PREC_MIXED,
 * summed in FP64 and rounded once). Device pointers.
 */
int em_csc_to_csc(em_csc_map_t map, const void* data_coo, void* data_csc, int32_t precision,
                  void* stream);
/*
 * Fused to_csc(dataE - k0^2 * dataB) (assembler.py:494) without forming the COO sum.
 * Pointers may be NULL only for a map without entries (a plan with no assembled tets).
 */
int em_csc_to_csc_combine(em_csc_map_t map, const void* dataE, const void* dataB, double k0,
                          void* data_csc, int32_t precision, void* stream);
/*
 * Both of the above with flags (dataB NULL: to_csc of dataE). EM_CSC_BLOCK_SYMMETRIC
 * promises that every 20x20 block of the COO values is symmetric; each contributor is
 * then read from its mirrored entry, so a column's contributors are contiguous rows of
 * the blocks (far fewer cache lines per column). The assembly with scalar or diagonal
 * tensors makes entries (ii, jj) and (jj, ii) bit-identical, and the result is then the
 * same bit for bit; symmetric full tensors give blocks symmetric to FP64 rounding, and a
 * result equal to that level.
 */
#define EM_CSC_BLOCK_SYMMETRIC 1
/*
 * A map sums each column from its tet incidences (20 lanes per column, three columns per
 * wavefront, in COO order); EM_CSC_SEGMENT_GATHER selects the per-entry gather through
 * perm/seg instead. Both give the same result bit for bit; the flag exists to compare
 * them.
 */
#define EM_CSC_SEGMENT_GATHER 2
int em_csc_to_csc_ex(em_csc_map_t map, const void* dataE, const void* dataB, double k0,
                     void* data_csc, int32_t precision, int32_t flags, void* stream);
/*
 * Direct CSC assembly: the CSC values of A = K - k0^2 M (which = 0), of K (which = 1) or of
 * M (which = 2) in the pattern of `map` -- a map created from this plan, or restricted
 * from one -- without the COO arrays. It fuses tet_mass_stiffness_matrices
 * (curlcurl.py:222-270) with CSCMapping.to_csc (csc_cast.py:55-63) as
 * Assembler.assemble_freq_matrix applies them (assembler.py:485-494:
 * to_csc(Evec - Bvec * K0**2)) and assemble_eig_matrix (:1151-1152: to_csc(K), to_csc(M)).
 * Equal bit for bit to em_ned2_assemble_a (which 0) or em_ned2_assemble_km (1, 2) followed
 * by em_csc_to_csc of that COO array. precision EM_PREC_FP64 (complex128 out) or
 * EM_PREC_MIXED (complex64). The plan keeps a scratch of per-tet records (128 B per tet
 * mixed, 256 B FP64; 160 / 320 B with full tensors), allocated on the first call and
 * whenever a larger one is needed (so make one call before capturing a graph); like every
 * launch on a plan, calls from several streams at once race on it.
 */
int em_ned2_assemble_csc(em_ned2_plan_t plan, em_csc_map_t map, const em_material_t* er,
                         const em_material_t* ur, double k0, int32_t which, void* data_csc,
                         int32_t precision, void* stream);
/*
 * to_csc of blocks that are not symmetric (full anisotropic tensors) without the plain
 * gather's strided column reads: the contributions dataE - k0^2 dataB (dataB may be NULL)
 * are first written block-transposed in FP64 to the caller's workspace work (complex128,
 * nnz_coo entries of the map; em_csc_map_sizes), then summed by the mirrored column
 * gather. The result equals em_csc_to_csc_ex's bit for bit. No counterpart in the
 * reference (same operation as CSCMapping.to_csc, csc_cast.py:55-63, 162-177).
 */
int em_csc_to_csc_staged(em_csc_map_t map, const void* dataE, const void* dataB, double k0,
                         void* work, void* data_csc, int32_t precision, void* stream);

/*
 * A_f = K - k0[f]^2 M for CSC values K and M of one pattern (n entries each, e.g. the
 * em_csc_to_csc results of one map): the per-frequency to_csc(Evec - Bvec * K0**2) of
 * assembler.py:494 with K and M summed once per mesh instead of once per frequency (each
 * entry is summed in another order: sum(E) - k0^2 sum(B); equal up to rounding). out holds
 * n_k0 value arrays, array f starting at element f * out_stride (out_stride >= n, even for
 * complex64). k0 is a HOST array; its values travel as kernel arguments (one launch per 128
 * wavenumbers), so the call allocates nothing, never synchronises and can be captured into
 * a HIP graph. complex128 (EM_PREC_FP64) or complex64 (EM_PREC_MIXED: FP64 arithmetic,
 * rounded once); all pointers 16-byte aligned; runs on the current device.
 */
int em_csc_freq_combine(int64_t n, const void* K, const void* M, const double* k0, int32_t n_k0,
                        void* out, int64_t out_stride, int32_t precision, void* stream);
/*
 * The same with a lossy medium's conductivity term: A_f = K - k0_f^2 M + j g_f C for the
 * CSC values K, M, C of one pattern and host arrays k0, g of n_k0 values. With M the mass
 * values of er (1 - j tand), C those of cond and g_f = k0_f^2 / (w_f EPS0) this is
 * to_csc(Evec - Bvec * K0**2) of assemble_freq_matrix (assembler.py:453, :494) for
 * frequency-independent er, tand, cond and ur: er' = er (1 - j tand) - j cond / (w EPS0)
 * is affine in 1/w, so the reference's per-frequency reassembly of lossy media becomes
 * one streaming pass per frequency (the same matrix up to rounding). No counterpart
 * function in the reference.
 */
int em_csc_freq_combine_lossy(int64_t n, const void* K, const void* M, const void* C,
                              const double* k0, const double* g, int32_t n_k0, void* out,
                              int64_t out_stride, int32_t precision, void* stream);

/*
 * PEC elimination and the solver hand-off. All pointers are DEVICE pointers; functions
 * that take no plan run on the calling thread's current HIP device.
 *
 * em_ned2_pec_mask: the prescribed DOFs of Assembler.assemble_freq_matrix
 * (assembler.py:533-564): keep[d] = 0 for the 20 DOFs of every conductor tet (plan tet
 * ids) and for the DOFs of every PEC face: the face's column of the field's tri_to_field
 * table (int64 [8][n_tris], nedelec2.py:70-79: its 3 edges' 2 DOFs each, then its 2 face
 * DOFs; the edge DOFs are the reference's edge_to_field of the face's edges), keep[d] = 1
 * otherwise. Out-of-range ids give EM_EINVAL. Blocks.
 */
int em_ned2_pec_mask(em_ned2_plan_t plan, const int64_t* conductor_tets, int64_t n_cond,
                     const int64_t* tri_to_field, int64_t n_tris, const int64_t* pec_tris,
                     int64_t n_pec, uint8_t* keep, void* stream);
/*
 * solve_ids = the kept DOFs ascending (assembler.py:742); new_index[d] = position of d in
 * solve_ids, or -1. solve_ids needs room for n entries; *n_solve gets the count. Blocks.
 */
int em_dof_compact(int64_t n, const uint8_t* keep, int64_t* new_index, int64_t* solve_ids,
                   int64_t* n_solve, void* stream);
/*
 * The free-DOF block A[:, solve_ids][solve_ids, :] of an n x n CSC matrix with sorted row
 * indices (solver.py:2071; solver.py:2229-2230 np.ix_ is the same block). Two calls:
 * _count writes sel_indptr [n_solve+1] and *nnz_sel (blocks); _fill writes sel_indices
 * [nnz_sel] and, when sel_data is not NULL, sel_data (complex128 or complex64 per
 * precision). drop_zeros also drops explicit zeros (K.eliminate_zeros(), assembler.py:760),
 * which needs data in both calls.
 */
int em_csc_select_count(int64_t n, const int64_t* indptr, const int64_t* indices,
                        const void* data, int32_t precision, int32_t drop_zeros,
                        const int64_t* new_index, int64_t n_solve, int64_t* sel_indptr,
                        int64_t* nnz_sel, void* stream);
int em_csc_select_fill(int64_t n, const int64_t* indptr, const int64_t* indices,
                       const void* data, int32_t precision, int32_t drop_zeros,
                       const int64_t* new_index, const int64_t* sel_indptr, int64_t* sel_indices,
                       void* sel_data, void* stream);
/*
 * A CSC map of the free-DOF block: em_csc_to_csc[_combine] on *sub read the same COO value
 * arrays as on `map` and write the block's values directly (the reference's
 * to_csc(Evec - Bvec K0^2) followed by A[:, solve_ids][solve_ids, :], fused). em_csc_map_sizes
 * gives n = n_solve and the block's nnz; em_csc_map_export gives its indptr and indices
 * only. Destroy with em_csc_map_destroy. Blocks.
 */
int em_csc_map_restrict(em_csc_map_t map, const int64_t* new_index, int64_t n_solve,
                        em_csc_map_t* sub, void* stream);
/*
 * Row gather dst[i, :] = src[ids[i], :] (bsel = b[solve_ids, :], solver.py:2072) and
 * scatter dst[ids[i], :] = src[i, :] (solution[solve_ids, :] = x, solver.py:2120, 2240-2241); row-major
 * [rows][ncols], complex128 or complex64 per precision. Rows of dst not named are untouched.
 */
int em_dof_gather(const int64_t* ids, int64_t n_ids, int64_t ncols, const void* src, void* dst,
                  int32_t precision, void* stream);
int em_dof_scatter(const int64_t* ids, int64_t n_ids, int64_t ncols, const void* src, void* dst,
                   int32_t precision, void* stream);

/*
 * Robin / port surface assembly on boundary triangles (physics/microwave/assembly/
 * robinbc.py). The plan supplies the node coordinates; tris is the mesh's face table
 * (int64 [3][n_tris], sorted node ids per face, mesh3d.py:707-732) and tri_to_field the
 * field's face DOF table (int64 [8][n_tris], nedelec2.py:70-79), both DEVICE pointers,
 * as are surf (face ids) and every array below. The first three block (they check ids).
 *
 * em_ned2_robin_blocks: ned2_tri_stiff + compute_bc_entries (robinbc.py:316-402):
 * the 8x8 block gamma * int f_i . f_j dA of every listed face (row-major, complex128),
 * ADDED to out at face * 64 (EM_TRI_BY_FACE: out is the reference's Bmat of n_tris * 64)
 * or at list position * 64 (EM_TRI_COMPACT: n_surf * 64). A face listed twice in one call
 * races, as in the reference's prange loop.
 */
#define EM_TRI_BY_FACE 0
#define EM_TRI_COMPACT 1
int em_ned2_robin_blocks(em_ned2_plan_t plan, const int64_t* tris, int64_t n_tris,
                         const int64_t* surf, int64_t n_surf, double gamma_re, double gamma_im,
                         void* out, int32_t layout, void* stream);
/*
 * ned2_tri_force + compute_force_entries (robinbc.py:220-313): bvec[tri_to_field[i, f]] +=
 * -A sum_w w_w f_i(x_w) . U(x_w) with the reference's 6-point rule; U is the caller's mode
 * field at the rule's points, complex128 [3][6][n_surf] (U_global_all, the points of
 * generate_points_3d, robinbc.py:111-135). With U_curl (same layout) and normals (double
 * [3][n_surf]) it is the scattered-field form ned2_tri_force_scat (robinbc.py:453-569).
 * bvec: complex128 [n_field], accumulated atomically.
 */
int em_ned2_robin_force(em_ned2_plan_t plan, const int64_t* tris, int64_t n_tris,
                        const int64_t* tri_to_field, const int64_t* surf, int64_t n_surf,
                        const void* U, const void* U_curl, const double* normals, void* bvec,
                        void* stream);
/* rows/cols (int64 [n_surf * 64]) of the listed faces' blocks (femdata.py:75-94). */
int em_tri_rowcol(const int64_t* tri_to_field, int64_t n_tris, const int64_t* surf,
                  int64_t n_surf, int64_t* rows, int64_t* cols, void* stream);
/*
 * K += generate_csc(B) (assembler.py:635; femdata.py:104-119) on the device: the CSC
 * position of each (rows[i], cols[i]) in the map's pattern, -1 if absent (new_index, when
 * not NULL, renumbers rows/cols first: for a restricted map, removed DOFs give -1); then
 * data_csc[pos[i]] += scale * values[i] (values complex128; data complex128 or complex64
 * per precision; atomic, entries of pos -1 skipped).
 */
int em_csc_map_find(em_csc_map_t map, const int64_t* rows, const int64_t* cols, int64_t n,
                    const int64_t* new_index, int64_t* pos, void* stream);
int em_csc_add(void* data_csc, int32_t precision, const int64_t* pos, const void* values,
               int64_t n, double scale_re, double scale_im, void* stream);

/*
 * Point location (compiled/base/index_inter.py:37-298, index_interp): for every point of
 * coords (DEVICE double [3][n]) the largest id among the candidate tets (tetids, DEVICE
 * int64, or all plan tets when NULL) whose barycentric coordinates at the point are all
 * >= -1e-8 and sum to <= 1 + 1e-8; -1 when none. Blocks.
 */
int em_ned2_locate(em_ned2_plan_t plan, const double* coords, int64_t n, const int64_t* tetids,
                   int64_t n_tetids, int64_t* tet_of, void* stream);
/*
 * Interpolation of a solution (complex128 [n_field]) at points with known tets
 * (compiled/base/interp.py:286-387 ned2_tet_interp; with EM_INTERP_CURL :410-519
 * ned2_tet_interp_curl, the curl times c_tet[tet] when c_tet (complex128 per plan tet) is
 * not NULL). out: complex128 [3][n]; NaN where tet_of is -1.
 */
#define EM_INTERP_E 0
#define EM_INTERP_CURL 1
int em_ned2_interp(em_ned2_plan_t plan, const double* coords, int64_t n, const int64_t* tet_of,
                   const void* solution, int32_t which, const void* c_tet, void* out,
                   void* stream);

/*
 * Heat conduction, second-order Lagrange elements
 * (physics/heatconduction/assembly/grad.py). A heat plan holds the node coordinates, the
 * tets and the Legrange2 DOF map (tet_to_field int64 [10][T], elements/leg2.py:43-50:
 * the 4 node DOFs then the 6 edge DOFs), copied to the device and checked once (ids in
 * range; a zero-volume tet gives EM_EDEGENERATE).
 */
int em_heat_plan_create(em_heat_plan_t* plan, int device, const double* nodes, int64_t n_nodes,
                        const int64_t* tets, int64_t n_tets, const int64_t* tet_to_field,
                        int64_t n_field);
int em_heat_plan_destroy(em_heat_plan_t plan);
/*
 * _stiff_matrix_builder / _mass_matrix_builder (grad.py:346-414) values: K = conductance
 * blocks (leg2_tet_stiff, grad.py:199-261) from the conductivity cond (DEVICE float64,
 * value (t, p*3+q) at cond[t * cond_tet_stride + (p*3+q) * cond_comp_stride] for
 * EM_MAT_FULL; EM_MAT_DIAG reads 3 components, EM_MAT_SCALAR 1; the reference's (3,3,T)
 * array is EM_MAT_FULL with strides 1, T) and M = capacity blocks (leg2_tet_mass,
 * grad.py:265-338) from rho_cp[t * rho_tet_stride]; float64 [100 T] each, row-major 10x10
 * per tet; either may be NULL. The reference's 11-point rule (its 8-digit table) is
 * reproduced exactly, including its |V| normalisation on negatively oriented tets.
 */
int em_heat_assemble(em_heat_plan_t plan, const double* cond, int32_t cond_kind,
                     int64_t cond_tet_stride, int64_t cond_comp_stride, const double* rho_cp,
                     int64_t rho_tet_stride, double* K, double* M, void* stream);
/* rows/cols int64 [100 T] of the blocks (grad.py:366-370). */
int em_heat_rowcol(em_heat_plan_t plan, int64_t* rows, int64_t* cols, void* stream);
/* CSCMapping.from_rowcol of the heat pattern (csc_cast.py:51-53, 65-160). Blocks.
 * Fails with EM_EUNSUPPORTED if some DOF lies in more than 6,553 tets. */
int em_csc_map_create_heat(em_csc_map_t* map, em_heat_plan_t plan, void* stream);
/*
 * CSCMapping.to_csc for float64 data (scatter_to_csc_f8, csc_cast.py:180-194):
 * data_csc[q] = sum of data[k] (+ scale * data_b[k] when data_b is not NULL, e.g. a
 * transient K + M / dt) over the entry's COO contributors, in COO order. flags: 0 (the
 * column-incidence gather on heat maps), EM_CSC_BLOCK_SYMMETRIC (the same, reading each
 * entry from its mirror: valid when every 10x10 block is bit-symmetric, as the capacity
 * blocks and the conductance blocks of a scalar or diagonal conductivity are) or
 * EM_CSC_SEGMENT_GATHER (the per-entry gather); all give the same result bit for bit.
 */
int em_csc_to_csc_real(em_csc_map_t map, const double* data, const double* data_b, double scale,
                       double* data_csc, int flags, void* stream);

/*
 * HIP graphs of assembly launches. The reference re-runs its Numba builder per call
 * (microwave_3d.py:1457 sweep loop) and has no counterpart: these record the launches
 * issued on `stream` between begin and end (hipStreamBeginCapture, thread-local mode)
 * and replay them with one graph launch. Only calls that allocate nothing and never
 * synchronise may be captured: em_ned2_assemble_km/_a, em_ned2_rowcol,
 * em_csc_to_csc[_combine]. `stream` must not be the default stream.
 */
int em_graph_begin_capture(void* stream);
int em_graph_end_capture(void* stream, em_graph_t* graph);
int em_graph_launch(em_graph_t graph, void* stream);
int em_graph_destroy(em_graph_t graph);

#ifdef __cplusplus
}
#endif

#endif /* EMERGE_AMD_H */
