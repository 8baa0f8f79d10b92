// C-ABI of libemerge_amd.so (declared in include/emerge_amd.h).
//
// A plan owns the device copy of one mesh: padded node coordinates, int32 tet node
// ids, the int32 DOF map and the assembly compaction. It replaces the table arguments
// that the reference passes to _matrix_builder on every call
// (src/emerge/_emerge/physics/microwave/assembly/curlcurl.py:387-417).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "em_internal.h"

namespace em {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }

}  // namespace em

using em::set_error;

struct em_ned2_plan_s {
    int device = 0;
    long long n_nodes = 0, n_tets = 0, n_assy = 0, n_field = 0;
    double4* xyz = nullptr;
    int4* tets = nullptr;
    int* t2f = nullptr;
    int* tetids = nullptr;
    int tile = 64;  // tets per workgroup tile of the assembly kernel (autotuned)
    long long window = 1024;        // tets per code-grouping window (autotuned)
    std::vector<unsigned char> code;  // vertex-rank code of every assembly index (host)
    void* csc_rec = nullptr;        // per-tet records of the direct CSC assembly
    size_t csc_rec_bytes = 0;
    int* order = nullptr;           // assembly indices grouped by vertex-rank code
    int2* tiles[3] = {nullptr, nullptr, nullptr};
    long long ntiles[3] = {0, 0, 0};
    std::vector<int2> htiles[3];    // host copies of the tile lists
    std::vector<int> htetids;       // host copy of tetids (empty: every tet assembled)
    // prefetched launches: tets per chunk (0: one launch, no prefetch) and the chunks of
    // each tile list
    long long prefetch_tets = 0;
    std::vector<em::PrefetchChunk> chunks[3];
    unsigned long long* sink = nullptr;  // the prefetch kernel's never-taken store

    void tilesets(em::TileSet ts[3]) const {
        for (int i = 0; i < 3; ++i) {
            ts[i].order = order;
            ts[i].tiles = tiles[i];
            ts[i].ntiles = ntiles[i];
            const bool pf = prefetch_tets > 0 && !chunks[i].empty();
            ts[i].chunks = pf ? chunks[i].data() : nullptr;
            ts[i].nchunks = pf ? (int)chunks[i].size() : 0;
            ts[i].n_nodes = n_nodes;
            ts[i].sink = sink;
        }
    }

    em::MeshDev dev() const {
        em::MeshDev m;
        m.xyz = xyz;
        m.tets = tets;
        m.t2f = t2f;
        m.tetids = tetids;
        m.n_assy = n_assy;
        return m;
    }
};

#define EM_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
            return e_ == hipErrorOutOfMemory ? EM_ENOMEM : EM_EHIP;                   \
        }                                                                             \
    } while (0)

namespace {

struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) == hipSuccess && hipSetDevice(dev) == hipSuccess) ok = true;
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <typename T>
int dmalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return EM_OK;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? EM_ENOMEM : EM_EHIP;
    }
    return EM_OK;
}

void free_plan(em_ned2_plan_s* p) {
    if (!p) return;
    (void)hipFree(p->xyz);
    (void)hipFree(p->tets);
    (void)hipFree(p->t2f);
    (void)hipFree(p->tetids);
    (void)hipFree(p->order);
    for (int i = 0; i < 3; ++i) (void)hipFree(p->tiles[i]);
    (void)hipFree(p->csc_rec);
    (void)hipFree(p->sink);
    delete p;
}

int check_material(const em_material_t* m, long long n_tets, const char* name) {
    if (!m || !m->data) {
        set_error(std::string(name) + ": null material");
        return EM_EINVAL;
    }
    if (m->kind < EM_MAT_SCALAR || m->kind > EM_MAT_FULL) {
        set_error(std::string(name) + ": unknown material kind");
        return EM_EINVAL;
    }
    if (m->tet_stride < 0 || m->comp_stride < 0) {
        set_error(std::string(name) + ": negative stride");
        return EM_EINVAL;
    }
    (void)n_tets;
    return EM_OK;
}

em::MatDesc to_desc(const em_material_t* m, bool is_mu) {
    em::MatDesc d;
    d.p = reinterpret_cast<const double2*>(m->data);
    d.kind = m->kind;
    d.invert = (is_mu && !(m->flags & EM_MAT_IS_INVERSE)) ? 1 : 0;
    d.ts = m->tet_stride;
    d.cs = m->comp_stride;
    return d;
}

// A non-symmetric tensor needs the antisymmetric contraction terms; the host cannot
// see device data cheaply, so FULL tensors always take that path.
int needs_anti(const em_material_t* er, const em_material_t* ur) {
    return (er->kind == EM_MAT_FULL || ur->kind == EM_MAT_FULL) ? 1 : 0;
}

}  // namespace

namespace {

// Tiles for the assembly kernels. The assembly range is cut into windows of `window`
// consecutive tets (neighbours share nodes); inside each window the tets are grouped by
// vertex-rank code (stable counting sort) and each group is cut into tiles of 32, 64 and
// 128 tets. Workgroups are dispatched round-robin over the XCDs, so the tile list is
// interleaved such that every tile of a window runs on the same XCD and its node/material
// reads hit that XCD's L2: tile j*kXcds + x is the j-th tile of XCD x, which takes the
// windows w with w % kXcds == x (shorter lists padded with empty tiles). Small windows
// keep the blocks written at the same time close together in memory, but windows so
// small that a code's group holds a few tets leave tiles nearly empty: on a 2.8M-tet mesh
// 1,024-tet windows launch 4% faster than 16,384-tet ones and 512-tet windows 6%, on an
// 8.3M-tet mesh 512-tet windows are 30% slower (tiles of ~20 tets). The autotuner picks
// the window per mesh (em_ned2_plan_set_window); plan creation uses 1,024.
//
// Giving tets to XCDs by where their output lands instead (tet -> the XCD x whose 4 KiB
// output chunks, chunk index % 8 == x, hold most of its block) was tried: stores alone
// run 6% faster when each XCD writes chunks of one residue, but the kernel gained nothing
// (2.8M tets, complex64: within +-1%, also with uniform materials): a stores-only replica
// of its tile walk is as fast with 80% of the bytes on the matching XCD as without, and
// the 8 tets sharing a 128-byte line of the per-tet inputs then run on ~6 XCDs (1,125 MB
// fetched per launch instead of 300 MB).
void build_tiles(const std::vector<unsigned char>& code, long long window, std::vector<int>& horder,
                 std::vector<std::vector<int2>>& htiles) {
    constexpr int kXcds = 8;
    const long long n_assy = (long long)code.size();
    horder.assign((size_t)n_assy, 0);
    htiles.assign(3, std::vector<int2>());
    const long long nwin = (n_assy + window - 1) / window;
    std::vector<std::vector<std::vector<int2>>> per_xcd(3, std::vector<std::vector<int2>>(kXcds));
    for (long long w = 0; w < nwin; ++w) {
        const long long w0 = w * window, w1 = std::min<long long>(n_assy, w0 + window);
        long long cnt[65] = {0};
        for (long long ia = w0; ia < w1; ++ia) cnt[code[ia] + 1]++;
        for (int c = 0; c < 64; ++c) cnt[c + 1] += cnt[c];
        std::vector<long long> cur(cnt, cnt + 64);
        for (long long ia = w0; ia < w1; ++ia) horder[w0 + cur[code[ia]]++] = (int)ia;
        for (int k = 0; k < 3; ++k) {
            const int ts = em::kTileSizes[k];
            for (int c = 0; c < 64; ++c)
                for (long long s0 = cnt[c]; s0 < cnt[c + 1]; s0 += ts) {
                    const long long n = std::min<long long>(ts, cnt[c + 1] - s0);
                    per_xcd[k][w % kXcds].push_back(
                        make_int2((int)(w0 + s0), (int)(n | ((long long)c << 16))));
                }
        }
    }
    for (int k = 0; k < 3; ++k) {
        size_t longest = 0;
        for (int x = 0; x < kXcds; ++x) longest = std::max(longest, per_xcd[k][x].size());
        htiles[k].assign(longest * kXcds, make_int2(0, 0));  // empty tiles: count 0
        for (int x = 0; x < kXcds; ++x)
            for (size_t j = 0; j < per_xcd[k][x].size(); ++j)
                htiles[k][j * kXcds + x] = per_xcd[k][x][j];
    }
}

constexpr long long kPrefetchDefaultMinTets = 1ll << 21;
constexpr long long kPrefetchDefaultChunk = 1ll << 20;

// Chunks of a plan's tile lists for prefetched launches (em_ned2_plan_set_prefetch): about
// prefetch_tets assembly indices each, cut at whole rounds of the XCD interleave (8 tiles).
// A tile's tets lie in its window [w W, (w + 1) W) with w = (first order position) / W, so
// a chunk reads order positions and assembly indices in [min w W, (max w + 1) W) and the
// tet ids that tetids maps those indices to. Host work only; O(tiles + assembly indices).
void build_chunks(em_ned2_plan_s* plan) {
    for (int k = 0; k < 3; ++k) plan->chunks[k].clear();
    if (plan->prefetch_tets <= 0 || plan->n_assy == 0) return;
    const long long na = plan->n_assy, W = plan->window;
    const long long nch = std::min<long long>(4096, (na + plan->prefetch_tets - 1) / plan->prefetch_tets);
    for (int k = 0; k < 3; ++k) {
        const std::vector<int2>& tl = plan->htiles[k];
        const long long rounds = ((long long)tl.size() + 7) / 8;
        const long long c_n = std::max<long long>(1, std::min(nch, rounds));
        for (long long c = 0; c < c_n; ++c) {
            em::PrefetchChunk ch;
            const long long b = std::min<long long>((long long)tl.size(), 8 * (rounds * c / c_n));
            const long long e = std::min<long long>((long long)tl.size(), 8 * (rounds * (c + 1) / c_n));
            ch.tile_begin = b;
            ch.tile_count = e - b;
            long long wlo = -1, whi = -1;
            for (long long i = b; i < e; ++i) {
                if ((tl[i].y & 0xffff) == 0) continue;  // padding tile
                const long long w = tl[i].x / W;
                wlo = wlo < 0 ? w : std::min(wlo, w);
                whi = std::max(whi, w);
            }
            if (wlo >= 0) {
                ch.ia_lo = wlo * W;
                ch.ia_hi = std::min(na, (whi + 1) * W);
                if (plan->htetids.empty()) {
                    ch.t_lo = ch.ia_lo;
                    ch.t_hi = ch.ia_hi;
                } else {
                    int lo = plan->htetids[ch.ia_lo], hi = lo;
                    for (long long i = ch.ia_lo; i < ch.ia_hi; ++i) {
                        lo = std::min(lo, plan->htetids[i]);
                        hi = std::max(hi, plan->htetids[i]);
                    }
                    ch.t_lo = lo;
                    ch.t_hi = (long long)hi + 1;
                }
            }
            plan->chunks[k].push_back(ch);
        }
    }
}

// Rebuild a plan's tile lists for another window: upload, wait for launches still
// reading the old lists, swap.
int rebuild_tiles(em_ned2_plan_s* plan, long long window) {
    std::vector<int> horder;
    std::vector<std::vector<int2>> htiles;
    build_tiles(plan->code, window, horder, htiles);
    DeviceGuard g(plan->device);
    int* order = nullptr;
    int2* tiles[3] = {nullptr, nullptr, nullptr};
    int st;
    if ((st = dmalloc(&order, horder.size())) || (st = dmalloc(&tiles[0], htiles[0].size())) ||
        (st = dmalloc(&tiles[1], htiles[1].size())) || (st = dmalloc(&tiles[2], htiles[2].size()))) {
        (void)hipFree(order);
        for (int k = 0; k < 3; ++k) (void)hipFree(tiles[k]);
        return st;
    }
    hipError_t e = horder.empty() ? hipSuccess
                                  : hipMemcpy(order, horder.data(), horder.size() * sizeof(int),
                                              hipMemcpyHostToDevice);
    for (int k = 0; k < 3 && e == hipSuccess; ++k)
        if (!htiles[k].empty())
            e = hipMemcpy(tiles[k], htiles[k].data(), htiles[k].size() * sizeof(int2),
                          hipMemcpyHostToDevice);
    // launches already queued may still read the previous lists
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        (void)hipFree(order);
        for (int k = 0; k < 3; ++k) (void)hipFree(tiles[k]);
        set_error(std::string("tile lists: ") + hipGetErrorString(e));
        return EM_EHIP;
    }
    (void)hipFree(plan->order);
    plan->order = order;
    for (int k = 0; k < 3; ++k) {
        (void)hipFree(plan->tiles[k]);
        plan->tiles[k] = tiles[k];
        plan->ntiles[k] = (long long)htiles[k].size();
        plan->htiles[k] = std::move(htiles[k]);
    }
    plan->window = window;
    build_chunks(plan);
    return EM_OK;
}

}  // namespace

extern "C" {

const char* em_last_error(void) { return em::g_err.c_str(); }

const char* em_version(void) { return "emerge_amd 0.1.0 (gfx950)"; }

int em_ned2_plan_create(em_ned2_plan_t* plan, int device, const double* nodes, int64_t n_nodes,
                        const int64_t* tets, int64_t n_tets, const int64_t* tris,
                        int64_t n_tris, const int64_t* edges, int64_t n_edges,
                        const int64_t* tet_to_field, const int64_t* tet_to_edge,
                        const int64_t* tet_to_tri, int64_t n_field, const int64_t* tetids,
                        int64_t n_assy, const int64_t* dofcodes, int32_t ndof) {
    if (!plan) {
        set_error("plan: null output pointer");
        return EM_EINVAL;
    }
    *plan = nullptr;
    if (!nodes || !tets || !tris || !edges || !tet_to_field || !tet_to_edge || !tet_to_tri ||
        !dofcodes || n_nodes <= 0 || n_tets < 0 || n_tris < 0 || n_edges < 0 || n_field <= 0) {
        set_error("plan_create: null table or non-positive size");
        return EM_EINVAL;
    }
    if (n_nodes >= (1ll << 31) || n_field >= (1ll << 31) || n_tets >= (1ll << 31)) {
        set_error("plan_create: ids must fit in int32");
        return EM_EINVAL;
    }
    static const int64_t kCodes[20] = {64, 64, 64, 64, 64, 64, 66, 66, 66, 66,
                                       66, 66, 128, 128, 128, 128, 129, 129, 129, 129};
    if (ndof != 20 || std::memcmp(dofcodes, kCodes, sizeof(kCodes)) != 0) {
        set_error("plan_create: only the SECOND_MIXED_VOLAKIS DOF set "
                  "[64]*6+[66]*6+[128]*4+[129]*4 is supported");
        return EM_EUNSUPPORTED;
    }
    if (tetids == nullptr) n_assy = n_tets;
    if (n_assy < 0 || n_assy > n_tets) {
        set_error("plan_create: n_assy out of range");
        return EM_EINVAL;
    }
    // host-side conversion and range checks
    std::vector<double4> hxyz((size_t)n_nodes);
    for (long long n = 0; n < n_nodes; ++n)
        hxyz[n] = make_double4(nodes[n], nodes[n_nodes + n], nodes[2 * n_nodes + n], 0.0);
    std::vector<int4> htets((size_t)n_tets);
    for (long long t = 0; t < n_tets; ++t) {
        long long v[4];
        for (int k = 0; k < 4; ++k) {
            v[k] = tets[k * n_tets + t];
            if (v[k] < 0 || v[k] >= n_nodes) {
                set_error("plan_create: tet node id out of range");
                return EM_EINVAL;
            }
        }
        if (v[0] == v[1] || v[0] == v[2] || v[0] == v[3] || v[1] == v[2] || v[1] == v[3] ||
            v[2] == v[3]) {
            set_error("plan_create: tet with repeated node");
            return EM_EINVAL;
        }
        htets[t] = make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
    }
    std::vector<int> ht2f((size_t)n_tets * 20);
    for (long long t = 0; t < n_tets; ++t)
        for (int d = 0; d < 20; ++d) {
            long long f = tet_to_field[d * n_tets + t];
            if (f < 0 || f >= n_field) {
                set_error("plan_create: tet_to_field entry out of range");
                return EM_EINVAL;
            }
            ht2f[t * 20 + d] = (int)f;
        }
    std::vector<int> hids;
    if (tetids) {
        hids.resize((size_t)n_assy);
        for (long long i = 0; i < n_assy; ++i) {
            if (tetids[i] < 0 || tetids[i] >= n_tets) {
                set_error("plan_create: tetids entry out of range");
                return EM_EINVAL;
            }
            hids[i] = (int)tetids[i];
        }
    }

    // vertex-rank code of every assembly index (the tet's node-id order), then the tiles
    std::vector<unsigned char> code((size_t)n_assy);
    for (long long ia = 0; ia < n_assy; ++ia) {
        const int4 v = htets[tetids ? hids[ia] : ia];
        const int n[4] = {v.x, v.y, v.z, v.w};
        int r[4];
        for (int a = 0; a < 4; ++a)
            r[a] = (n[0] < n[a]) + (n[1] < n[a]) + (n[2] < n[a]) + (n[3] < n[a]);
        code[ia] = (unsigned char)(r[0] * 16 + r[1] * 4 + r[2]);
    }
    constexpr long long kWindow = 1024;
    std::vector<int> horder;
    std::vector<std::vector<int2>> htiles;
    build_tiles(code, kWindow, horder, htiles);

    DeviceGuard g(device);
    if (!g.ok) {
        set_error("plan_create: cannot select device");
        return EM_EHIP;
    }
    auto* p = new em_ned2_plan_s();
    p->device = device;
    p->n_nodes = n_nodes;
    p->n_tets = n_tets;
    p->n_assy = n_assy;
    p->n_field = n_field;
    p->window = kWindow;
    int st;
    if ((st = dmalloc(&p->xyz, hxyz.size())) || (st = dmalloc(&p->tets, htets.size())) ||
        (st = dmalloc(&p->t2f, ht2f.size())) || (st = dmalloc(&p->tetids, hids.size())) ||
        (st = dmalloc(&p->order, horder.size())) || (st = dmalloc(&p->tiles[0], htiles[0].size())) ||
        (st = dmalloc(&p->tiles[1], htiles[1].size())) ||
        (st = dmalloc(&p->tiles[2], htiles[2].size())) || (st = dmalloc(&p->sink, 1))) {
        free_plan(p);
        return st;
    }
    for (int k = 0; k < 3; ++k) p->ntiles[k] = (long long)htiles[k].size();
    auto cp = [&](void* d, const void* h, size_t bytes) -> int {
        if (bytes == 0) return EM_OK;
        hipError_t e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error(std::string("hipMemcpy: ") + hipGetErrorString(e));
            return EM_EHIP;
        }
        return EM_OK;
    };
    if ((st = cp(p->xyz, hxyz.data(), hxyz.size() * sizeof(double4))) ||
        (st = cp(p->tets, htets.data(), htets.size() * sizeof(int4))) ||
        (st = cp(p->t2f, ht2f.data(), ht2f.size() * sizeof(int))) ||
        (st = cp(p->tetids, hids.data(), hids.size() * sizeof(int))) ||
        (st = cp(p->order, horder.data(), horder.size() * sizeof(int))) ||
        (st = cp(p->tiles[0], htiles[0].data(), htiles[0].size() * sizeof(int2))) ||
        (st = cp(p->tiles[1], htiles[1].data(), htiles[1].size() * sizeof(int2))) ||
        (st = cp(p->tiles[2], htiles[2].data(), htiles[2].size() * sizeof(int2)))) {
        free_plan(p);
        return st;
    }

    // Validate the edge/face tables and the volumes on the device.
    long long *d_edges = nullptr, *d_tris = nullptr, *d_t2e = nullptr, *d_t2t = nullptr;
    unsigned long long* d_bad = nullptr;
    unsigned long long bad[2] = {0, 0};
    int vst = EM_OK;
    if ((vst = dmalloc(&d_edges, (size_t)2 * n_edges)) || (vst = dmalloc(&d_tris, (size_t)3 * n_tris)) ||
        (vst = dmalloc(&d_t2e, (size_t)6 * n_tets)) || (vst = dmalloc(&d_t2t, (size_t)4 * n_tets)) ||
        (vst = dmalloc(&d_bad, 2))) {
    } else if ((vst = cp(d_edges, edges, sizeof(long long) * 2 * n_edges)) ||
               (vst = cp(d_tris, tris, sizeof(long long) * 3 * n_tris)) ||
               (vst = cp(d_t2e, tet_to_edge, sizeof(long long) * 6 * n_tets)) ||
               (vst = cp(d_t2t, tet_to_tri, sizeof(long long) * 4 * n_tets))) {
    } else {
        hipError_t e = hipMemset(d_bad, 0, 2 * sizeof(unsigned long long));
        if (e == hipSuccess)
            e = em::launch_check_tables(p->tets, n_tets, d_edges, n_edges, d_tris, n_tris, d_t2e,
                                        d_t2t, d_bad, nullptr);
        if (e == hipSuccess)
            e = em::launch_check_volumes(p->xyz, p->tets, n_tets, d_bad + 1, nullptr);
        if (e == hipSuccess)
            e = hipMemcpy(bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error(std::string("plan validation: ") + hipGetErrorString(e));
            vst = EM_EHIP;
        }
    }
    (void)hipFree(d_edges);
    (void)hipFree(d_tris);
    (void)hipFree(d_t2e);
    (void)hipFree(d_t2t);
    (void)hipFree(d_bad);
    if (vst != EM_OK) {
        free_plan(p);
        return vst;
    }
    if (bad[0]) {
        set_error("plan_create: " + std::to_string(bad[0]) +
                  " edge/face slots differ from the sorted-tuple tables of the tets");
        free_plan(p);
        return EM_ETABLES;
    }
    if (bad[1]) {
        set_error("plan_create: " + std::to_string(bad[1]) + " tets with zero volume");
        free_plan(p);
        return EM_EDEGENERATE;
    }
    p->code = std::move(code);
    for (int k = 0; k < 3; ++k) p->htiles[k] = std::move(htiles[k]);
    p->htetids = std::move(hids);
    // Prefetched launches by default on meshes whose inputs approach the memory-side
    // cache (2^21 tets and up): there they gain or cost nothing warm and gain cold; on
    // small meshes the extra launch would cost more than the reads it saves.
    if (n_assy >= kPrefetchDefaultMinTets) {
        p->prefetch_tets = kPrefetchDefaultChunk;
        build_chunks(p);
    }
    *plan = p;
    return EM_OK;
}

int em_ned2_plan_destroy(em_ned2_plan_t plan) {
    if (!plan) return EM_OK;
    DeviceGuard g(plan->device);
    free_plan(plan);
    return EM_OK;
}

int em_ned2_plan_sizes(em_ned2_plan_t plan, int64_t* n_assy, int64_t* n_field,
                       int64_t* nnz_coo) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    if (n_assy) *n_assy = plan->n_assy;
    if (n_field) *n_field = plan->n_field;
    if (nnz_coo) *nnz_coo = plan->n_assy * 400;
    return EM_OK;
}

int em_ned2_assemble_km(em_ned2_plan_t plan, const em_material_t* er, const em_material_t* ur,
                        void* dataE, void* dataB, int32_t precision, void* stream) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    int st;
    if ((st = check_material(er, plan->n_tets, "er")) || (st = check_material(ur, plan->n_tets, "ur")))
        return st;
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED &&
        precision != EM_PREC_FP64_REF) {
        set_error("unknown precision");
        return EM_EINVAL;
    }
    if (plan->n_assy > 0 && (!dataE || !dataB)) {
        set_error("null output");
        return EM_EINVAL;
    }
    if (((uintptr_t)dataE & 15) || ((uintptr_t)dataB & 15)) {
        set_error("outputs must be 16-byte aligned");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    if (precision == EM_PREC_FP64_REF) {
        EM_HIP(em::launch_assemble_ref(plan->dev(), to_desc(er, false), to_desc(ur, true),
                                       em::MODE_KM, nullptr, 0, dataE, dataB, 0,
                                       (hipStream_t)stream));
        return EM_OK;
    }
    em::TileSet ts[3];
    plan->tilesets(ts);
    EM_HIP(em::launch_assemble(plan->dev(), ts, plan->tile, to_desc(er, false), to_desc(ur, true),
                               em::MODE_KM, precision, needs_anti(er, ur), nullptr, 0, dataE,
                               dataB, 0, (hipStream_t)stream));
    return EM_OK;
}

int em_ned2_assemble_a(em_ned2_plan_t plan, const em_material_t* er, const em_material_t* ur,
                       const double* k0, int32_t n_k0, void* out, int64_t freq_stride,
                       int32_t precision, void* stream) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    int st;
    if ((st = check_material(er, plan->n_tets, "er")) || (st = check_material(ur, plan->n_tets, "ur")))
        return st;
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED &&
        precision != EM_PREC_FP64_REF) {
        set_error("unknown precision");
        return EM_EINVAL;
    }
    if (!k0 || n_k0 <= 0) {
        set_error("assemble_a: need at least one k0");
        return EM_EINVAL;
    }
    if (n_k0 > 1 && freq_stride < plan->n_assy * 400) {
        set_error("assemble_a: freq_stride smaller than one COO array");
        return EM_EINVAL;
    }
    if (plan->n_assy > 0 && !out) {
        set_error("null output");
        return EM_EINVAL;
    }
    if (((uintptr_t)out & 15) || (precision == EM_PREC_MIXED && n_k0 > 1 && (freq_stride & 1))) {
        set_error("output must be 16-byte aligned (complex64 sweeps need an even freq_stride)");
        return EM_EINVAL;
    }
    std::vector<double> k0sq((size_t)n_k0);
    for (int f = 0; f < n_k0; ++f) {
        if (!std::isfinite(k0[f])) {
            set_error("assemble_a: non-finite k0");
            return EM_EINVAL;
        }
        k0sq[f] = k0[f] * k0[f];
    }
    DeviceGuard g(plan->device);
    if (precision == EM_PREC_FP64_REF) {
        EM_HIP(em::launch_assemble_ref(plan->dev(), to_desc(er, false), to_desc(ur, true),
                                       n_k0 == 1 ? em::MODE_A : em::MODE_AF, k0sq.data(), n_k0,
                                       out, nullptr, freq_stride, (hipStream_t)stream));
        return EM_OK;
    }
    em::TileSet ts[3];
    plan->tilesets(ts);
    EM_HIP(em::launch_assemble(plan->dev(), ts, plan->tile, to_desc(er, false), to_desc(ur, true),
                               n_k0 == 1 ? em::MODE_A : em::MODE_AF, precision,
                               needs_anti(er, ur), k0sq.data(), n_k0, out, nullptr, freq_stride,
                               (hipStream_t)stream));
    return EM_OK;
}

int em_ned2_backward(em_ned2_plan_t plan, const void* grad_coo, double k0, int32_t eps_kind,
                     void* grad_eps, int32_t mu_kind, void* grad_mu_inv, void* stream) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    if (plan->n_assy > 0 && !grad_coo) {
        set_error("backward: null upstream gradient");
        return EM_EINVAL;
    }
    if ((grad_eps && (eps_kind < EM_MAT_SCALAR || eps_kind > EM_MAT_FULL)) ||
        (grad_mu_inv && (mu_kind < EM_MAT_SCALAR || mu_kind > EM_MAT_FULL))) {
        set_error("backward: unknown leaf kind");
        return EM_EINVAL;
    }
    if (!std::isfinite(k0)) {
        set_error("backward: non-finite k0");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    EM_HIP(em::launch_backward(plan->dev(), (const double2*)grad_coo, k0 * k0, eps_kind, mu_kind,
                               (double2*)grad_eps, (double2*)grad_mu_inv, (hipStream_t)stream));
    return EM_OK;
}

struct em_graph_s {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int em_graph_begin_capture(void* stream) {
    if (!stream) {
        set_error("graph capture needs a non-default stream");
        return EM_EINVAL;
    }
    EM_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return EM_OK;
}

int em_graph_end_capture(void* stream, em_graph_t* graph) {
    if (!stream || !graph) {
        set_error("null argument");
        return EM_EINVAL;
    }
    *graph = nullptr;
    auto* gr = new em_graph_s();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &gr->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (gr->graph) (void)hipGraphDestroy(gr->graph);
        delete gr;
        set_error(std::string("graph capture: ") + hipGetErrorString(e));
        return EM_EHIP;
    }
    *graph = gr;
    return EM_OK;
}

int em_graph_launch(em_graph_t graph, void* stream) {
    if (!graph) {
        set_error("null graph");
        return EM_EINVAL;
    }
    EM_HIP(hipGraphLaunch(graph->exec, (hipStream_t)stream));
    return EM_OK;
}

int em_graph_destroy(em_graph_t graph) {
    if (!graph) return EM_OK;
    (void)hipGraphExecDestroy(graph->exec);
    (void)hipGraphDestroy(graph->graph);
    delete graph;
    return EM_OK;
}

int em_ned2_plan_set_tuning(em_ned2_plan_t plan, int32_t tile) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    if (tile != 32 && tile != 64 && tile != 128) {
        set_error("tuning: tets per tile must be 32, 64 or 128");
        return EM_EINVAL;
    }
    plan->tile = tile;
    return EM_OK;
}

int em_ned2_plan_set_window(em_ned2_plan_t plan, int64_t window) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    if (window < 32 || window > (1ll << 24)) {
        set_error("tuning: the window must hold 32 to 16,777,216 tets");
        return EM_EINVAL;
    }
    if (window == plan->window) return EM_OK;
    return rebuild_tiles(plan, window);
}

int em_ned2_plan_set_prefetch(em_ned2_plan_t plan, int64_t chunk_tets) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    if (chunk_tets != 0 && (chunk_tets < 4096 || chunk_tets > (1ll << 31))) {
        set_error("prefetch: the chunk must hold 4,096 to 2^31 tets, or 0 (off)");
        return EM_EINVAL;
    }
    plan->prefetch_tets = chunk_tets;
    build_chunks(plan);
    return EM_OK;
}

int em_ned2_plan_get_prefetch(em_ned2_plan_t plan, int64_t* chunk_tets) {
    if (!plan || !chunk_tets) {
        set_error("null argument");
        return EM_EINVAL;
    }
    *chunk_tets = plan->prefetch_tets;
    return EM_OK;
}

int em_ned2_plan_get_window(em_ned2_plan_t plan, int64_t* window) {
    if (!plan || !window) {
        set_error("null argument");
        return EM_EINVAL;
    }
    *window = plan->window;
    return EM_OK;
}

int em_ned2_plan_get_tuning(em_ned2_plan_t plan, int32_t* tile) {
    if (!plan || !tile) {
        set_error("null argument");
        return EM_EINVAL;
    }
    *tile = plan->tile;
    return EM_OK;
}

int em_ned2_rowcol(em_ned2_plan_t plan, int64_t* rows, int64_t* cols, void* stream) {
    if (!plan) {
        set_error("null plan");
        return EM_EINVAL;
    }
    if (plan->n_assy > 0 && (!rows || !cols)) {
        set_error("null output");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    EM_HIP(em::launch_rowcol(plan->dev(), (long long*)rows, (long long*)cols, (hipStream_t)stream));
    return EM_OK;
}

struct em_csc_map_s {
    int device = 0;
    long long N = 0, nnz_coo = 0;
    em::CscBuild b;
    bool restricted = false;  // made by em_csc_map_restrict: no perm/target/col_offsets export
    long long n_perm = 0;     // COO contributors of the entries (nnz_coo for a full map)
};

static void free_csc(em_csc_map_s* c) {
    if (!c) return;
    (void)hipFree(c->b.indptr);
    (void)hipFree(c->b.indices);
    (void)hipFree(c->b.perm);
    (void)hipFree(c->b.seg);
    (void)hipFree(c->b.inc_ptr);
    (void)hipFree(c->b.inc);
    (void)hipFree(c->b.pos);
    (void)hipFree(c->b.cols);
    delete c;
}

int em_csc_map_create(em_csc_map_t* map, em_ned2_plan_t plan, void* stream) {
    if (!map || !plan) {
        set_error("csc_map_create: null argument");
        return EM_EINVAL;
    }
    *map = nullptr;
    if (plan->n_assy * 20 >= (1ll << 31)) {
        set_error("csc_map_create: more than 2^31 tet-DOF incidences");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    auto* c = new em_csc_map_s();
    c->device = plan->device;
    c->N = plan->n_field;
    c->nnz_coo = plan->n_assy * 400;
    long long max_inc = 0;
    hipError_t e = em::csc_build(plan->dev(), c->N, c->b, max_inc, (hipStream_t)stream);
    if (e != hipSuccess) {
        delete c;
        if (e == hipErrorInvalidValue && max_inc * 20 > em::kCscMaxColumnKeys) {
            set_error("csc_map_create: a DOF lies in " + std::to_string(max_inc) + " tets (at most " +
                      std::to_string(em::kCscMaxColumnKeys / 20) + " supported)");
            return EM_EUNSUPPORTED;
        }
        set_error(std::string("csc_map_create: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? EM_ENOMEM : EM_EHIP;
    }
    *map = c;
    return EM_OK;
}

int em_csc_map_destroy(em_csc_map_t map) {
    if (!map) return EM_OK;
    DeviceGuard g(map->device);
    free_csc(map);
    return EM_OK;
}

int em_csc_map_sizes(em_csc_map_t map, int64_t* n, int64_t* nnz, int64_t* nnz_coo) {
    if (!map) {
        set_error("null csc map");
        return EM_EINVAL;
    }
    if (n) *n = map->N;
    if (nnz) *nnz = map->b.nnz;
    if (nnz_coo) *nnz_coo = map->nnz_coo;
    return EM_OK;
}

int em_csc_map_export(em_csc_map_t map, int64_t* indptr, int64_t* indices, int64_t* perm,
                      int64_t* target_csc, int64_t* col_offsets, void* stream) {
    if (!map) {
        set_error("null csc map");
        return EM_EINVAL;
    }
    if (map->restricted && (perm || target_csc || col_offsets)) {
        set_error("csc_map_export: a restricted map exports indptr and indices only");
        return EM_EUNSUPPORTED;
    }
    DeviceGuard g(map->device);
    hipStream_t s = (hipStream_t)stream;
    const hipMemcpyKind k = hipMemcpyDeviceToDevice;
    if (indptr) EM_HIP(hipMemcpyAsync(indptr, map->b.indptr, sizeof(int64_t) * (map->N + 1), k, s));
    if (indices && map->b.nnz)
        EM_HIP(hipMemcpyAsync(indices, map->b.indices, sizeof(int64_t) * map->b.nnz, k, s));
    if (perm && map->nnz_coo)
        EM_HIP(hipMemcpyAsync(perm, map->b.perm, sizeof(int64_t) * map->nnz_coo, k, s));
    if (target_csc) EM_HIP(em::csc_target(map->b.seg, map->b.nnz, (long long*)target_csc, s));
    if (col_offsets)
        EM_HIP(em::csc_col_offsets(map->b.inc_ptr, map->N, map->b.ndof, (long long*)col_offsets, s));
    return EM_OK;
}

static int csc_apply(em_csc_map_t map, const void* E, const void* B, double k0, void* out,
                     int32_t precision, int32_t flags, void* stream) {
    if (flags & ~(EM_CSC_BLOCK_SYMMETRIC | EM_CSC_SEGMENT_GATHER)) {
        set_error("unknown to_csc flags");
        return EM_EINVAL;
    }
    if (!map) {
        set_error("null csc map");
        return EM_EINVAL;
    }
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED) {
        set_error("unknown precision");
        return EM_EINVAL;
    }
    if (map->b.nnz > 0 && (!E || !out)) {
        set_error("null data");
        return EM_EINVAL;
    }
    if (map->b.ndof != 20) {
        set_error("to_csc: complex values need a Nedelec (20-DOF) map; use em_csc_to_csc_real");
        return EM_EUNSUPPORTED;
    }
    DeviceGuard g(map->device);
    const int c64 = precision == EM_PREC_MIXED ? 1 : 0;
    const int mirror = (flags & EM_CSC_BLOCK_SYMMETRIC) ? 1 : 0;
    if (map->b.inc && map->b.pos && !(flags & EM_CSC_SEGMENT_GATHER))
        EM_HIP(em::csc_gather_inc(map->b, map->N, E, B, k0 * k0, out, c64, mirror,
                                  (hipStream_t)stream));
    else
        EM_HIP(em::csc_gather(map->b.seg, map->b.perm, map->b.nnz, E, B, k0 * k0, out, c64,
                              mirror, (hipStream_t)stream));
    return EM_OK;
}

int em_csc_to_csc(em_csc_map_t map, const void* data_coo, void* data_csc, int32_t precision,
                  void* stream) {
    return csc_apply(map, data_coo, nullptr, 0.0, data_csc, precision, 0, stream);
}

int em_csc_to_csc_combine(em_csc_map_t map, const void* dataE, const void* dataB, double k0,
                          void* data_csc, int32_t precision, void* stream) {
    // a map without entries (no assembled tets) reads nothing: empty inputs may be null
    if (map && map->b.nnz > 0 && !dataB) {
        set_error("to_csc_combine: null dataB");
        return EM_EINVAL;
    }
    return csc_apply(map, dataE, dataB, k0, data_csc, precision, 0, stream);
}

int em_csc_to_csc_ex(em_csc_map_t map, const void* dataE, const void* dataB, double k0,
                     void* data_csc, int32_t precision, int32_t flags, void* stream) {
    return csc_apply(map, dataE, dataB, k0, data_csc, precision, flags, stream);
}

int em_ned2_assemble_csc(em_ned2_plan_t plan, em_csc_map_t map, const em_material_t* er,
                         const em_material_t* ur, double k0, int32_t which, void* data_csc,
                         int32_t precision, void* stream) {
    if (!plan || !map) {
        set_error("assemble_csc: null plan or map");
        return EM_EINVAL;
    }
    int st;
    if ((st = check_material(er, plan->n_tets, "er")) || (st = check_material(ur, plan->n_tets, "ur")))
        return st;
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED) {
        set_error("assemble_csc: precision must be FP64 or MIXED");
        return EM_EINVAL;
    }
    if (which < 0 || which > 2) {
        set_error("assemble_csc: which must be 0 (A), 1 (K) or 2 (M)");
        return EM_EINVAL;
    }
    if (!std::isfinite(k0)) {
        set_error("assemble_csc: non-finite k0");
        return EM_EINVAL;
    }
    if (map->b.ndof != 20 || !map->b.inc || !map->b.pos || map->device != plan->device ||
        map->nnz_coo != plan->n_assy * 400) {
        set_error("assemble_csc: the map must be a Nedelec map of this plan (or restricted "
                  "from one)");
        return EM_EINVAL;
    }
    if (map->b.nnz > 0 && !data_csc) {
        set_error("null output");
        return EM_EINVAL;
    }
    if ((uintptr_t)data_csc & 15) {
        set_error("output must be 16-byte aligned");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    const int full = needs_anti(er, ur);
    const size_t need = em::assemble_csc_record_bytes(plan->n_assy, precision, full);
    if (need > plan->csc_rec_bytes) {
        // launches still queued may read the old records: hipFree waits for the device
        (void)hipFree(plan->csc_rec);
        plan->csc_rec = nullptr;
        plan->csc_rec_bytes = 0;
        const hipError_t e = hipMalloc(&plan->csc_rec, need);
        if (e != hipSuccess) {
            plan->csc_rec = nullptr;
            set_error(std::string("assemble_csc: records: ") + hipGetErrorString(e));
            return e == hipErrorOutOfMemory ? EM_ENOMEM : EM_EHIP;
        }
        plan->csc_rec_bytes = need;
    }
    EM_HIP(em::launch_assemble_csc(plan->dev(), map->b, map->N, to_desc(er, false),
                                   to_desc(ur, true), which, precision, full, k0 * k0,
                                   plan->csc_rec, data_csc, (hipStream_t)stream));
    return EM_OK;
}

int em_csc_to_csc_staged(em_csc_map_t map, const void* dataE, const void* dataB, double k0,
                         void* work, void* data_csc, int32_t precision, void* stream) {
    if (!map) {
        set_error("null csc map");
        return EM_EINVAL;
    }
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED) {
        set_error("unknown precision");
        return EM_EINVAL;
    }
    if (map->b.nnz > 0 && (!dataE || !data_csc || !work)) {
        set_error("to_csc_staged: null data, output or workspace");
        return EM_EINVAL;
    }
    if (map->b.ndof != 20 || !map->b.inc || !map->b.pos) {
        set_error("to_csc_staged: needs a Nedelec (20-DOF) map with incidence lists");
        return EM_EUNSUPPORTED;
    }
    DeviceGuard g(map->device);
    EM_HIP(em::csc_gather_inc_staged(map->b, map->N, map->nnz_coo / 400, dataE, dataB, k0 * k0,
                                     (double2*)work, data_csc,
                                     precision == EM_PREC_MIXED ? 1 : 0, (hipStream_t)stream));
    return EM_OK;
}

int em_csc_freq_combine(int64_t n, const void* K, const void* M, const double* k0, int32_t n_k0,
                        void* out, int64_t out_stride, int32_t precision, void* stream) {
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED) {
        set_error("csc_freq_combine: unknown precision");
        return EM_EINVAL;
    }
    if (n < 0 || n_k0 <= 0 || !k0 || (n > 0 && (!K || !M || !out)) ||
        (n_k0 > 1 && out_stride < n)) {
        set_error("csc_freq_combine: bad argument");
        return EM_EINVAL;
    }
    if (n == 0) return EM_OK;  // nothing read or written: no layout to check
    if (((uintptr_t)K & 15) || ((uintptr_t)M & 15) || ((uintptr_t)out & 15) ||
        (precision == EM_PREC_MIXED && n_k0 > 1 && (out_stride & 1))) {
        set_error("csc_freq_combine: pointers must be 16-byte aligned (complex64: even stride)");
        return EM_EINVAL;
    }
    std::vector<double> k0sq((size_t)n_k0);
    for (int f = 0; f < n_k0; ++f) {
        if (!std::isfinite(k0[f])) {
            set_error("csc_freq_combine: non-finite k0");
            return EM_EINVAL;
        }
        k0sq[f] = k0[f] * k0[f];
    }
    EM_HIP(em::csc_freq_combine(n, K, M, nullptr, k0sq.data(), nullptr, n_k0, out, out_stride,
                                precision == EM_PREC_MIXED ? 1 : 0, (hipStream_t)stream));
    return EM_OK;
}

int em_csc_freq_combine_lossy(int64_t n, const void* K, const void* M, const void* C,
                              const double* k0, const double* g, int32_t n_k0, void* out,
                              int64_t out_stride, int32_t precision, void* stream) {
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED) {
        set_error("csc_freq_combine_lossy: unknown precision");
        return EM_EINVAL;
    }
    if (n < 0 || n_k0 <= 0 || !k0 || !g || (n > 0 && (!K || !M || !C || !out)) ||
        (n_k0 > 1 && out_stride < n)) {
        set_error("csc_freq_combine_lossy: bad argument");
        return EM_EINVAL;
    }
    if (n == 0) return EM_OK;
    if (((uintptr_t)K & 15) || ((uintptr_t)M & 15) || ((uintptr_t)C & 15) ||
        ((uintptr_t)out & 15) || (precision == EM_PREC_MIXED && n_k0 > 1 && (out_stride & 1))) {
        set_error("csc_freq_combine_lossy: pointers must be 16-byte aligned (complex64: even "
                  "stride)");
        return EM_EINVAL;
    }
    std::vector<double> k0sq((size_t)n_k0);
    for (int f = 0; f < n_k0; ++f) {
        if (!std::isfinite(k0[f]) || !std::isfinite(g[f])) {
            set_error("csc_freq_combine_lossy: non-finite k0 or g");
            return EM_EINVAL;
        }
        k0sq[f] = k0[f] * k0[f];
    }
    EM_HIP(em::csc_freq_combine(n, K, M, C, k0sq.data(), g, n_k0, out, out_stride,
                                precision == EM_PREC_MIXED ? 1 : 0, (hipStream_t)stream));
    return EM_OK;
}

// ---------------------------------------------------------------------------------------
// PEC elimination and the solver hand-off
// ---------------------------------------------------------------------------------------

static int check_prec(int32_t precision) {
    if (precision != EM_PREC_FP64 && precision != EM_PREC_MIXED) {
        set_error("unknown precision");
        return EM_EINVAL;
    }
    return EM_OK;
}

int em_ned2_pec_mask(em_ned2_plan_t plan, const int64_t* conductor_tets, int64_t n_cond,
                     const int64_t* tri_to_field, int64_t n_tris, const int64_t* pec_tris,
                     int64_t n_pec, uint8_t* keep, void* stream) {
    if (!plan || !keep || n_cond < 0 || n_pec < 0 || n_tris < 0 || (n_cond && !conductor_tets) ||
        (n_pec && (!pec_tris || !tri_to_field))) {
        set_error("pec_mask: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    long long bad = 0;
    EM_HIP(em::pec_mark(plan->t2f, plan->n_tets, (const long long*)conductor_tets, n_cond,
                        (const long long*)tri_to_field, n_tris, (const long long*)pec_tris, n_pec,
                        plan->n_field, keep, bad, (hipStream_t)stream));
    if (bad) {
        set_error(std::string("pec_mask: out-of-range ") +
                  ((bad & 1) ? "conductor tet id" : (bad & 2) ? "PEC face id" : "face DOF"));
        return EM_EINVAL;
    }
    return EM_OK;
}

}  // extern "C"

namespace {

// Device scratch of em_material_prep, per device: grown when a call needs more, reused by
// every later call, held until the process ends (about T/64 * 8 + 9T/256 + T/32 bytes for
// the largest T seen). The lock serialises the calls on one device; each call blocks until its work
// is done, so the buffer is free again when the lock is released.
struct PrepScratchCache {
    std::mutex mu;
    void* p = nullptr;
    size_t bytes = 0;
};
std::mutex g_prep_caches_mu;
std::map<int, std::unique_ptr<PrepScratchCache>> g_prep_caches;

PrepScratchCache& prep_cache(int device) {
    std::lock_guard<std::mutex> lk(g_prep_caches_mu);
    auto& c = g_prep_caches[device];
    if (!c) c.reset(new PrepScratchCache());
    return *c;
}

}  // namespace

namespace em {

// Shared argument checks of the material preparation entry points.
int material_prep_check_args(const void* er, const void* tand, const void* cond, int64_t n_tets,
                             double omega, const void* er_out, const int64_t* conductor_tets,
                             const int64_t* n_conductor, const int32_t* frequency_dependent) {
    if (n_tets < 0 || !n_conductor || !frequency_dependent ||
        (n_tets > 0 && (!er || !tand || !cond || !er_out || !conductor_tets))) {
        set_error("material_prep: bad argument");
        return EM_EINVAL;
    }
    if (!(omega > 0.0)) {
        set_error("material_prep: omega must be positive");
        return EM_EINVAL;
    }
    if (9 * n_tets > (1ll << 40) - 256) {  // one thread per component, 2^32 workgroups
        set_error("material_prep: too many tets");
        return EM_EINVAL;
    }
    return EM_OK;
}

}  // namespace em

extern "C" {

int em_material_prep(const void* er, const void* tand, const void* cond, int64_t n_tets,
                     double omega, double pec_limit, double surf_limit, void* er_out,
                     int64_t* conductor_tets, int64_t* n_conductor,
                     int32_t* frequency_dependent, void* stream) {
    if (int st = em::material_prep_check_args(er, tand, cond, n_tets, omega, er_out,
                                              conductor_tets, n_conductor, frequency_dependent))
        return st;
    *n_conductor = 0;
    *frequency_dependent = 0;
    if (n_tets == 0) return EM_OK;
    int dev = 0;
    EM_HIP(hipGetDevice(&dev));
    PrepScratchCache& c = prep_cache(dev);
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t need = em::material_prep_scratch_bytes(n_tets);
    if (need > c.bytes) {
        (void)hipFree(c.p);
        c.p = nullptr;
        c.bytes = 0;
        EM_HIP(hipMalloc(&c.p, need));
        c.bytes = need;
    }
    long long nc = 0;
    bool fd = false;
    EM_HIP(em::material_prep((const double2*)er, (const double2*)tand, (const double2*)cond,
                             n_tets, omega, pec_limit, surf_limit, (double2*)er_out,
                             (long long*)conductor_tets, c.p, nc, fd, (hipStream_t)stream));
    *n_conductor = nc;
    *frequency_dependent = fd ? 1 : 0;
    return EM_OK;
}

int em_dof_compact(int64_t n, const uint8_t* keep, int64_t* new_index, int64_t* solve_ids,
                   int64_t* n_solve, void* stream) {
    if (n < 0 || !n_solve || (n > 0 && (!keep || !new_index || !solve_ids))) {
        set_error("dof_compact: bad argument");
        return EM_EINVAL;
    }
    long long ns = 0;
    EM_HIP(em::dof_compact(n, keep, (long long*)new_index, (long long*)solve_ids, ns,
                           (hipStream_t)stream));
    *n_solve = ns;
    return EM_OK;
}

int em_csc_select_count(int64_t n, const int64_t* indptr, const int64_t* indices,
                        const void* data, int32_t precision, int32_t drop_zeros,
                        const int64_t* new_index, int64_t n_solve, int64_t* sel_indptr,
                        int64_t* nnz_sel, void* stream) {
    if (int st = check_prec(precision)) return st;
    if (n < 0 || n_solve < 0 || n_solve > n || !indptr || !sel_indptr || !nnz_sel ||
        (n > 0 && (!indices || !new_index)) || (drop_zeros && !data)) {
        set_error("csc_select_count: bad argument");
        return EM_EINVAL;
    }
    long long nz = 0;
    bool bad = false;
    EM_HIP(em::csc_select_count(n, (const long long*)indptr, (const long long*)indices, data,
                                precision == EM_PREC_MIXED, drop_zeros != 0,
                                (const long long*)new_index, n_solve, (long long*)sel_indptr, nz,
                                bad, (hipStream_t)stream));
    if (bad) {
        set_error("csc_select_count: row index out of range");
        return EM_EINVAL;
    }
    *nnz_sel = nz;
    return EM_OK;
}

int em_csc_select_fill(int64_t n, const int64_t* indptr, const int64_t* indices,
                       const void* data, int32_t precision, int32_t drop_zeros,
                       const int64_t* new_index, const int64_t* sel_indptr, int64_t* sel_indices,
                       void* sel_data, void* stream) {
    if (int st = check_prec(precision)) return st;
    if (n < 0 || !indptr || !sel_indptr || (n > 0 && (!indices || !new_index || !sel_indices)) ||
        (drop_zeros && !data) || (sel_data && !data)) {
        set_error("csc_select_fill: bad argument");
        return EM_EINVAL;
    }
    EM_HIP(em::csc_select_fill(n, (const long long*)indptr, (const long long*)indices,
                               (sel_data || drop_zeros) ? data : nullptr,
                               precision == EM_PREC_MIXED,
                               drop_zeros != 0, (const long long*)new_index,
                               (const long long*)sel_indptr, (long long*)sel_indices, sel_data,
                               (hipStream_t)stream));
    return EM_OK;
}

int em_csc_map_restrict(em_csc_map_t map, const int64_t* new_index, int64_t n_solve,
                        em_csc_map_t* sub, void* stream) {
    if (!map || !sub || n_solve < 0 || n_solve > map->N || (map->N > 0 && !new_index)) {
        set_error("csc_map_restrict: bad argument");
        return EM_EINVAL;
    }
    *sub = nullptr;
    if (map->restricted) {
        set_error("csc_map_restrict: the map is already restricted");
        return EM_EUNSUPPORTED;
    }
    DeviceGuard g(map->device);
    auto* c = new em_csc_map_s();
    c->device = map->device;
    c->N = n_solve;
    c->nnz_coo = map->nnz_coo;
    c->restricted = true;
    long long n_perm = 0;
    hipError_t e = em::csc_restrict(map->b, map->N, (const long long*)new_index, n_solve, c->b,
                                    n_perm, (hipStream_t)stream);
    if (e != hipSuccess) {
        delete c;
        set_error(std::string("csc_map_restrict: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? EM_ENOMEM : EM_EHIP;
    }
    c->n_perm = n_perm;
    *sub = c;
    return EM_OK;
}

static int dof_move(const int64_t* ids, int64_t n_ids, int64_t ncols, const void* src, void* dst,
                    int32_t precision, int scatter, void* stream) {
    if (int st = check_prec(precision)) return st;
    if (n_ids < 0 || ncols < 0 || (n_ids * ncols > 0 && (!ids || !src || !dst))) {
        set_error("dof gather/scatter: bad argument");
        return EM_EINVAL;
    }
    EM_HIP(em::dof_move((const long long*)ids, n_ids, ncols, src, dst,
                        precision == EM_PREC_MIXED, scatter, (hipStream_t)stream));
    return EM_OK;
}

int em_dof_gather(const int64_t* ids, int64_t n_ids, int64_t ncols, const void* src, void* dst,
                  int32_t precision, void* stream) {
    return dof_move(ids, n_ids, ncols, src, dst, precision, 0, stream);
}

int em_dof_scatter(const int64_t* ids, int64_t n_ids, int64_t ncols, const void* src, void* dst,
                   int32_t precision, void* stream) {
    return dof_move(ids, n_ids, ncols, src, dst, precision, 1, stream);
}

// ---------------------------------------------------------------------------------------
// Robin / port surface assembly
// ---------------------------------------------------------------------------------------

int em_ned2_robin_blocks(em_ned2_plan_t plan, const int64_t* tris, int64_t n_tris,
                         const int64_t* surf, int64_t n_surf, double gamma_re, double gamma_im,
                         void* out, int32_t layout, void* stream) {
    if (!plan || n_surf < 0 || n_tris < 0 || (n_surf && (!tris || !surf || !out)) ||
        (layout != EM_TRI_BY_FACE && layout != EM_TRI_COMPACT)) {
        set_error("robin_blocks: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    long long bad = 0;
    EM_HIP(em::robin_blocks(plan->xyz, plan->n_nodes, (const long long*)tris, n_tris,
                            (const long long*)surf, n_surf, gamma_re, gamma_im, (double2*)out,
                            layout == EM_TRI_BY_FACE, bad, (hipStream_t)stream));
    if (bad) {
        set_error("robin_blocks: face or node id out of range");
        return EM_EINVAL;
    }
    return EM_OK;
}

int em_ned2_robin_force(em_ned2_plan_t plan, const int64_t* tris, int64_t n_tris,
                        const int64_t* tri_to_field, const int64_t* surf, int64_t n_surf,
                        const void* U, const void* U_curl, const double* normals, void* bvec,
                        void* stream) {
    if (!plan || n_surf < 0 || n_tris < 0 ||
        (n_surf && (!tris || !tri_to_field || !surf || !U || !bvec)) ||
        (U_curl && !normals)) {
        set_error("robin_force: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    long long bad = 0;
    EM_HIP(em::robin_force(plan->xyz, plan->n_nodes, (const long long*)tris, n_tris,
                           (const long long*)tri_to_field, plan->n_field, (const long long*)surf,
                           n_surf, (const double2*)U, (const double2*)U_curl, normals,
                           (double2*)bvec, bad, (hipStream_t)stream));
    if (bad) {
        set_error((bad & 1) ? "robin_force: face or node id out of range"
                            : "robin_force: tri_to_field entry out of range");
        return EM_EINVAL;
    }
    return EM_OK;
}

int em_tri_rowcol(const int64_t* tri_to_field, int64_t n_tris, const int64_t* surf,
                  int64_t n_surf, int64_t* rows, int64_t* cols, void* stream) {
    if (n_surf < 0 || n_tris < 0 || (n_surf && (!tri_to_field || !surf || !rows || !cols))) {
        set_error("tri_rowcol: bad argument");
        return EM_EINVAL;
    }
    long long bad = 0;
    EM_HIP(em::tri_rowcol((const long long*)tri_to_field, n_tris, (const long long*)surf, n_surf,
                          (long long*)rows, (long long*)cols, bad, (hipStream_t)stream));
    if (bad) {
        set_error("tri_rowcol: face id out of range");
        return EM_EINVAL;
    }
    return EM_OK;
}

int em_csc_map_find(em_csc_map_t map, const int64_t* rows, const int64_t* cols, int64_t n,
                    const int64_t* new_index, int64_t* pos, void* stream) {
    if (!map || n < 0 || (n && (!rows || !cols || !pos))) {
        set_error("csc_map_find: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(map->device);
    EM_HIP(em::csc_find(map->b.indptr, map->b.indices, map->N, (const long long*)rows,
                        (const long long*)cols, (const long long*)new_index, n, (long long*)pos,
                        (hipStream_t)stream));
    return EM_OK;
}

int em_csc_add(void* data_csc, int32_t precision, const int64_t* pos, const void* values,
               int64_t n, double scale_re, double scale_im, void* stream) {
    if (int st = check_prec(precision)) return st;
    if (n < 0 || (n && (!data_csc || !pos || !values))) {
        set_error("csc_add: bad argument");
        return EM_EINVAL;
    }
    EM_HIP(em::csc_add(data_csc, precision == EM_PREC_MIXED, (const long long*)pos,
                       (const double2*)values, n, scale_re, scale_im, (hipStream_t)stream));
    return EM_OK;
}

// ---------------------------------------------------------------------------------------
// Point location and field interpolation
// ---------------------------------------------------------------------------------------

int em_ned2_locate(em_ned2_plan_t plan, const double* coords, int64_t n, const int64_t* tetids,
                   int64_t n_tetids, int64_t* tet_of, void* stream) {
    if (!plan || n < 0 || n_tetids < 0 || (n && (!coords || !tet_of))) {
        set_error("locate: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    long long bad = 0;
    const long long n_cand = tetids ? n_tetids : plan->n_tets;
    EM_HIP(em::locate_points(plan->xyz, plan->tets, plan->n_tets, coords, n,
                             (const long long*)tetids, n_cand, (long long*)tet_of, bad,
                             (hipStream_t)stream));
    if (bad) {
        set_error("locate: candidate tet id out of range");
        return EM_EINVAL;
    }
    return EM_OK;
}

int em_ned2_interp(em_ned2_plan_t plan, const double* coords, int64_t n, const int64_t* tet_of,
                   const void* solution, int32_t which, const void* c_tet, void* out,
                   void* stream) {
    if (!plan || n < 0 || (n && (!coords || !tet_of || !solution || !out)) ||
        (which != EM_INTERP_E && which != EM_INTERP_CURL)) {
        set_error("interp: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    EM_HIP(em::ned2_interp(plan->xyz, plan->tets, plan->t2f, plan->n_tets, coords, n,
                           (const long long*)tet_of, (const double2*)solution,
                           which == EM_INTERP_CURL, (const double2*)c_tet, (double2*)out,
                           (hipStream_t)stream));
    return EM_OK;
}

// ---------------------------------------------------------------------------------------
// Heat conduction (Lagrange-2)
// ---------------------------------------------------------------------------------------

struct em_heat_plan_s {
    int device = 0;
    long long n_nodes = 0, n_tets = 0, n_field = 0;
    double4* xyz = nullptr;
    int4* tets = nullptr;
    int* t2f = nullptr;  // [T][10]
};

static void free_heat(em_heat_plan_s* p) {
    if (!p) return;
    (void)hipFree(p->xyz);
    (void)hipFree(p->tets);
    (void)hipFree(p->t2f);
    delete p;
}

int em_heat_plan_create(em_heat_plan_t* plan, int device, const double* nodes, int64_t n_nodes,
                        const int64_t* tets, int64_t n_tets, const int64_t* tet_to_field,
                        int64_t n_field) {
    if (!plan) {
        set_error("heat_plan: null output pointer");
        return EM_EINVAL;
    }
    *plan = nullptr;
    if (!nodes || !tets || !tet_to_field || n_nodes <= 0 || n_tets < 0 || n_field <= 0 ||
        n_nodes >= (1ll << 31) || n_field >= (1ll << 31) || n_tets * 10 >= (1ll << 31)) {
        set_error("heat_plan: bad argument or size");
        return EM_EINVAL;
    }
    std::vector<double4> hxyz((size_t)n_nodes);
    for (long long n = 0; n < n_nodes; ++n)
        hxyz[n] = make_double4(nodes[n], nodes[n_nodes + n], nodes[2 * n_nodes + n], 0.0);
    std::vector<int4> ht((size_t)n_tets);
    std::vector<int> hf((size_t)n_tets * 10);
    for (long long t = 0; t < n_tets; ++t) {
        long long v[4];
        for (int k = 0; k < 4; ++k) {
            v[k] = tets[k * n_tets + t];
            if (v[k] < 0 || v[k] >= n_nodes) {
                set_error("heat_plan: tet node id out of range");
                return EM_EINVAL;
            }
        }
        ht[t] = make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
        double J[3][3];
        for (int k = 0; k < 3; ++k) {
            J[k][0] = hxyz[v[k + 1]].x - hxyz[v[0]].x;
            J[k][1] = hxyz[v[k + 1]].y - hxyz[v[0]].y;
            J[k][2] = hxyz[v[k + 1]].z - hxyz[v[0]].z;
        }
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (det == 0.0 || !std::isfinite(det)) {
            set_error("heat_plan: zero-volume tet " + std::to_string(t));
            return EM_EDEGENERATE;
        }
        for (int d = 0; d < 10; ++d) {
            const long long f = tet_to_field[d * n_tets + t];
            if (f < 0 || f >= n_field) {
                set_error("heat_plan: tet_to_field entry out of range");
                return EM_EINVAL;
            }
            hf[t * 10 + d] = (int)f;
        }
    }
    DeviceGuard g(device);
    auto* p = new em_heat_plan_s();
    p->device = device;
    p->n_nodes = n_nodes;
    p->n_tets = n_tets;
    p->n_field = n_field;
    int st;
    if ((st = dmalloc(&p->xyz, hxyz.size())) || (st = dmalloc(&p->tets, ht.size())) ||
        (st = dmalloc(&p->t2f, hf.size()))) {
        free_heat(p);
        return st;
    }
    hipError_t e = hipMemcpy(p->xyz, hxyz.data(), hxyz.size() * sizeof(double4),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess && !ht.empty())
        e = hipMemcpy(p->tets, ht.data(), ht.size() * sizeof(int4), hipMemcpyHostToDevice);
    if (e == hipSuccess && !hf.empty())
        e = hipMemcpy(p->t2f, hf.data(), hf.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_error(std::string("heat_plan: ") + hipGetErrorString(e));
        free_heat(p);
        return EM_EHIP;
    }
    *plan = p;
    return EM_OK;
}

int em_heat_plan_destroy(em_heat_plan_t plan) {
    if (!plan) return EM_OK;
    DeviceGuard g(plan->device);
    free_heat(plan);
    return EM_OK;
}

int em_heat_assemble(em_heat_plan_t plan, const double* cond, int32_t cond_kind,
                     int64_t cond_tet_stride, int64_t cond_comp_stride, const double* rho_cp,
                     int64_t rho_tet_stride, double* K, double* M, void* stream) {
    if (!plan || (K && !cond) || (M && !rho_cp) || cond_tet_stride < 0 ||
        cond_comp_stride < 0 || rho_tet_stride < 0 ||
        (K && (cond_kind < EM_MAT_SCALAR || cond_kind > EM_MAT_FULL))) {
        set_error("heat_assemble: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    EM_HIP(em::heat_assemble(plan->xyz, plan->tets, plan->n_tets, cond, cond_kind,
                             cond_tet_stride, cond_comp_stride, rho_cp, rho_tet_stride, K, M,
                             (hipStream_t)stream));
    return EM_OK;
}

int em_heat_rowcol(em_heat_plan_t plan, int64_t* rows, int64_t* cols, void* stream) {
    if (!plan || (plan->n_tets > 0 && (!rows || !cols))) {
        set_error("heat_rowcol: bad argument");
        return EM_EINVAL;
    }
    DeviceGuard g(plan->device);
    EM_HIP(em::heat_rowcol(plan->t2f, plan->n_tets, (long long*)rows, (long long*)cols,
                           (hipStream_t)stream));
    return EM_OK;
}

int em_csc_map_create_heat(em_csc_map_t* map, em_heat_plan_t plan, void* stream) {
    if (!map || !plan) {
        set_error("csc_map_create_heat: null argument");
        return EM_EINVAL;
    }
    *map = nullptr;
    DeviceGuard g(plan->device);
    auto* c = new em_csc_map_s();
    c->device = plan->device;
    c->N = plan->n_field;
    c->nnz_coo = plan->n_tets * 100;
    c->n_perm = c->nnz_coo;
    long long max_inc = 0;
    hipError_t e = em::csc_build_ndof10(plan->t2f, plan->n_tets, c->N, c->b, max_inc,
                                        (hipStream_t)stream);
    if (e != hipSuccess) {
        delete c;
        if (e == hipErrorInvalidValue && max_inc * 10 > em::kCscMaxColumnKeys) {
            set_error("csc_map_create_heat: a DOF lies in " + std::to_string(max_inc) + " tets (at most " +
                      std::to_string(em::kCscMaxColumnKeys / 10) + " supported)");
            return EM_EUNSUPPORTED;
        }
        set_error(std::string("csc_map_create_heat: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? EM_ENOMEM : EM_EHIP;
    }
    *map = c;
    return EM_OK;
}

int em_csc_to_csc_real(em_csc_map_t map, const double* data, const double* data_b, double scale,
                       double* data_csc, int flags, void* stream) {
    if (!map || (map->b.nnz > 0 && (!data || !data_csc))) {
        set_error("to_csc_real: bad argument");
        return EM_EINVAL;
    }
    if (flags & ~(EM_CSC_SEGMENT_GATHER | EM_CSC_BLOCK_SYMMETRIC)) {
        set_error("to_csc_real: unknown flags");
        return EM_EINVAL;
    }
    if ((flags & EM_CSC_SEGMENT_GATHER) && (flags & EM_CSC_BLOCK_SYMMETRIC)) {
        set_error("to_csc_real: EM_CSC_BLOCK_SYMMETRIC applies to the incidence gather only");
        return EM_EINVAL;
    }
    DeviceGuard g(map->device);
    if (map->b.ndof == 10 && map->b.inc && map->b.pos && !(flags & EM_CSC_SEGMENT_GATHER)) {
        EM_HIP(em::csc_gather_inc_real(map->b, map->N, data, data_b, scale,
                                       (flags & EM_CSC_BLOCK_SYMMETRIC) != 0, data_csc,
                                       (hipStream_t)stream));
    } else {
        EM_HIP(em::csc_gather_real(map->b.seg, map->b.perm, map->b.nnz, data, data_b, scale,
                                   data_csc, (hipStream_t)stream));
    }
    return EM_OK;
}

}  // extern "C"
