// Host runtime of the gfx950 Barnes-Hut library: the C ABI of include/galaxy_bh.h.
//
// Drop-in objects (device, state, runtime, tree) mirror NbodyGpu and
// ParallelTreeRuntime (runtime/src/nbody_gpu.rs, runtime/src/nbody_parallel.rs).
// All of them, and the batched and serving forms, drive the same kernels through
// a SimSet: a table of simulations (SimDesc) plus per-kernel tile tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/galaxy_bh.h"
#include "bh_kernels.h"

namespace gbh {

thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

struct HipError {
    std::string msg;
};

#define HIP_OK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)};              \
    } while (0)

template <class F>
int guarded(F&& f) {
    try {
        return f();
    } catch (const HipError& e) {
        set_error(e.msg);
        return -1;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return -1;
    } catch (const std::exception& e) {  // no C++ exception crosses the C ABI
        set_error(std::string("internal error: ") + e.what());
        return -1;
    }
}

double now_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static uint32_t div_up(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Upper bound of the cell count for n bodies: cells at depth d are distinct
// 2d-bit key prefixes (<= 4^d) and disjoint non-empty ranges (<= n).
constexpr uint64_t cell_capacity(uint32_t n) {
    uint64_t cap = 0, p4 = 1;
    for (uint32_t d = 0; d < kLevels; ++d) {
        cap += p4 < n ? p4 : n;
        p4 *= 4;
    }
    return cap;
}

// The node stream (at most cell_capacity(n) + n nodes) must fit the subtree-size
// field of the hot node word for every tree the library builds.
static_assert(sizeof(NodeHot) == 16 && sizeof(NodeCold) == 16, "node records are 16 bytes");
static_assert(cell_capacity(GBH_MAX_LARGE_TREE_BODIES) + GBH_MAX_LARGE_TREE_BODIES < kMaxNodes,
              "the node stream of the largest tree fits the hot word");

// (for either topology tile size: at most 16 spanning cells per tile edge)
uint32_t span_capacity(uint32_t n) { return 16u * div_up(n, kTopoTileMulti) + 16u; }

// Layout of one simulation's device arena.  With tree == false only the
// resident state is placed.
size_t carve(SimDesc& d, char* base, uint32_t n, bool tree) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        off = align_up(off, 256);
        char* p = base ? base + off : nullptr;
        off += bytes;
        return p;
    };
    d.pos = (float2*)take(8ull * n);
    d.mass = (float*)take(4ull * n);
    d.vel = (float2*)take(8ull * n);
    d.acc = (float2*)take(8ull * n);
    if (!tree) return align_up(off, 256);
    const uint64_t cap = cell_capacity(n);
    const uint32_t scap = span_capacity(n);
    d.key0 = (uint32_t*)take(4ull * n);
    d.key1 = (uint32_t*)take(4ull * n);
    d.val0 = (uint32_t*)take(4ull * n);
    d.val1 = (uint32_t*)take(4ull * n);
    d.sbody = (float4*)take(16ull * n);
    d.cbase = (uint32_t*)take(4ull * n);
    d.cparent = (uint32_t*)take(4ull * n);
    d.cdmin = (uint8_t*)take(1ull * n);
    d.ccount = (uint8_t*)take(1ull * n);
    d.cells = (CellT*)take(sizeof(CellT) * cap);
    d.cend = (uint32_t*)take(4ull * cap);
    d.cchild = (uint4*)take(16ull * cap);
    d.cagg = (Agg*)take(sizeof(Agg) * cap);
    d.span = (uint2*)take(8ull * scap);
    d.spanbuf = (uint2*)take(8ull * scap);
    d.nodes = (NodeHot*)take(sizeof(NodeHot) * (cap + n));
    d.ncold = (NodeCold*)take(sizeof(NodeCold) * (cap + n));
    d.nbase = (uint32_t*)take(4ull * n);
    d.bnode = (uint32_t*)take(4ull * n);
    d.bpart = (float4*)take(16ull * div_up(n, kTile));
    d.meta = (SimMeta*)take(sizeof(SimMeta));
    d.cell_cap = (uint32_t)cap;
    d.span_cap = scap;
    d.node_cap = (uint32_t)(cap + n);
    d.wlayout = d.node_cap <= kNarrowNodes ? 1u : 0u;
    return align_up(off, 256);
}

}  // namespace gbh

using namespace gbh;

// ============================================================================
// device
// ============================================================================
struct gbh_device {
    int index = 0;
    int count = 0;
    hipStream_t stream = nullptr;
    std::string info;
    int cu_count = 0;
    // Device address of this GPU's copy of `scratch` (below), from
    // hipGetSymbolAddress with the device current; the array's name in host
    // code is not a device address.
    unsigned char* scratch = nullptr;
};

namespace gbh {

// Scratch for the entry points that allocate nothing (gbh_force_evidence_noalloc):
// the per-target interaction evidence of up to kScratchTargets targets, staged
// here between the traversal and the copy to the caller.  One copy per GPU, held
// until the process ends; every call on a GPU shares it, so calls that overlap
// on one GPU race.
constexpr uint32_t kScratchTargets = GBH_NOALLOC_MAX_BODIES;
constexpr size_t kScratchStatsBytes = 12ull * kScratchTargets;
constexpr size_t kScratchBytes = kScratchStatsBytes + 8ull * kScratchTargets;
static_assert(kScratchStatsBytes % 256 == 0, "the hash part of the scratch stays 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

}  // namespace gbh

// ============================================================================
// SimSet: the simulations one group of launches advances
// ============================================================================
namespace gbh {

struct SimSet {
    gbh_device* dev = nullptr;
    std::vector<SimDesc> desc;
    SimDesc* d_desc = nullptr;
    std::vector<TileRef> t256, tsort, ttopo, twave;
    TileRef *d_t256 = nullptr, *d_tsort = nullptr, *d_ttopo = nullptr, *d_twave = nullptr;
    uint32_t* d_ctl = nullptr;
    uint32_t* d_sflags = nullptr;
    unsigned long long* d_tflags = nullptr;
    uint32_t* d_tsum = nullptr;  // per topology tile boundary summaries (k_topo_pre)
    uint32_t* d_psum = nullptr;  // and their scans over the simulation's tiles (k_topo_scan)
    uint32_t* d_errs = nullptr;  // collected error words (k_collect_errors)
    uint32_t* h_errs = nullptr;  // their pinned host copy
    size_t cap_herrs = 0;
    size_t cap_desc = 0, cap_t256 = 0, cap_tsort = 0, cap_ttopo = 0, cap_twave = 0;
    size_t cap_sflags = 0, cap_tflags = 0, cap_tsum = 0, cap_psum = 0, cap_errs = 0;
    uint32_t mode = 0;                // traversal work distribution (kCtlMode)
    uint32_t ranges[2 * kXcds] = {};  // traversal items per XCD: start[8], length[8]
    uint32_t sort_kpt = 8;
    uint32_t trav_block = 256;        // traversal workgroup size (64, 128 or 256)
    uint32_t trav_waves_per_cu = 32;  // persistent traversal wavefronts per CU
    uint32_t traverse_blocks = 0;
    uint32_t lane_targets = 1;        // traversal targets per lane (build_tables)
    uint32_t topo_tile = kTopoTile;   // topology tile positions (build_tables)
    uint32_t topo_tile_multi = kTopoTileMulti;  // the tile of a set of several simulations

    ~SimSet() { release(); }

    void release() {
        for (void* p : {(void*)d_desc, (void*)d_t256, (void*)d_tsort, (void*)d_ttopo,
                        (void*)d_twave, (void*)d_ctl, (void*)d_sflags, (void*)d_tflags,
                        (void*)d_tsum, (void*)d_psum, (void*)d_errs})
            if (p) (void)hipFree(p);
        d_desc = nullptr;
        d_t256 = d_tsort = d_ttopo = d_twave = nullptr;
        d_ctl = d_sflags = d_tsum = d_psum = d_errs = nullptr;
        d_tflags = nullptr;
        cap_desc = cap_t256 = cap_tsort = cap_ttopo = cap_twave = cap_sflags = cap_tflags = 0;
        cap_tsum = cap_psum = cap_errs = 0;
        if (h_errs) (void)hipHostFree(h_errs);
        h_errs = nullptr;
        cap_herrs = 0;
    }

    template <class T>
    static void ensure(T*& p, size_t& cap, size_t count) {
        count = std::max<size_t>(count, 1);
        if (count <= cap) return;
        if (p) HIP_OK(hipFree(p));
        p = nullptr;
        HIP_OK(hipMalloc((void**)&p, sizeof(T) * count));
        cap = count;
    }

    uint32_t sort_tile() const { return kSortThreads * sort_kpt; }

    // Tile tables from desc[].n; fills the per-simulation first-tile indices.
    void build_tables() {
        t256.clear();
        tsort.clear();
        ttopo.clear();
        twave.clear();
        const uint32_t st = sort_tile();
        topo_tile = desc.size() == 1 ? kTopoTile : topo_tile_multi;
        for (uint32_t s = 0; s < desc.size(); ++s) {
            SimDesc& d = desc[s];
            d.first_tile = (uint32_t)t256.size();
            for (uint32_t b = 0; b < d.n; b += kTile) t256.push_back({s, b});
            d.first_sort_tile = (uint32_t)tsort.size();
            for (uint32_t b = 0; b < d.n; b += st) tsort.push_back({s, b});
            d.first_topo_tile = (uint32_t)ttopo.size();
            for (uint32_t b = 0; b < d.n; b += topo_tile) ttopo.push_back({s, b});
        }
        // Traversal work items.  Mode 0 (many simulations): one contiguous range
        // per XCD; simulations are dealt to the XCDs round-robin in size order
        // (similar work per XCD, each simulation's items together) and the
        // sequence of the XCDs' lists is cut into kXcds equal ranges.  Mode 1
        // (a simulation larger than one XCD's share): one range in size order.
        // Within a range, larger simulations come first (longer walks) for a
        // short tail.
        // Targets per lane: two (items of 128 targets, the multi-target walk)
        // when the set has items enough to keep the persistent grid busy twice
        // over; else one, so that a small set still spreads over the device.
        {
            uint64_t items2 = 0;
            for (const SimDesc& d : desc) items2 += div_up(d.n, 2u * 64u);
            const uint64_t grid_waves = (uint64_t)(dev ? dev->cu_count : 256) * trav_waves_per_cu;
            lane_targets = items2 >= 2ull * grid_waves ? 2u : 1u;
        }
        const uint32_t per_item = 64u * lane_targets;
        std::vector<uint32_t> order(desc.size());
        for (uint32_t s = 0; s < order.size(); ++s) order[s] = s;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return desc[a].n > desc[b].n; });
        uint64_t total = 0, largest = 0;
        for (const SimDesc& d : desc) {
            const uint64_t items = div_up(d.n, per_item);
            total += items;
            largest = std::max(largest, items);
        }
        mode = largest * kXcds > total ? 1u : 0u;
        if (mode == 0u) {
            for (uint32_t x = 0; x < kXcds; ++x)
                for (size_t k = x; k < order.size(); k += kXcds)
                    for (uint32_t b = 0; b < desc[order[k]].n; b += per_item) twave.push_back({order[k], b});
        } else {
            for (uint32_t s : order)
                for (uint32_t b = 0; b < desc[s].n; b += per_item) twave.push_back({s, b});
        }
        const uint32_t nt = (uint32_t)twave.size(), q = nt / kXcds, r = nt % kXcds;
        for (uint32_t x = 0; x < kXcds; ++x) {
            ranges[x] = x * q + std::min(x, r);
            ranges[kXcds + x] = q + (x < r ? 1u : 0u);
        }
    }

    void upload(hipStream_t stream) {
        ensure(d_desc, cap_desc, desc.size());
        ensure(d_t256, cap_t256, t256.size());
        ensure(d_tsort, cap_tsort, tsort.size());
        ensure(d_ttopo, cap_ttopo, ttopo.size());
        ensure(d_twave, cap_twave, twave.size());
        size_t ctl_cap = d_ctl ? kCtlWords : 0;
        ensure(d_ctl, ctl_cap, kCtlWords);
        ensure(d_sflags, cap_sflags, 4ull * kRadix * tsort.size());
        ensure(d_tflags, cap_tflags, ttopo.size());
        ensure(d_tsum, cap_tsum, kTopoSum * ttopo.size());
        ensure(d_psum, cap_psum, kTopoSum * ttopo.size());
        prepare_errors();
        HIP_OK(hipMemcpyAsync(d_desc, desc.data(), sizeof(SimDesc) * desc.size(),
                              hipMemcpyHostToDevice, stream));
        auto put = [&](TileRef* dst, const std::vector<TileRef>& v) {
            if (!v.empty())
                HIP_OK(hipMemcpyAsync(dst, v.data(), sizeof(TileRef) * v.size(),
                                      hipMemcpyHostToDevice, stream));
        };
        put(d_t256, t256);
        put(d_tsort, tsort);
        put(d_ttopo, ttopo);
        put(d_twave, twave);
        HIP_OK(hipMemsetAsync(d_ctl, 0, sizeof(uint32_t) * kCtlWords, stream));
        HIP_OK(hipMemcpyAsync(d_ctl + kCtlMode, &mode, sizeof(mode), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemcpyAsync(d_ctl + kCtlRanges, ranges, sizeof(ranges), hipMemcpyHostToDevice, stream));
        HIP_OK(hipMemsetAsync(d_sflags, 0, sizeof(uint32_t) * 4ull * kRadix * std::max<size_t>(tsort.size(), 1), stream));
        HIP_OK(hipMemsetAsync(d_tflags, 0, sizeof(unsigned long long) * std::max<size_t>(ttopo.size(), 1), stream));
        // persistent traversal grid: trav_waves_per_cu wavefronts per CU, no more
        // workgroups than there are work items to share
        const uint32_t waves = (uint32_t)twave.size();
        const uint32_t per_block = trav_block / 64;
        const uint32_t want = std::max<uint32_t>(1, (uint32_t)dev->cu_count * trav_waves_per_cu / per_block);
        traverse_blocks = std::max<uint32_t>(1, std::min<uint32_t>(want, div_up(waves, per_block)));
    }

    void upload_desc(hipStream_t stream) {
        HIP_OK(hipMemcpyAsync(d_desc, desc.data(), sizeof(SimDesc) * desc.size(),
                              hipMemcpyHostToDevice, stream));
    }

    // ---- launches ----
    void bounds(int mode, hipStream_t s, bool clear_errors = false) {
        hipLaunchKernelGGL(k_bounds, dim3((uint32_t)t256.size()), dim3(kTile), 0, s, d_desc,
                           d_t256, d_ctl, d_tflags, (uint32_t)ttopo.size(), mode, clear_errors ? 1u : 0u);
        HIP_OK(hipGetLastError());
    }
    // Sticky error words of the set and its simulations back to zero (a kernel,
    // so it may also run where a caller captures the stream).
    void clear_errors(hipStream_t s) {
        const uint32_t nsims = (uint32_t)desc.size();
        hipLaunchKernelGGL(k_clear_errors, dim3(div_up(nsims, 256u)), dim3(256), 0, s, d_desc, nsims, d_ctl);
        HIP_OK(hipGetLastError());
    }
    void keys(hipStream_t s) {
        if (sort_kpt == 8)
            hipLaunchKernelGGL(k_keys<8>, dim3((uint32_t)tsort.size()), dim3(kSortThreads), 0, s,
                               d_desc, d_tsort, d_sflags, (uint32_t)tsort.size());
        else
            hipLaunchKernelGGL(k_keys<4>, dim3((uint32_t)tsort.size()), dim3(kSortThreads), 0, s,
                               d_desc, d_tsort, d_sflags, (uint32_t)tsort.size());
        HIP_OK(hipGetLastError());
    }
    void sort(hipStream_t s) {
        const dim3 grid((uint32_t)tsort.size()), block(kSortThreads);
        const uint32_t nt = (uint32_t)tsort.size();
        if (nt == 1) {  // one tile: the four passes in one launch
            if (sort_kpt == 8)
                hipLaunchKernelGGL((k_sort_single<8>), dim3(1), block, 0, s, d_desc, d_tsort, d_ctl, d_sflags);
            else
                hipLaunchKernelGGL((k_sort_single<4>), dim3(1), block, 0, s, d_desc, d_tsort, d_ctl, d_sflags);
            HIP_OK(hipGetLastError());
            return;
        }
        for (uint32_t pass = 0; pass < 4; ++pass) {
            if (sort_kpt == 8)
                hipLaunchKernelGGL((k_sort_pass<8>), grid, block, 0, s, d_desc, d_tsort, nt, pass, d_ctl, d_sflags);
            else
                hipLaunchKernelGGL((k_sort_pass<4>), grid, block, 0, s, d_desc, d_tsort, nt, pass, d_ctl, d_sflags);
            HIP_OK(hipGetLastError());
        }
    }
    // A single simulation of at most one topology tile: the whole topology and
    // the error gather in one launch (k_topo_single), no k_agg_top.
    bool single_topo() const { return desc.size() == 1 && ttopo.size() == 1; }
    void topo(hipStream_t s) {
        if (single_topo()) {
            hipLaunchKernelGGL(k_topo_single, dim3(1), dim3(kTopoTile), 0, s, d_desc, d_ttopo, d_ctl, d_tflags,
                               d_tsum, d_psum, d_errs);
            HIP_OK(hipGetLastError());
            return;
        }
        if (topo_tile == kTopoTileMulti) topo_tiled<kTopoTileMulti>(s);
        else topo_tiled<kTopoTile>(s);
    }
    // Tiles of kTT positions; the boundary pass and the column pass run them on
    // kTopoPassThreads threads (two positions per thread in 1,024-position
    // tiles), the link pass on kTT.
    template <uint32_t kTT>
    void topo_tiled(hipStream_t s) {
        constexpr uint32_t kPT = kTopoPassThreads;
        const dim3 grid((uint32_t)ttopo.size());
        hipLaunchKernelGGL((k_topo_pre<kTT, kPT>), grid, dim3(kPT), 0, s, d_desc, d_ttopo, d_tsum);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_topo_scan, dim3((uint32_t)desc.size(), 2 * kLevels + 1), dim3(64), 0, s, d_desc, d_tsum,
                           d_psum, kTT);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL((k_topo_count<kTT, kPT>), grid, dim3(kPT), 0, s, d_desc, d_ttopo, d_ctl, d_tflags,
                           d_psum);
        HIP_OK(hipGetLastError());
        hipLaunchKernelGGL(k_topo_link<kTT>, grid, dim3(kTT), 0, s, d_desc, d_ttopo);
        HIP_OK(hipGetLastError());
    }
    void agg(hipStream_t s) {
        if (single_topo()) return;  // done by k_topo_single
        // a set of one simulation stages more spanning cells in LDS (the small
        // footprint matters only beside another group's kernels)
        if (desc.size() == 1)
            hipLaunchKernelGGL((k_agg_top<kTopLdsCellsOne, kTopThreadsOne>), dim3(1), dim3(kTopThreadsOne), 0, s, d_desc, d_ctl,
                               d_errs);
        else
            hipLaunchKernelGGL((k_agg_top<kTopLdsCells, kTopThreads>), dim3((uint32_t)desc.size()), dim3(kTopThreads), 0, s,
                               d_desc, d_ctl, d_errs);
        HIP_OK(hipGetLastError());
    }
    void build(hipStream_t s) {
        keys(s);
        sort(s);
        topo(s);
        agg(s);
    }
    // Kick/drift (mode as in k_bounds) and the whole build, in launch order.  A
    // single simulation of at most one topology tile is one launch
    // (k_build_small); the per-phase launches above stay for phase timing.
    void bounds_build(int mode, hipStream_t s, bool clear_errors = false) {
        if (single_topo()) {
            hipLaunchKernelGGL(k_build_small, dim3(1), dim3(kTopoTile), 0, s, d_desc, d_ttopo, d_ctl, d_tflags,
                               d_tsum, d_psum, d_errs, mode, clear_errors ? 1u : 0u);
            HIP_OK(hipGetLastError());
            return;
        }
        bounds(mode, s, clear_errors);
        build(s);
    }
    void force(hipStream_t s, bool reset_queue, uint32_t* ev_stats = nullptr,
               unsigned long long* ev_hash = nullptr, uint32_t ev_sim = 0) {
        if (reset_queue) {
            hipLaunchKernelGGL(k_reset_queues, dim3(1), dim3(64), 0, s, d_ctl);
            HIP_OK(hipGetLastError());
        }
        if (ev_stats || ev_hash)
            hipLaunchKernelGGL(k_traverse_nodes<true>, dim3(traverse_blocks), dim3(trav_block), 0, s,
                               d_desc, d_twave, (uint32_t)twave.size(), d_ctl, ev_stats, ev_hash, ev_sim,
                               lane_targets);
        else
            hipLaunchKernelGGL(k_traverse_nodes<false>, dim3(traverse_blocks), dim3(trav_block), 0, s,
                               d_desc, d_twave, (uint32_t)twave.size(), d_ctl, ev_stats, ev_hash,
                               0xffffffffu, lane_targets);
        HIP_OK(hipGetLastError());
    }
    // body-node cold records for an evidence traversal (the step leaves them unwritten)
    void fill_body_cold(hipStream_t s) {
        hipLaunchKernelGGL(k_fill_body_cold, dim3((uint32_t)t256.size()), dim3(kTile), 0, s, d_desc, d_t256);
        HIP_OK(hipGetLastError());
        // and those of a lean build's internal cells
        bool lean = false;
        for (const SimDesc& d : desc) lean = lean || d.lean != 0u;
        if (lean) {
            hipLaunchKernelGGL(k_fill_cell_cold, dim3(64, (uint32_t)desc.size()), dim3(256), 0, s, d_desc);
            HIP_OK(hipGetLastError());
        }
    }
    void final_kick(hipStream_t s) {
        hipLaunchKernelGGL(k_final_kick, dim3((uint32_t)t256.size()), dim3(kTile), 0, s, d_desc, d_t256);
        HIP_OK(hipGetLastError());
    }
    void kick_drift(hipStream_t s) {
        hipLaunchKernelGGL(k_kick_drift, dim3((uint32_t)t256.size()), dim3(kTile), 0, s, d_desc, d_t256);
        HIP_OK(hipGetLastError());
    }

    // Sticky error words of the set and of each simulation.
    // One gather kernel and one copy, whatever the number of simulations.
    // enqueue_errors puts the gather and its copy into pinned host memory on the
    // stream; take_errors reads the words once the stream has been synchronized.
    // The error gather buffers (sized by upload, so nothing is allocated while a
    // caller captures the stream).
    void prepare_errors() {
        const uint32_t nsims = (uint32_t)desc.size();
        ensure(d_errs, cap_errs, 1ull + nsims);
        if (cap_herrs < 1ull + nsims) {
            if (h_errs) HIP_OK(hipHostFree(h_errs));
            h_errs = nullptr;
            cap_herrs = 0;
            HIP_OK(hipHostMalloc((void**)&h_errs, 4ull * (1ull + nsims), hipHostMallocDefault));
            cap_herrs = 1ull + nsims;
        }
    }
    void enqueue_errors(hipStream_t s) {
        const uint32_t nsims = (uint32_t)desc.size();
        prepare_errors();
        hipLaunchKernelGGL(k_collect_errors, dim3(div_up(nsims + 1u, 256u)), dim3(256), 0, s, d_desc, nsims,
                           d_ctl, d_errs);
        HIP_OK(hipGetLastError());
        copy_errors(s);
    }
    // Copy of the gathered words to pinned host memory: after a build, k_agg_top
    // has gathered them (no kernel after it in a build sets error words).
    void copy_errors(hipStream_t s) {
        HIP_OK(hipMemcpyAsync(h_errs, d_errs, 4ull * (1ull + desc.size()), hipMemcpyDeviceToHost, s));
    }
    uint32_t take_errors(std::vector<uint32_t>* per_sim = nullptr) const {
        const uint32_t nsims = (uint32_t)desc.size();
        uint32_t all = 0;
        for (uint32_t i = 0; i <= nsims; ++i) all |= h_errs[i];
        if (per_sim) {  // a set-wide error (word 0) taints every simulation of the set
            per_sim->resize(nsims);
            for (uint32_t i = 0; i < nsims; ++i) (*per_sim)[i] = h_errs[1 + i] | h_errs[0];
        }
        return all;
    }
    uint32_t read_errors(hipStream_t s, std::vector<uint32_t>* per_sim = nullptr) {
        enqueue_errors(s);
        HIP_OK(hipStreamSynchronize(s));
        return take_errors(per_sim);
    }
};

std::string error_text(uint32_t bits) {
    std::string s;
    if (bits & kErrLookback) s += "look-back wait exceeded its spin limit; ";
    if (bits & kErrCellCapacity) s += "cell capacity exceeded; ";
    if (bits & kErrSpanCapacity) s += "spanning-cell capacity exceeded; ";
    return s.empty() ? "unknown device error" : s;
}

bool finite_f32(double v, float* out) {
    float f = (float)v;
    if (!std::isfinite(f)) return false;
    *out = f;
    return true;
}

// Config::validate (nbody/src/lib.rs:48-67)
bool validate_config(const gbh_config* c) {
    if (!c) {
        set_error("config is null");
        return false;
    }
    if (!std::isfinite(c->theta) || c->theta < 0.0) {
        set_error("theta must be finite and >= 0");
        return false;
    }
    if (!std::isfinite(c->softening) || c->softening < 0.0) {
        set_error("softening_kpc must be finite and >= 0");
        return false;
    }
    if (!std::isfinite(c->g) || c->g <= 0.0) {
        set_error("g must be positive and finite");
        return false;
    }
    if (c->bucket == 0 || c->bucket > 1024) {
        set_error("bucket must be in 1..=1024");
        return false;
    }
    if (c->max_depth == 0 || c->max_depth > 128) {
        set_error("max_depth must be in 1..=128");
        return false;
    }
    return true;
}

// Physics fields of a SimDesc from a Config; false with the reference's
// message when they are not finite f32 (nbody_parallel.rs:848-850).
bool apply_config(SimDesc& d, const gbh_config* c, const char* what) {
    float th, ep, g;
    if (!finite_f32(c->theta, &th) || !finite_f32(c->softening, &ep) || !finite_f32(c->g, &g)) {
        set_error(std::string(what));
        return false;
    }
    d.theta = th;
    d.eps = ep;
    d.G = g;
    d.bucket = c->bucket;
    d.maxd = std::min<uint32_t>(c->max_depth, 16u);
    return true;
}

bool validate_states(const GbhEvolveState* s, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* pm = s[i].position_mass;
        const float* v = s[i].velocity;
        if (!std::isfinite(pm[0]) || !std::isfinite(pm[1]) || !std::isfinite(v[0]) ||
            !std::isfinite(v[1]) || !std::isfinite(pm[2]) || pm[2] <= 0.0f) {
            set_error("body " + std::to_string(i) + " contains invalid evolving state");
            return false;
        }
    }
    return true;
}

// Upload frozen records into SoA state through a staging buffer.
// Device staging buffer for the record <-> SoA state transfers, grown on demand
// and kept by its owner (hipMalloc is slow and hipFree synchronizes the device,
// so a server does neither per request).
struct Staging {
    void* p = nullptr;
    size_t cap = 0;
    void* get(size_t bytes) {
        if (bytes > cap) {
            if (p) HIP_OK(hipFree(p));
            p = nullptr;
            cap = 0;
            HIP_OK(hipMalloc(&p, bytes));
            cap = bytes;
        }
        return p;
    }
    ~Staging() {
        if (p) (void)hipFree(p);
    }
};

void upload_state(const SimDesc& d, const GbhEvolveState* host, uint32_t n, hipStream_t s,
                  Staging* keep = nullptr) {
    Staging local;
    GbhEvolveState* staging = (GbhEvolveState*)(keep ? keep : &local)->get(sizeof(GbhEvolveState) * n);
    HIP_OK(hipMemcpyAsync(staging, host, sizeof(GbhEvolveState) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_unpack_state, dim3(div_up(n, 256)), dim3(256), 0, s, staging, n, d.pos,
                       d.mass, d.vel, d.acc);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(s));
}

void download_state(const SimDesc& d, GbhEvolveState* out, GbhGpuAccel* acc, uint32_t n,
                    hipStream_t s, Staging* keep = nullptr) {
    size_t bytes = std::max(sizeof(GbhEvolveState), sizeof(GbhGpuAccel)) * (size_t)n;
    Staging local;
    void* staging = (keep ? keep : &local)->get(bytes);
    if (out) {
        hipLaunchKernelGGL(k_pack_state, dim3(div_up(n, 256)), dim3(256), 0, s,
                           (GbhEvolveState*)staging, n, d.pos, d.mass, d.vel);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out, staging, sizeof(GbhEvolveState) * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
    if (acc) {
        hipLaunchKernelGGL(k_pack_accel, dim3(div_up(n, 256)), dim3(256), 0, s,
                           (GbhGpuAccel*)staging, n, d.acc);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(acc, staging, sizeof(GbhGpuAccel) * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
    }
}

// Host copy of a built tree: everything evidence readbacks need.
struct TreeHost {
    uint32_t n = 0, ncells = 0;
    std::vector<CellT> cells;
    std::vector<float4> geom;
    std::vector<uint4> child;
    std::vector<uint32_t> key, val, pos_of;
    SimMeta meta;
};

void read_tree(const SimDesc& d, TreeHost& h, hipStream_t s) {
    h.n = d.n;
    HIP_OK(hipMemcpyAsync(&h.meta, d.meta, sizeof(SimMeta), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    if (h.meta.err) throw HipError{"tree build failed: " + error_text(h.meta.err)};
    h.ncells = h.meta.ncells;
    if (h.ncells == 0 || h.ncells > d.cell_cap) throw HipError{"tree has no valid cells"};
    h.cells.resize(h.ncells);
    h.geom.resize(h.ncells);
    h.child.resize(h.ncells);
    h.key.resize(d.n);
    h.val.resize(d.n);
    h.pos_of.resize(d.n);
    HIP_OK(hipMemcpyAsync(h.cells.data(), d.cells, sizeof(CellT) * h.ncells, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h.child.data(), d.cchild, 16ull * h.ncells, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h.key.data(), d.key0, 4ull * d.n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(h.val.data(), d.val0, 4ull * d.n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    // target positions (bh_parallel_tree_positions): the inverse of the sort permutation
    for (uint32_t p = 0; p < d.n; ++p) {
        if (h.val[p] >= d.n) throw HipError{"tree has an invalid sort permutation"};
        h.pos_of[h.val[p]] = p;
    }
    // cell geometry: the f64 root box halved along the quadrant path of the
    // cell's start key to its depth (child_bounds, flat.rs:96-103), rounded once
    for (uint32_t k = 0; k < h.ncells; ++k) {
        const CellT& c = h.cells[k];
        const uint32_t depth = c.info & 255u;
        if (c.info & kInfoLeaf) h.child[k] = make_uint4(kNoChild, kNoChild, kNoChild, kNoChild);  // not stored
        if (c.start >= d.n || depth >= kLevels) throw HipError{"tree has an invalid cell"};
        const uint32_t key = h.key[c.start];
        double cx = h.meta.root[0], cy = h.meta.root[1], hh = h.meta.root[2];
        for (uint32_t l = 0; l < depth; ++l) {
            const uint32_t q = (key >> (30u - 2u * l)) & 3u;
            hh = hh * 0.5;
            cx = cx + ((q & 1u) ? hh : -hh);
            cy = cy + ((q & 2u) ? hh : -hh);
        }
        h.geom[k] = make_float4((float)cx, (float)cy, (float)hh, 0.0f);
    }
}

void flat_cell(const TreeHost& h, uint32_t k, GbhGpuCell& c) {
    const CellT& t = h.cells[k];
    const float4 g = h.geom[k];
    c.center_half_mass[0] = g.x;
    c.center_half_mass[1] = g.y;
    c.center_half_mass[2] = g.z;
    c.center_half_mass[3] = t.mass;
    c.com[0] = t.comx;
    c.com[1] = t.comy;
    c.com[2] = 0.f;
    c.com[3] = 0.f;
    c.range_depth[0] = t.start;
    c.range_depth[1] = t.end;
    c.range_depth[2] = t.info & 255u;
    c.range_depth[3] = 0;
    const uint4 ch = h.child[k];
    c.children[0] = ch.x;
    c.children[1] = ch.y;
    c.children[2] = ch.z;
    c.children[3] = ch.w;
}

}  // namespace gbh

// ============================================================================
// drop-in objects
// ============================================================================
struct gbh_state {
    gbh_device* dev = nullptr;
    uint32_t n = 0;
    char* arena = nullptr;
    size_t bytes = 0;
    SimSet kdk;  // one-simulation set over the state (kick/drift, final kick)
};

struct gbh_runtime {
    gbh_device* dev = nullptr;
};

struct gbh_tree {
    gbh_runtime* rt = nullptr;
    gbh_state* st = nullptr;
    uint32_t n = 0;
    char* arena = nullptr;
    size_t bytes = 0;
    bool built = false;
    // The traversal queue counter is zero: a rebuild's k_bounds cleared it and
    // no traversal ran since, so the next force needs no reset kernel.
    bool queue_zero = false;
    SimSet set;
    hipEvent_t ev[8] = {};
    double last_force_ms = 0.0;
    hipGraphExec_t rebuild_graph = nullptr;  // the build, captured at the first rebuild
};

extern "C" {

const char* gbh_last_error(void) { return g_error.c_str(); }

int gbh_device_open(const char* selector, int allow_software, gbh_device** out) {
    (void)allow_software;
    return guarded([&] {
        if (!out) {
            set_error("output handle is null");
            return -1;
        }
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
            set_error("No matching hardware compute adapter. --allow-software permits a software adapter for validation.");
            return -1;
        }
        int index = -1;
        if (selector && *selector) {
            char* end = nullptr;
            long v = strtol(selector, &end, 10);
            if (end && *end == 0 && v >= 0 && v < count) index = (int)v;
            if (index < 0) {
                std::string q(selector);
                std::transform(q.begin(), q.end(), q.begin(), ::tolower);
                for (int i = 0; i < count && index < 0; ++i) {
                    hipDeviceProp_t p;
                    HIP_OK(hipGetDeviceProperties(&p, i));
                    std::string nm = std::string(p.name) + " " + p.gcnArchName;
                    std::transform(nm.begin(), nm.end(), nm.begin(), ::tolower);
                    if (nm.find(q) != std::string::npos) index = i;
                }
            }
        } else {
            index = 0;
        }
        if (index < 0) {
            set_error("No matching hardware compute adapter. --allow-software permits a software adapter for validation.");
            return -1;
        }
        HIP_OK(hipSetDevice(index));
        // this device's copy of the no-allocation scratch (the device is current)
        unsigned char* scratch_ptr = nullptr;
        HIP_OK(hipGetSymbolAddress((void**)&scratch_ptr, HIP_SYMBOL(gbh::scratch)));
        auto* d = new gbh_device();
        d->index = index;
        d->count = count;
        d->scratch = scratch_ptr;
        HIP_OK(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
        hipDeviceProp_t p;
        HIP_OK(hipGetDeviceProperties(&p, index));
        d->cu_count = p.multiProcessorCount;
        int rt_ver = 0, drv_ver = 0;
        (void)hipRuntimeGetVersion(&rt_ver);
        (void)hipDriverGetVersion(&drv_ver);
        // adapter name: the marketing name when the runtime reports one, else the ISA
        std::string arch(p.gcnArchName);
        std::string name(p.name);
        if (name.empty()) name = "AMD " + arch.substr(0, arch.find(':'));
        for (char& ch : name)
            if (ch == '"' || ch == '\\') ch = ' ';
        char buf[1024];
        snprintf(buf, sizeof buf,
                 "{\"index\": %d, \"adapter_count\": %d, \"name\": \"%s\", \"backend\": \"ROCm-HIP\", "
                 "\"device_type\": \"DiscreteGpu\", \"driver\": \"HIP %d\", \"driver_info\": \"%s, "
                 "%d CUs, driver %d\", \"software\": false}",
                 index, count, name.c_str(), rt_ver, p.gcnArchName, p.multiProcessorCount, drv_ver);
        d->info = buf;
        *out = d;
        return 0;
    });
}

void gbh_device_close(gbh_device* dev) {
    if (!dev) return;
    if (dev->stream) (void)hipStreamDestroy(dev->stream);
    delete dev;
}

int gbh_device_info_json(gbh_device* dev, char* buf, size_t cap, size_t* len) {
    if (!dev) {
        set_error("device is null");
        return -1;
    }
    if (len) *len = dev->info.size();
    if (buf && cap) {
        size_t k = std::min(cap - 1, dev->info.size());
        memcpy(buf, dev->info.data(), k);
        buf[k] = 0;
    }
    return 0;
}

// ---------------------------------------------------------------------------- state
int gbh_state_create(gbh_device* dev, const GbhEvolveState* states, uint32_t n, gbh_state** out,
                     gbh_stage_timing* timing) {
    return guarded([&] {
        if (!dev || !states || !out) {
            set_error("null argument");
            return -1;
        }
        if (n < 2) {
            set_error("GPU evolving state requires 2..=u32::MAX resident bodies");
            return -1;
        }
        if (!validate_states(states, n)) return -1;
        double t0 = now_seconds();
        auto st = std::make_unique<gbh_state>();
        st->dev = dev;
        st->n = n;
        SimDesc d{};
        d.n = n;
        st->bytes = carve(d, nullptr, n, false);
        HIP_OK(hipMalloc((void**)&st->arena, st->bytes));
        carve(d, st->arena, n, false);
        st->kdk.dev = dev;
        st->kdk.desc = {d};
        st->kdk.build_tables();
        st->kdk.upload(dev->stream);
        upload_state(d, states, n, dev->stream);
        if (timing) *timing = gbh_stage_timing{now_seconds() - t0, 0, 0};
        *out = st.release();
        return 0;
    });
}

void gbh_state_destroy(gbh_state* st) {
    if (!st) return;
    if (st->arena) (void)hipFree(st->arena);
    delete st;
}

uint32_t gbh_state_count(const gbh_state* st) { return st ? st->n : 0; }

uint64_t gbh_state_bytes(const gbh_state* st) {
    // state + acceleration records of the reference layout (32 + 16 bytes/body)
    return st ? (uint64_t)st->n * 48ull : 0;
}

int gbh_state_read(gbh_state* st, GbhEvolveState* out, gbh_stage_timing* timing) {
    return guarded([&] {
        if (!st || !out) {
            set_error("null argument");
            return -1;
        }
        double t0 = now_seconds();
        download_state(st->kdk.desc[0], out, nullptr, st->n, st->dev->stream);
        for (uint32_t i = 0; i < st->n; ++i) {
            const float* pm = out[i].position_mass;
            const float* v = out[i].velocity;
            if (!std::isfinite(pm[0]) || !std::isfinite(pm[1]) || !std::isfinite(v[0]) ||
                !std::isfinite(v[1]) || !std::isfinite(pm[2]) || pm[2] <= 0.0f) {
                set_error("GPU produced invalid evolving body state at " + std::to_string(i));
                return -1;
            }
        }
        if (timing) *timing = gbh_stage_timing{0, 0, now_seconds() - t0};
        return 0;
    });
}

int gbh_accel_read(gbh_state* st, GbhGpuAccel* out, gbh_stage_timing* timing) {
    return guarded([&] {
        if (!st || !out) {
            set_error("null argument");
            return -1;
        }
        double t0 = now_seconds();
        download_state(st->kdk.desc[0], nullptr, out, st->n, st->dev->stream);
        for (uint32_t i = 0; i < st->n; ++i) {
            if (!std::isfinite(out[i].value[0]) || !std::isfinite(out[i].value[1])) {
                set_error("GPU produced invalid evolving acceleration at " + std::to_string(i));
                return -1;
            }
        }
        if (timing) *timing = gbh_stage_timing{0, 0, now_seconds() - t0};
        return 0;
    });
}

static int integrate(gbh_state* st, double dt, bool final_kick, gbh_stage_timing* timing) {
    return guarded([&] {
        if (!st) {
            set_error("state is null");
            return -1;
        }
        if (!std::isfinite(dt) || dt <= 0.0) {
            set_error("dt_myr must be positive and finite");
            return -1;
        }
        float dtf;
        if (!finite_f32(dt, &dtf)) {
            set_error("dt_myr cannot be represented as finite f32");
            return -1;
        }
        hipStream_t s = st->dev->stream;
        double t0 = now_seconds();
        SimSet& set = st->kdk;
        if (set.desc[0].dt != dtf) {
            set.desc[0].dt = dtf;
            set.upload_desc(s);
        }
        double t1 = now_seconds();
        if (final_kick) set.final_kick(s); else set.kick_drift(s);
        HIP_OK(hipStreamSynchronize(s));
        if (timing) *timing = gbh_stage_timing{t1 - t0, now_seconds() - t1, 0};
        return 0;
    });
}

int gbh_kick_drift(gbh_state* st, double dt, gbh_stage_timing* timing) {
    return integrate(st, dt, false, timing);
}

int gbh_final_kick(gbh_state* st, double dt, gbh_stage_timing* timing) {
    return integrate(st, dt, true, timing);
}

// ---------------------------------------------------------------------------- runtime/tree
int gbh_runtime_create(gbh_device* dev, gbh_runtime** out) {
    return guarded([&] {
        if (!dev || !out) {
            set_error("null argument");
            return -1;
        }
        auto* rt = new gbh_runtime();
        rt->dev = dev;
        *out = rt;
        return 0;
    });
}

void gbh_runtime_destroy(gbh_runtime* rt) { delete rt; }

static int tree_create(gbh_runtime* rt, gbh_state* st, gbh_tree** out, uint32_t cap) {
    return guarded([&] {
        if (!rt || !st || !out) {
            set_error("null argument");
            return -1;
        }
        if (st->n < 2 || st->n > cap) {
            set_error(cap == GBH_MAX_PARALLEL_TREE_BODIES
                          ? "BH #2D parallel GPU tree supports 2..=65536 resident bodies"
                          : "large GPU tree supports 2..=" + std::to_string(cap) + " resident bodies");
            return -1;
        }
        auto t = std::make_unique<gbh_tree>();
        t->rt = rt;
        t->st = st;
        t->n = st->n;
        SimDesc d = st->kdk.desc[0];
        SimDesc tmp{};
        t->bytes = carve(tmp, nullptr, st->n, true);
        HIP_OK(hipMalloc((void**)&t->arena, t->bytes));
        carve(tmp, t->arena, st->n, true);
        // state pointers stay the state's own; tree buffers come from the tree arena
        tmp.pos = d.pos;
        tmp.mass = d.mass;
        tmp.vel = d.vel;
        tmp.acc = d.acc;
        tmp.n = st->n;
        tmp.dt = d.dt;
        t->set.dev = rt->dev;
        t->set.desc = {tmp};
        t->set.build_tables();
        t->set.upload(rt->dev->stream);
        HIP_OK(hipMemsetAsync(tmp.meta, 0, sizeof(SimMeta), rt->dev->stream));
        // timing-only events (read after a stream synchronize): no system-scope
        // fence at each record, which would stall the stream between stages
        for (auto& e : t->ev) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        HIP_OK(hipStreamSynchronize(rt->dev->stream));
        *out = t.release();
        return 0;
    });
}

int gbh_tree_create(gbh_runtime* rt, gbh_state* st, gbh_tree** out) {
    return tree_create(rt, st, out, GBH_MAX_PARALLEL_TREE_BODIES);
}

int gbh_tree_create_large(gbh_runtime* rt, gbh_state* st, gbh_tree** out) {
    return tree_create(rt, st, out, GBH_MAX_LARGE_TREE_BODIES);
}

void gbh_tree_destroy(gbh_tree* t) {
    if (!t) return;
    for (auto& e : t->ev)
        if (e) (void)hipEventDestroy(e);
    if (t->rebuild_graph) (void)hipGraphExecDestroy(t->rebuild_graph);
    if (t->arena) (void)hipFree(t->arena);
    delete t;
}

static float ms_between(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

int gbh_tree_rebuild(gbh_runtime* rt, gbh_state* st, gbh_tree* t, const gbh_config* cfg,
                     gbh_build_timing* timing) {
    return guarded([&] {
        if (!rt || !st || !t) {
            set_error("null argument");
            return -1;
        }
        if (t->n != st->n || t->st != st) {
            set_error("BH #2D tree count does not match evolving state");
            return -1;
        }
        if (!validate_config(cfg)) return -1;
        SimDesc& d = t->set.desc[0];
        SimDesc prev = d;
        if (!apply_config(d, cfg, "BH #2D physics parameters must remain finite")) return -1;
        hipStream_t s = rt->dev->stream;
        if (memcmp(&prev, &d, sizeof d) != 0) t->set.upload_desc(s);
        double t0 = now_seconds();
        // a failed rebuild leaves no usable tree until the next one succeeds
        t->built = false;
        // The whole build is one HIP graph, captured at the first rebuild: bounds
        // (which also clears the previous rebuild's error words and the traversal
        // queue), keys, the radix passes, topology, top aggregates with the error
        // gather, and the copy of the error words to pinned host memory (a tree
        // of at most 1,024 bodies: one fused kernel and the copy).  The kernels
        // read the configuration from the simulation table, so a changed
        // configuration needs no new capture.
        if (!t->rebuild_graph) {
            hipGraph_t gr = nullptr;
            HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            t->set.bounds_build(0, s, true);
            t->set.copy_errors(s);
            HIP_OK(hipStreamEndCapture(s, &gr));
            const hipError_t e = hipGraphInstantiate(&t->rebuild_graph, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            HIP_OK(e);
        }
        HIP_OK(hipGraphLaunch(t->rebuild_graph, s));
        t->queue_zero = true;
        HIP_OK(hipStreamSynchronize(s));
        uint32_t err = t->set.take_errors();
        if (err) {
            set_error("BH #2D rebuild failed on the device: " + error_text(err));
            return -1;
        }
        t->built = true;
        // the stages run as one graph: the call's wall time is reported as the
        // bounds stage, the slot of the first stage
        if (timing) *timing = gbh_build_timing{now_seconds() - t0, 0.0, 0.0, 0.0, 0.0, 0.0};
        return 0;
    });
}

// noalloc: stage the evidence in the device's scratch array instead of device
// memory allocated for the call (n <= GBH_NOALLOC_MAX_BODIES).
static int force_impl(gbh_runtime* rt, gbh_state* st, gbh_tree* t, const gbh_config* cfg,
                      gbh_stage_timing* timing, uint32_t* stats, uint64_t* hash,
                      bool noalloc = false) {
    return guarded([&] {
        if (!rt || !st || !t) {
            set_error("null argument");
            return -1;
        }
        if (noalloc && t->n > kScratchTargets) {
            set_error("gbh_force_evidence_noalloc supports at most " + std::to_string(kScratchTargets) +
                      " bodies; this tree has " + std::to_string(t->n));
            return -1;
        }
        if (noalloc && !rt->dev->scratch) {
            set_error("the device has no scratch address");
            return -1;
        }
        if (t->n != st->n || t->st != st) {
            set_error("BH #2D force/tree count does not match evolving state");
            return -1;
        }
        if (!validate_config(cfg)) return -1;
        if (!t->built) {
            set_error("BH #2D tree has not been built");
            return -1;
        }
        hipStream_t s = rt->dev->stream;
        double t0 = now_seconds();
        SimDesc& d = t->set.desc[0];
        float th, ep, g;
        if (!finite_f32(cfg->theta, &th) || !finite_f32(cfg->softening, &ep) ||
            !finite_f32(cfg->g, &g)) {
            set_error("BH #2D force settings are not representable as finite f32");
            return -1;
        }
        bool theta_changed = th != d.theta;
        if (th != d.theta || ep != d.eps || g != d.G) {
            d.theta = th;
            d.eps = ep;
            d.G = g;
            t->set.upload_desc(s);
        }
        if (theta_changed) {
            hipLaunchKernelGGL(k_set_theta, dim3(1), dim3(64), 0, s, t->set.d_desc, 1u);
            HIP_OK(hipGetLastError());
            hipLaunchKernelGGL(k_rewrite_T, dim3(256), dim3(256), 0, s, t->set.d_desc, 0u);
            HIP_OK(hipGetLastError());
        }
        double t1 = now_seconds();
        // device staging of the evidence: allocated for this call (freed on every
        // exit), or the device's scratch array
        struct Owned {
            void* p = nullptr;
            ~Owned() {
                if (p) (void)hipFree(p);
            }
        } own_stats, own_hash;
        uint32_t* d_stats = nullptr;
        unsigned long long* d_hash = nullptr;
        if (noalloc) {
            if (stats) d_stats = (uint32_t*)rt->dev->scratch;
            if (hash) d_hash = (unsigned long long*)(rt->dev->scratch + kScratchStatsBytes);
        } else {
            if (stats) HIP_OK(hipMalloc(&own_stats.p, 12ull * t->n));
            if (hash) HIP_OK(hipMalloc(&own_hash.p, 8ull * t->n));
            d_stats = (uint32_t*)own_stats.p;
            d_hash = (unsigned long long*)own_hash.p;
        }
        if (d_stats || d_hash) t->set.fill_body_cold(s);
        HIP_OK(hipEventRecord(t->ev[6], s));
        t->set.force(s, !t->queue_zero, d_stats, d_hash);
        t->queue_zero = false;
        HIP_OK(hipEventRecord(t->ev[7], s));
        HIP_OK(hipStreamSynchronize(s));
        double t2 = now_seconds();
        t->last_force_ms = ms_between(t->ev[6], t->ev[7]);
        if (stats) HIP_OK(hipMemcpyAsync(stats, d_stats, 12ull * t->n, hipMemcpyDeviceToHost, s));
        if (hash) HIP_OK(hipMemcpyAsync(hash, d_hash, 8ull * t->n, hipMemcpyDeviceToHost, s));
        if (stats || hash) HIP_OK(hipStreamSynchronize(s));
        if (timing) *timing = gbh_stage_timing{t1 - t0, t2 - t1, 0};
        return 0;
    });
}

int gbh_force(gbh_runtime* rt, gbh_state* st, gbh_tree* t, const gbh_config* cfg,
              gbh_stage_timing* timing) {
    return force_impl(rt, st, t, cfg, timing, nullptr, nullptr);
}

int gbh_force_evidence(gbh_runtime* rt, gbh_state* st, gbh_tree* t, const gbh_config* cfg,
                       uint32_t* stats, uint64_t* hash) {
    if (!stats && !hash) {
        set_error("evidence outputs are both null");
        return -1;
    }
    return force_impl(rt, st, t, cfg, nullptr, stats, hash);
}

int gbh_force_evidence_noalloc(gbh_runtime* rt, gbh_state* st, gbh_tree* t, const gbh_config* cfg,
                               uint32_t* stats, uint64_t* hash) {
    if (!stats && !hash) {
        set_error("evidence outputs are both null");
        return -1;
    }
    return force_impl(rt, st, t, cfg, nullptr, stats, hash, true);
}

int gbh_tree_counts(gbh_tree* t, uint32_t* active, uint32_t* leaves, uint32_t* max_depth) {
    return guarded([&] {
        if (!t || !t->built) {
            set_error("BH #2D tree has not been built");
            return -1;
        }
        TreeHost h;
        read_tree(t->set.desc[0], h, t->rt->dev->stream);
        uint32_t lf = 0, md = 0;
        for (const CellT& c : h.cells) {
            lf += (c.info & kInfoLeaf) ? 1u : 0u;
            md = std::max(md, c.info & 255u);
        }
        if (active) *active = h.ncells;
        if (leaves) *leaves = lf;
        if (max_depth) *max_depth = md;
        return 0;
    });
}

int gbh_tree_read_evidence(gbh_tree* t, GbhTreeMeta* meta, GbhGpuEntry* entries, GbhGpuCell* cells,
                           uint32_t cell_capacity, uint32_t* active, uint32_t* leaves,
                           uint32_t* max_depth) {
    return guarded([&] {
        if (!t || !t->built) {
            set_error("BH #2D tree has not been built");
            return -1;
        }
        TreeHost h;
        read_tree(t->set.desc[0], h, t->rt->dev->stream);
        const uint32_t n = h.n;
        if (cells && cell_capacity < h.ncells) {
            set_error("evidence cell buffer holds fewer records than active cells");
            return -1;
        }
        if (meta) {
            memset(meta, 0, sizeof *meta);
            for (int i = 0; i < 4; ++i) meta->bounds[i] = h.meta.bounds_f32[i];
        }
        if (entries)
            for (uint32_t p = 0; p < n; ++p) {
                entries[p].data[0] = h.key[p];
                entries[p].data[1] = h.val[p];
                entries[p].data[2] = entries[p].data[3] = 0;
            }
        uint32_t lf = 0, md = 0;
        if (cells) {
            std::vector<std::pair<uint64_t, uint32_t>> order(h.ncells);
            for (uint32_t k = 0; k < h.ncells; ++k)
                order[k] = {(uint64_t)(h.cells[k].info & 255u) * n + h.cells[k].start, k};
            std::sort(order.begin(), order.end());
            for (uint32_t i = 0; i < h.ncells; ++i) {
                const uint32_t k = order[i].second;
                GbhGpuCell& c = cells[i];
                flat_cell(h, k, c);
                c.range_depth[3] = 1;
                for (int q = 0; q < 4; ++q) {
                    uint32_t ch = c.children[q];
                    c.children[q] = ch == kNoChild ? kNoChild
                                                   : (h.cells[ch].info & 255u) * n + h.cells[ch].start;
                }
            }
        }
        for (const CellT& c : h.cells) {
            lf += (c.info & kInfoLeaf) ? 1u : 0u;
            md = std::max(md, c.info & 255u);
        }
        if (active) *active = h.ncells;
        if (leaves) *leaves = lf;
        if (max_depth) *max_depth = md;
        return 0;
    });
}

int gbh_tree_read_flat(gbh_tree* t, GbhGpuEntry* entries, GbhGpuCell* cells, uint32_t cell_capacity,
                       uint32_t* ncells, uint32_t* position_by_body) {
    return guarded([&] {
        if (!t || !t->built) {
            set_error("BH #2D tree has not been built");
            return -1;
        }
        TreeHost h;
        read_tree(t->set.desc[0], h, t->rt->dev->stream);
        if (ncells) *ncells = h.ncells;
        if (cells) {
            if (cell_capacity < h.ncells) {
                set_error("flat cell buffer holds fewer records than active cells");
                return -1;
            }
            for (uint32_t k = 0; k < h.ncells; ++k) flat_cell(h, k, cells[k]);
        }
        if (entries)
            for (uint32_t p = 0; p < h.n; ++p) {
                entries[p].data[0] = h.key[p];
                entries[p].data[1] = h.val[p];
                entries[p].data[2] = entries[p].data[3] = 0;
            }
        if (position_by_body) memcpy(position_by_body, h.pos_of.data(), 4ull * h.n);
        return 0;
    });
}

double gbh_tree_last_force_ms(gbh_tree* t) { return t ? t->last_force_ms : 0.0; }

int gbh_copy_calibration(gbh_device* dev, uint64_t bytes, double* ms) {
    return guarded([&] {
        if (!dev || !ms || bytes < 16) {
            set_error("copy calibration needs a device, an output and >= 16 bytes");
            return -1;
        }
        hipStream_t s = dev->stream;
        const size_t n = bytes / 16;
        float4 *src = nullptr, *dst = nullptr;
        HIP_OK(hipMalloc((void**)&src, n * 16));
        HIP_OK(hipMalloc((void**)&dst, n * 16));
        HIP_OK(hipMemsetAsync(src, 0, n * 16, s));
        hipEvent_t a, b;
        HIP_OK(hipEventCreate(&a));
        HIP_OK(hipEventCreate(&b));
        HIP_OK(hipEventRecord(a, s));
        hipLaunchKernelGGL(k_copy_calibration, dim3(8192), dim3(256), 0, s, src, dst, n);
        HIP_OK(hipGetLastError());
        HIP_OK(hipEventRecord(b, s));
        HIP_OK(hipStreamSynchronize(s));
        *ms = ms_between(a, b);
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        HIP_OK(hipFree(src));
        HIP_OK(hipFree(dst));
        return 0;
    });
}

int gbh_diag_direct(gbh_state* st, const gbh_config* cfg, uint32_t count, const uint32_t* index,
                    double* ax, double* ay) {
    return guarded([&] {
        if (!st || !cfg || (count && (!index || !ax || !ay))) {
            set_error("null argument");
            return -1;
        }
        if (!validate_config(cfg)) return -1;
        for (uint32_t k = 0; k < count; ++k)
            if (index[k] >= st->n) {
                set_error("body index is out of range");
                return -1;
            }
        if (count == 0) return 0;
        hipStream_t s = st->dev->stream;
        const SimDesc& d = st->kdk.desc[0];
        uint32_t* d_idx = nullptr;
        double* d_out = nullptr;
        HIP_OK(hipMalloc((void**)&d_idx, 4ull * count));
        HIP_OK(hipMalloc((void**)&d_out, 16ull * count));
        HIP_OK(hipMemcpyAsync(d_idx, index, 4ull * count, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_diag_direct, dim3(div_up(count, 256)), dim3(256), 0, s, d.pos, d.mass,
                           st->n, d_idx, count, cfg->softening * cfg->softening, cfg->g, d_out,
                           d_out + count);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(ax, d_out, 8ull * count, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(ay, d_out + count, 8ull * count, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        HIP_OK(hipFree(d_idx));
        HIP_OK(hipFree(d_out));
        return 0;
    });
}

int gbh_diag_energy(gbh_state* st, double softening, double g, double out[3]) {
    return guarded([&] {
        if (!st || !out) {
            set_error("null argument");
            return -1;
        }
        if (!std::isfinite(softening) || !std::isfinite(g)) {
            set_error("softening and G must be finite");
            return -1;
        }
        hipStream_t s = st->dev->stream;
        const SimDesc& d = st->kdk.desc[0];
        const uint32_t n = st->n;
        double* d_pot = nullptr;
        HIP_OK(hipMalloc((void**)&d_pot, 8ull * n));
        hipLaunchKernelGGL(k_diag_potential, dim3(div_up(n, 256)), dim3(256), 0, s, d.pos, d.mass, n,
                           softening * softening, g, d_pot);
        HIP_OK(hipGetLastError());
        std::vector<double> pot(n);
        std::vector<float2> pos(n), vel(n);
        std::vector<float> mass(n);
        HIP_OK(hipMemcpyAsync(pot.data(), d_pot, 8ull * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(pos.data(), d.pos, 8ull * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(vel.data(), d.vel, 8ull * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(mass.data(), d.mass, 4ull * n, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        HIP_OK(hipFree(d_pot));
        double kin = 0.0, potential = 0.0, lz = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            const double m = mass[i], x = pos[i].x, y = pos[i].y, vx = vel[i].x, vy = vel[i].y;
            kin += 0.5 * m * (vx * vx + vy * vy);
            potential += pot[i];
            lz += m * (x * vy - y * vx);
        }
        out[0] = kin + potential;
        out[1] = lz;
        out[2] = kin;
        return 0;
    });
}

int gbh_check_thresholds(gbh_device* dev, const float* half, const float* theta, uint32_t count, float* out) {
    return guarded([&] {
        if (!dev || (count && (!half || !theta || !out))) {
            set_error("null argument");
            return -1;
        }
        if (count == 0) return 0;
        hipStream_t s = dev->stream;
        float* d = nullptr;
        HIP_OK(hipMalloc((void**)&d, 12ull * count));
        HIP_OK(hipMemcpyAsync(d, half, 4ull * count, hipMemcpyHostToDevice, s));
        HIP_OK(hipMemcpyAsync(d + count, theta, 4ull * count, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_check_thresholds, dim3(div_up(count, 256u)), dim3(256), 0, s, d, d + count, count,
                           d + 2ull * count);
        const hipError_t e = hipGetLastError();
        if (e == hipSuccess) (void)hipMemcpyAsync(out, d + 2ull * count, 4ull * count, hipMemcpyDeviceToHost, s);
        const hipError_t e2 = hipStreamSynchronize(s);
        (void)hipFree(d);
        HIP_OK(e);
        HIP_OK(e2);
        return 0;
    });
}

int gbh_check_rounding(gbh_device* dev, uint64_t mismatches[3], uint32_t first_bad[3]) {
    return guarded([&] {
        if (!dev || !mismatches || !first_bad) {
            set_error("null argument");
            return -1;
        }
        hipStream_t s = dev->stream;
        unsigned long long* counts = nullptr;
        uint32_t* first = nullptr;
        HIP_OK(hipMalloc((void**)&counts, 24));
        HIP_OK(hipMalloc((void**)&first, 12));
        HIP_OK(hipMemsetAsync(counts, 0, 24, s));
        HIP_OK(hipMemsetAsync(first, 0xff, 12, s));
        hipLaunchKernelGGL(k_check_rounding, dim3(4096), dim3(256), 0, s, counts, first);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(mismatches, counts, 24, hipMemcpyDeviceToHost, s));
        HIP_OK(hipMemcpyAsync(first_bad, first, 12, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        HIP_OK(hipFree(counts));
        HIP_OK(hipFree(first));
        return 0;
    });
}

uint64_t gbh_tree_buffer_bytes(gbh_tree* t) { return t ? (uint64_t)t->bytes : 0; }

int gbh_fnv_mix64(uint64_t seed, const uint32_t* words, uint64_t count, uint64_t* out) {
    if (!out || (count && !words)) {
        set_error("null argument");
        return -1;
    }
    // checksum_word (galaxy-bh-gpu-tree-parallel.rs:145-149) over each word widened to 64 bits
    uint64_t h = seed;
    for (uint64_t i = 0; i < count; ++i) {
        const uint64_t w = words[i];
        h ^= w;
        h *= 0x00000100000001b3ull;
        h ^= (w << 23) | (w >> 41);
    }
    *out = h;
    return 0;
}

uint64_t gbh_sparse_layout_bytes(uint32_t n) {
    // nbody_parallel.rs:485-491: bodies + 2 entry buffers + 17N cells + meta
    // + 2 bounds scratch + histogram + offsets
    uint64_t blocks = (n + 127u) / 128u;
    return 32ull * n + 2ull * 16ull * n + 64ull * 17ull * n + 32ull + 2ull * 16ull * blocks +
           2ull * 4ull * 16ull * blocks;
}

}  // extern "C"

#include "bh_packed.inc"
#include "bh_gpu_tree.inc"
#include "bh_host_tree.inc"
#include "bh_serve.inc"
