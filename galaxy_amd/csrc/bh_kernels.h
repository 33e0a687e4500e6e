// gfx950 HIP kernels of the resident Barnes-Hut step (tree build, traversal, KDK).
//
// One code path serves one simulation (the drop-in) and many (batched/served):
// every kernel walks a table of "tiles", each tile belongs to exactly one
// simulation, and every simulation owns its own device buffers described by a
// SimDesc.  The build reproduces the BH #2A flat tree (nbody/src/flat.rs)
// bit-for-bit: root bounds, Morton keys, cell geometry and aggregates are
// computed in f64 in the reference's operation order and rounded to f32 once,
// as pack_flat_tree does (runtime/src/nbody_gpu.rs:190-229).  Cells are laid
// out in flat.rs preorder, which lets the traversal run stackless.
//
// Arithmetic discipline: built with -ffp-contract=off and correctly rounded
// f32 division / square root, so every f32 and f64 operation below is one
// IEEE round-to-nearest operation in source order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbh {

constexpr uint32_t kNoChild = 0xffffffffu;
constexpr uint32_t kTile = 256;          // bodies per tile: bounds, KDK
constexpr uint32_t kTopoTile = 1024;     // topology tile of a one-simulation set; limit of the one-workgroup builds
constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kRadix = 256;
constexpr uint32_t kMaxLaneTargets = 2;  // traversal targets per lane (work items of 64 or 128 targets)
constexpr uint32_t kLevels = 17;         // Morton depths 0..16
constexpr uint32_t kInfoLeaf = 1u << 8;
constexpr uint32_t kInfoSpanShift = 9;   // (span slot + 1) << 9 for internal cells that cross tiles

// Control words shared by the kernels of one simulation set (tile tickets are
// per simulation, SimMeta::ticket): the sticky error bits, the traversal's work
// distribution (mode, the item range of each XCD: start[8], length[8]; set by
// the host) and one work-queue counter per XCD, each in its own 128-byte line.
constexpr uint32_t kXcds = 8;
enum : uint32_t {
    kCtlError = 6,
    kCtlMode = 7,  // 0: one item range per XCD; 1: one range for the whole grid
    kCtlRanges = 16,
    kCtlQueue0 = 32,
    kQueueStride = 32,
    kCtlWords = kCtlQueue0 + kXcds * kQueueStride,
};

// Error bits (sticky in kCtlError and SimMeta::err).
enum : uint32_t {
    kErrLookback = 1u,
    kErrCellCapacity = 2u,
    kErrSpanCapacity = 4u,
};

// Cell record, 32 bytes, in flat.rs preorder: f32 aggregates, the opening
// threshold, the sorted range [start, end) and info (depth, leaf flag, span
// slot).  The evidence readback and the aggregation read it; the traversal
// reads the node stream instead.  `skip` (the next cell after the subtree) is
// filled only for host-built trees, where the packing uses it to lay out their
// node stream; device builds leave it 0.
struct __attribute__((aligned(32))) CellT {
    float comx, comy, mass, T;  // T: smallest d2 that passes the opening test
    uint32_t start, end, skip, info;
};

// Node stream: the preorder of the internal cells with every leaf replaced by
// its bodies (leaves are never accepted whole, nbody.wgsl:71-84, so a leaf
// contributes exactly its bodies' direct terms, in sorted order).  One rule
// serves both kinds of node: a target interacts with node c iff its own body
// node lies outside c's subtree [c, c + size) and d2 >= T, and then continues at
// c + size, else at c + 1.  A body node's subtree is itself (size 1, T = 0), so
// it interacts with every target but its own body.
//
// The traversal reads only the 16-byte hot record; T comes from the
// simulation's 16-entry code table (SimMeta::tab), with the code stored
// pre-scaled by 4 so that it is the table's byte offset.  Code 0 has T = 0 and
// marks bodies and the root (the root contains every target, so its threshold
// is never used); codes 1..15 are depths 1..15 (depth-16 cells are always
// leaves, flat.rs:127-177, so they never appear as internal nodes).  The cold
// record carries the sorted range, the node's own threshold and the evidence
// tag, for the evidence traversal and for host-built trees whose ranges or
// thresholds do not follow the preorder (general mode).
struct __attribute__((aligned(16))) NodeHot {
    float x, y, m;
    uint32_t w;  // subtree size - 1 and threshold code (node_word)
};
struct __attribute__((aligned(16))) NodeCold {
    uint32_t start, count;
    float T;
    uint32_t tag;  // kTagBody | body index, or the depth
};
// Two layouts of the hot word, chosen per simulation (SimDesc::wlayout):
//  - 0: ((size - 1) << 6) | (code << 2), for any node count below kMaxNodes;
//  - 1 (device trees of at most 2^22 node records): ((size - 1) << 10) |
//    (code == 0 ? 1 << 9 : 0) | (code << 2).  Then w << 22 is the opening
//    test's exponent offset: code << 24 for the depth codes and 2^31 for code 0
//    (bodies and the root, threshold 0), which every d2 >= 0 passes (see
//    pow4_thresholds).
// Both keep the code-table byte offset in bits 2..5.  The word holds size - 1,
// so a walk advances by at least one node per step whatever the word holds.
constexpr uint32_t kSizeShift = 6;
constexpr uint32_t kSizeShiftNarrow = 10;
constexpr uint32_t kNarrowNodes = 1u << (32 - kSizeShiftNarrow);
constexpr uint32_t kCodes = 16;
constexpr uint32_t kCodeField = (kCodes - 1u) << 2;          // byte offset into the code table
constexpr uint32_t kMaxNodes = 1u << (32 - kSizeShift);      // subtree sizes fit the hot word
__host__ __device__ constexpr uint32_t node_word(uint32_t size, uint32_t code, uint32_t layout = 0u) {
    return layout != 0u ? (((size - 1u) << kSizeShiftNarrow) | (code == 0u ? 1u << 9 : 0u) | (code << 2))
                        : (((size - 1u) << kSizeShift) | (code << 2));
}
constexpr uint32_t kTagBody = 1u << 31;
constexpr uint32_t kTagFirst = 1u << 30;  // first body of its leaf
constexpr uint32_t kTagValue = kTagFirst - 1u;

struct Agg {  // f64 aggregates of a cell (mass, centre of mass)
    double m, x, y, pad;
};

struct SimMeta {
    double root[4];         // f64 root (cx, cy, half) of root_bounds
    float bounds_f32[4];    // TreeMeta.bounds
    float T[20];            // per-depth acceptance thresholds (17 used)
    float tab[kCodes];      // traversal code table: 0 -> 0, code k -> threshold of code k
    uint32_t mode;          // 0: preorder-consistent nodes and T table; 1: general (cold records)
    uint32_t geom_ok;       // every position lies in a box of half-width < 1e18
    uint32_t ncells;
    uint32_t nnodes;        // internal cells + bodies
    uint32_t nspan;
    uint32_t err;
    uint32_t theta_bits;    // f32 theta the thresholds were computed for
    uint32_t pow4;          // T[d] = T[0] * 4^-d exactly for d = 1..15 (see pow4_thresholds)
    uint32_t t0_bits;       // bit pattern of T[0]
    uint32_t contain_rejected;  // the opening test rejects every cell of depth 1..15 holding the target
    uint32_t hist[4 * 256]; // per-pass digit histograms of the Morton keys
    // tile tickets of the simulation: radix passes 0..3, topology (4)
    uint32_t ticket[8];
};

struct __attribute__((aligned(16))) SimDesc {
    uint32_t n, cell_cap, bucket, maxd;
    float theta, eps, G, dt;
    uint32_t first_tile, first_sort_tile, first_topo_tile, span_cap;
    uint32_t node_cap;  // node records allocated (the traversal reads at most index node_cap - 1)
    uint32_t wlayout;   // hot-word layout (node_word): 1 when node_cap <= kNarrowNodes (device trees)
    // 1 (batched and served sets): a step leaves the internal cells' cold records,
    // which only an evidence traversal reads, unwritten; they are filled when one
    // asks (k_fill_cell_cold)
    uint32_t lean;
    uint32_t pad_;
    // resident state (SoA)
    float2* pos;
    float* mass;
    float2* vel;
    float2* acc;
    // radix sort ping-pong (local body index values)
    uint32_t* key0;
    uint32_t* key1;
    uint32_t* val0;
    uint32_t* val1;
    // sorted bodies
    float4* sbody;     // (x, y, m, 0) in sorted order
    // topology per sorted position
    uint32_t* cbase;   // preorder index of the column's first cell
    uint32_t* cparent; // start of the parent group of the column head
    uint8_t* cdmin;    // depth of the column head (17: no cell starts here)
    uint8_t* ccount;   // cells in the column
    // cells (preorder)
    CellT* cells;
    uint32_t* cend;    // end of each cell's range (written by k_topo_count)
    uint4* cchild;     // children (kNoChild when absent)
    Agg* cagg;         // f64 aggregates
    uint2* span;       // spanning internal cells (span_cap entries): (cell, depth << 26 | node)
    uint2* spanbuf;    // the same bucketed by depth (k_agg_top, listed mode)
    NodeHot* nodes;    // traversal node stream (cell_cap + n entries)
    NodeCold* ncold;   // its cold records
    uint32_t* nbase;   // per sorted position: node index of the column's first cell
    uint32_t* bnode;   // per sorted position: node index of that body
    float4* bpart;     // bounds partials, one per kTile bodies
    SimMeta* meta;
};

struct TileRef {
    uint32_t sim;
    uint32_t base;
};

// The simulation descriptors are read-only while a kernel runs.  Reading them
// through the constant address space lets the compiler treat the buffer pointers
// they hold as global memory: global_* loads and stores instead of flat_* ones,
// which would also count against the LDS/scalar wait counter and serialize store
// latency into every LDS wait.
typedef __attribute__((address_space(4))) const SimDesc* SimTable;
__device__ __forceinline__ const SimDesc& sim_at(const SimDesc* sims, uint32_t i) {
    return *(const SimDesc*)&((SimTable)sims)[i];
}

// ----------------------------------------------------------------------------
// wavefront primitives (wave64, DPP within 16-lane rows + lane reads across rows)
// ----------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Tile of workgroup b among n when the hardware deals workgroups to the 8 XCDs
// round-robin (b to XCD b % 8): each XCD gets one contiguous run of tiles, so
// the simulations whose data its L2 holds are few.  A bijection on [0, n)
// whatever the dealing is.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t n) {
    const uint32_t q = n / kXcds, r = n % kXcds, x = b % kXcds;
    return x * q + min(x, r) + b / kXcds;
}

// Workgroup barrier that orders LDS accesses only (no wait for outstanding
// global stores): for phases that exchange data through LDS alone.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__device__ __forceinline__ uint32_t dpp_shr_u32(uint32_t v, uint32_t identity, int ctrl_sel) {
    switch (ctrl_sel) {
    case 1: return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, 0x111, 0xf, 0xf, false);
    case 2: return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, 0x112, 0xf, 0xf, false);
    case 4: return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, 0x114, 0xf, 0xf, false);
    default: return (uint32_t)__builtin_amdgcn_update_dpp((int)identity, (int)v, 0x118, 0xf, 0xf, false);
    }
}

// Inclusive prefix sum over the 64 lanes (every lane must be active).
__device__ __forceinline__ uint32_t wave_incl_sum_u32(uint32_t v) {
    v += dpp_shr_u32(v, 0u, 1);
    v += dpp_shr_u32(v, 0u, 2);
    v += dpp_shr_u32(v, 0u, 4);
    v += dpp_shr_u32(v, 0u, 8);
    uint32_t r0 = __builtin_amdgcn_readlane(v, 15);
    uint32_t r1 = __builtin_amdgcn_readlane(v, 31);
    uint32_t r2 = __builtin_amdgcn_readlane(v, 47);
    uint32_t row = lane_id() >> 4;
    uint32_t add = (row > 0 ? r0 : 0u) + (row > 1 ? r1 : 0u) + (row > 2 ? r2 : 0u);
    return v + add;
}

// Block-wide exclusive prefix sum; returns the block total.  `scratch` holds one
// word per wavefront.  Every thread of the block must call it.
template <uint32_t kThreads>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* scratch, uint32_t& total) {
    constexpr uint32_t kWaves = kThreads / 64;
    uint32_t incl = wave_incl_sum_u32(v);
    uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 63) scratch[w] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t i = 0; i < kWaves; ++i) {
        uint32_t s = scratch[i];
        before += i < w ? s : 0u;
        all += s;
    }
    __syncthreads();
    total = all;
    return before + incl - v;
}

// ----------------------------------------------------------------------------
// decoupled look-back flags: status in the top two bits, value in the low 30
// ----------------------------------------------------------------------------
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagInc = 2u << 30;
constexpr uint32_t kFlagValue = (1u << 30) - 1;
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ void flag_store(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t flag_load(uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of tile `t` over earlier tiles of the same segment, whose
// flags are at flags[k * stride].  The segment's first tile publishes an
// inclusive value without looking back, so every wait ends.  Eight
// predecessor flags are read per round (independent loads), so a walk over m
// aggregate-only predecessors costs about m/8 memory latencies.
constexpr uint32_t kLookBatch = 8;

__device__ uint32_t lookback(uint32_t* flags, uint32_t stride, uint32_t t, uint32_t* err) {
    uint32_t excl = 0;
    uint32_t k = t;  // flags [0, k) not yet consumed
    uint32_t spins = 0;
    while (k > 0) {
        uint32_t f[kLookBatch];
#pragma unroll
        for (uint32_t j = 0; j < kLookBatch; ++j)
            f[j] = flag_load(flags + (size_t)(j < k ? k - 1u - j : k - 1u) * stride);
        uint32_t used = 0;
        bool done = false;
#pragma unroll
        for (uint32_t j = 0; j < kLookBatch; ++j) {
            if (done || used != j || j >= k || f[j] == 0u) continue;
            excl += f[j] & kFlagValue;
            used = j + 1u;
            if (f[j] & kFlagInc) done = true;
        }
        if (done) return excl;
        k -= used;
        if (used == 0u) {  // nearest predecessor not published yet
            if (++spins >= kSpinLimit) {
                atomicOr(err, kErrLookback);
                return excl;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    return excl;
}

// The same look-back over 64-bit flags that carry two 31-bit counts (status in
// bits 62-63), done by a whole wavefront: lane l reads the flag of tile
// k - 1 - l, so 64 predecessors cost one memory latency.  Every lane of the
// calling wavefront must be active; all lanes return the prefix.
constexpr unsigned long long kFlag64Agg = 1ull << 62;
constexpr unsigned long long kFlag64Inc = 2ull << 62;

__device__ __forceinline__ unsigned long long pack2(uint32_t a, uint32_t b) {
    return (unsigned long long)a | ((unsigned long long)b << 31);
}

__device__ unsigned long long lookback2_wave(unsigned long long* flags, uint32_t t, uint32_t* err) {
    const uint32_t lane = lane_id();
    uint32_t a = 0, b = 0;
    uint32_t k = t;
    uint32_t spins = 0;
    while (k > 0) {
        unsigned long long f = kFlag64Inc;  // lanes past tile 0 read as an inclusive zero
        if (lane < k)
            f = __hip_atomic_load(flags + (k - 1u - lane), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long inc = __ballot((f & kFlag64Inc) != 0ull);
        const unsigned long long pending = __ballot(f == 0ull);
        const uint32_t first_inc = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;
        const unsigned long long need = first_inc >= 63u ? ~0ull : ((2ull << first_inc) - 1ull);
        if (pending & need) {
            if (++spins >= kSpinLimit) {
                if (lane == 0) atomicOr(err, kErrLookback);
                return pack2(a, b);
            }
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        const bool take = lane <= first_inc;
        const uint32_t va = wave_incl_sum_u32(take ? (uint32_t)(f & 0x7fffffffull) : 0u);
        const uint32_t vb = wave_incl_sum_u32(take ? (uint32_t)((f >> 31) & 0x7fffffffull) : 0u);
        a += __builtin_amdgcn_readlane(va, 63);
        b += __builtin_amdgcn_readlane(vb, 63);
        if (first_inc < 64u) return pack2(a, b);
        k -= 64u;
    }
    return pack2(a, b);
}

// ----------------------------------------------------------------------------
// Morton keys (flat.rs:67-94) and geometry (flat.rs:96-103)
// ----------------------------------------------------------------------------
__device__ __forceinline__ uint32_t spread16(uint32_t v) {
    v &= 0x0000ffffu;
    v = (v | (v << 8)) & 0x00ff00ffu;
    v = (v | (v << 4)) & 0x0f0f0f0fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

__device__ __forceinline__ uint32_t quantize_axis(double value, double center, double half) {
    double low = center - half;
    double width = half * 2.0;
    double scaled = floor((value - low) / width * 65536.0);
    if (!(scaled >= 0.0)) return 0u;  // clamp: NaN and negatives -> 0 (Rust `as u16`)
    if (scaled > 65535.0) return 65535u;
    return (uint32_t)scaled;
}

// Number of leading 2-bit Morton digits shared by two keys (16 when equal).
__device__ __forceinline__ uint32_t lcp2(uint32_t a, uint32_t b) {
    uint32_t x = a ^ b;
    return x == 0 ? 16u : (uint32_t)__clz((int)x) >> 1;
}

// prefix at depth d: the top 2d bits
__device__ __forceinline__ bool same_prefix(uint32_t a, uint32_t b, uint32_t d) {
    return d == 0 || ((a ^ b) >> (32u - 2u * d)) == 0u;
}

// f32 opening threshold: the smallest positive d2 with fl(w / fl(sqrt(d2))) < theta
// (nbody.wgsl:105-108 with w = 2*half).  NaN when no d2 passes (theta <= 0).
__device__ float accept_threshold(float half, float theta) {
    const float w = 2.0f * half;
    auto pass = [&](uint32_t bits) {
        float d2 = __uint_as_float(bits);
        float s = __builtin_sqrtf(d2);
        return (w / s) < theta;
    };
    const uint32_t inf_bits = 0x7f800000u;
    if (!pass(inf_bits)) return __uint_as_float(0x7fc00000u);
    if (pass(1u)) return __uint_as_float(1u);
    uint32_t lo = 1u, hi = inf_bits;  // pass(lo) false, pass(hi) true
    // pass() is monotone in d2 (a correctly rounded square root and quotient
    // are), so the answer is the one bit pattern where it turns true.  Start
    // next to (w / theta)^2 and step a few patterns either way, keeping the
    // bracket; the bisection below closes what the steps leave open.
    const double est = ((double)w / (double)theta) * ((double)w / (double)theta);
    if (est > 0.0 && est < 3.0e38) {
        const uint32_t b = min(max(__float_as_uint((float)est), lo + 1u), hi - 1u);
        if (pass(b)) {
            hi = b;
            for (int k = 0; k < 4 && hi - lo > 1u; ++k) {
                if (!pass(hi - 1u)) {
                    lo = hi - 1u;
                    break;
                }
                --hi;
            }
        } else {
            lo = b;
            for (int k = 0; k < 4 && hi - lo > 1u; ++k) {
                if (pass(lo + 1u)) {
                    hi = lo + 1u;
                    break;
                }
                ++lo;
            }
        }
    }
    while (hi - lo > 1u) {
        uint32_t mid = lo + ((hi - lo) >> 1);
        if (pass(mid)) hi = mid; else lo = mid;
    }
    return __uint_as_float(hi);
}

// Thresholds of given (half, theta) pairs (gbh_check_thresholds).
__global__ void k_check_thresholds(const float* __restrict__ half, const float* __restrict__ theta,
                                   uint32_t count, float* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = accept_threshold(half[i], theta[i]);
}

// Whether a simulation's opening thresholds scale by powers of four: T[0] is a
// finite normal f32 and T[d] = T[0] * 4^-d exactly (both ways) for d = 1..15.
// The thresholds are defined by w_d / fl(sqrt(d2)) < theta with w_d = w_0 2^-d,
// and scaling d2 by 4^d scales fl(sqrt(d2)) by 2^d exactly while all values stay
// normal, so this holds except near the ends of the f32 range.  Then, for every
// finite d2 >= 0, d2 >= T[d] <=> bits(d2) + d * 2^24 >= bits(T[0]) as unsigned
// integers: for normal d2 the sum is bits(d2 4^d) (or exceeds bits(FLT_MAX) when
// d2 4^d overflows, and then d2 > T[d] since T[0] <= FLT_MAX); for subnormal d2
// both sides are false since T[0] >= 2^-126 4^d.  Called by a whole wavefront
// whose lane d < kCodes holds T[d]; returns 1 when the property holds.
__device__ __forceinline__ uint32_t pow4_thresholds(float Td, float T0) {
    const int d = (int)lane_id();
    bool ok = true;
    if (d < (int)kCodes) {
        const bool normal = T0 >= 1.17549435e-38f && T0 < __builtin_inff();
        ok = normal && (d == 0 || (Td == __builtin_ldexpf(T0, -2 * d) && __builtin_ldexpf(Td, 2 * d) == T0));
    }
    return __ballot(!ok) == 0ull ? 1u : 0u;
}

// Whether the opening test alone rejects every internal cell of depth 1..15 for
// the targets the cell holds, so that a walk needs no containment test below
// the root.  A target in the cell and the cell's centre of mass both lie in the
// cell's box of side w = 2 half (the f32 centre of mass rounds into the range
// of its bodies' coordinates), so d2 <= 2 w^2 up to a few relative roundings
// of 2^-24, and d2 < T[d] whenever T[d] >= 2.25 w^2 (theta <= 2/3).  Called by
// a whole wavefront whose lane d < kCodes holds T[d] and the depth's f32
// half-width; returns 1 when it holds for every d = 1..15.
__device__ __forceinline__ uint32_t contained_rejected(float Td, float half_d) {
    const uint32_t d = lane_id();
    bool ok = true;
    if (d >= 1u && d < kCodes) {
        const double w = 2.0 * (double)half_d;
        ok = (double)Td >= 2.25 * w * w;  // (false for a NaN threshold)
    }
    return __ballot(!ok) == 0ull ? 1u : 0u;
}

// Correctly rounded f32 square root and reciprocal from the f64 hardware
// approximations: the f32 result is the f64 approximation rounded once.  This is
// exact when the f64 error is smaller than the distance from the true value to
// the nearest f32 rounding midpoint, which k_check_rounding verifies
// exhaustively for every positive f32 input against the compiler's correctly
// rounded sqrtf and division.
// Square root of a finite positive f32: one Goldschmidt step in f64 from the
// hardware reciprocal square root (g ~ sqrt(x), h ~ 1/(2 sqrt(x))), rounded once
// to f32.
__device__ __forceinline__ float sqrt_rn_pos(float xf) {
    const double x = (double)xf;
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    const double h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    return (float)g;
}

// Square root of a non-negative f32; zero, +inf and NaN pass through as sqrtf
// returns them.
__device__ __forceinline__ float sqrt_rn(float xf) {
    const float g = sqrt_rn_pos(xf);
    const bool finite_pos = xf > 0.0f && xf < __builtin_inff();
    return finite_pos ? g : (xf < 0.0f ? __builtin_nanf("") : xf);
}

// Reciprocal of a finite positive f32: f64 hardware reciprocal plus one Newton
// step, rounded once to f32.
__device__ __forceinline__ double rcp_f64_step(float sf) {
    const double s = (double)sf;
    const double y = __builtin_amdgcn_rcp(s);
    const double e = __builtin_fma(-s, y, 1.0);
    return __builtin_fma(e, y, y);
}

__device__ __forceinline__ float rcp_rn_pos(float sf) { return (float)rcp_f64_step(sf); }

// Reciprocal of a normal f32 below 2^126 (1/s normal): f32 hardware reciprocal
// plus one Newton step with an exact residual (Markstein).  k_check_rounding
// verifies it equals the correctly rounded division for every such input.
constexpr uint32_t kRcpNormalLo = 0x00800000u, kRcpNormalHi = 0x7e800000u;  // [2^-126, 2^126)
__device__ __forceinline__ float rcp_rn_normal(float s) {
    const float y = __builtin_amdgcn_rcpf(s);
    const float e = __builtin_fmaf(-s, y, 1.0f);
    return __builtin_fmaf(e, y, y);
}

// Reciprocal of a non-negative f32; 1/0 = +inf and 1/inf = 0 as for the division.
__device__ __forceinline__ float rcp_rn(float sf) {
    const bool finite_pos = sf > 0.0f && sf < __builtin_inff();
    return finite_pos ? rcp_rn_pos(sf) : (float)__builtin_amdgcn_rcp((double)sf);
}

// point_mass (nbody.wgsl:35-41) with inverseSqrt = 1 / sqrt, each correctly
// rounded, from the offset (dx, dy) and d2 = dx*dx + dy*dy.
__device__ __forceinline__ void point_mass_d(float dx, float dy, float d2, float m, float eps2,
                                             float G, float& ax, float& ay) {
    float r2 = d2 + eps2;
    float s = sqrt_rn(r2);
    float inv_r = rcp_rn(s);
    float scale = G * m;
    scale = scale * inv_r;
    scale = scale * inv_r;
    scale = scale * inv_r;
    ax = ax + dx * scale;
    ay = ay + dy * scale;
}

// The same where r2 = d2 + eps2 is finite and positive: then sqrt_rn takes its
// finite_pos branch and s lies in [2^-75, 2^64], where rcp_rn_normal is the
// correctly rounded reciprocal, so the result is identical.
// kUnitG: G == 1 (then fl(G m) = m, and the product is left out).
template <bool kUnitG = false>
__device__ __forceinline__ void point_mass_pos(float dx, float dy, float d2, float m, float eps2,
                                               float G, float& ax, float& ay) {
    float r2 = d2 + eps2;
    float s = sqrt_rn_pos(r2);
    float inv_r = rcp_rn_normal(s);
    float scale = kUnitG ? m : G * m;
    scale = scale * inv_r;
    scale = scale * inv_r;
    scale = scale * inv_r;
    ax = ax + dx * scale;
    ay = ay + dy * scale;
}

__device__ __forceinline__ void point_mass(float tx, float ty, float sx, float sy, float m,
                                           float eps2, float G, float& ax, float& ay) {
    const float dx = sx - tx;
    const float dy = sy - ty;
    const float dd = dx * dx;
    const float ee = dy * dy;
    point_mass_d(dx, dy, dd + ee, m, eps2, G, ax, ay);
}

// Exhaustive check of the traversal's rounding primitives against the
// compiler's correctly rounded sqrtf and division: sqrt_rn and rcp_rn over every
// non-negative f32 bit pattern (+0 to +inf), rcp_rn_normal over [2^-126, 2^126).
// counts[0..2]: mismatches; first[0..2]: smallest mismatching bit pattern.
__global__ void k_check_rounding(unsigned long long* __restrict__ counts,
                                 uint32_t* __restrict__ first) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= 0x7f800000u; b += stride) {
        const float x = __uint_as_float(b);
        if (__float_as_uint(sqrt_rn(x)) != __float_as_uint(__builtin_sqrtf(x))) {
            atomicAdd(&counts[0], 1ull);
            atomicMin(&first[0], b);
        }
        const float want = 1.0f / x;
        if (__float_as_uint(rcp_rn(x)) != __float_as_uint(want)) {
            atomicAdd(&counts[1], 1ull);
            atomicMin(&first[1], b);
        }
        if (b >= kRcpNormalLo && b < kRcpNormalHi &&
            __float_as_uint(rcp_rn_normal(x)) != __float_as_uint(want)) {
            atomicAdd(&counts[2], 1ull);
            atomicMin(&first[2], b);
        }
        if (b > 0x7f800000u - stride) break;
    }
}

__device__ __forceinline__ void mix64(unsigned long long& h, unsigned long long w) {
    h ^= w;
    h *= 0x00000100000001b3ull;
    h ^= (w << 23) | (w >> 41);
}

// ----------------------------------------------------------------------------
// K1: kick/drift (+ optional preceding final kick) fused with the bounds partials
// mode 0: bounds only; 1: kick/drift then bounds; 2: final kick + kick/drift then bounds
// (nbody_evolve.wgsl:19-46, nbody_tree_parallel_bounds.wgsl:51-76).
// clear_errors: also clear the set's and every simulation's sticky error word
// (the first build of a call; no kernel of this launch sets them).
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kTile) void k_bounds(const SimDesc* __restrict__ sims,
                                                  const TileRef* __restrict__ tiles,
                                                  uint32_t* __restrict__ ctl,
                                                  unsigned long long* __restrict__ tflags,
                                                  uint32_t ntopo, int mode, uint32_t clear_errors) {
    __shared__ float s_red[4][kTile / 64];
    const TileRef tr = tiles[blockIdx.x];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t i = tr.base + threadIdx.x;
    const bool valid = i < S.n;
    // the body's loads first (a load issued after a store waits for it)
    const uint32_t ic = min(i, S.n - 1u);
    float2 p = S.pos[ic];
    float2 v = make_float2(0.f, 0.f), a = make_float2(0.f, 0.f);
    if (mode != 0) {
        v = S.vel[ic];
        a = S.acc[ic];
    }
    if (blockIdx.x == 0 && threadIdx.x < kXcds) ctl[kCtlQueue0 + threadIdx.x * kQueueStride] = 0u;
    // one look-back flag of the topology scan per block (there are fewer
    // topology tiles than bounds tiles)
    if (threadIdx.x == 0 && blockIdx.x < ntopo) tflags[blockIdx.x] = 0ull;
    if (tr.base == 0) {
        for (uint32_t k = threadIdx.x; k < 4 * 256; k += kTile) S.meta->hist[k] = 0u;
        if (threadIdx.x < 8) S.meta->ticket[threadIdx.x] = 0u;
        if (clear_errors && threadIdx.x == 0) S.meta->err = 0u;
    }
    if (clear_errors && blockIdx.x == 0 && threadIdx.x == 0) ctl[kCtlError] = 0u;
    float x = 0.f, y = 0.f;
    if (valid) {
        if (mode != 0) {
            const float dt = S.dt;
            const float h = 0.5f * dt;
            if (mode == 2) {  // bh_final_kick of the previous step
                v.x = v.x + h * a.x;
                v.y = v.y + h * a.y;
            }
            float hvx = v.x + h * a.x;  // bh_kick_drift
            float hvy = v.y + h * a.y;
            p.x = p.x + dt * hvx;
            p.y = p.y + dt * hvy;
            S.pos[i] = p;
            S.vel[i] = make_float2(hvx, hvy);
        }
        x = p.x;
        y = p.y;
    }
    float mnx = valid ? x : __builtin_inff(), mny = valid ? y : __builtin_inff();
    float mxx = valid ? x : -__builtin_inff(), mxy = valid ? y : -__builtin_inff();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, off));
        mny = fminf(mny, __shfl_xor(mny, off));
        mxx = fmaxf(mxx, __shfl_xor(mxx, off));
        mxy = fmaxf(mxy, __shfl_xor(mxy, off));
    }
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_red[0][w] = mnx;
        s_red[1][w] = mny;
        s_red[2][w] = mxx;
        s_red[3][w] = mxy;
    }
    lds_barrier();  // the reduction reads LDS only: no wait for the state stores
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < kTile / 64; ++k) {
            mnx = fminf(mnx, s_red[0][k]);
            mny = fminf(mny, s_red[1][k]);
            mxx = fmaxf(mxx, s_red[2][k]);
            mxy = fmaxf(mxy, s_red[3][k]);
        }
        S.bpart[tr.base / kTile] = make_float4(mnx, mny, mxx, mxy);
    }
}

__global__ void k_reset_queues(uint32_t* __restrict__ ctl) {
    if (threadIdx.x < kXcds) ctl[kCtlQueue0 + threadIdx.x * kQueueStride] = 0u;
}

// Clears the sticky error words of a set and of each of its simulations (the
// start of a batched or served call: each call reports its own errors).
__global__ __launch_bounds__(256) void k_clear_errors(const SimDesc* __restrict__ sims, uint32_t nsims,
                                                      uint32_t* __restrict__ ctl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) ctl[kCtlError] = 0u;
    if (i < nsims && sims[i].meta) sims[i].meta->err = 0u;
}

// Sticky error words of a set into one array: out[0] = the set's control error
// word, out[1 + s] = simulation s's (so the host reads them with one copy).
__global__ __launch_bounds__(256) void k_collect_errors(const SimDesc* __restrict__ sims, uint32_t nsims,
                                                        const uint32_t* __restrict__ ctl,
                                                        uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) out[0] = ctl[kCtlError];
    if (i < nsims) {
        const SimMeta* m = sims[i].meta;
        out[1 + i] = m ? m->err : 0u;
    }
}

// ----------------------------------------------------------------------------
// K2: root bounds (root_bounds, lib.rs:137-154), Morton keys (flat.rs:224-231),
// per-pass digit histograms, look-back flag reset, per-depth thresholds.
// ----------------------------------------------------------------------------
template <uint32_t kKPT>
__global__ __launch_bounds__(kSortThreads) void k_keys(const SimDesc* __restrict__ sims,
                                                       const TileRef* __restrict__ stiles,
                                                       uint32_t* __restrict__ sflags,
                                                       uint32_t nstiles) {
    constexpr uint32_t kSortTile = kSortThreads * kKPT;
    __shared__ float s_red[4][kSortThreads / 64];
    __shared__ double s_root[3];
    __shared__ uint32_t s_hist[4][kRadix];
    const uint32_t tb = xcd_tile(blockIdx.x, gridDim.x);
    const TileRef tr = stiles[tb];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t n = S.n;
    // Every load of the kernel is issued before its first store: the positions
    // of this tile, then the bounds partials (a load issued after a store
    // waits for the store).
    // (Unconditional loads at a clamped index: a conditional load would be
    // waited for inside its branch.)
    float2 pos[kKPT];
#pragma unroll
    for (uint32_t j = 0; j < kKPT; ++j) pos[j] = S.pos[min(tr.base + j * kSortThreads + threadIdx.x, n - 1u)];
    // global f32 min/max of the simulation from the bounds partials
    const uint32_t nparts = (n + kTile - 1) / kTile;
    float mnx = __builtin_inff(), mny = __builtin_inff(), mxx = -__builtin_inff(), mxy = -__builtin_inff();
    for (uint32_t k = threadIdx.x; k < nparts; k += kSortThreads) {
        float4 b = S.bpart[k];
        mnx = fminf(mnx, b.x);
        mny = fminf(mny, b.y);
        mxx = fmaxf(mxx, b.z);
        mxy = fmaxf(mxy, b.w);
    }
    // reset this tile's look-back flags for the four radix passes
    for (uint32_t k = threadIdx.x; k < 4 * kRadix; k += kSortThreads) {
        uint32_t pass = k / kRadix, d = k % kRadix;
        sflags[((size_t)pass * nstiles + tb) * kRadix + d] = 0u;
        s_hist[pass][d] = 0u;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, off));
        mny = fminf(mny, __shfl_xor(mny, off));
        mxx = fmaxf(mxx, __shfl_xor(mxx, off));
        mxy = fmaxf(mxy, __shfl_xor(mxy, off));
    }
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_red[0][w] = mnx;
        s_red[1][w] = mny;
        s_red[2][w] = mxx;
        s_red[3][w] = mxy;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t k = 1; k < kSortThreads / 64; ++k) {
            mnx = fminf(mnx, s_red[0][k]);
            mny = fminf(mny, s_red[1][k]);
            mxx = fmaxf(mxx, s_red[2][k]);
            mxy = fmaxf(mxy, s_red[3][k]);
        }
        // root_bounds in f64 over the widened f32 positions (min/max are exact)
        double min_x = mnx, min_y = mny, max_x = mxx, max_y = mxy;
        double sx = max_x - min_x, sy = max_y - min_y;
        double span = sy > sx ? sy : sx;
        span = 1e-12 > span ? 1e-12 : span;
        s_root[0] = (min_x + max_x) * 0.5;
        s_root[1] = (min_y + max_y) * 0.5;
        s_root[2] = span * 0.500000000001 + 1e-12;
    }
    __syncthreads();
    const double cx = s_root[0], cy = s_root[1], half = s_root[2];
    if (tr.base == 0 && threadIdx.x < 64) {
        SimMeta* M = S.meta;
        if (threadIdx.x == 0) {
            M->root[0] = cx;
            M->root[1] = cy;
            M->root[2] = half;
            M->root[3] = 0.0;
            M->bounds_f32[0] = (float)cx;
            M->bounds_f32[1] = (float)cy;
            M->bounds_f32[2] = (float)half;
            M->bounds_f32[3] = 0.0f;
            M->nspan = 0u;
            M->theta_bits = __float_as_uint(S.theta);
            M->mode = 0u;
            // |dx|, |dy| <= 2 * half between any two positions (bodies and
            // centres of mass lie in the root box), so d2 stays finite
            M->geom_ok = half < 1.0e18 ? 1u : 0u;
        }
        float T = 0.0f, hf = 0.0f;
        if (threadIdx.x < kLevels) {
            // half at depth d is half * 2^-d exactly in f64; one rounding to f32
            double hd = half;
            for (uint32_t l = 0; l < threadIdx.x; ++l) hd = hd * 0.5;
            hf = (float)hd;
            T = accept_threshold(hf, S.theta);
            M->T[threadIdx.x] = T;
            if (threadIdx.x < kCodes) M->tab[threadIdx.x] = threadIdx.x == 0 ? 0.0f : T;
        }
        const float T0 = __shfl(T, 0);
        const uint32_t p4 = pow4_thresholds(T, T0);
        const uint32_t cr = contained_rejected(T, hf);
        if (threadIdx.x == 0) {
            M->pow4 = p4;
            M->t0_bits = __float_as_uint(T0);
            M->contain_rejected = cr;
        }
    }
    // Morton keys of this tile
#pragma unroll
    for (uint32_t j = 0; j < kKPT; ++j) {
        uint32_t i = tr.base + j * kSortThreads + threadIdx.x;
        if (i < n) {
            const float2 p = pos[j];
            uint32_t code = spread16(quantize_axis((double)p.x, cx, half)) |
                            (spread16(quantize_axis((double)p.y, cy, half)) << 1);
            S.key0[i] = code;
            S.val0[i] = i;
            atomicAdd(&s_hist[0][code & 255u], 1u);
            atomicAdd(&s_hist[1][(code >> 8) & 255u], 1u);
            atomicAdd(&s_hist[2][(code >> 16) & 255u], 1u);
            atomicAdd(&s_hist[3][code >> 24], 1u);
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < 4 * kRadix; k += kSortThreads) {
        uint32_t c = s_hist[k / kRadix][k % kRadix];
        if (c) atomicAdd(&S.meta->hist[k], c);
    }
    (void)kSortTile;
}

// ----------------------------------------------------------------------------
// K3: one stable LSD radix pass over 8-bit digits (onesweep: local ranking with
// wavefront ballots, LDS-staged scatter, decoupled look-back across tiles).
// Equals the stable sort_by_key of flat.rs:233-234 after four passes.
// ----------------------------------------------------------------------------
// One pass over the tile of workgroup `block` of `grid` (every thread of the
// workgroup calls it).
template <uint32_t kKPT>
__device__ __forceinline__ void sort_pass_body(const SimDesc* __restrict__ sims,
                                               const TileRef* __restrict__ stiles, uint32_t nstiles,
                                               uint32_t pass, uint32_t* __restrict__ ctl,
                                               uint32_t* __restrict__ sflags, uint32_t block,
                                               uint32_t grid) {
    constexpr uint32_t kSortTile = kSortThreads * kKPT;
    constexpr uint32_t kWaves = kSortThreads / 64;
    constexpr uint32_t kChunk = 64 * kKPT;  // keys per wavefront
    __shared__ uint32_t s_keys[kSortTile];
    __shared__ uint32_t s_vals[kSortTile];
    __shared__ uint32_t s_wcnt[kWaves][kRadix];
    __shared__ uint32_t s_start[kRadix];  // tile-local start of each digit bucket
    __shared__ uint32_t s_glob[kRadix];   // destination of each digit bucket
    __shared__ uint32_t s_scan[kWaves];
    __shared__ uint32_t s_ticket;

    // The workgroup's simulation comes from its block index; its tile within
    // the simulation from a ticket of that simulation, so tiles take tickets
    // in order and a look-back only waits for tiles whose workgroups run.
    // (One ticket counter per simulation: a single counter for every tile
    // serialises on one address.)
    const uint32_t sim = stiles[xcd_tile(block, grid)].sim;
    const SimDesc& S = sim_at(sims, sim);
    if (threadIdx.x == 0) s_ticket = atomicAdd(&S.meta->ticket[pass], 1u);
    for (uint32_t k = threadIdx.x; k < kWaves * kRadix; k += kSortThreads) (&s_wcnt[0][0])[k] = 0u;
    __syncthreads();
    // the simulation's tiles are consecutive and of kTileKeys keys, so the
    // ticket gives the tile without a load of the tile table
    constexpr uint32_t kTileKeys = kSortThreads * kKPT;
    const uint32_t local = s_ticket;
    if (local >= (S.n + kTileKeys - 1u) / kTileKeys) {  // tickets not reset: no tile of this simulation is left
        if (threadIdx.x == 0) atomicOr(&S.meta->err, kErrLookback);
        return;
    }
    const uint32_t t = S.first_sort_tile + local;
    const TileRef tr{sim, local * kTileKeys};
    const uint32_t n = S.n;
    const uint32_t shift = 8u * pass;
    const uint32_t* __restrict__ kin = (pass & 1u) ? S.key1 : S.key0;
    const uint32_t* __restrict__ vin = (pass & 1u) ? S.val1 : S.val0;
    uint32_t* __restrict__ kout = (pass & 1u) ? S.key0 : S.key1;
    uint32_t* __restrict__ vout = (pass & 1u) ? S.val0 : S.val1;
    uint32_t* flags = sflags + (size_t)pass * nstiles * kRadix;

    const uint32_t w = threadIdx.x >> 6;
    const uint32_t lane = lane_id();
    const unsigned long long lt_mask = __lanemask_lt();
    uint32_t key[kKPT], val[kKPT], rank[kKPT];
    // all loads first: the ranking below is ordered by wavefront barriers,
    // which a load issued inside the loop could not pass
#pragma unroll
    for (uint32_t j = 0; j < kKPT; ++j) {
        const uint32_t i = tr.base + w * kChunk + j * 64 + lane;
        const bool valid = i < n;
        key[j] = valid ? kin[i] : 0u;
        val[j] = valid ? vin[i] : 0u;
    }
#pragma unroll
    for (uint32_t j = 0; j < kKPT; ++j) {
        const uint32_t i = tr.base + w * kChunk + j * 64 + lane;
        const bool valid = i < n;
        const uint32_t d = (key[j] >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
        const uint32_t prior = s_wcnt[w][d];  // read by every peer before the leader writes
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) s_wcnt[w][d] = prior + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rank[j] = valid ? prior + below : 0xffffffffu;
    }
    __syncthreads();
    // per digit: offsets of each wavefront inside the digit bucket, tile total
    const uint32_t d = threadIdx.x;  // kSortThreads == kRadix
    uint32_t tile_count = 0;
#pragma unroll
    for (uint32_t k = 0; k < kWaves; ++k) {
        uint32_t c = s_wcnt[k][d];
        s_wcnt[k][d] = tile_count;
        tile_count += c;
    }
    uint32_t total;
    const uint32_t start = block_excl_sum<kSortThreads>(tile_count, s_scan, total);
    s_start[d] = start;
    // publish and look back
    const bool seg_first = (t == S.first_sort_tile);
    if (seg_first) {
        // digit base of this simulation: exclusive scan of its histogram
        uint32_t tot2;
        const uint32_t base = block_excl_sum<kSortThreads>(S.meta->hist[pass * kRadix + d], s_scan, tot2);
        flag_store(flags + (size_t)t * kRadix + d, kFlagInc | (base + tile_count));
        s_glob[d] = base;
    } else {
        flag_store(flags + (size_t)t * kRadix + d, kFlagAgg | tile_count);
        uint32_t excl = lookback(flags + d, kRadix, t, &ctl[kCtlError]);
        flag_store(flags + (size_t)t * kRadix + d, kFlagInc | (excl + tile_count));
        s_glob[d] = excl;
    }
    __syncthreads();
    // LDS-staged reorder: digit buckets become contiguous in LDS
#pragma unroll
    for (uint32_t j = 0; j < kKPT; ++j) {
        if (rank[j] != 0xffffffffu) {
            const uint32_t dj = (key[j] >> shift) & 255u;
            const uint32_t lpos = s_start[dj] + s_wcnt[w][dj] + rank[j];
            s_keys[lpos] = key[j];
            s_vals[lpos] = val[j];
        }
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < total; l += kSortThreads) {
        const uint32_t kk = s_keys[l];
        const uint32_t vv = s_vals[l];
        const uint32_t dd = (kk >> shift) & 255u;
        const uint32_t g = s_glob[dd] + (l - s_start[dd]);
        kout[g] = kk;
        vout[g] = vv;
    }
}

template <uint32_t kKPT>
__global__ __launch_bounds__(kSortThreads) void k_sort_pass(const SimDesc* __restrict__ sims,
                                                            const TileRef* __restrict__ stiles,
                                                            uint32_t nstiles, uint32_t pass,
                                                            uint32_t* __restrict__ ctl,
                                                            uint32_t* __restrict__ sflags) {
    sort_pass_body<kKPT>(sims, stiles, nstiles, pass, ctl, sflags, blockIdx.x, gridDim.x);
}

// All four passes of a set with a single sort tile (one simulation of at most
// one tile's keys), in one workgroup: the passes follow one another after a
// barrier, which also makes each pass's global stores visible to the next.
template <uint32_t kKPT>
__global__ __launch_bounds__(kSortThreads) void k_sort_single(const SimDesc* __restrict__ sims,
                                                              const TileRef* __restrict__ stiles,
                                                              uint32_t* __restrict__ ctl,
                                                              uint32_t* __restrict__ sflags) {
    for (uint32_t pass = 0; pass < 4; ++pass) {
        sort_pass_body<kKPT>(sims, stiles, 1u, pass, ctl, sflags, 0u, 1u);
        __syncthreads();
    }
}

// ----------------------------------------------------------------------------
// K4: cell columns.  Position p starts cells at depths dmin(p)..dleaf(p): dmin
// from the common prefix with its predecessor, existence from the parent group's
// size, leaf depth from the key at p + bucket.  An exclusive scan of column
// lengths (decoupled look-back) gives every cell its preorder index.  Writes
// ranges, geometry (f64, flat.rs:96-103), first-child links, leaf aggregates
// (fill_aggregates leaf branch, flat.rs:181-195) and thresholds.
// ----------------------------------------------------------------------------
// A position q starts a group at depth d iff dmin(q) <= d (dmin(0) = 0: the
// simulation's first position starts a group at every depth; equal keys give
// dmin = 17, no group).  So the end of the depth-d group containing p is the
// next position after p with dmin <= d (or n), and its start the last position
// at or before p with dmin <= d.  These queries are answered without searching
// the keys: per-wavefront ballot masks of "dmin <= d" in LDS, per-wavefront
// first/last boundaries within the tile, and per-tile summaries over the other
// tiles of the simulation (k_topo_pre, k_topo_scan).
// The tiled kernels take the tile size kTT as a template argument: one position
// per thread, kTT threads per workgroup.  Sets of several simulations use
// kTopoTileMulti positions (workgroups small enough to run beside the other
// group's traversal), sets of one simulation kTopoTile.
constexpr uint32_t kTopoTileMulti = 512;
constexpr uint32_t kTopoPassThreads = 512;  // threads of the tiled boundary and column passes
constexpr uint32_t kTopoSum = 2 * kLevels + 2;  // per-tile summary words: boundaries by depth, cells, columns
constexpr uint32_t kLocalBucket = 32;           // leaf sizes whose columns the boundary pass counts
constexpr uint32_t kNone = 0xffffffffu;

// Per topology tile: dmin of every position (cdmin) and, per depth d, the first
// boundary position in the tile and one past the last (tsum[t][d], tsum[t][17+d];
// kNone / 0 when the tile has none).  Also the sorted body copy, sbody[p] =
// (x, y, m) of the body at sorted position p: its reads go to random body
// indices (inside the last radix pass their latency held every tile's
// workgroup), here they overlap the boundary work.
//
// With a leaf size of at most kLocalBucket it also counts each position's
// column of cells (ccount) and the tile's cells and columns (tsum[t][34],
// tsum[t][35]), so that k_topo_scan can give every tile its place in the
// preorder and the column pass needs no look-back.  The column at p exists iff
// its head depth dmin <= maxd and, for dmin >= 1, the depth dmin - 1 group
// holding p has more than `bucket` positions; that group is [pl, e) with
// pl < p < e, and it has more than `bucket` positions iff some window
// [q, q + bucket] containing p lies inside it, i.e. iff
// lcp(key[q], key[q + bucket]) >= dmin - 1 for some q in
// [max(0, p - bucket), min(p, n - 1 - bucket)] (sorted keys: the two ends of a
// window share a prefix iff all of it does).  The leaf depth comes from the key
// at p + bucket as in the column pass.
template <uint32_t kTT, uint32_t kThreads = kTT>
__device__ __forceinline__ void topo_pre_body(const SimDesc* __restrict__ sims,
                                              const TileRef* __restrict__ ttiles,
                                              uint32_t* __restrict__ tsum, uint32_t block, uint32_t grid) {
    // kTT positions on kThreads threads: position j * kThreads + threadIdx.x of
    // the tile for j < kTopoPer, so a wavefront covers 64 consecutive positions
    // in every slice j.
    constexpr uint32_t kTopoThreads = kThreads, kTopoPer = kTT / kThreads;
    constexpr uint32_t kHalo = kLocalBucket;
    static_assert(kTT % kThreads == 0 && kThreads % 64 == 0, "whole wavefronts per slice");
    // per wavefront and depth: first boundary, last boundary + 1 (no LDS atomics)
    __shared__ uint32_t s_first[kTopoThreads / 64][kLevels], s_last[kTopoThreads / 64][kLevels];
    __shared__ uint32_t s_k[kTT + 2 * kHalo];  // keys at base - kHalo .. base + kTT + kHalo - 1
    __shared__ uint32_t s_tot[kTopoThreads / 64][2];
    const uint32_t t = xcd_tile(block, grid);
    const TileRef tr = ttiles[t];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t n = S.n, bucket = S.bucket, maxd = S.maxd;
    const bool local = bucket <= kLocalBucket;
    const uint32_t w = threadIdx.x >> 6;
    // every load before the first store: the keys around the tile, then the body gather
    for (uint32_t i = threadIdx.x; i < kTT + 2 * kHalo; i += kTopoThreads) {
        const uint32_t q = tr.base + i - kHalo;
        s_k[i] = tr.base + i >= kHalo && q < n ? S.key0[q] : 0u;
    }
    float2 bq[kTopoPer];
    float bm[kTopoPer];
#pragma unroll
    for (uint32_t j = 0; j < kTopoPer; ++j) {
        const uint32_t p = tr.base + j * kTopoThreads + threadIdx.x;
        const uint32_t v = S.val0[min(p, n - 1u)];
        bq[j] = S.pos[v];
        bm[j] = S.mass[v];
    }
    lds_barrier();
    auto K = [&](uint32_t q) { return s_k[q - tr.base + kHalo]; };  // q in [base - kHalo, base + kTT + kHalo)
    uint32_t dmin[kTopoPer];
#pragma unroll
    for (uint32_t j = 0; j < kTopoPer; ++j) {
        const uint32_t p = tr.base + j * kTopoThreads + threadIdx.x;
        const uint32_t kp = K(p);
        dmin[j] = p >= n ? 17u : p == 0 ? 0u : K(p - 1u) == kp ? 17u : lcp2(K(p - 1u), kp) + 1u;
    }
    for (uint32_t d = 0; d < kLevels; ++d) {
        uint32_t first = kNone, last = 0u;
#pragma unroll
        for (uint32_t j = 0; j < kTopoPer; ++j) {
            const uint32_t p = tr.base + j * kTopoThreads + threadIdx.x;
            const uint32_t wbase = p - lane_id();
            const unsigned long long m = __ballot(p < n && dmin[j] <= d);
            if (m) {
                first = min(first, wbase + (uint32_t)__builtin_ctzll(m));
                last = max(last, wbase + 64u - (uint32_t)__builtin_clzll(m));
            }
        }
        if (lane_id() == 0) {
            s_first[w][d] = first;
            s_last[w][d] = last;
        }
    }
    uint32_t cells = 0, cols = 0;
    if (local) {
#pragma unroll
        for (uint32_t j = 0; j < kTopoPer; ++j) {
            const uint32_t p = tr.base + j * kTopoThreads + threadIdx.x;
            uint32_t count = 0;
            if (p < n && dmin[j] <= maxd) {
                const uint32_t kp = K(p);
                bool exists = dmin[j] == 0u;
                if (!exists && n > bucket) {
                    const uint32_t dd = dmin[j] - 1u;
                    const uint32_t qhi = min(p, n - 1u - bucket);
                    for (uint32_t q = p >= bucket ? p - bucket : 0u; q <= qhi; ++q) {
                        if (lcp2(K(q), K(q + bucket)) >= dd) {
                            exists = true;
                            break;
                        }
                    }
                }
                if (exists) {
                    const int L = (p + bucket < n) ? (int)lcp2(kp, K(p + bucket)) : -1;
                    int dl = L + 1 < (int)maxd ? L + 1 : (int)maxd;
                    if (dl < (int)dmin[j]) dl = (int)dmin[j];
                    count = (uint32_t)dl - dmin[j] + 1;
                }
            }
            if (p < n) S.ccount[p] = (uint8_t)count;
            cells += count;
            cols += count > 0u ? 1u : 0u;
        }
        cells = wave_incl_sum_u32(cells);
        cols = wave_incl_sum_u32(cols);
        if (lane_id() == 63) {
            s_tot[w][0] = cells;
            s_tot[w][1] = cols;
        }
    }
    // the per-position store comes after the ballots; the barrier below waits for it
#pragma unroll
    for (uint32_t j = 0; j < kTopoPer; ++j) {
        const uint32_t p = tr.base + j * kTopoThreads + threadIdx.x;
        if (p < n) S.cdmin[p] = (uint8_t)dmin[j];
    }
    lds_barrier();  // (orders LDS only: the cdmin stores need not complete)
    if (threadIdx.x < kLevels) {
        uint32_t first = kNone, last = 0u;
        for (uint32_t k = 0; k < kTopoThreads / 64; ++k) {
            first = min(first, s_first[k][threadIdx.x]);
            last = max(last, s_last[k][threadIdx.x]);
        }
        tsum[(size_t)t * kTopoSum + threadIdx.x] = first;
        tsum[(size_t)t * kTopoSum + kLevels + threadIdx.x] = last;
    } else if (local && threadIdx.x < kLevels + 2) {
        const uint32_t c = threadIdx.x - kLevels;
        uint32_t v = 0;
        for (uint32_t k = 0; k < kTopoThreads / 64; ++k) v += s_tot[k][c];
        tsum[(size_t)t * kTopoSum + 2 * kLevels + c] = v;
    }
#pragma unroll
    for (uint32_t j = 0; j < kTopoPer; ++j) {
        const uint32_t p = tr.base + j * kTopoThreads + threadIdx.x;
        if (p < n) S.sbody[p] = make_float4(bq[j].x, bq[j].y, bm[j], 0.0f);
    }
}

template <uint32_t kTT, uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_topo_pre(const SimDesc* __restrict__ sims,
                                                       const TileRef* __restrict__ ttiles,
                                                       uint32_t* __restrict__ tsum) {
    topo_pre_body<kTT, kThreads>(sims, ttiles, tsum, blockIdx.x, gridDim.x);
}

// Inclusive wavefront scan with max / min (all lanes active).
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, o);
        if (lane_id() >= (uint32_t)o) v = max(v, u);
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_incl_min_rev(uint32_t v) {  // min over lanes >= self
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = (uint32_t)__shfl_down((int)v, o);
        if (lane_id() + (uint32_t)o < 64u) v = min(v, u);
    }
    return v;
}

// Per simulation (one workgroup): for every tile and depth, one past the last
// boundary before the tile (psum[t][d], 0: none) and the first boundary after it
// (psum[t][17+d], n: none).  Wavefront w handles depths w, w + 16, ...
__global__ __launch_bounds__(64) void k_topo_scan(const SimDesc* __restrict__ sims,
                                                  const uint32_t* __restrict__ tsum,
                                                  uint32_t* __restrict__ psum, uint32_t tile) {
    // One wavefront per (simulation, depth, direction): blockIdx.x is the
    // simulation, blockIdx.y < kLevels the forward scan of depth blockIdx.y and
    // the rest the backward scans.  The scan walks the simulation's tiles 64 at
    // a time; the loads of up to kScanBatch chunks are issued before their scans.
    constexpr uint32_t kScanBatch = 8;
    const SimDesc& S = sim_at(sims, blockIdx.x);
    const uint32_t T = (S.n + tile - 1) / tile;  // the set's topology tiles of this simulation
    const uint32_t t0 = S.first_topo_tile;
    const uint32_t lane = lane_id();
    const uint32_t chunks = (T + 63) / 64;
    if (blockIdx.y == 2 * kLevels) {
        // The cells and columns before each tile (exclusive prefix of the tile
        // counts of k_topo_pre; simulations with larger leaves take theirs from
        // the column pass's look-back).
        if (S.bucket > kLocalBucket) return;
        uint32_t cc = 0, cl = 0;
        for (uint32_t c = 0; c < chunks; ++c) {
            const uint32_t i = c * 64 + lane;
            const size_t at = (size_t)(t0 + min(i, T - 1u)) * kTopoSum + 2 * kLevels;
            const uint32_t vc = i < T ? tsum[at] : 0u, vl = i < T ? tsum[at + 1] : 0u;
            const uint32_t ic = wave_incl_sum_u32(vc), il = wave_incl_sum_u32(vl);
            if (i < T) {
                psum[at] = cc + ic - vc;
                psum[at + 1] = cl + il - vl;
            }
            cc += __builtin_amdgcn_readlane(ic, 63);
            cl += __builtin_amdgcn_readlane(il, 63);
        }
        return;
    }
    const bool forward = blockIdx.y < kLevels;
    const uint32_t d = forward ? blockIdx.y : blockIdx.y - kLevels;
    uint32_t carry = forward ? 0u : S.n;  // max of "last + 1" before / min of "first" after
    for (uint32_t c0 = 0; c0 < chunks; c0 += kScanBatch) {
        const uint32_t nb = min(kScanBatch, chunks - c0);
        uint32_t v[kScanBatch];
#pragma unroll
        for (uint32_t k = 0; k < kScanBatch; ++k) {
            const uint32_t c = forward ? c0 + k : chunks - 1u - (c0 + k);
            const uint32_t i = c * 64 + lane;
            const bool in = k < nb && i < T;
            v[k] = forward ? (in ? tsum[(size_t)(t0 + i) * kTopoSum + kLevels + d] : 0u)
                           : (in ? tsum[(size_t)(t0 + i) * kTopoSum + d] : kNone);
        }
#pragma unroll
        for (uint32_t k = 0; k < kScanBatch; ++k) {
            if (k >= nb) break;  // wave-uniform
            const uint32_t c = forward ? c0 + k : chunks - 1u - (c0 + k);
            const uint32_t i = c * 64 + lane;
            if (forward) {
                const uint32_t incl = max(wave_incl_max(v[k]), carry);
                const uint32_t prev = (uint32_t)__shfl_up((int)incl, 1);
                if (i < T) psum[(size_t)(t0 + i) * kTopoSum + d] = lane == 0 ? carry : prev;
                carry = __builtin_amdgcn_readlane(incl, 63);
            } else {
                const uint32_t incl = min(wave_incl_min_rev(v[k]), carry);
                const uint32_t next = (uint32_t)__shfl_down((int)incl, 1);
                if (i < T) psum[(size_t)(t0 + i) * kTopoSum + kLevels + d] = lane == 63 ? carry : next;
                carry = __builtin_amdgcn_readlane(incl, 0);
            }
        }
    }
}

// Boundary queries of one topology tile of kTT positions (see k_topo_pre).
template <uint32_t kTT>
struct TileBounds {
    static constexpr uint32_t kTopoWaves = kTT / 64;  // 64-position chunks of the tile
    unsigned long long mask[kTopoWaves][kLevels];  // bit l: position wbase + l has dmin <= d
    uint32_t before[kTopoWaves][kLevels];          // one past the last boundary in earlier waves (0: none)
    uint32_t after[kTopoWaves][kLevels];           // first boundary in later waves (kNone: none)
    uint32_t pre[kLevels], post[kLevels];          // tile summaries of the other tiles

    // The other tiles' summaries; called before the barrier that precedes build.
    __device__ void load_summaries(const uint32_t* __restrict__ ps) {
        if (threadIdx.x < kLevels) {
            pre[threadIdx.x] = ps[threadIdx.x];
            post[threadIdx.x] = ps[kLevels + threadIdx.x];
        }
    }

    // Every thread of the workgroup calls this (block-wide barriers), after a
    // barrier that publishes dmin[] (LDS, one byte per position of the tile)
    // and the summaries of load_summaries.
    __device__ void build(uint32_t base, uint32_t n, const uint8_t* dmin) {
        for (uint32_t c = threadIdx.x >> 6; c < kTopoWaves; c += blockDim.x >> 6) {
            const uint32_t o = c * 64u + lane_id();
            const uint32_t dm = base + o < n ? dmin[o] : 17u;
            for (uint32_t d = 0; d < kLevels; ++d) {
                const unsigned long long m = __ballot(dm <= d);
                if (lane_id() == 0) mask[c][d] = m;
            }
        }
        lds_barrier();
        if (threadIdx.x < kLevels) {
            const uint32_t d = threadIdx.x;
            uint32_t b = 0;
            for (uint32_t k = 0; k < kTopoWaves; ++k) {
                before[k][d] = b;
                const unsigned long long m = mask[k][d];
                if (m) b = base + k * 64u + 64u - (uint32_t)__builtin_clzll(m);
            }
            uint32_t a = kNone;
            for (uint32_t k = kTopoWaves; k > 0; --k) {
                after[k - 1][d] = a;
                const unsigned long long m = mask[k - 1][d];
                if (m) a = base + (k - 1) * 64u + (uint32_t)__builtin_ctzll(m);
            }
        }
        lds_barrier();
    }

    // Last boundary at depth d at or before position x (x in the tile).
    __device__ __forceinline__ uint32_t prev(uint32_t base, uint32_t x, uint32_t d) const {
        const uint32_t o = x - base, w = o >> 6, l = o & 63u;
        const unsigned long long m = mask[w][d] & (l == 63u ? ~0ull : ((2ull << l) - 1ull));
        if (m) return base + w * 64u + 63u - (uint32_t)__builtin_clzll(m);
        const uint32_t b = before[w][d] ? before[w][d] : pre[d];
        return b - 1u;  // position 0 is a boundary at every depth, so b >= 1
    }

    // First boundary at depth d after position x (x in the tile), or n.
    __device__ __forceinline__ uint32_t next(uint32_t base, uint32_t x, uint32_t d) const {
        const uint32_t o = x - base, w = o >> 6, l = o & 63u;
        const unsigned long long m = l == 63u ? 0ull : (mask[w][d] & (~0ull << (l + 1u)));
        if (m) return base + w * 64u + (uint32_t)__builtin_ctzll(m);
        return after[w][d] != kNone ? after[w][d] : post[d];
    }
};

template <uint32_t kTT, uint32_t kThreads = kTT>
__device__ __forceinline__ void topo_count_body(const SimDesc* __restrict__ sims,
                                                const TileRef* __restrict__ ttiles,
                                                uint32_t* __restrict__ ctl,
                                                unsigned long long* __restrict__ tflags,
                                                const uint32_t* __restrict__ psum, uint32_t block,
                                                uint32_t grid) {
    // kTT positions on kThreads threads, kPer per thread (position j * kThreads
    // + threadIdx.x of the tile for j < kPer) for the columns, their scan and
    // the body nodes; the cells are then written cell-parallel: the tile's
    // leaves (one per column, listed in position order) and its internal cells
    // (listed in preorder, in passes of kTT) each get a thread of their own, so
    // a column of many cells does not hold its wavefront for every other lane.
    // More positions per thread spread the workgroup's fixed latency chain
    // (ticket, loads, look-back) over more of them.
    constexpr uint32_t kPer = kTT / kThreads;
    static_assert(kTT % kThreads == 0 && kThreads % 64 == 0, "whole wavefronts per slice");
    __shared__ uint32_t s_scan[kThreads / 64];
    __shared__ uint32_t s_ticket;
    __shared__ uint32_t s_excl, s_excl_cols;
    __shared__ TileBounds<kTT> tb;
    __shared__ uint32_t s_key[kTT];   // keys of the tile
    __shared__ uint32_t s_mine[kTT];  // exclusive (columns << 16 | cells) prefix in the tile
    __shared__ float4 s_body[kTT];    // (x, y, m) at the tile's positions
    __shared__ uint16_t s_list[kTT];  // column heads, then internal cells (tile offsets of their columns)
    __shared__ uint8_t s_dmin[kTT], s_count[kTT];
    __shared__ float s_T[kLevels];  // per-depth thresholds (a global load among the stores would wait for them)
    // Leaf sizes up to kLocalBucket: the columns were counted by k_topo_pre and
    // the tile's place in the preorder comes from k_topo_scan, so tiles are
    // independent and the block index is the tile.  Larger leaves: the columns
    // are counted here and the place comes from a decoupled look-back, the
    // simulation from the block index and the tile from a ticket of that
    // simulation (as in k_sort_pass), so a look-back only waits for tiles whose
    // workgroups run.
    TileRef tr = ttiles[block];
    const uint32_t sim = tr.sim;
    const SimDesc& S = sim_at(sims, sim);
    const bool counted = S.bucket <= kLocalBucket;
    uint32_t t = block;
    if (!counted) {  // (uniform)
        if (threadIdx.x == 0) s_ticket = atomicAdd(&S.meta->ticket[4], 1u);
        lds_barrier();
        // the simulation's tiles are consecutive and of kTT positions, so the
        // ticket gives the tile without a load of the tile table
        const uint32_t local = s_ticket;
        if (local >= (S.n + kTT - 1u) / kTT) {  // tickets not reset: no tile of this simulation is left
            if (threadIdx.x == 0) atomicOr(&S.meta->err, kErrLookback);
            return;
        }
        t = S.first_topo_tile + local;
        tr.base = local * kTT;
    }
    if (threadIdx.x < kLevels) s_T[threadIdx.x] = S.meta->T[threadIdx.x];
    const uint32_t n = S.n, bucket = S.bucket, maxd = S.maxd;
    const uint32_t* __restrict__ K = S.key0;
    const uint32_t base = tr.base;
    const uint32_t tile_end = min(base + kTT, n);
    float4 body[kPer];
    uint32_t count[kPer], pl[kPer], mine[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t o = j * kThreads + threadIdx.x, p = base + o;
        s_key[o] = p < n ? K[p] : 0u;
        s_dmin[o] = p < n ? S.cdmin[p] : (uint8_t)17u;
        count[j] = counted && p < n ? S.ccount[p] : 0u;
        body[j] = S.sbody[min(p, n - 1u)];
        s_body[o] = body[j];
    }
    tb.load_summaries(psum + (size_t)t * kTopoSum);
    lds_barrier();
    tb.build(base, n, s_dmin);
    auto key_at = [&](uint32_t q) { return q - base < tile_end - base ? s_key[q - base] : K[q]; };
    // columns: existence from the parent group's size, leaf depth from the key
    // at p + bucket (or the count of k_topo_pre); the parent group's start
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t o = j * kThreads + threadIdx.x, p = base + o;
        pl[j] = 0;
        const uint32_t dmin = s_dmin[o];
        if (counted) {
            if (count[j] > 0u && dmin >= 1u) pl[j] = tb.prev(base, p, dmin - 1u);
        } else if (p < n && dmin <= maxd) {
            count[j] = 0;
            const uint32_t kp = s_key[o];
            bool exists = true;
            if (dmin >= 1) {
                const uint32_t dd = dmin - 1;
                pl[j] = tb.prev(base, p, dd);
                const uint32_t q = pl[j] + bucket;
                exists = (q < n) && same_prefix(key_at(q), kp, dd);
            }
            if (exists) {
                const int L = (p + bucket < n) ? (int)lcp2(kp, key_at(p + bucket)) : -1;
                int dl = L + 1 < (int)maxd ? L + 1 : (int)maxd;
                if (dl < (int)dmin) dl = (int)dmin;
                count[j] = (uint32_t)dl - dmin + 1;
            }
        }
        s_count[o] = (uint8_t)count[j];
    }
    // cells (low 16 bits) and columns (high 16 bits) in position order, slice by
    // slice: a tile holds at most 17 kTT cells
    uint32_t total = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        uint32_t part;
        mine[j] = total + block_excl_sum<kThreads>(count[j] | ((count[j] > 0 ? 1u : 0u) << 16), s_scan, part);
        s_mine[j * kThreads + threadIdx.x] = mine[j];
        total += part;
    }
    if (counted) {
        if (threadIdx.x == 0) {
            s_excl = psum[(size_t)t * kTopoSum + 2 * kLevels];
            s_excl_cols = psum[(size_t)t * kTopoSum + 2 * kLevels + 1];
            if (base + kTT >= n) {
                const uint32_t ncells = s_excl + (total & 0xffffu);
                const uint32_t ncols = s_excl_cols + (total >> 16);
                S.meta->ncells = ncells;
                S.meta->nnodes = ncells - ncols + n;  // internal cells + bodies
            }
        }
    } else if (threadIdx.x < 64) {  // wavefront 0 publishes and looks back
        unsigned long long* f = tflags + t;
        const unsigned long long tot = pack2(total & 0xffffu, total >> 16);
        unsigned long long excl;
        if (t == S.first_topo_tile) {
            excl = 0;
            if (threadIdx.x == 0)
                __hip_atomic_store(f, kFlag64Inc | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (threadIdx.x == 0)
                __hip_atomic_store(f, kFlag64Agg | tot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            excl = lookback2_wave(tflags, t, &ctl[kCtlError]);
            if (threadIdx.x == 0)
                __hip_atomic_store(f, kFlag64Inc | (excl + tot), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        s_excl = (uint32_t)(excl & 0x7fffffffull);
        s_excl_cols = (uint32_t)(excl >> 31);
        if (threadIdx.x == 0 && base + kTT >= n) {
            const uint32_t ncells = s_excl + (total & 0xffffu);
            const uint32_t ncols = s_excl_cols + (total >> 16);
            S.meta->ncells = ncells;
            S.meta->nnodes = ncells - ncols + n;  // internal cells + bodies
        }
    }
    // (the barriers from here on order LDS only: no wait for the global stores)
    lds_barrier();
    const uint32_t excl = s_excl, cell_cap = S.cell_cap;
    // Per position: column bookkeeping and the body's node.  The node stream is
    // the preorder of internal cells with each leaf replaced by its bodies, so
    // the body at p follows every internal cell of the columns at or before p:
    // node nb + count - 1 for a column head, nb (= that of its leaf's head plus
    // its offset in the leaf) for a position inside a leaf.
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        const uint32_t o = j * kThreads + threadIdx.x, p = base + o;
        if (p >= n) break;
        const uint32_t cbase = excl + (mine[j] & 0xffffu);
        const uint32_t nb = cbase - (s_excl_cols + (mine[j] >> 16)) + p;  // internal cells and bodies before p
        S.cbase[p] = cbase;
        S.nbase[p] = nb;
        S.ccount[p] = (uint8_t)count[j];
        S.cparent[p] = pl[j];
        const uint32_t bn = nb + (count[j] > 0u ? count[j] - 1u : 0u);
        if (bn < S.node_cap) {
            S.nodes[bn] = NodeHot{body[j].x, body[j].y, body[j].z, node_word(1u, 0u, S.wlayout)};
            S.bnode[p] = bn;
        }
        if (count[j] > 0u) s_list[mine[j] >> 16] = (uint16_t)o;
    }
    lds_barrier();
    // Leaves, one thread each: the bodies [p, e) summed in f64 in order
    // (fill_aggregates leaf branch, flat.rs:181-195), from LDS inside the tile.
    // (cell geometry is not stored: evidence readbacks derive it from the
    // cell's start key and depth, as child_bounds does, flat.rs:96-103)
    const uint32_t ncols = total >> 16, nint = (total & 0xffffu) - ncols;
    for (uint32_t jl = threadIdx.x; jl < ncols; jl += kThreads) {
        const uint32_t oc = s_list[jl];
        const uint32_t pc = base + oc;
        const uint32_t cnt = s_count[oc], dm = s_dmin[oc];
        const uint32_t dleaf = dm + cnt - 1u;
        const uint32_t idx = excl + (s_mine[oc] & 0xffffu) + cnt - 1u;
        if (idx < cell_cap) {
            const uint32_t e = tb.next(base, pc, dleaf);
            double mass = 0.0, wx = 0.0, wy = 0.0;
            for (uint32_t q = pc; q < e; ++q) {
                const float4 bq = q < tile_end ? s_body[q - base] : S.sbody[q];
                const double bm = bq.z;
                mass = mass + bm;
                wx = wx + bm * (double)bq.x;
                wy = wy + bm * (double)bq.y;
            }
            const double comx = wx / mass, comy = wy / mass;
            S.cagg[idx] = Agg{mass, comx, comy, 0.0};
            S.cells[idx] = CellT{(float)comx, (float)comy, (float)mass, s_T[dleaf], pc, e, 0u, dleaf | kInfoLeaf};
        } else {
            atomicOr(&S.meta->err, kErrCellCapacity);
        }
    }
    // Internal cells, one thread each: range end and first child (the same
    // column, depth d + 1).  Their records are written whole by k_topo_link
    // once their aggregates are known.  The internal cells of the column at
    // offset o are the tile's internal cells iex .. iex + count - 2 in preorder.
    for (uint32_t c0 = 0; c0 < nint; c0 += kTT) {
        lds_barrier();  // the previous list is consumed
#pragma unroll
        for (uint32_t j = 0; j < kPer; ++j) {
            if (count[j] > 1u) {
                const uint32_t iex = (mine[j] & 0xffffu) - (mine[j] >> 16);
                const uint32_t lo = max(iex, c0), hi = min(iex + count[j] - 1u, c0 + kTT);
                for (uint32_t i = lo; i < hi; ++i) s_list[i - c0] = (uint16_t)(j * kThreads + threadIdx.x);
            }
        }
        lds_barrier();
        const uint32_t m = min(kTT, nint - c0);
        for (uint32_t jc = threadIdx.x; jc < m; jc += kThreads) {
            const uint32_t oc = s_list[jc];
            const uint32_t mc = s_mine[oc];
            const uint32_t k = c0 + jc - ((mc & 0xffffu) - (mc >> 16));
            const uint32_t d = s_dmin[oc] + k;
            const uint32_t idx = excl + (mc & 0xffffu) + k;
            if (idx < cell_cap) {
                const uint32_t q0 = (s_key[oc] >> (30u - 2u * d)) & 3u;  // first child's quadrant
                S.cend[idx] = tb.next(base, base + oc, d);
                S.cchild[idx] = make_uint4(q0 == 0u ? idx + 1u : kNoChild, q0 == 1u ? idx + 1u : kNoChild,
                                           q0 == 2u ? idx + 1u : kNoChild, q0 == 3u ? idx + 1u : kNoChild);
            } else {
                atomicOr(&S.meta->err, kErrCellCapacity);
            }
        }
    }
}

template <uint32_t kTT, uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_topo_count(const SimDesc* __restrict__ sims,
                                                         const TileRef* __restrict__ ttiles,
                                                         uint32_t* __restrict__ ctl,
                                                         unsigned long long* __restrict__ tflags,
                                                         const uint32_t* __restrict__ psum) {
    topo_count_body<kTT, kThreads>(sims, ttiles, ctl, tflags, psum, blockIdx.x, gridDim.x);
}

// Internal-cell aggregate from its children in quadrant order
// (fill_aggregates internal branch, flat.rs:196-213).
__device__ __forceinline__ Agg combine_children(const Agg* __restrict__ cagg, uint4 ch) {
    double mass = 0.0, wx = 0.0, wy = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t c = (&ch.x)[q];
        if (c != kNoChild) {
            const Agg a = cagg[c];
            mass = mass + a.m;
            wx = wx + a.m * a.x;
            wy = wy + a.m * a.y;
        }
    }
    return Agg{mass, wx / mass, wy / mass, 0.0};
}

// Aggregate of an internal cell lying inside its tile from its children's
// aggregates in global memory (tiles whose cells exceed the LDS stage).
__device__ __forceinline__ Agg combine_children_if(const SimDesc& S, uint32_t idx, bool inside) {
    return inside ? combine_children(S.cagg, S.cchild[idx]) : Agg{0.0, 0.0, 0.0, 0.0};
}

// Records of the internal cell idx (column head position p, depth d, node ni,
// range end e).  Inside its tile: aggregate, cell, hot node and, from LDS, child
// links.  Crossing a tile edge: a span-list slot, the cell and node without
// aggregates (k_agg_top fills them) and the child links held in LDS (links to
// children in other tiles were written by those tiles).  w = 0 means the node
// word is still to be computed from the skip target.
__device__ __forceinline__ void link_write_cell(const SimDesc& S, uint32_t idx, uint32_t ni, uint32_t p, uint32_t d,
                                                bool inside, const Agg& a, uint32_t e, uint32_t w,
                                                const uint32_t* lds_ch, uint32_t slot = kNone,
                                                bool span_node = true) {
    const float T = S.meta->T[d];
    if (w == 0u && (inside || span_node)) w = node_word((e < S.n ? S.nbase[e] : S.meta->nnodes) - ni, d, S.wlayout);
    if (inside) {
        if (lds_ch) S.cchild[idx] = make_uint4(lds_ch[0], lds_ch[1], lds_ch[2], lds_ch[3]);
        S.cagg[idx] = a;
        S.cells[idx] = CellT{(float)a.x, (float)a.y, (float)a.m, T, p, e, 0u, d};
        S.nodes[ni] = NodeHot{(float)a.x, (float)a.y, (float)a.m, w};
    } else {
        uint32_t info = d;
        if (slot == kNone) slot = atomicAdd(&S.meta->nspan, 1u);
        if (slot < S.span_cap) {
            S.span[slot] = make_uint2(idx, (d << 26) | ni);  // node indices are below kMaxNodes = 2^26
            info |= (slot + 1u) << kInfoSpanShift;
        } else {
            atomicOr(&S.meta->err, kErrSpanCapacity);
        }
        // aggregates follow in k_agg_top (which writes the whole node); the
        // cell's other fields are final
        S.cells[idx] = CellT{0.f, 0.f, 0.f, T, p, e, 0u, info};
        if (span_node) S.nodes[ni] = NodeHot{0.f, 0.f, 0.f, w};
        if (lds_ch) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (lds_ch[q] != kNoChild) (&S.cchild[idx].x)[q] = lds_ch[q];
        }
    }
}

// ----------------------------------------------------------------------------
// K5: skip links (next cell after each subtree), the remaining child links, and
// bottom-up aggregates of every internal cell whose range lies inside this tile.
// Internal cells that cross a tile boundary go to the simulation's span list.
// The tile's cells are a contiguous preorder range [cb0, cb1).  When it fits
// kLinkCells (1.5 cells per position), the work is cell-parallel with the state
// in LDS: each internal cell gets a thread for its loads, node word and cold
// record; the tile's internal cells are bucketed by depth, so each round of the
// bottom-up pass runs exactly the cells of its depth, one per thread, on
// children links and f64 aggregates held in LDS; and every internal cell's
// records (cell, hot node, child links, aggregate) are then written whole, once,
// one cell per thread: no partial updates of lines written by another kernel.
// Larger tiles run the same rounds per column through global memory.
// ----------------------------------------------------------------------------
constexpr uint32_t kLinkBatch = 4;  // cells of a column whose loads are issued together
constexpr uint16_t kNoLocal = 0xffffu;  // child link in LDS: absent or outside the tile's cells

// The global-memory form (tiles whose cells exceed the LDS stage), one column
// per thread: p, its column (count cells from preorder index base, node index
// nb, head depth dmin, key kp) and the parent loads (pl_base, pl_dmin).
template <uint32_t kTT>
__device__ __forceinline__ void topo_link_columns(const SimDesc& S, uint32_t p, uint32_t tile_end, uint32_t count,
                                                  uint32_t dmin, uint32_t base, uint32_t nb, uint32_t kp,
                                                  uint32_t pl_base, uint32_t pl_dmin) {
    __shared__ uint32_t s_dlo, s_dhi;
    const uint32_t n = S.n, ncells = S.meta->ncells, nnodes = S.meta->nnodes;
    if (threadIdx.x == 0) {
        s_dlo = 16u;
        s_dhi = 0u;
    }
    __syncthreads();  // (before the atomics below)
    uint32_t inside = 0;  // bit k: cell base + k is internal and lies inside the tile
    if (count > 0u) {
        // The column's cells, kLinkBatch at a time: each batch loads its range
        // ends, then their skip targets and thresholds, then stores.  (The
        // compiler cannot move a load across a store to another array, so a
        // load-store loop would pay two memory latencies per cell.)
        // (Loads at clamped indices, unconditional: a conditional load would be
        // waited for inside its branch, one cell after the other.)
        for (uint32_t k0 = 0; k0 < count; k0 += kLinkBatch) {
            uint32_t ev[kLinkBatch], sk[kLinkBatch];
            float tv[kLinkBatch];
#pragma unroll
            for (uint32_t j = 0; j < kLinkBatch; ++j) {
                const uint32_t e = S.cend[base + min(k0 + j, count - 1u)];
                ev[j] = k0 + j + 1u < count ? e : 0u;
            }
            uint32_t nbe[kLinkBatch];
            float trr[kLinkBatch];
#pragma unroll
            for (uint32_t j = 0; j < kLinkBatch; ++j) {
                nbe[j] = S.nbase[min(ev[j], n - 1u)];
                trr[j] = S.meta->T[min(dmin + k0 + j, kLevels - 1u)];
            }
            // The values are consumed here, all at once: otherwise the compiler
            // moves each load under the condition that selects it and waits for
            // it there, one cell after the other.
            static_assert(kLinkBatch == 4, "the register list below names four cells");
            asm volatile("" : "+v"(nbe[0]), "+v"(nbe[1]), "+v"(nbe[2]), "+v"(nbe[3]), "+v"(trr[0]), "+v"(trr[1]),
                         "+v"(trr[2]), "+v"(trr[3]));
#pragma unroll
            for (uint32_t j = 0; j < kLinkBatch; ++j) {
                const bool internal = k0 + j + 1u < count;
                sk[j] = internal ? (ev[j] < n ? nbe[j] : nnodes) : 0u;
                tv[j] = internal ? trr[j] : 0.0f;
            }
#pragma unroll
            for (uint32_t j = 0; j < kLinkBatch; ++j) {
                const uint32_t k = k0 + j;
                if (k + 1u >= count) break;
                const uint32_t e = ev[j];
                if (!S.lean) S.ncold[nb + k] = NodeCold{p, e - p, tv[j], dmin + k};
                if (e <= tile_end) inside |= 1u << k;
            }
        }
        if (dmin >= 1) {
            const uint32_t parent = pl_base + (dmin - 1u - pl_dmin);
            const uint32_t q = (kp >> (30u - 2u * (dmin - 1u))) & 3u;
            if (pl_dmin <= dmin - 1u && parent < ncells && parent < S.cell_cap) (&S.cchild[parent].x)[q] = base;
            else atomicOr(&S.meta->err, kErrCellCapacity);
        }
    }
    // depth range of the tile's internal cells: reduced per wavefront first (an
    // LDS atomic per thread would serialize on two words)
    {
        const uint32_t lo = wave_incl_min_rev(count > 1u ? dmin : 16u);
        const uint32_t hi = wave_incl_max(count > 1u ? dmin + count - 2u : 0u);
        if (lane_id() == 0) atomicMin(&s_dlo, lo);
        if (lane_id() == 63) atomicMax(&s_dhi, hi);
    }
    __syncthreads();
    const uint32_t dlo = s_dlo, dhi = s_dhi;
    for (int d = (int)dhi; d >= (int)dlo; --d) {
        const uint32_t k = (uint32_t)d - dmin;
        if (count > 1u && (uint32_t)d >= dmin && k + 1u < count) {
            const uint32_t idx = base + k;
            const bool in = (inside & (1u << k)) != 0u;
            link_write_cell(S, idx, nb + k, p, (uint32_t)d, in, combine_children_if(S, idx, in), S.cend[idx], 0u,
                            nullptr);
        }
        __syncthreads();  // the next depth reads this depth's global stores
    }
}

template <uint32_t kTT>
__device__ __forceinline__ void topo_link_body(const SimDesc* __restrict__ sims,
                                               const TileRef* __restrict__ ttiles, uint32_t block,
                                               uint32_t grid) {
    // The LDS stage holds 2 cells per position (trees with one body per leaf
    // have about 1.7), kept under 40 KB for 512-position tiles so that four
    // workgroups share a CU: cells are 16-bit offsets from cb0, range ends
    // 16-bit offsets from the tile base (kSpanEnd: past the tile), node words
    // are formed when the records are written.
    constexpr uint32_t kLinkCells = 2 * kTT;
    constexpr uint16_t kSpanEnd = 0xffffu;
    static_assert(kLinkCells < kNoLocal && kTT <= 1024, "LDS cell offsets and bucket entries are 16-bit");
    // per cell of the tile (offset from cb0)
    __shared__ double s_m[kLinkCells], s_x[kLinkCells], s_y[kLinkCells];  // f64 aggregates
    __shared__ uint16_t s_ch[kLinkCells][4];                               // children (offsets)
    __shared__ uint16_t s_e[kLinkCells];     // range end - tile base (kSpanEnd: beyond the tile); first a rank
    __shared__ uint16_t s_bcol[kLinkCells];  // internal cells bucketed by depth: (column offset << 5) | k
    // per position of the tile
    __shared__ uint16_t s_pcb[kTT], s_pnb[kTT];  // the column's first cell - cb0, its node - nb0
    __shared__ uint8_t s_pdm[kTT];
    __shared__ uint32_t s_dcnt[kLevels], s_doff[kLevels + 1];
    __shared__ float s_T[kLevels];
    __shared__ uint32_t s_cb0, s_cb1, s_nb0, s_nbend, s_maxb;
    const TileRef tr = ttiles[xcd_tile(block, grid)];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t n = S.n;
    const uint32_t o = threadIdx.x;
    const uint32_t p = tr.base + o;
    const bool valid = p < n;
    const uint32_t tile_end = tr.base + kTT;
    const uint32_t last = min(n, tile_end) - 1u;
    uint32_t count = 0, dmin = 17, base = 0, nb = 0;
    if (valid) {
        count = S.ccount[p];
        dmin = S.cdmin[p];
        base = S.cbase[p];
        nb = S.nbase[p];
    }
    if (o == 0) {
        s_cb0 = base;  // cells before the tile
        s_nb0 = nb;    // nodes before the tile
        s_nbend = tile_end < n ? S.nbase[tile_end] : S.meta->nnodes;  // node index after the tile
    }
    if (p == last) s_cb1 = base + count;
    if (o < kLevels) {
        s_dcnt[o] = 0u;
        s_T[o] = S.meta->T[o];
    }
    bool ok = count > 0 && base + count <= S.cell_cap;
    // loads that do not depend on the column's cells, issued before any store
    uint32_t pl_base = 0, pl_dmin = 0;
    if (ok && dmin >= 1) {
        const uint32_t pl = S.cparent[p];
        pl_base = S.cbase[pl];
        pl_dmin = S.cdmin[pl];
    }
    const uint32_t kp = ok ? S.key0[p] : 0u;
    lds_barrier();  // (LDS only: the parent loads stay in flight until the links below)
    const uint32_t cb0 = s_cb0, cb1 = s_cb1, nb0 = s_nb0;
    const uint32_t ntile = cb1 - cb0;
    if (!(cb0 <= cb1 && ntile <= kLinkCells)) {  // (uniform) the tile's cells exceed the LDS stage
        topo_link_columns<kTT>(S, p, tile_end, ok ? count : 0u, dmin, base, nb, kp, pl_base, pl_dmin);
        return;
    }
    const uint32_t ncells = S.meta->ncells;
    // a consistent scan puts every column of the tile inside [cb0, cb1), and its
    // node indices within the tile's positions and internal cells
    if (ok && (base < cb0 || base + count > cb1 || nb < nb0 || nb - nb0 > kTT + ntile)) {
        atomicOr(&S.meta->err, kErrCellCapacity);
        ok = false;
    }
    if (!ok) count = 0;
    s_pcb[o] = (uint16_t)(base - cb0);
    s_pnb[o] = (uint16_t)(nb - nb0);
    s_pdm[o] = (uint8_t)dmin;
    for (uint32_t i = o; i < ntile; i += kTT) {
#pragma unroll
        for (int q = 0; q < 4; ++q) s_ch[i][q] = kNoLocal;
    }
    Agg leaf{0.0, 0.0, 0.0, 0.0};
    if (count > 0u) leaf = S.cagg[base + count - 1u];  // the leaf's aggregate (k_topo_count)
    // rank of each internal cell among the tile's cells of its depth (held in s_e until the list is built)
    for (uint32_t k = 0; k + 1u < count; ++k)
        s_e[base - cb0 + k] = (uint16_t)atomicAdd(&s_dcnt[dmin + k], 1u);
    // (the barriers from here on order LDS only: no wait for the global stores)
    lds_barrier();
    if (o < 64) {  // bucket offsets: exclusive prefix of the per-depth counts
        const uint32_t c = o < kLevels ? s_dcnt[o] : 0u;
        const uint32_t incl = wave_incl_sum_u32(c);
        if (o <= kLevels) s_doff[o] = incl - c;
        const uint32_t mx = wave_incl_max(c);
        if (o == 63) s_maxb = mx;  // the largest bucket
    }
    lds_barrier();
    // Per column: the leaf's aggregate, the column's internal cells into their
    // depth buckets with the link to their first child (the same column, depth
    // d + 1), and the link from the parent group's cell to the column head.
    if (count > 0u) {
        const uint32_t r0 = base - cb0;
        const uint32_t r = r0 + count - 1u;
        s_m[r] = leaf.m;
        s_x[r] = leaf.x;
        s_y[r] = leaf.y;
        for (uint32_t k = 0; k + 1u < count; ++k) {
            const uint32_t d = dmin + k;
            s_bcol[s_doff[d] + s_e[r0 + k]] = (uint16_t)((o << 5) | k);
            s_ch[r0 + k][(kp >> (30u - 2u * d)) & 3u] = (uint16_t)(r0 + k + 1u);
        }
        if (dmin >= 1) {
            const uint32_t parent = pl_base + (dmin - 1u - pl_dmin);
            const uint32_t q = (kp >> (30u - 2u * (dmin - 1u))) & 3u;
            if (pl_dmin <= dmin - 1u && parent < ncells && parent < S.cell_cap) {
                if (parent >= cb0 && parent < cb1) s_ch[parent - cb0][q] = (uint16_t)r0;
                else (&S.cchild[parent].x)[q] = base;  // a parent outside the tile's cells
            } else {
                atomicOr(&S.meta->err, kErrCellCapacity);
            }
        }
    }
    lds_barrier();
    const uint32_t nint = s_doff[kLevels];
    // Internal cells, one per thread: range end and cold record.
    for (uint32_t j = o; j < nint; j += kTT) {
        const uint32_t ent = s_bcol[j];
        const uint32_t oc = ent >> 5, k = ent & 31u;
        const uint32_t r = s_pcb[oc] + k, d = s_pdm[oc] + k, pc = tr.base + oc;
        const uint32_t e = S.cend[cb0 + r];
        if (!S.lean) S.ncold[nb0 + s_pnb[oc] + k] = NodeCold{pc, e - pc, s_T[d], d};
        s_e[r] = e <= tile_end ? (uint16_t)(e - tr.base) : kSpanEnd;
    }
    lds_barrier();
    // Bottom-up rounds, deepest depth first, over the cells of each depth that
    // lie inside the tile (fill_aggregates internal branch, flat.rs:196-213):
    // LDS traffic only.  When no depth has more than 64 internal cells,
    // wavefront 0 runs every round alone: a wavefront's LDS accesses take
    // effect in order, so the rounds need no workgroup barriers.
    auto round_cell = [&](uint32_t j) {
        const uint32_t ent = s_bcol[j];
        const uint32_t r = s_pcb[ent >> 5] + (ent & 31u);
        if (s_e[r] == kSpanEnd) return;  // spanning: k_agg_top
        double mass = 0.0, wx = 0.0, wy = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t c = s_ch[r][q];
            if (c != kNoLocal) {
                const double cm = s_m[c];
                mass = mass + cm;
                wx = wx + cm * s_x[c];
                wy = wy + cm * s_y[c];
            }
        }
        s_m[r] = mass;
        s_x[r] = wx / mass;
        s_y[r] = wy / mass;
    };
    if (s_maxb <= 64u) {  // (uniform)
        if (o < 64u) {
            for (int d = 15; d >= 0; --d) {
                const uint32_t lo = s_doff[d], hi = s_doff[d + 1];
                if (lo == hi) continue;  // (uniform) no internal cell of this depth
                if (lo + o < hi) round_cell(lo + o);
                // (the next round's loads stay behind this round's stores)
                asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
            }
        }
        lds_barrier();
    } else {
        for (int d = 15; d >= 0; --d) {
            const uint32_t lo = s_doff[d], hi = s_doff[d + 1];
            if (lo == hi) continue;  // (uniform) no internal cell of this depth
            for (uint32_t j = lo + o; j < hi; j += kTT) round_cell(j);
            lds_barrier();
        }
    }
    // Records, one internal cell per thread.  A cell inside the tile skips to
    // the node of a position in the tile (or the one after it).  A tile-crossing
    // cell takes a span slot (one atomic per wavefront) and the links to its
    // children that lie in the tile (those in later tiles were or will be
    // written by their tiles); k_agg_top writes its aggregate and whole node.
    const uint32_t nbend = s_nbend, lim = min(tile_end, n);
    for (uint32_t j0 = 0; j0 < nint; j0 += kTT) {  // (uniform trip count)
        const uint32_t j = j0 + o;
        const bool have = j < nint;
        uint32_t r = 0, oc = 0, k = 0, eo = 0;
        if (have) {
            const uint32_t ent = s_bcol[j];
            oc = ent >> 5;
            k = ent & 31u;
            r = s_pcb[oc] + k;
            eo = s_e[r];
        }
        const bool in = eo != kSpanEnd;
        const unsigned long long span = __ballot(have && !in);
        uint32_t slot = kNone;
        if (span) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(span);
            uint32_t first = 0u;
            if (lane_id() == leader) first = atomicAdd(&S.meta->nspan, (uint32_t)__builtin_popcountll(span));
            first = __shfl(first, (int)leader);
            slot = first + (uint32_t)__builtin_popcountll(span & ((1ull << lane_id()) - 1ull));
        }
        if (!have) continue;
        const uint32_t idx = cb0 + r, ni = nb0 + s_pnb[oc] + k, d = s_pdm[oc] + k;
        uint32_t e, w = 0;
        if (in) {
            e = tr.base + eo;
            const uint32_t skip = e < lim ? nb0 + s_pnb[eo] : nbend;
            w = node_word(skip - ni, d, S.wlayout);  // internal depths are 0..15
        } else {
            e = S.cend[idx];
        }
        uint32_t ch[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) ch[q] = s_ch[r][q] == kNoLocal ? kNoChild : cb0 + s_ch[r][q];
        const Agg a{s_m[r], s_x[r], s_y[r], 0.0};
        link_write_cell(S, idx, ni, tr.base + oc, d, in, a, e, w, ch, in ? kNone : slot, false);
    }
}

template <uint32_t kTT>
__global__ __launch_bounds__(kTT) void k_topo_link(const SimDesc* __restrict__ sims,
                                                   const TileRef* __restrict__ ttiles) {
    topo_link_body<kTT>(sims, ttiles, blockIdx.x, gridDim.x);
}

// The whole topology of a set with a single topology tile (one simulation of
// at most kTopoTile bodies) in one workgroup: boundary pass, the scan of one
// tile (no boundary before it, none after it but n), columns and cells,
// links and aggregates, then the error gather of k_agg_top.  No cell crosses
// a tile edge, so there are no spanning cells to aggregate.  Barriers between
// the phases also order their global stores before the next phase's loads.
__device__ __forceinline__ void topo_single_body(const SimDesc* __restrict__ sims,
                                                 const TileRef* __restrict__ ttiles, uint32_t* __restrict__ ctl,
                                                 unsigned long long* __restrict__ tflags,
                                                 uint32_t* __restrict__ tsum, uint32_t* __restrict__ psum,
                                                 uint32_t* __restrict__ errs) {
    const SimDesc& S = sim_at(sims, ttiles[0].sim);
    topo_pre_body<kTopoTile>(sims, ttiles, tsum, 0u, 1u);
    if (threadIdx.x < kLevels) {
        psum[threadIdx.x] = 0u;
        psum[kLevels + threadIdx.x] = S.n;
    }
    if (threadIdx.x < 2) psum[2 * kLevels + threadIdx.x] = 0u;  // no cells or columns before the tile
    __syncthreads();
    topo_count_body<kTopoTile>(sims, ttiles, ctl, tflags, psum, 0u, 1u);
    __syncthreads();
    topo_link_body<kTopoTile>(sims, ttiles, 0u, 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (S.meta->nspan != 0u) atomicOr(&S.meta->err, kErrSpanCapacity);  // cannot happen
        errs[0] = ctl[kCtlError];
        errs[1] = S.meta->err;
    }
}

__global__ __launch_bounds__(kTopoTile) void k_topo_single(const SimDesc* __restrict__ sims,
                                                           const TileRef* __restrict__ ttiles,
                                                           uint32_t* __restrict__ ctl,
                                                           unsigned long long* __restrict__ tflags,
                                                           uint32_t* __restrict__ tsum,
                                                           uint32_t* __restrict__ psum,
                                                           uint32_t* __restrict__ errs) {
    topo_single_body(sims, ttiles, ctl, tflags, tsum, psum, errs);
}

// The whole build of a set of one simulation of at most kTopoTile bodies in one
// workgroup, one body per thread: the kick/drift (mode as in k_bounds) and the
// resets of k_bounds, root bounds, thresholds and Morton keys as in k_keys, the
// four stable radix passes with the ranking of sort_pass_body held in LDS (no
// look-back: the set has one tile), the sorted keys and indices to key0/val0,
// then the topology of k_topo_single.  The sort leaves the per-pass digit
// histograms and look-back flags unused.
__global__ __launch_bounds__(kTopoTile) void k_build_small(const SimDesc* __restrict__ sims,
                                                           const TileRef* __restrict__ ttiles,
                                                           uint32_t* __restrict__ ctl,
                                                           unsigned long long* __restrict__ tflags,
                                                           uint32_t* __restrict__ tsum,
                                                           uint32_t* __restrict__ psum,
                                                           uint32_t* __restrict__ errs, int mode,
                                                           uint32_t clear_errors) {
    constexpr uint32_t kWaves = kTopoTile / 64;
    __shared__ float s_red[4][kWaves];
    __shared__ double s_root[3];
    __shared__ uint32_t s_key[kTopoTile], s_val[kTopoTile];
    __shared__ uint32_t s_wcnt[kWaves][kRadix];
    __shared__ uint32_t s_start[kRadix];
    __shared__ uint32_t s_scan[kWaves];
    const SimDesc& S = sim_at(sims, ttiles[0].sim);
    const uint32_t n = S.n;
    const uint32_t i = threadIdx.x;
    const bool valid = i < n;
    const uint32_t w = i >> 6;
    // kick/drift and bounds (k_bounds), the body's loads first
    const uint32_t ic = min(i, n - 1u);
    float2 p = S.pos[ic];
    float2 v = make_float2(0.f, 0.f), a = make_float2(0.f, 0.f);
    if (mode != 0) {
        v = S.vel[ic];
        a = S.acc[ic];
    }
    if (i < kXcds) ctl[kCtlQueue0 + i * kQueueStride] = 0u;
    if (i == 0) tflags[0] = 0ull;
    if (i < 8) S.meta->ticket[i] = 0u;
    if (clear_errors && i == 0) {
        S.meta->err = 0u;
        ctl[kCtlError] = 0u;
    }
    if (valid && mode != 0) {
        const float dt = S.dt;
        const float h = 0.5f * dt;
        if (mode == 2) {  // bh_final_kick of the previous step
            v.x = v.x + h * a.x;
            v.y = v.y + h * a.y;
        }
        const float hvx = v.x + h * a.x;  // bh_kick_drift
        const float hvy = v.y + h * a.y;
        p.x = p.x + dt * hvx;
        p.y = p.y + dt * hvy;
        S.pos[i] = p;
        S.vel[i] = make_float2(hvx, hvy);
    }
    float mnx = valid ? p.x : __builtin_inff(), mny = valid ? p.y : __builtin_inff();
    float mxx = valid ? p.x : -__builtin_inff(), mxy = valid ? p.y : -__builtin_inff();
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, off));
        mny = fminf(mny, __shfl_xor(mny, off));
        mxx = fmaxf(mxx, __shfl_xor(mxx, off));
        mxy = fmaxf(mxy, __shfl_xor(mxy, off));
    }
    if (lane_id() == 0) {
        s_red[0][w] = mnx;
        s_red[1][w] = mny;
        s_red[2][w] = mxx;
        s_red[3][w] = mxy;
    }
    __syncthreads();
    // root_bounds in f64 over the widened f32 positions (min/max are exact, so
    // the order of the reduction does not matter)
    if (i == 0) {
        for (uint32_t k = 1; k < kWaves; ++k) {
            mnx = fminf(mnx, s_red[0][k]);
            mny = fminf(mny, s_red[1][k]);
            mxx = fmaxf(mxx, s_red[2][k]);
            mxy = fmaxf(mxy, s_red[3][k]);
        }
        double min_x = mnx, min_y = mny, max_x = mxx, max_y = mxy;
        double sx = max_x - min_x, sy = max_y - min_y;
        double span = sy > sx ? sy : sx;
        span = 1e-12 > span ? 1e-12 : span;
        s_root[0] = (min_x + max_x) * 0.5;
        s_root[1] = (min_y + max_y) * 0.5;
        s_root[2] = span * 0.500000000001 + 1e-12;
    }
    __syncthreads();
    const double cx = s_root[0], cy = s_root[1], half = s_root[2];
    if (w == 0) {
        SimMeta* M = S.meta;
        if (i == 0) {
            M->root[0] = cx;
            M->root[1] = cy;
            M->root[2] = half;
            M->root[3] = 0.0;
            M->bounds_f32[0] = (float)cx;
            M->bounds_f32[1] = (float)cy;
            M->bounds_f32[2] = (float)half;
            M->bounds_f32[3] = 0.0f;
            M->nspan = 0u;
            M->theta_bits = __float_as_uint(S.theta);
            M->mode = 0u;
            M->geom_ok = half < 1.0e18 ? 1u : 0u;
        }
        float T = 0.0f, hf = 0.0f;
        if (i < kLevels) {
            double hd = half;
            for (uint32_t l = 0; l < i; ++l) hd = hd * 0.5;
            hf = (float)hd;
            T = accept_threshold(hf, S.theta);
            M->T[i] = T;
            if (i < kCodes) M->tab[i] = i == 0 ? 0.0f : T;
        }
        const float T0 = __shfl(T, 0);
        const uint32_t p4 = pow4_thresholds(T, T0);
        const uint32_t cr = contained_rejected(T, hf);
        if (i == 0) {
            M->pow4 = p4;
            M->t0_bits = __float_as_uint(T0);
            M->contain_rejected = cr;
        }
    }
    // Morton keys, then four stable 8-bit passes: element k of the sequence is
    // held by thread k (waves in order, lanes in order within a wave), ranked
    // among its wavefront's equal digits by ballots, and placed after the
    // smaller digits and after the earlier wavefronts' equal digits.
    uint32_t key = 0u, val = i;
    if (valid)
        key = spread16(quantize_axis((double)p.x, cx, half)) | (spread16(quantize_axis((double)p.y, cy, half)) << 1);
    const unsigned long long lt_mask = __lanemask_lt();
    for (uint32_t pass = 0; pass < 4; ++pass) {
        for (uint32_t k = i; k < kWaves * kRadix; k += kTopoTile) (&s_wcnt[0][0])[k] = 0u;
        __syncthreads();
        const uint32_t d = (key >> (8u * pass)) & 255u;
        unsigned long long peers = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t below = (uint32_t)__popcll(peers & lt_mask);
        if (valid && below == 0) s_wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t tot = 0;
        if (i < kRadix) {
            for (uint32_t k = 0; k < kWaves; ++k) {
                const uint32_t c = s_wcnt[k][i];
                s_wcnt[k][i] = tot;
                tot += c;
            }
        }
        uint32_t all;
        const uint32_t start = block_excl_sum<kTopoTile>(tot, s_scan, all);
        if (i < kRadix) s_start[i] = start;
        __syncthreads();
        if (valid) {
            const uint32_t dst = s_start[d] + s_wcnt[w][d] + below;
            s_key[dst] = key;
            s_val[dst] = val;
        }
        __syncthreads();
        if (valid) {
            key = s_key[i];
            val = s_val[i];
        }
    }
    if (valid) {
        S.key0[i] = key;
        S.val0[i] = val;
    }
    __syncthreads();
    topo_single_body(sims, ttiles, ctl, tflags, tsum, psum, errs);
}

// ----------------------------------------------------------------------------
// K6: aggregates of the spanning internal cells (the top of the tree), one
// workgroup per simulation, bottom-up by depth.  The footprint is small (256
// threads, about 40 KB of LDS) so that these workgroups run beside the other
// group's persistent traversal instead of waiting for it to end.
//  - Up to kTopLdsCells spanning cells ("LDS mode"): every global read happens
//    before the depth rounds; each spanning cell's children become references
//    to either another spanning cell's running value or a preloaded aggregate
//    of a non-spanning child, all in LDS, so a round is an LDS round trip.
//  - More ("listed mode"): the span entries are bucketed by depth into the
//    simulation's spanbuf (a counting sort), and each depth round aggregates its
//    own bucket through global memory.
// Each cell's children are summed in quadrant order, absent ones skipped
// (fill_aggregates internal branch, flat.rs:196-213).  The loads of a level
// (the four children's span slots, their aggregates) are made unconditional at
// a clamped index and issued together: a conditional load would be waited for
// inside its branch, one child after the other.
// ----------------------------------------------------------------------------
constexpr uint32_t kTopThreads = 256;
constexpr uint32_t kTopLdsCells = 256;      // spanning cells aggregated in LDS (sets of several simulations)
constexpr uint32_t kTopLdsCellsOne = 768;   // the same for a set of one simulation (nothing runs beside it)
constexpr uint32_t kTopThreadsOne = 1024;   // threads for a set of one simulation
constexpr uint32_t kSpanNodeMask = (1u << 26) - 1u;  // span entry: (cell, depth << 26 | node)

// Sum of present children in quadrant order (c[q] == kNoChild: absent).
__device__ __forceinline__ Agg sum_children(const uint32_t c[4], const double m[4], const double x[4],
                                            const double y[4]) {
    double mass = 0.0, wx = 0.0, wy = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (c[q] == kNoChild) continue;
        mass = mass + m[q];
        wx = wx + m[q] * x[q];
        wy = wy + m[q] * y[q];
    }
    return Agg{mass, wx / mass, wy / mass, 0.0};
}

// Aggregate of spanning cell idx from its children's aggregates in global memory.
__device__ __forceinline__ Agg agg_span_global(const SimDesc& S, uint32_t idx) {
    const uint4 ch = S.cchild[idx];
    uint32_t c[4];
    double m[4], x[4], y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        c[q] = (&ch.x)[q];
        const Agg a = S.cagg[c[q] == kNoChild ? idx : c[q]];
        m[q] = a.m;
        x[q] = a.x;
        y[q] = a.y;
    }
    return sum_children(c, m, x, y);
}

// A spanning cell's aggregate and its whole node (w: the node word).
__device__ __forceinline__ void store_span(const SimDesc& S, uint32_t idx, uint32_t node, const Agg& a, uint32_t w) {
    S.cagg[idx] = a;
    S.cells[idx].comx = (float)a.x;
    S.cells[idx].comy = (float)a.y;
    S.cells[idx].mass = (float)a.m;
    S.nodes[node] = NodeHot{(float)a.x, (float)a.y, (float)a.m, w};
}

// Node word of a spanning cell: its subtree runs to the node of its range end.
__device__ __forceinline__ uint32_t span_word(const SimDesc& S, uint32_t end, uint32_t node, uint32_t depth) {
    return node_word((end < S.n ? S.nbase[end] : S.meta->nnodes) - node, depth, S.wlayout);
}

// Also gathers the sticky error words as k_collect_errors does (errs[0] = the
// set's, errs[1 + s] = simulation s's): it runs last in a build, and no kernel
// after it in a step sets them.
template <uint32_t kTopLdsCells, uint32_t kThreads>
__global__ __launch_bounds__(kThreads) void k_agg_top(const SimDesc* __restrict__ sims,
                                                         const uint32_t* __restrict__ ctl,
                                                         uint32_t* __restrict__ errs) {
    constexpr uint32_t kTopLds = 5 * kTopLdsCells;  // the cells' slots, then preloaded children
    __shared__ double s_m[kTopLds], s_x[kTopLds], s_y[kTopLds];
    __shared__ uint4 s_ref[kTopLdsCells];
    __shared__ uint32_t s_idx[kTopLdsCells], s_node[kTopLdsCells];  // cell and node index
    __shared__ uint32_t s_w[kTopLdsCells];                           // node word
    __shared__ uint16_t s_list[kTopLdsCells];                        // slots bucketed by depth
    __shared__ uint8_t s_depth[kTopLdsCells];
    __shared__ uint32_t s_next, s_cnt[kLevels], s_off[kLevels + 1];
    static_assert(kTopLdsCells <= 65536, "LDS slots are 16-bit");
    const SimDesc& S = sim_at(sims, blockIdx.x);
    if (threadIdx.x == 0) {
        errs[1 + blockIdx.x] = S.meta->err;
        if (blockIdx.x == 0) errs[0] = ctl[kCtlError];
    }
    const uint32_t nspan = min(S.meta->nspan, S.span_cap);
    if (nspan == 0) return;
    if (threadIdx.x < kLevels) s_cnt[threadIdx.x] = 0u;
    if (threadIdx.x == 0) s_next = nspan;
    __syncthreads();
    if (nspan <= kTopLdsCells) {
        for (uint32_t i = threadIdx.x; i < nspan; i += kThreads) {
            // three levels of gathers: the span entry, the cell's children, then
            // the children's span slots and aggregates
            const uint2 se = S.span[i];
            const uint32_t idx = se.x, depth = min(se.y >> 26, kLevels - 1u);
            s_idx[i] = idx;
            s_node[i] = se.y & kSpanNodeMask;
            s_depth[i] = (uint8_t)depth;
            atomicAdd(&s_cnt[depth], 1u);
            const uint4 ch = S.cchild[idx];
            const uint32_t end = S.cells[idx].end;
            uint32_t c[4], info[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                c[q] = (&ch.x)[q];
                info[q] = S.cells[c[q] == kNoChild ? idx : c[q]].info;
            }
            Agg ca[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) ca[q] = S.cagg[c[q] == kNoChild ? idx : c[q]];
            s_w[i] = span_word(S, end, se.y & kSpanNodeMask, depth);
            uint4 ref;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t slot = c[q] != kNoChild ? info[q] >> kInfoSpanShift : 0u;
                const bool preload = c[q] != kNoChild && slot == 0u;
                // preload slots: one LDS atomic per wavefront, ranks from the ballot
                const unsigned long long m = __ballot(preload);
                uint32_t r = c[q] == kNoChild ? kNoChild : slot - 1u;
                if (m) {
                    const uint32_t leader = (uint32_t)__builtin_ctzll(m);
                    uint32_t first = 0u;
                    if (lane_id() == leader) first = atomicAdd(&s_next, (uint32_t)__builtin_popcountll(m));
                    first = __shfl(first, (int)leader);
                    if (preload) {
                        r = first + (uint32_t)__builtin_popcountll(m & ((1ull << lane_id()) - 1ull));
                        s_m[r] = ca[q].m;
                        s_x[r] = ca[q].x;
                        s_y[r] = ca[q].y;
                    }
                }
                (&ref.x)[q] = r;
            }
            s_ref[i] = ref;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t acc = 0;
            for (uint32_t d = 0; d < kLevels; ++d) {
                s_off[d] = acc;
                acc += s_cnt[d];
                s_cnt[d] = 0u;
            }
            s_off[kLevels] = acc;
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nspan; i += kThreads) {
            const uint32_t d = s_depth[i];
            s_list[s_off[d] + atomicAdd(&s_cnt[d], 1u)] = (uint16_t)i;
        }
        __syncthreads();
        for (int d = 15; d >= 0; --d) {
            const uint32_t lo = s_off[d], hi = s_off[d + 1];
            if (lo == hi) continue;  // (uniform) no spanning cell at this depth
            for (uint32_t k = lo + threadIdx.x; k < hi; k += kThreads) {
                const uint32_t i = s_list[k];
                const uint4 ref = s_ref[i];
                uint32_t c[4];
                double m[4], x[4], y[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    c[q] = (&ref.x)[q];
                    const uint32_t r = c[q] == kNoChild ? i : c[q];
                    m[q] = s_m[r];
                    x[q] = s_x[r];
                    y[q] = s_y[r];
                }
                const Agg a = sum_children(c, m, x, y);
                s_m[i] = a.m;
                s_x[i] = a.x;
                s_y[i] = a.y;
                store_span(S, s_idx[i], s_node[i], a, s_w[i]);
            }
            // the next depth reads only LDS: no wait for this depth's global stores
            lds_barrier();
        }
        return;
    }
    // listed mode: counting sort of the span entries by depth into spanbuf
    for (uint32_t i = threadIdx.x; i < nspan; i += kThreads)
        atomicAdd(&s_cnt[min(S.span[i].y >> 26, kLevels - 1u)], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (uint32_t d = 0; d < kLevels; ++d) {
            s_off[d] = acc;
            acc += s_cnt[d];
            s_cnt[d] = 0u;
        }
        s_off[kLevels] = acc;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nspan; i += kThreads) {
        const uint2 se = S.span[i];
        const uint32_t d = min(se.y >> 26, kLevels - 1u);
        S.spanbuf[s_off[d] + atomicAdd(&s_cnt[d], 1u)] = se;
    }
    __syncthreads();  // (also orders the list's global stores before the rounds' loads)
    for (int d = 15; d >= 0; --d) {
        if (s_off[d] == s_off[d + 1]) continue;  // (uniform) no spanning cell at this depth
        for (uint32_t k = s_off[d] + threadIdx.x; k < s_off[d + 1]; k += kThreads) {
            const uint2 se = S.spanbuf[k];
            const uint32_t node = se.y & kSpanNodeMask;
            store_span(S, se.x, node, agg_span_global(S, se.x), span_word(S, S.cells[se.x].end, node, se.y >> 26));
        }
        __syncthreads();  // the next depth reads this depth's global stores
    }
}

// Thresholds for a new theta (force called with a theta other than the build's).
__global__ void k_set_theta(const SimDesc* __restrict__ sims, uint32_t nsims) {
    const uint32_t s = blockIdx.x;
    if (s >= nsims) return;
    const SimDesc& S = sim_at(sims, s);
    float T = 0.0f, hf = 0.0f;
    if (threadIdx.x < kLevels) {
        double hd = S.meta->root[2];
        for (uint32_t l = 0; l < threadIdx.x; ++l) hd = hd * 0.5;
        hf = (float)hd;
        T = accept_threshold(hf, S.theta);
        S.meta->T[threadIdx.x] = T;
        if (threadIdx.x < kCodes) S.meta->tab[threadIdx.x] = threadIdx.x == 0 ? 0.0f : T;
    }
    const float T0 = __shfl(T, 0);
    const uint32_t p4 = pow4_thresholds(T, T0);  // (one wavefront: blockDim.x == 64)
    const uint32_t cr = contained_rejected(T, hf);
    if (threadIdx.x == 0) {
        S.meta->theta_bits = __float_as_uint(S.theta);
        S.meta->pow4 = p4;
        S.meta->t0_bits = __float_as_uint(T0);
        S.meta->contain_rejected = cr;
    }
}

// Thresholds of a device-built tree for a new theta: every cell, and the cold
// record of every internal cell's node (the node of the cell at depth d of the
// column at position p is nbase[p] + d - cdmin[p]).
__global__ void k_rewrite_T(const SimDesc* __restrict__ sims, uint32_t sim) {
    const SimDesc& S = sim_at(sims, sim);
    const uint32_t nc = min(S.meta->ncells, S.cell_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
        const CellT c = S.cells[i];
        const uint32_t d = c.info & 255u;
        const float T = S.meta->T[d < kLevels ? d : kLevels - 1u];
        S.cells[i].T = T;
        if (!(c.info & kInfoLeaf) && c.start < S.n) {
            const uint32_t ni = S.nbase[c.start] + (d - S.cdmin[c.start]);
            if (ni < S.node_cap) S.ncold[ni].T = T;
        }
    }
}

// Cold records of the body nodes of device-built trees, which the step does not
// write (only the evidence traversal reads them): body at sorted position p,
// first of its leaf iff a column of cells starts at p.
__global__ __launch_bounds__(kTile) void k_fill_body_cold(const SimDesc* __restrict__ sims,
                                                          const TileRef* __restrict__ tiles) {
    const TileRef tr = tiles[blockIdx.x];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t p = tr.base + threadIdx.x;
    if (p >= S.n) return;
    const uint32_t ni = S.bnode[p];
    if (ni < S.node_cap)
        S.ncold[ni] = NodeCold{p, 1u, 0.0f, kTagBody | (S.ccount[p] > 0u ? kTagFirst : 0u) | (S.val0[p] & kTagValue)};
}

// Cold records of the internal cells of a lean build, which its steps do not
// write (only the evidence traversal reads them): from each internal cell's
// record, at its node (the node of the cell at depth d of the column at p is
// nbase[p] + d - cdmin[p]).  blockIdx.y is the simulation.
__global__ __launch_bounds__(256) void k_fill_cell_cold(const SimDesc* __restrict__ sims) {
    const SimDesc& S = sim_at(sims, blockIdx.y);
    if (!S.lean) return;
    const uint32_t nc = min(S.meta->ncells, S.cell_cap);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
        const CellT c = S.cells[i];
        if ((c.info & kInfoLeaf) || c.start >= S.n) continue;
        const uint32_t d = c.info & 255u;
        const uint32_t ni = S.nbase[c.start] + (d - S.cdmin[c.start]);
        if (ni < S.node_cap) S.ncold[ni] = NodeCold{c.start, c.end - c.start, c.T, d};
    }
}

// ----------------------------------------------------------------------------
// K7: force traversal (bh_traverse, nbody.wgsl:43-123), stackless over the
// node stream with an independent cursor per lane.  Work is pulled from a
// global queue of (simulation, 64 consecutive sorted targets) items, larger
// simulations first.  Each lane walks its own target's sequence one node per
// iteration with the single interaction rule of the node stream, so every
// iteration is the same straight-line code for all lanes (no per-kind
// branches) and the interaction is evaluated for every lane that has one.  The
// sequence is the reference's per-target DFS: internal cells in preorder, a
// leaf as its bodies in sorted order, self skipped, so the interaction list and
// the f32 summation order are the reference's.
//
// Per iteration a lane gathers one 16-byte hot node; the opening threshold of
// its depth comes from the wavefront's copy of the simulation's table in LDS,
// and containment is a comparison of node indices (the target's own body node
// against the subtree [c, c + size)).  General-mode trees (host-built trees whose
// cell ranges or thresholds do not follow the node stream) read the cold record
// for the range and the threshold instead, exactly as bh_traverse tests them.
// ----------------------------------------------------------------------------
typedef float F4v __attribute__((ext_vector_type(4)));
typedef uint32_t U4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) F4v* GlobalF4;
typedef const __attribute__((address_space(1))) U4v* GlobalU4;

struct WalkEvidence {
    uint32_t visited = 0, accepted = 0, direct = 0;
    unsigned long long h = 0xcbf29ce484222325ull;
};

// kFast: every r2 = d2 + eps2 of the walk is finite and positive (the
// simulation's positions lie in a box of half-width < 1e18 and
// 0 < eps2 <= FLT_MAX), where point_mass_pos equals point_mass_d bit for bit.
// The walk is software-pipelined: the next node's gather is issued before the
// interaction of the current one is evaluated.  Whatever the node words hold,
// every step advances the cursor by at least one node, so a walk always ends.
// The hot records are read through a structured buffer resource (16-byte
// records, the record index in the instruction: no address arithmetic in the
// loop) whose size is the node stream's: a gather at or beyond the stream's end
// (the early gather past the last node) returns zeros instead of touching
// memory.  `tab` is the byte address of the wavefront's code table in LDS.
//
// A zero record is a body of mass 0 at the origin with subtree size 1.  In the
// fast walks (no evidence) a lane whose cursor has reached the end may take one
// more step on it, so the end test runs once per two steps: its interaction adds
// G 0 r^-3 d = +-0 to the lane's sums (r^-3 is finite there, as for every node:
// positions lie within 1e18 and 0 < eps2), which leaves them unchanged because a
// sum that starts at +0 is never -0, and the cursor moves on by one.
constexpr int kRawBufferWord3 = 0x00020000;  // gfx9 buffer word 3 (32-bit data format)
constexpr short kNodeStride = 16;
// The arguments must be wave-uniform (they are made explicitly so, which keeps
// the resource in scalar registers).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t node_buffer(const void* p, uint32_t records) {
    const uint64_t a = (uint64_t)(uintptr_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    const uint32_t nrec = __builtin_amdgcn_readfirstlane(records);
    void* base = (void*)(uintptr_t)(((uint64_t)hi << 32) | lo);
    return __builtin_amdgcn_make_buffer_rsrc(base, kNodeStride, (int)nrec, kRawBufferWord3);
}
// llvm.amdgcn.struct.ptr.buffer.load (buffer_load_dwordx4 ... idxen): record
// `vindex` of the resource, `voffset` bytes into it.
__device__ F4v struct_buffer_load_v4f32(__amdgpu_buffer_rsrc_t rsrc, int vindex, int voffset, int soffset,
                                        int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v4f32");
__device__ __forceinline__ F4v load_node(__amdgpu_buffer_rsrc_t r, uint32_t index) {
    return struct_buffer_load_v4f32(r, (int)index, 0, 0, 0);
}

// One target per lane.  The threshold comes from the code table in LDS (or,
// in general mode, from the node's cold record) and the size from the word
// layout's shift.
template <bool kEvidence, bool kGeneral, bool kFast>
struct NodeWalk {
    __amdgpu_buffer_rsrc_t nh;
    const __attribute__((address_space(1))) U4v* nc;
    uint32_t tab, nn, p, self;
    float tx, ty, eps2, G;
    uint32_t wshift = kSizeShift;  // the word layout's size shift
    float ax = 0.f, ay = 0.f;
    WalkEvidence ev;

    // Evaluates node c (record g), issues the gather of the next node into gn
    // and returns the next cursor.
    __device__ __forceinline__ uint32_t step(uint32_t c, const F4v& g, F4v& gn) {
        const uint32_t w = __float_as_uint(g.w);
        const float dx = g.x - tx;
        const float dy = g.y - ty;
        const float dd = dx * dx;
        const float ee = dy * dy;
        const float d2 = dd + ee;
        const uint32_t size_m1 = w >> wshift;  // subtree [c, c + size_m1]
        bool inside;
        bool open_ok;  // d2 >= the node's threshold
        U4v l;
        if (kGeneral || kEvidence) l = nc[c];  // start, count, T, tag
        if (kGeneral) {
            inside = (p - l.x) < l.y;
            open_ok = d2 >= __uint_as_float(l.z);
        } else {
            inside = (self - c) <= size_m1;
            open_ok = d2 >= *(const __attribute__((address_space(3))) float*)(uintptr_t)((w & kCodeField) | tab);
        }
        const bool acc = !inside && open_ok;
        const uint32_t next = c + 1u + (acc ? size_m1 : 0u);  // at least one node per step
        gn = load_node(nh, next);
        if (kEvidence) {
            if (!(l.w & kTagBody) || (l.w & kTagFirst)) ++ev.visited;  // a cell (a leaf: its first body)
        }
        if (acc) {
            if (kFast) point_mass_pos(dx, dy, d2, g.z, eps2, G, ax, ay);
            else point_mass_d(dx, dy, d2, g.z, eps2, G, ax, ay);
            if (kEvidence) {
                unsigned long long word;
                if (l.w & kTagBody) {
                    ++ev.direct;
                    word = (1ull << 60) | (unsigned long long)(l.w & kTagValue);
                } else {
                    ++ev.accepted;
                    word = (2ull << 60) | ((unsigned long long)(l.w & 255u) << 32) | l.x;
                }
                mix64(ev.h, word);
            }
        }
        return next;
    }

    // Two steps per trip with the records alternating between g0 and g1 (the
    // fast walks test the end once per trip: see the zero record above).
    __device__ __forceinline__ void run(uint32_t c) {
        F4v g0 = load_node(nh, c), g1;
        while (c < nn) {
            c = step(c, g0, g1);
            if (kGeneral || kEvidence || !kFast) {
                if (c >= nn) break;
            }
            c = step(c, g1, g0);
        }
    }
};

// K targets per lane, sorted neighbours (positions K l .. K l + K - 1
// of the item): their walks visit mostly the same nodes, so the lane walks the
// union of them with one gather per node.  Each target keeps its own cursor,
// the next node it has to look at; the lane's cursor is the smallest one, and a
// node is evaluated for the targets whose cursor it is, so every target sees its
// own nodes in its own (increasing) order and its sums equal its own walk's.
// It serves simulations whose thresholds scale by powers of four (SimMeta::
// pow4) and whose words have layout 1, so the opening test is one integer
// comparison against T[0] (pow4_thresholds): bits(d2) + (w << 22) >= bits(T[0]),
// where w << 22 is the code's exponent offset (node_word).  As in the fast
// NodeWalk, the end test runs once per two steps (zero record).
// kUnitG: G == 1, where fl(G m) = m for every m, so the product is left out.
// kContain = false: no containment test (SimMeta::contain_rejected and a walk
// that starts below the root, see self_terms_vanish): the opening test rejects
// every cell holding the target, and the target's own body node is accepted,
// which adds exactly +0 to its sums.
template <uint32_t K, bool kUnitG, bool kContain = true>
struct MultiWalk {
    __amdgpu_buffer_rsrc_t nh;
    uint32_t nn, t0_bits;
    float eps2, G;
    uint32_t self[K], cur[K];  // cur: each target's cursor (nn: none)
    float tx[K], ty[K];
    float ax[K] = {}, ay[K] = {};

    // The opening test is evaluated as bits(d2) >= t0 - (w << 22) in signed 32-bit
    // arithmetic (one subtraction per node, one comparison per target).  It is the
    // same integer inequality: w << 22 is code << 24 (at most 15 << 24) plus 2^31
    // for code 0, so without the flag t0 - (w << 22) lies in (-2^31, 2^31) and the
    // unsigned sum bits(d2) + (w << 22) <= 0x7f800000 + (15 << 24) does not wrap;
    // with it the difference wraps to t0 - 2^31 < 0 and both forms hold for every
    // d2 >= +0 (bits(d2) < 2^31).  The cursor update uses that every cursor is at
    // least c (c is their minimum): a target not at c keeps its cursor, which
    // already exceeds c + 1.
    __device__ __forceinline__ uint32_t step(uint32_t c, const F4v& g, F4v& gn) {
        const uint32_t w = __float_as_uint(g.w);
        const uint32_t size_m1 = w >> kSizeShiftNarrow;
        const uint32_t c1 = c + 1u;
        uint32_t skip = c1 + size_m1;
        asm volatile("" : "+v"(skip));  // one sum for all targets: a select per target, not an add
        const int tnode = (int)(t0_bits - (w << 22));
        float dx[K], dy[K], d2[K];
        bool acc[K];
        uint32_t next = 0xffffffffu;
#pragma unroll
        for (uint32_t k = 0; k < K; ++k) {
            dx[k] = g.x - tx[k];
            dy[k] = g.y - ty[k];
            const float dd = dx[k] * dx[k];
            const float ee = dy[k] * dy[k];
            d2[k] = dd + ee;
            // (bitwise, not short-circuit: straight-line selects, no branches)
            const bool at = K == 1 || cur[k] == c;
            acc[k] = at & ((int)__float_as_uint(d2[k]) >= tnode);
            if (kContain) acc[k] = acc[k] & ((self[k] - c) > size_m1);
            cur[k] = max(cur[k], acc[k] ? skip : c1);
            next = min(next, cur[k]);  // greater than c
        }
        gn = load_node(nh, next);
#pragma unroll
        for (uint32_t k = 0; k < K; ++k)
            if (acc[k]) point_mass_pos<kUnitG>(dx[k], dy[k], d2[k], g.z, eps2, G, ax[k], ay[k]);
        return next;
    }

    __device__ __forceinline__ void run() {
        uint32_t c = cur[0];
#pragma unroll
        for (uint32_t k = 1; k < K; ++k) c = min(c, cur[k]);
        F4v g0 = load_node(nh, c), g1;
        while (c < nn) {
            c = step(c, g0, g1);
            c = step(c, g1, g0);
        }
    }
};

// Whether every target of the item (targets base .. base + 64 K - 1) would add
// exactly +0 to its own sums by interacting with its own body node: there dx =
// dy = +0 (the node holds the target's own coordinates), r2 = eps2, and the
// term is +0 * scale = +0 when scale = G m / r^3, rounded as point_mass_pos
// rounds it, is finite (a sum that starts at +0 is never -0, so adding +0
// leaves it unchanged).  Called by whole wavefronts.
template <uint32_t K, bool kUnitG>
__device__ __forceinline__ bool self_terms_vanish(const SimDesc& S, uint32_t base, uint32_t n, float eps2,
                                                  float G) {
    const float inv_r = rcp_rn_normal(sqrt_rn_pos(eps2));
    bool ok = true;
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t q = base + K * lane_id() + k;
        if (q < n) {
            const float m = S.sbody[q].z;
            float scale = kUnitG ? m : G * m;
            scale = scale * inv_r;
            scale = scale * inv_r;
            scale = scale * inv_r;
            ok = ok && __builtin_fabsf(scale) < __builtin_inff();
        }
    }
    return __ballot(!ok) == 0ull;
}

// The multi-target walk of one work item (targets base .. base + 64 K - 1),
// starting at node `start`: 0, or 1 when the walk has no containment test and
// node 0 is the root (the root holds every target and is always opened).
template <uint32_t K, bool kUnitG, bool kContain = true>
__device__ __forceinline__ void multi_walk(const SimDesc& S, uint32_t base, uint32_t n, __amdgpu_buffer_rsrc_t nh,
                                           uint32_t nn, uint32_t t0_bits, float eps2, float G,
                                           uint32_t start = 0u) {
    MultiWalk<K, kUnitG, kContain> wk;
    wk.nh = nh;
    wk.nn = nn;
    wk.t0_bits = t0_bits;
    wk.eps2 = eps2;
    wk.G = G;
    const uint32_t p0 = base + K * lane_id();
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const uint32_t q = min(p0 + k, n - 1u);
        const float4 b = S.sbody[q];
        wk.tx[k] = b.x;
        wk.ty[k] = b.y;
        if (kContain) wk.self[k] = S.bnode[q];
        wk.cur[k] = p0 + k < n ? start : nn;
    }
    wk.run();
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (uint32_t k = 0; k < K; ++k)
        if (p0 + k < n) S.acc[S.val0[p0 + k]] = make_float2(wk.ax[k], wk.ay[k]);
}

// Traversal work items by XCD (mode 0).  For a set of many simulations the host
// lays the items out as one contiguous, balanced range per XCD (simulations
// dealt to the XCDs in size order), so an XCD's L2 holds the trees of few
// simulations.  A wavefront of workgroup b runs on XCD b % 8 (the hardware deals
// workgroups round-robin; were it otherwise, only locality would suffer): its
// first item is its index among that XCD's wavefronts, later ones come from that
// XCD's queue counter, and when the range is exhausted it takes items from
// another XCD's queue.  A set whose largest simulation would have to be split
// across XCDs uses mode 1 instead: one range, first items by grid index, one
// counter (measured faster there: the whole grid walks one window of targets).
__device__ __forceinline__ uint32_t xcd_waves(uint32_t x, uint32_t wpb) {
    return ((gridDim.x + kXcds - 1u - x) / kXcds) * wpb;  // wavefronts on XCD x
}

// Next item from queue q (or, once q is exhausted, from any queue that still
// holds items); nitems when none does.  Every failed claim leaves its queue
// exhausted for good (counters only grow), so the loop ends within kXcds + 1
// claims.  Mode 1: the grid's one counter.  Called by whole wavefronts.
__device__ uint32_t claim_item(uint32_t* ctl, uint32_t mode, uint32_t& q, uint32_t wpb, uint32_t nitems) {
    const uint32_t lane = lane_id();
    if (mode != 0u) {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(&ctl[kCtlQueue0], 1u);
        k = __builtin_amdgcn_readlane(k, 0);
        const unsigned long long item = (unsigned long long)gridDim.x * wpb + k;
        return item < nitems ? (uint32_t)item : nitems;
    }
    for (;;) {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(&ctl[kCtlQueue0 + q * kQueueStride], 1u);
        k = __builtin_amdgcn_readlane(k, 0);
        const unsigned long long j = (unsigned long long)xcd_waves(q, wpb) + k;
        if (j < ctl[kCtlRanges + kXcds + q]) return ctl[kCtlRanges + q] + (uint32_t)j;
        bool open = false;
        if (lane < kXcds) {
            const uint32_t c = flag_load(&ctl[kCtlQueue0 + lane * kQueueStride]);
            open = (unsigned long long)xcd_waves(lane, wpb) + c < ctl[kCtlRanges + kXcds + lane];
        }
        const unsigned long long m = __ballot(open);
        if (m == 0ull) return nitems;
        q = (uint32_t)__builtin_ctzll(m);
    }
}

// Work items of 64 lane_targets targets (lane_targets 1 or 2, the set's
// choice): the multi-target walk where it applies, else lane_targets parts of
// 64 targets one after the other, one target per lane.
template <bool kEvidence>
__global__ __launch_bounds__(256) void k_traverse_nodes(const SimDesc* __restrict__ sims,
                                                        const TileRef* __restrict__ wtiles,
                                                        uint32_t nwtiles, uint32_t* __restrict__ ctl,
                                                        uint32_t* __restrict__ ev_stats,
                                                        unsigned long long* __restrict__ ev_hash,
                                                        uint32_t ev_sim, uint32_t lane_targets) {
    __shared__ __attribute__((aligned(64))) float s_tab[256 / 64][kCodes];
    float* tabp = s_tab[threadIdx.x >> 6];
    const uint32_t tab = __builtin_amdgcn_readfirstlane(
        (uint32_t)(uintptr_t)(__attribute__((address_space(3))) float*)tabp);
    const uint32_t lane = lane_id();
    // first item: the wavefront's index among its XCD's wavefronts (so the
    // launch does not start with every wavefront on an atomic)
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t mode = __builtin_amdgcn_readfirstlane(ctl[kCtlMode]);
    uint32_t queue = blockIdx.x % kXcds;
    uint32_t next_item;
    if (mode != 0u) {
        next_item = blockIdx.x * wpb + (threadIdx.x >> 6);
    } else {
        const uint32_t j0 = (blockIdx.x / kXcds) * wpb + (threadIdx.x >> 6);
        next_item = j0 < ctl[kCtlRanges + kXcds + queue] ? ctl[kCtlRanges + queue] + j0
                                                          : claim_item(ctl, mode, queue, wpb, nwtiles);
    }
    for (;;) {
        const uint32_t item = next_item;
        if (item >= nwtiles) break;
        const TileRef tr = wtiles[item];
        const SimDesc& S = sim_at(sims, tr.sim);
        const uint32_t n = S.n;
        // a failed build (its error bits are reported by the host) cannot send
        // the walk past the node arrays
        const uint32_t nn = __builtin_amdgcn_readfirstlane(min(S.meta->nnodes, S.node_cap - 1u));
        const uint32_t general = __builtin_amdgcn_readfirstlane(S.meta->mode);
        const uint32_t geom_ok = __builtin_amdgcn_readfirstlane(S.meta->geom_ok);
        const uint32_t pow4 = __builtin_amdgcn_readfirstlane(S.meta->pow4);
        const uint32_t t0_bits = __builtin_amdgcn_readfirstlane(S.meta->t0_bits);
        const uint32_t narrow = __builtin_amdgcn_readfirstlane(S.wlayout);
        const uint32_t wshift = narrow != 0u ? kSizeShiftNarrow : kSizeShift;
        const float eps = S.eps;
        const float eps2 = eps * eps;
        const float G = S.G;
        const bool fast = geom_ok != 0u && eps2 > 0.0f && eps2 <= 3.4028234663852886e38f;
        const __amdgpu_buffer_rsrc_t nh = node_buffer(S.nodes, nn);
        if (!kEvidence && !general && fast && pow4 && narrow != 0u) {
            // no containment test where the thresholds reject every cell that
            // holds its target and the item's self terms vanish; the walk then
            // starts below the root (node 0 when the root is a leaf: then every
            // node is a body)
            const uint32_t cr = __builtin_amdgcn_readfirstlane(S.meta->contain_rejected);
            const uint32_t start = __builtin_amdgcn_readfirstlane(S.meta->nnodes) > n ? 1u : 0u;
            if (G == 1.0f) {
                if (lane_targets == 2u) {
                    if (cr && self_terms_vanish<2, true>(S, tr.base, n, eps2, G))
                        multi_walk<2, true, false>(S, tr.base, n, nh, nn, t0_bits, eps2, G, start);
                    else
                        multi_walk<2, true>(S, tr.base, n, nh, nn, t0_bits, eps2, G);
                } else {
                    if (cr && self_terms_vanish<1, true>(S, tr.base, n, eps2, G))
                        multi_walk<1, true, false>(S, tr.base, n, nh, nn, t0_bits, eps2, G, start);
                    else
                        multi_walk<1, true>(S, tr.base, n, nh, nn, t0_bits, eps2, G);
                }
            } else {
                if (lane_targets == 2u) {
                    if (cr && self_terms_vanish<2, false>(S, tr.base, n, eps2, G))
                        multi_walk<2, false, false>(S, tr.base, n, nh, nn, t0_bits, eps2, G, start);
                    else
                        multi_walk<2, false>(S, tr.base, n, nh, nn, t0_bits, eps2, G);
                } else {
                    if (cr && self_terms_vanish<1, false>(S, tr.base, n, eps2, G))
                        multi_walk<1, false, false>(S, tr.base, n, nh, nn, t0_bits, eps2, G, start);
                    else
                        multi_walk<1, false>(S, tr.base, n, nh, nn, t0_bits, eps2, G);
                }
            }
        } else {
            // this wavefront's copy of the code table (LDS is in order per wavefront)
            if (lane < kCodes) tabp[lane] = S.meta->tab[lane];
            __builtin_amdgcn_wave_barrier();
            const GlobalU4 nc = (GlobalU4)(const void*)S.ncold;
            // the item's 64-target parts, one target per lane
            for (uint32_t half = 0; half < lane_targets; ++half) {
                const uint32_t p = tr.base + 64u * half + lane;
                const bool valid = p < n;
                float tx = 0.f, ty = 0.f;
                uint32_t self = 0;
                if (valid) {
                    const float4 b = S.sbody[p];
                    tx = b.x;
                    ty = b.y;
                    if (!general) self = S.bnode[p];
                }
                float ax = 0.f, ay = 0.f;
                WalkEvidence ev;
                const uint32_t c0 = valid ? 0u : nn;
                if (general) {
                    NodeWalk<kEvidence, true, false> wk{nh, nc, tab, nn, p, self, tx, ty, eps2, G, wshift};
                    wk.run(c0);
                    ax = wk.ax, ay = wk.ay, ev = wk.ev;
                } else if (fast) {
                    NodeWalk<kEvidence, false, true> wk{nh, nc, tab, nn, p, self, tx, ty, eps2, G, wshift};
                    wk.run(c0);
                    ax = wk.ax, ay = wk.ay, ev = wk.ev;
                } else {
                    NodeWalk<kEvidence, false, false> wk{nh, nc, tab, nn, p, self, tx, ty, eps2, G, wshift};
                    wk.run(c0);
                    ax = wk.ax, ay = wk.ay, ev = wk.ev;
                }
                __builtin_amdgcn_wave_barrier();
                if (valid) {
                    const uint32_t body = S.val0[p];
                    S.acc[body] = make_float2(ax, ay);
                    if (kEvidence && tr.sim == ev_sim) {
                        if (ev_stats) {
                            ev_stats[3 * (size_t)body + 0] = ev.visited;
                            ev_stats[3 * (size_t)body + 1] = ev.accepted;
                            ev_stats[3 * (size_t)body + 2] = ev.direct;
                        }
                        if (ev_hash) ev_hash[body] = ev.h;
                    }
                }
            }
        }
        // the next item is claimed when this one is done (a claim made earlier
        // holds the item back from wavefronts that finish first)
        next_item = claim_item(ctl, mode, queue, wpb, nwtiles);
    }
}

// ----------------------------------------------------------------------------
// K8: final kick (bh_final_kick, nbody_evolve.wgsl:35-46)
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(kTile) void k_final_kick(const SimDesc* __restrict__ sims,
                                                      const TileRef* __restrict__ tiles) {
    const TileRef tr = tiles[blockIdx.x];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t i = tr.base + threadIdx.x;
    if (i >= S.n) return;
    const float h = 0.5f * S.dt;
    float2 v = S.vel[i];
    const float2 a = S.acc[i];
    v.x = v.x + h * a.x;
    v.y = v.y + h * a.y;
    S.vel[i] = v;
}

// bh_kick_drift alone (nbody_evolve.wgsl:19-33), for the drop-in kick_drift call
__global__ __launch_bounds__(kTile) void k_kick_drift(const SimDesc* __restrict__ sims,
                                                      const TileRef* __restrict__ tiles) {
    const TileRef tr = tiles[blockIdx.x];
    const SimDesc& S = sim_at(sims, tr.sim);
    const uint32_t i = tr.base + threadIdx.x;
    if (i >= S.n) return;
    const float dt = S.dt;
    const float h = 0.5f * dt;
    float2 v = S.vel[i];
    float2 p = S.pos[i];
    const float2 a = S.acc[i];
    const float hvx = v.x + h * a.x;
    const float hvy = v.y + h * a.y;
    p.x = p.x + dt * hvx;
    p.y = p.y + dt * hvy;
    S.pos[i] = p;
    S.vel[i] = make_float2(hvx, hvy);
}

// FP64 direct-summation diagnostics over the resident state (measurement only,
// not on the step path).  k_diag_direct: the acceleration of each listed body,
// summed over all other bodies in index order with the operation order of
// add_point_mass / direct_acceleration_for (nbody/src/lib.rs:247-255, 315-329).
// k_diag_potential: per body i, -G m_i sum_{j > i} m_j / sqrt(r2 + eps2), the
// j-sum in index order.  f64 sqrt and division are correctly rounded.
__global__ __launch_bounds__(256) void k_diag_direct(const float2* __restrict__ pos,
                                                     const float* __restrict__ mass, uint32_t n,
                                                     const uint32_t* __restrict__ index,
                                                     uint32_t count, double eps2, double G,
                                                     double* __restrict__ ax,
                                                     double* __restrict__ ay) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t i = index[k];
    const double tx = (double)pos[i].x, ty = (double)pos[i].y;
    double sx = 0.0, sy = 0.0;
    for (uint32_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const float2 q = pos[j];
        const double dx = (double)q.x - tx;
        const double dy = (double)q.y - ty;
        const double r2 = dx * dx + dy * dy + eps2;
        const double inv_r = 1.0 / __builtin_sqrt(r2);
        const double scale = G * (double)mass[j] * inv_r * inv_r * inv_r;
        sx += dx * scale;
        sy += dy * scale;
    }
    ax[k] = sx;
    ay[k] = sy;
}

__global__ __launch_bounds__(256) void k_diag_potential(const float2* __restrict__ pos,
                                                        const float* __restrict__ mass, uint32_t n,
                                                        double eps2, double G,
                                                        double* __restrict__ pot) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = (double)pos[i].x, yi = (double)pos[i].y;
    double s = 0.0;
    for (uint32_t j = i + 1; j < n; ++j) {
        const float2 q = pos[j];
        const double dx = (double)q.x - xi;
        const double dy = (double)q.y - yi;
        s += (double)mass[j] / __builtin_sqrt(dx * dx + dy * dy + eps2);
    }
    pot[i] = -G * (double)mass[i] * s;
}

// Streaming copy used to calibrate the HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy_calibration(const float4* __restrict__ src,
                                                          float4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ABI <-> SoA conversion of the frozen records
__global__ void k_unpack_state(const GbhEvolveState* __restrict__ in, uint32_t n, float2* pos,
                               float* mass, float2* vel, float2* acc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GbhEvolveState s = in[i];
    pos[i] = make_float2(s.position_mass[0], s.position_mass[1]);
    mass[i] = s.position_mass[2];
    vel[i] = make_float2(s.velocity[0], s.velocity[1]);
    acc[i] = make_float2(0.f, 0.f);
}

__global__ void k_pack_state(GbhEvolveState* __restrict__ out, uint32_t n, const float2* pos,
                             const float* mass, const float2* vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GbhEvolveState s;
    s.position_mass[0] = pos[i].x;
    s.position_mass[1] = pos[i].y;
    s.position_mass[2] = mass[i];
    s.position_mass[3] = 0.f;
    s.velocity[0] = vel[i].x;
    s.velocity[1] = vel[i].y;
    s.velocity[2] = 0.f;
    s.velocity[3] = 0.f;
    out[i] = s;
}

__global__ void k_pack_accel(GbhGpuAccel* __restrict__ out, uint32_t n, const float2* acc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    GbhGpuAccel a;
    a.value[0] = acc[i].x;
    a.value[1] = acc[i].y;
    a.value[2] = 0.f;
    a.value[3] = 0.f;
    out[i] = a;
}

}  // namespace gbh
