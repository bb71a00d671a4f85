// Device-side layout and wave64 helpers for the matching path on gfx950.
//
// Layout in HBM (one labelled data graph shared by every standing query):
//   adj[v]   = {offset, degree}: vertex v's neighbour keys live at arena[offset .. offset+degree),
//              sorted ascending; capacity cap[v] >= degree (slack for merge-in-place growth).
//   key      = (label_slot << id_bits) | neighbour id, label_slot = edge label rank * #vertex labels
//              + neighbour label rank, so each (edge label, neighbour label) run is contiguous and
//              sorted by id: the neighbours of v along one query relation are one sorted run.
//   delta    = the current batch, per direction: seg[v] = {ins_off, ins_len, del_off, del_len}
//              into ins_keys / del_keys (same key format, sorted per vertex); zero when untouched.
// Directed graphs keep a second adjacency (in-arcs) and a second delta; undirected graphs alias
// both directions to the same arrays.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "../common/csm_common.h"

namespace csm {

enum ErrBit : uint32_t {
    kErrArenaFull = 1u,      // adjacency arena exhausted
    kErrRingFull = 2u,       // partial-result ring too small (handled by falling back to DFS)
    kErrPrecondition = 4u,   // inserted edge already present / deleted edge absent
    kErrSegmentTooLong = 8u, // one vertex received more batch arcs than the segment sorter holds
    kErrProductOverflow = 16u,
    kErrScratchFull = 32u,
    kErrLabelOverflow = 64u,
};

struct DevAdj {
    uint2* adj;      // [dv] {offset into arena, degree}
    uint32_t* cap;   // [dv]
};

struct DevDelta {
    uint4* seg;       // [dv] {ins_off, ins_len, del_off, del_len}
    uint32_t* ins;    // insertion keys
    uint32_t* del;    // deletion keys
};

struct DevGraph {
    uint32_t dv;
    uint32_t id_bits;
    uint32_t id_mask;
    uint32_t nvl;         // vertex-label ranks
    const uint32_t* vrank;  // [dv] label rank of each vertex (kNoVertex if undeclared)
    uint32_t* arena;
    unsigned long long* arena_top;  // bump pointer (words) for relocated adjacency lists
    unsigned long long arena_words;
    DevAdj dir[2];        // 0 = out (or undirected), 1 = in
    DevDelta delta[2];
    uint32_t* err;        // error bits
    // Label-slot directory (optional, nullptr when disabled): row v of sdir[d] holds nslots + 1
    // absolute arena offsets, row[s] = first position of v's direction-d list whose key has slot
    // >= s, so the run of slot s is [row[s], row[s + 1]).
    uint32_t nslots;
    uint32_t* sdir[2];
    // Free arena blocks by size class (a block of class c holds 2^c words): the blocks that
    // relocated lists leave behind.  A relocation takes a block of its class from `blk` before it
    // bumps arena_top; blocks freed during a batch wait in `pend` and join `blk` only after the
    // batch's merge (a block being copied out of must not be handed out in the same launch).
    uint32_t* fl_blk;        // class c entries at fl_blk[fl_cls[c].x ..], at most fl_cls[c].y of them
    uint32_t* fl_pend;       // same layout
    int* fl_top;             // [kFreeClasses] available entries (may dip below 0 during a batch)
    uint32_t* fl_pend_n;     // [kFreeClasses] blocks freed in the current batch
    const uint2* fl_cls;     // [kFreeClasses] {first entry, capacity}
};
constexpr uint32_t kFreeClasses = 32;

// One backward neighbour of the vertex matched at some depth: the partner is at position `pos`
// of the matching order, its list along the query relation is the run `slot` of its adjacency in
// direction `dir`, seen through `view` (relation views of the current batch step).
struct BackEntry {
    uint32_t slot;
    uint8_t pos;
    uint8_t view;
    uint8_t dir;
    uint8_t pad;
};

// Everything one (query, batch step, phase) needs on the device; passed by value.
struct StepDesc {
    uint32_t n;                       // query vertices
    uint32_t phase;                   // 0 insertion (delta = ins), 1 deletion (delta = del)
    uint32_t order[kMaxQV];           // query vertex matched at each depth
    uint32_t vrank[kMaxQV];           // data label rank required at each depth
    uint8_t boff[kMaxQV + 1];         // backward entries of depth d: be[boff[d] .. boff[d+1])
    uint8_t pivot[kMaxQV];            // entry whose list generates the candidates of depth d
    uint8_t pad0[3];
    // Matched positions that a candidate of depth d must be compared with for injectivity: the
    // ones of the same label rank (candidates carry their query vertex's label) that are not
    // backward neighbours of d (a neighbour of f[q] is never f[q]: data graphs have no self-loops).
    uint16_t inj[kMaxQV];
    BackEntry be[kMaxQE];
    const uint32_t* filter[kMaxQV];   // candidate bitmap of the vertex at depth d (local or global)
};

// ---- wave64 helpers --------------------------------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// Flattened item processing (the local-index kernels, the initial partial results): G items per
// wave, the smallest power of two that covers `items` with the grid's waves.
__device__ __forceinline__ uint32_t items_per_wave(uint32_t items) {
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    uint32_t G = 1;
    while (G < 64 && static_cast<uint64_t>(G) * waves < items) G <<= 1;
    return G;
}

// Wave-wide exclusive prefix of x; returns it and sets total.
__device__ __forceinline__ uint32_t wave_prefix(uint32_t x, uint32_t& total) {
    const uint32_t lane = lane_id();
    uint32_t inc = x;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o, 64));
        if (lane >= o) inc += y;
    }
    total = static_cast<uint32_t>(__shfl(static_cast<int>(inc), 63, 64));
    return inc - x;
}

// The item of the wave whose run holds flattened element e (largest k with pre[k] <= e).
__device__ __forceinline__ uint32_t owner_of(const uint32_t* pre, uint32_t e) {
    uint32_t lo = 0, hi = 63;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= e) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// First index in a[lo, hi) whose value is >= key (hi if none): search.cuh:7-25 semantics.
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t key) {
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ bool contains_u32(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t key) {
    const uint32_t p = lower_bound_u32(a, lo, hi, key);
    return p < hi && a[p] == key;
}

// The run of `slot` inside the sorted list a[lo, hi).
__device__ __forceinline__ void slot_range(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t slot, uint32_t id_bits,
                                           uint32_t& rlo, uint32_t& rhi) {
    const uint64_t k0 = static_cast<uint64_t>(slot) << id_bits;
    const uint64_t k1 = static_cast<uint64_t>(slot + 1) << id_bits;
    rlo = lower_bound_u32(a, lo, hi, static_cast<uint32_t>(k0));
    rhi = k1 > 0xFFFFFFFFull ? hi : lower_bound_u32(a, rlo, hi, static_cast<uint32_t>(k1));
}

__device__ __forceinline__ bool bit_test(const uint32_t* bm, uint32_t v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

}  // namespace csm
