// Shared host/device definitions for the continuous-subgraph-matching path.
//
// Limits follow the reference (utils/config.h:9-11): at most 16 query vertices and 42 query edges,
// minimum adjacency capacity 8.  Everything here is integer arithmetic.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define CSM_HD __host__ __device__ __forceinline__
#else
#define CSM_HD inline
#endif

namespace csm {

constexpr uint32_t kMaxQV = 16;          // config.h:9  MAX_QV_COUNT
constexpr uint32_t kMaxQE = 42;          // config.h:10 MAX_QE_COUNT
constexpr uint32_t kMaxRel = 2 * kMaxQE; // directed query edges (relations)
constexpr uint32_t kMinNbrSize = 8;      // config.h:11 MIN_NBR_SIZE
constexpr uint32_t kNoVertex = 0xFFFFFFFFu;
// config.h:47 MIN_NUM_RESULTS_TO_GPU, kept for the flattening rule it also sets (match_gpu.cu:2081);
// where the levels stop is the session's semi_bfs_threshold
constexpr uint32_t kSemiBfsThreshold = 1u << 16;

// Relation views of one directed query edge while step i of a batch is processed
// (technique_permutation.h:133-211 interleaving; deletions mirror it):
//   insertion phase:  edge j < i -> base U delta,  j == i -> delta (anchors),  j > i -> base
//   deletion phase:   edge j < i -> base \ delta,  j == i -> delta (anchors),  j > i -> base
enum View : uint8_t { kViewBase = 0, kViewBasePlusDelta = 1, kViewDelta = 2, kViewBaseMinusDelta = 3 };

// ---- embedding hash -------------------------------------------------------------------------
// fmix64 is the MurmurHash3 64-bit finaliser (public domain).  An embedding f is hashed as
//   H(f) = fmix64( sum_q fmix64((f(q) << 4 | q) + kHashSeed) )      (sums mod 2^64)
// i.e. a function of the tuple (f(0),...,f(n-1)) in query-vertex order, and a batch's embedding
// set hashes to sum_f H(f) mod 2^64, which does not depend on enumeration order.
constexpr uint64_t kHashSeed = 0x9E3779B97F4A7C15ull;

CSM_HD uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}
CSM_HD uint64_t pair_hash(uint32_t q, uint32_t v) {
    return fmix64(((static_cast<uint64_t>(v) << 4) | q) + kHashSeed);
}
CSM_HD uint64_t embedding_hash(uint64_t pair_sum) { return fmix64(pair_sum); }

// ---- packed adjacency keys ---------------------------------------------------------------------
// Every adjacency entry of the device graph is one u32 key = (label_slot << id_bits) | vertex id,
// where label_slot = edge_label_rank * num_vlabels + vertex_label_rank of the neighbour.  A vertex's
// list sorted ascending is therefore grouped by (edge label, neighbour label) and sorted by vertex id
// inside each group, so the neighbours of v along one query relation are one contiguous sorted run.
struct KeyLayout {
    uint32_t id_bits;      // bits for the vertex id
    uint32_t num_vlabels;  // vertex-label ranks in the data graph
    uint32_t num_elabels;  // edge-label ranks in the data graph
};
CSM_HD uint32_t key_of(const KeyLayout& kl, uint32_t slot, uint32_t v) { return (slot << kl.id_bits) | v; }
CSM_HD uint32_t id_of(uint32_t key, uint32_t id_mask) { return key & id_mask; }

// Reference float->uint32 conversion as compiled by g++ on x86-64 (cvttss2si to 64 bits, low half):
// NaN and |f| >= 2^63 give 0.  Used only to reproduce plan.cpp:543-566 bit for bit.
inline uint32_t f2u_x86(float f) {
    if (!(f == f)) return 0u;
    if (f >= 9.2233720368547758e18f || f < -9.2233720368547758e18f) return 0u;
    return static_cast<uint32_t>(static_cast<int64_t>(f));
}

}  // namespace csm
