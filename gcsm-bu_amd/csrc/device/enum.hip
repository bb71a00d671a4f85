// Edge-anchored incremental enumeration on wave64.
//
//   init_from_pairs / init_from_set   initial partial results of the first matching-order edge
//                                      (write_initial_partial_results, enumeration.cu:24-79)
//   extend                             one wave per partial result, depth-first down to end_depth
//                                      with a per-wave LDS stack of 64-candidate chunks; counts and
//                                      hashes complete embeddings or writes partial results of
//                                      end_depth (a BFS level when end_depth = d0 + 1).  Replaces
//                                      extendBFSDFSRegTwo (enumeration.cu:407-781) and gamma_enumerate
//                                      (gamma_enumeration.cu:359-880).
//   level_prepare / level_flat         one breadth-first level flattened over its candidates: per
//   (and the _chain variants)          record, the runs of every backward entry and the shortest as
//                                      the generator, then four candidates per lane tested in
//                                      lock-step; the _chain variants read their size from the
//                                      levels before them, so a step's levels run without the host.
//   cp_prepare / cp_enumerate          Cartesian-product flattening of trailing tree leaves
//                                      (GetNumTree + enumerateCartesianProductTree,
//                                      cartesian_product.cu:193-284).
// In the depth-first kernel, candidates of depth d come from the pivot backward neighbour's view
// run; each lane tests one candidate: candidate bitmap, injectivity against the matched prefix, and
// membership in every other backward neighbour's view run by binary search (search.cuh:7-25
// semantics).
#include "kernels.h"
#include "view.h"

namespace csm {
namespace {

constexpr uint32_t kWaves = 4;  // waves per block

// Per-wave DFS state in LDS.  Behind it (dynamic LDS, per wave): `rows` candidate buffers of 64
// words, one per depth d0 .. end_depth-2.
struct WaveHdr {
    unsigned long long ph[kMaxQV + 1];  // ph[d] = sum_{j<d} pair_hash(order[j], f[j])
    RunView run[kMaxQE];
    uint32_t f[kMaxQV];
    uint32_t cur[kMaxQV], tot[kMaxQV], nb[kMaxQV], bcnt[kMaxQV], bpos[kMaxQV];
};

__host__ __device__ constexpr uint32_t wave_lds_bytes(uint32_t rows) {
    return static_cast<uint32_t>((sizeof(WaveHdr) + rows * 256u + 15u) & ~15u);
}

__device__ __forceinline__ const uint32_t* entry_delta(const DevGraph& g, const StepDesc& d, const BackEntry& e) {
    return phase_delta(g, e.dir, d.phase);
}

// Computes the view runs of every backward entry of `depth` and resets its frame.
// Algorithmic bytes of a frame: every view run of the depth's backward lists read once.
__device__ __forceinline__ void setup_frame(const DevGraph& g, const StepDesc& d, WaveHdr& W, uint32_t depth, uint32_t lane,
                                            unsigned long long& bytes, bool count_bytes) {
    const uint32_t e0 = d.boff[depth], e1 = d.boff[depth + 1];
    if (lane < e1 - e0) {
        const BackEntry e = d.be[e0 + lane];
        W.run[e0 + lane] = make_run(g, d.phase, e.view, e.dir, W.f[e.pos], e.slot);
    }
    wave_sync();
    const uint32_t p = d.pivot[depth];
    if (lane == 0) {
        if (count_bytes)
            for (uint32_t e = e0; e < e1; e++) bytes += 4ull * ((W.run[e].hi - W.run[e].lo) + (W.run[e].dhi - W.run[e].dlo));
        const RunView r = W.run[p];
        W.nb[depth] = run_base_len(r, d.be[p].view);
        W.tot[depth] = run_seq_len(r, d.be[p].view);
        W.cur[depth] = 0;
        W.bcnt[depth] = 0;
        W.bpos[depth] = 0;
    }
    wave_sync();
}

__device__ __forceinline__ bool test_candidate(const DevGraph& g, const StepDesc& d, const WaveHdr& W, uint32_t depth, uint32_t key) {
    const uint32_t c = key & g.id_mask;
    if (!bit_test(d.filter[depth], c)) return false;
    for (uint32_t im = d.inj[depth]; im; im &= im - 1)
        if (W.f[__ffs(im) - 1] == c) return false;
    const uint32_t p = d.pivot[depth];
    const BackEntry pe = d.be[p];
    if (!run_elem_live(entry_delta(g, d, pe), W.run[p], pe.view, key)) return false;
    for (uint32_t e = d.boff[depth]; e < d.boff[depth + 1]; e++) {
        if (e == p) continue;
        const BackEntry be = d.be[e];
        if (!view_contains(g, d.phase, be.view, be.dir, be.slot, W.f[be.pos], (be.slot << g.id_bits) | c)) return false;
    }
    return true;
}

// Last record r with off[r] <= item, given off[0] <= item < off[n]: a 64-ary search in which every
// lane probes one of 64 evenly spaced offsets per round.
__device__ __forceinline__ uint32_t find_record(const unsigned long long* off, uint32_t n, unsigned long long item, uint32_t lane) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63) / 64;
        const uint32_t p = lo + lane * step;
        const uint64_t m = __ballot(p < hi && off[p] <= item);
        lo += (63u - static_cast<uint32_t>(__clzll(m))) * step;
        hi = min(hi, lo + step);
    }
    return uniform(lo);
}

// ---- Range donation (work stealing) of the depth-first kernel --------------------------------
// Each wave owns one word per launch: the candidate range [cur, end) of its current record's first
// depth still to be explored, packed as version (16 bits) | cur (24) | end (24), with the record's
// index in a second array.  The owner takes its next 64-candidate round by moving cur up; a wave
// out of work takes the upper half of another wave's range by moving that wave's end down to the
// midpoint.  Both are compare-and-swaps on the one word, so the rounds taken and the halves
// donated never overlap.  The version changes whenever the owner publishes a new record, so a
// thief's swap fails if the record index it read belongs to another record.  Ranges only shrink,
// so every wave's search ends.  (The reference's disabled design: gamma_enumeration.cu:209-357.)
constexpr uint32_t kRangeBits = 24;
constexpr uint32_t kRangeMask = (1u << kRangeBits) - 1;
constexpr uint32_t kStealRounds = 8;  // victim probes of 64 waves each before a wave gives up

__device__ __forceinline__ uint32_t range_cur(unsigned long long w) { return static_cast<uint32_t>(w >> kRangeBits) & kRangeMask; }
__device__ __forceinline__ uint32_t range_end(unsigned long long w) { return static_cast<uint32_t>(w) & kRangeMask; }

// Owner (one lane): the next round [t, t_end) of its published range, false when it is used up.
__device__ __forceinline__ bool claim_round(unsigned long long* slot, uint32_t& t, uint32_t& t_end) {
    unsigned long long w = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const uint32_t cur = range_cur(w), end = range_end(w);
        if (cur >= end) return false;
        const uint32_t nc = min(cur + 64u, end);
        const unsigned long long nw = (w & ~(static_cast<unsigned long long>(kRangeMask) << kRangeBits)) |
                                      (static_cast<unsigned long long>(nc) << kRangeBits);
        const unsigned long long seen = atomicCAS(slot, w, nw);
        if (seen == w) {
            t = cur;
            t_end = nc;
            return true;
        }
        w = seen;
    }
}

// Thief (whole wave): the upper half of some wave's range with at least two rounds left.
__device__ __forceinline__ bool steal_range(unsigned long long* donate, const uint32_t* donate_rec, uint32_t total_waves, uint32_t gw,
                                            uint32_t lane, uint32_t& rec, uint32_t& lo, uint32_t& hi) {
    for (uint32_t a = 0; a < kStealRounds; a++) {
        const uint32_t v = (gw + 1 + a * 64 + lane * (1 + a)) % total_waves;
        const unsigned long long w = v != gw ? __hip_atomic_load(donate + v, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : 0ull;
        const uint32_t cur = range_cur(w), end = range_end(w);
        const uint32_t rem = end > cur ? end - cur : 0u;
        uint32_t key = rem << 6 | lane;  // rem < 2^24: fits
        for (int o = 32; o > 0; o >>= 1) key = max(key, static_cast<uint32_t>(__shfl_xor(key, o, 64)));
        if ((key >> 6) < 128u) continue;
        const uint32_t src = key & 63u;
        uint32_t got = 0, r = 0, mid = 0;
        if (lane == src) {
            mid = cur + (((rem >> 1) + 63u) & ~63u);  // cur + 64 <= mid < end
            r = __hip_atomic_load(donate_rec + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned long long nw = (w & ~static_cast<unsigned long long>(kRangeMask)) | mid;
            got = atomicCAS(donate + v, w, nw) == w ? 1u : 0u;
        }
        if (__shfl(got, src, 64)) {
            rec = __shfl(r, src, 64);
            lo = __shfl(mid, src, 64);
            hi = __shfl(end, src, 64);
            return true;
        }
    }
    return false;
}

template <int kMode>
__global__ void __launch_bounds__(kWaves * 64) k_extend(DevGraph g, StepDesc d, uint32_t d0, uint32_t end_depth, Ring r, MatchOut out,
                                                         const unsigned long long* split, unsigned long long* work, int schedule,
                                                         unsigned long long* donate, uint32_t* donate_rec) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t rows = end_depth > d0 + 1 ? end_depth - 1 - d0 : 0u;
    unsigned char* wave_base = lds_raw + (threadIdx.x >> 6) * wave_lds_bytes(rows);
    WaveHdr& W = *reinterpret_cast<WaveHdr*>(wave_base);
    uint32_t* const buf = reinterpret_cast<uint32_t*>(wave_base + sizeof(WaveHdr));  // row (depth - d0)
    const uint32_t lane = lane_id();
    const uint32_t nrec = min(*r.in_count, r.in_max);
    const unsigned long long nitems = split ? split[nrec] : nrec;
    const uint32_t total_waves = gridDim.x * kWaves;
    const uint32_t gw = blockIdx.x * kWaves + (threadIdx.x >> 6);
    unsigned long long my_cnt = 0, my_hash = 0, my_bytes = 0;

    const bool donating = schedule == 2;
    uint32_t version = 0;
    unsigned long long item;
    if (schedule == 0) item = gw;
    else {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(work, 1ull);
        item = __shfl(t, 0, 64);
    }
    for (;;) {
        // the next item of the launch, else (donating) a range taken from another wave
        uint32_t rec_idx, chunk = 0, s_lo = 0, s_hi = 0;
        bool stolen = false;
        if (item < nitems) {
            rec_idx = static_cast<uint32_t>(item);
            if (split) {
                rec_idx = find_record(split, nrec, item, lane);
                chunk = static_cast<uint32_t>(item - split[rec_idx]);
            }
        } else if (donating && steal_range(donate, donate_rec, total_waves, gw, lane, rec_idx, s_lo, s_hi)) {
            stolen = true;
        } else {
            break;
        }
        if (kMode == kWriteAtEnd && uniform(__hip_atomic_load(r.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) break;
        if (lane < d0) W.f[lane] = r.in[static_cast<uint64_t>(rec_idx) * d0 + lane];
        wave_sync();
        if (lane == 0) {
            unsigned long long h = 0;
            W.ph[0] = 0;
            for (uint32_t j = 0; j < d0; j++) { h += pair_hash(d.order[j], W.f[j]); W.ph[j + 1] = h; }
        }
        wave_sync();
        uint32_t depth = d0;
        setup_frame(g, d, W, depth, lane, my_bytes, !stolen && chunk == 0);
        if (split && !stolen) {
            if (lane == 0) {
                const uint32_t k = static_cast<uint32_t>(split[rec_idx + 1] - split[rec_idx]);  // chunks of the record
                const uint32_t width = ((W.tot[d0] + k - 1) / k + 63u) & ~63u;
                W.cur[d0] = chunk * width;
                W.tot[d0] = min(W.tot[d0], chunk * width + width);
            }
            wave_sync();
        }
        // donating: publish the item's first-depth range and take its first round from the word
        // (a range past 2^24 candidates stays the wave's own)
        bool published = false;
        if (donating) {
            uint32_t lo = stolen ? s_lo : W.cur[d0], hi = stolen ? s_hi : W.tot[d0];
            published = hi <= kRangeMask;
            if (published) {
                version = (version + 1) & 0xFFFFu;
                uint32_t t = 0, t_end = 0, ok = 0;
                if (lane == 0) {
                    __hip_atomic_store(donate_rec + gw, rec_idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(donate + gw,
                                       static_cast<unsigned long long>(version) << (2 * kRangeBits) |
                                           static_cast<unsigned long long>(lo) << kRangeBits | hi,
                                       __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                    ok = claim_round(donate + gw, t, t_end) ? 1u : 0u;
                    W.cur[d0] = ok ? t : 0u;
                    W.tot[d0] = ok ? t_end : 0u;
                }
            } else if (stolen && lane == 0) {
                W.cur[d0] = lo;
                W.tot[d0] = hi;
            }
            wave_sync();
        }
        while (true) {
            if (W.bpos[depth] < W.bcnt[depth]) {
                const uint32_t c = buf[(depth - d0) * 64u + W.bpos[depth]];
                wave_sync();
                if (lane == 0) {
                    W.bpos[depth] += 1;
                    W.f[depth] = c;
                    W.ph[depth + 1] = W.ph[depth] + pair_hash(d.order[depth], c);
                }
                wave_sync();
                depth += 1;
                setup_frame(g, d, W, depth, lane, my_bytes, true);
                continue;
            }
            if (W.cur[depth] >= W.tot[depth]) {
                if (depth == d0) {
                    if (!published) break;
                    uint32_t ok = 0;
                    if (lane == 0) {
                        uint32_t t = 0, t_end = 0;
                        ok = claim_round(donate + gw, t, t_end) ? 1u : 0u;
                        if (ok) {
                            W.cur[d0] = t;
                            W.tot[d0] = t_end;
                        }
                    }
                    if (!__shfl(ok, 0, 64)) break;
                    wave_sync();
                    continue;
                }
                depth -= 1;
                continue;
            }
            const uint32_t t = W.cur[depth] + lane;
            const bool valid = t < W.tot[depth];
            const uint32_t p = d.pivot[depth];
            const uint32_t key = valid ? run_elem(g, entry_delta(g, d, d.be[p]), W.run[p], W.nb[depth], t) : 0u;
            const bool ok = valid && test_candidate(g, d, W, depth, key);
            const uint32_t c = key & g.id_mask;
            const uint64_t m = __ballot(ok);
            wave_sync();
            if (lane == 0) { W.cur[depth] += 64; my_bytes += 4ull * static_cast<uint32_t>(__popcll(m)); }
            if (depth + 1 == end_depth) {
                if (kMode == kCountAtEnd) {
                    if (ok) {
                        my_cnt += 1;
                        my_hash += embedding_hash(W.ph[depth] + pair_hash(d.order[depth], c));
                    }
                } else if (m) {
                    const uint32_t np = static_cast<uint32_t>(__popcll(m));
                    uint32_t base = 0;
                    if (lane == 0) base = atomicAdd(r.out_count, np);
                    base = __shfl(base, 0, 64);
                    if (static_cast<uint64_t>(base) + np > r.out_cap) {
                        if (lane == 0) atomicOr(r.overflow, 1u);
                    } else if (ok) {
                        uint32_t* o = r.out + static_cast<uint64_t>(base + lanes_below(m)) * end_depth;
                        for (uint32_t j = 0; j < depth; j++) o[j] = W.f[j];
                        o[depth] = c;
                    }
                }
            } else {
                if (ok) buf[(depth - d0) * 64u + lanes_below(m)] = c;
                if (lane == 0) { W.bcnt[depth] = static_cast<uint32_t>(__popcll(m)); W.bpos[depth] = 0; }
            }
            wave_sync();
        }
        if (stolen) continue;  // the launch's items are used up: steal again
        if (schedule == 0) item += total_waves;
        else {
            unsigned long long t2 = 0;
            if (lane == 0) t2 = atomicAdd(work, 1ull);
            item = __shfl(t2, 0, 64);
        }
    }
    if (kMode == kCountAtEnd) {  // one pair of device atomics per block
        __shared__ unsigned long long s_cnt[kWaves], s_hash[kWaves];
        my_cnt = wave_sum_u64(my_cnt);
        my_hash = wave_sum_u64(my_hash);
        if (lane == 0) { s_cnt[threadIdx.x >> 6] = my_cnt; s_hash[threadIdx.x >> 6] = my_hash; }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long c = 0, h = 0;
            for (uint32_t w = 0; w < kWaves; w++) { c += s_cnt[w]; h += s_hash[w]; }
            if (c) { atomicAdd(out.count, c); atomicAdd(out.hash, h); }
        }
    }
    if (out.bytes && lane == 0 && my_bytes) atomicAdd(out.bytes, my_bytes);
}

// ---- One level, flattened over the whole level ------------------------------------------------
// k_level_len: per record, the view runs of all backward entries of depth d0 (entry order, stored
// at runs[i * E ..], E = entries of the depth) and the entry whose view sequence is the shortest:
// that one generates the record's candidates, the others are probed (the candidate set is their
// intersection whichever entry generates it).  Its length goes to len[i] and, when counting, the
// record's prefix hash to ph[i].  After an exclusive scan, k_level_flat gives one thread per
// candidate of the whole level: thread t finds its record by binary search over the offsets, so
// skewed records spread over the grid, and probes the precomputed runs (no directory or delta
// segment lookups per candidate).
// Algorithmic bytes (as the DFS kernel's frames): every view run of the depth's backward lists
// read once per record, plus 4 per passing candidate (counted by k_level_flat).
// Membership of key[j] in the sorted a[lo[j], hi[j]) for the active j, N searches advancing in
// lock-step (each round issues all N loads before using any).  A present key is always probed: a
// lower-bound search probes the position it ends on unless that position is hi.
template <uint32_t N>
__device__ __forceinline__ void contains_many(const uint32_t* a, const uint32_t* lo0, const uint32_t* hi0, const uint32_t* key,
                                              const bool* act, bool* found) {
    uint32_t lo[N], n[N];
    bool more = false;
#pragma unroll
    for (uint32_t j = 0; j < N; j++) {
        lo[j] = lo0[j];
        n[j] = act[j] ? hi0[j] - lo0[j] : 0u;
        found[j] = false;
        more = more || n[j] != 0;
    }
    while (more) {
        uint32_t v[N];
#pragma unroll
        for (uint32_t j = 0; j < N; j++) v[j] = n[j] ? a[lo[j] + (n[j] >> 1)] : 0u;
        more = false;
#pragma unroll
        for (uint32_t j = 0; j < N; j++) {
            if (!n[j]) continue;
            const uint32_t h = n[j] >> 1;
            if (v[j] == key[j]) {
                found[j] = true;
                n[j] = 0;
            } else if (v[j] < key[j]) {
                lo[j] += h + 1;
                n[j] -= h + 1;
            } else {
                n[j] = h;
            }
            more = more || n[j] != 0;
        }
    }
}

constexpr uint32_t kRunsUnroll = 4;  // entries whose lookups are all issued before the first store

__device__ __forceinline__ uint32_t record_runs(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* f, uint4* runs,
                                                uint4* prun, uint8_t* piv, unsigned long long* ph, unsigned long long& bytes,
                                                bool count_bytes) {
    const uint32_t e0 = d.boff[d0], ne = d.boff[d0 + 1] - e0;
    uint32_t best = 0xFFFFFFFFu, tag = 0;
    uint4 brun = make_uint4(0, 0, 0, 0);
    auto take = [&](uint32_t e, const RunView& rv) {
        const BackEntry be = d.be[e0 + e];
        const uint4 x = make_uint4(rv.lo, rv.hi, rv.dlo, rv.dhi);
        if (ne > 1) runs[e] = x;
        const uint32_t len = run_seq_len(rv, be.view);
        if (len < best) {
            best = len;
            brun = x;
            tag = e | static_cast<uint32_t>(be.view) << 4 | static_cast<uint32_t>(be.dir) << 6;
        }
        if (count_bytes) bytes += 4ull * ((rv.hi - rv.lo) + (rv.dhi - rv.dlo));
    };
    // the stores below may alias the graph for the compiler: look the first entries up first, so
    // their dependent loads overlap
    RunView rv[kRunsUnroll];
#pragma unroll
    for (uint32_t e = 0; e < kRunsUnroll; e++) {
        if (e >= ne) break;
        const BackEntry be = d.be[e0 + e];
        rv[e] = make_run(g, d.phase, be.view, be.dir, f[be.pos], be.slot);
    }
    unsigned long long h = 0;
    if (ph)
        for (uint32_t j = 0; j < d0; j++) h += pair_hash(d.order[j], f[j]);
#pragma unroll
    for (uint32_t e = 0; e < kRunsUnroll; e++) {
        if (e >= ne) break;
        take(e, rv[e]);
    }
    for (uint32_t e = kRunsUnroll; e < ne; e++) {
        const BackEntry be = d.be[e0 + e];
        take(e, make_run(g, d.phase, be.view, be.dir, f[be.pos], be.slot));
    }
    *prun = brun;
    *piv = static_cast<uint8_t>(tag);
    if (ph) *ph = h;
    return best;
}

// Block-wide exclusive scan of one value per thread over T threads (s_wave holds T / 64 words);
// returns the exclusive prefix and the block total.
template <uint32_t T>
__device__ __forceinline__ unsigned long long block_scan(unsigned long long v, unsigned long long* s_wave, unsigned long long& total) {
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    unsigned long long x = v;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[wv] = x;
    __syncthreads();
    unsigned long long before = 0;
    total = 0;
    for (uint32_t k = 0; k < T / 64; k++) {
        const unsigned long long t = s_wave[k];
        if (k < wv) before += t;
        total += t;
    }
    __syncthreads();  // s_wave may be rewritten by the caller's next scan
    return before + x - v;
}

constexpr uint32_t kSmallThreads = 1024;
// Multi-block level preparation: 256-thread blocks (small enough to find room beside the other
// lanes' kernels), two records per thread.
constexpr uint32_t kLenThreads = 256;
constexpr uint32_t kLenItems = 2;
constexpr uint32_t kLenTile = kLenThreads * kLenItems;
static_assert(kLenTile == kLevelTile, "tile totals are sized by kLevelTile");
constexpr int kChainLenGrid = 512;    // blocks of a chained level's preparation (they loop over tiles)
constexpr int kChainFlatGrid = 1024;  // blocks of a chained level's candidate kernel (four per CU)

// Where level k of a device chain reads its records: false when the level does not run (an
// earlier level stopped the chain or overflowed, or its own input is empty, larger than one call
// takes, or large enough for the host's Cartesian-product decision).  Every thread computes the
// same answer from the host-known start and the counters of the levels before k.
__device__ __forceinline__ bool chain_at(const LevelChain& c, uint32_t k, uint64_t& off, uint32_t& n) {
    off = c.off0;
    n = c.n0;
    for (uint32_t j = 0;; j++) {
        if (n == 0 || n > c.cap[j] || n >= c.stop_at[j]) return false;
        if (j == k) return true;
        if (c.ctr[2 * j + 1]) return false;
        off += static_cast<uint64_t>(n) * (c.depth0 + j);
        n = c.ctr[2 * j];
    }
}

// Multi-block levels: lengths and runs of the records of tiles blockIdx.x, blockIdx.x + gridDim.x,
// ..., and each tile's candidate total in tile_sum[tile]; level_scan then turns the lengths into
// offsets.
__device__ __forceinline__ void level_len_tiles(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n,
                                                unsigned long long* len, uint4* prun, uint4* runs, uint8_t* piv, unsigned long long* ph,
                                                unsigned long long* tile_sum, unsigned long long* bytes_out) {
    __shared__ unsigned long long s_wave[kLenThreads / 64];
    const uint32_t ne = d.boff[d0 + 1] - d.boff[d0];
    const uint32_t tiles = (n + kLenTile - 1) / kLenTile;
    unsigned long long bytes = 0;
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        unsigned long long l[kLenItems], sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kLenItems; j++) {
            const uint32_t i = tile * kLenTile + j * kLenThreads + threadIdx.x;
            l[j] = 0;
            if (i >= n) continue;
            l[j] = record_runs(g, d, d0, rec + static_cast<uint64_t>(i) * d0, runs + static_cast<uint64_t>(i) * ne, prun + i, piv + i,
                               ph ? ph + i : nullptr, bytes, bytes_out != nullptr);
            sum += l[j];
        }
        if (tiles == 1) {  // a one-tile level is scanned here (item rows in record order); level_scan skips it
            unsigned long long carry = 0;
#pragma unroll
            for (uint32_t j = 0; j < kLenItems; j++) {
                unsigned long long row;
                const unsigned long long ex = block_scan<kLenThreads>(l[j], s_wave, row);
                const uint32_t i = j * kLenThreads + threadIdx.x;
                if (i < n) len[i] = carry + ex;
                carry += row;
            }
            if (threadIdx.x == 0) len[n] = carry;
            continue;
        }
#pragma unroll
        for (uint32_t j = 0; j < kLenItems; j++) {
            const uint32_t i = tile * kLenTile + j * kLenThreads + threadIdx.x;
            if (i < n) len[i] = l[j];
        }
        unsigned long long total;
        block_scan<kLenThreads>(sum, s_wave, total);
        if (threadIdx.x == 0) tile_sum[tile] = total;
    }
    if (bytes_out) {
        bytes = wave_sum_u64(bytes);
        if (lane_id() == 0 && bytes) atomicAdd(bytes_out, bytes);
    }
}

__global__ void __launch_bounds__(kLenThreads) k_level_len(DevGraph g, StepDesc d, uint32_t d0, const uint32_t* rec, uint32_t n,
                                                             unsigned long long* len, uint4* prun, uint4* runs, uint8_t* piv,
                                                             unsigned long long* ph, unsigned long long* tile_sum, uint32_t* zero,
                                                             unsigned long long* bytes_out) {
    if (zero && blockIdx.x == 0 && threadIdx.x == 0) { zero[0] = 0; zero[1] = 0; }
    level_len_tiles(g, d, d0, rec, n, len, prun, runs, piv, ph, tile_sum, bytes_out);
}

__global__ void __launch_bounds__(kLenThreads) k_level_len_chain(DevGraph g, StepDesc d, LevelChain c, uint32_t k, unsigned long long* len,
                                                                   uint4* prun, uint4* runs, uint8_t* piv, unsigned long long* ph,
                                                                   unsigned long long* tile_sum, unsigned long long* bytes_out) {
    if (k == 0 && blockIdx.x == 0 && threadIdx.x < 2 * kMaxQV) c.ctr[threadIdx.x] = 0;
    uint64_t off;
    uint32_t n;
    if (!chain_at(c, k, off, n)) return;
    level_len_tiles(g, d, c.depth0 + k, c.ring + off, n, len, prun, runs, piv, ph, tile_sum, bytes_out);
}

// Exclusive scan of the lengths in place (total at off[n]): the block of a tile adds the totals of
// the tiles before it (at most a few thousand words, read by the whole block) to the tile's scan.
__device__ __forceinline__ void level_scan_tiles(unsigned long long* off, uint32_t n, const unsigned long long* tile_sum) {
    __shared__ unsigned long long s_wave[kLenThreads / 64];
    const uint32_t tiles = (n + kLenTile - 1) / kLenTile;
    if (tiles == 1) return;  // scanned by level_len_tiles
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        unsigned long long carry = 0, total;
        for (uint32_t t = threadIdx.x; t < tile; t += kLenThreads) carry += tile_sum[t];
        carry = wave_sum_u64(carry);
        if (lane_id() == 0) s_wave[threadIdx.x >> 6] = carry;
        __syncthreads();
        carry = 0;
        for (uint32_t w = 0; w < kLenThreads / 64; w++) carry += s_wave[w];
        __syncthreads();
#pragma unroll
        for (uint32_t j = 0; j < kLenItems; j++) {
            const uint32_t i = tile * kLenTile + j * kLenThreads + threadIdx.x;
            const unsigned long long v = i < n ? off[i] : 0ull;
            const unsigned long long ex = block_scan<kLenThreads>(v, s_wave, total);
            if (i < n) off[i] = carry + ex;
            carry += total;
        }
        if (tile == tiles - 1 && threadIdx.x == 0) off[n] = carry;
    }
}

__global__ void __launch_bounds__(kLenThreads) k_level_scan(unsigned long long* off, uint32_t n, const unsigned long long* tile_sum) {
    level_scan_tiles(off, n, tile_sum);
}

__global__ void __launch_bounds__(kLenThreads) k_level_scan_chain(LevelChain c, uint32_t k, unsigned long long* off,
                                                                    const unsigned long long* tile_sum) {
    uint64_t o;
    uint32_t n;
    if (!chain_at(c, k, o, n)) return;
    level_scan_tiles(off, n, tile_sum);
}

// Levels of under kSmallLevel records (most of them): sizes and scan in one single-block launch.
constexpr uint32_t kSmallItems = 4;
constexpr uint32_t kSmallLevel = kSmallThreads * kSmallItems;

__global__ void __launch_bounds__(kSmallThreads) k_level_len_small(DevGraph g, StepDesc d, uint32_t d0, const uint32_t* rec, uint32_t n,
                                                                   unsigned long long* off, uint4* prun, uint4* runs, uint8_t* piv,
                                                                   unsigned long long* ph, uint32_t* zero, unsigned long long* bytes_out) {
    __shared__ unsigned long long s_wave[kSmallThreads / 64];
    const uint32_t tid = threadIdx.x;
    if (zero && tid == 0) { zero[0] = 0; zero[1] = 0; }
    const uint32_t ne = d.boff[d0 + 1] - d.boff[d0];
    unsigned long long len[kSmallItems], bytes = 0;
    // thread tid owns records j * kSmallThreads + tid (independent lookups in flight together)
#pragma unroll
    for (uint32_t j = 0; j < kSmallItems; j++) {
        const uint32_t i = j * kSmallThreads + tid;
        len[j] = 0;
        if (i >= n) continue;
        len[j] = record_runs(g, d, d0, rec + static_cast<uint64_t>(i) * d0, runs + static_cast<uint64_t>(i) * ne, prun + i, piv + i,
                             ph ? ph + i : nullptr, bytes, bytes_out != nullptr);
    }
    // exclusive scan in record order: one block scan per item row, carried over the rows
    unsigned long long carry = 0;
    for (uint32_t j = 0; j < kSmallItems && j * kSmallThreads < n; j++) {
        unsigned long long total;
        const unsigned long long ex = block_scan<kSmallThreads>(len[j], s_wave, total);
        const uint32_t i = j * kSmallThreads + tid;
        if (i < n) off[i] = carry + ex;
        carry += total;
    }
    if (tid == 0) off[n] = carry;
    if (bytes_out) {
        bytes = wave_sum_u64(bytes);
        if (lane_id() == 0 && bytes) atomicAdd(bytes_out, bytes);
    }
}

// Each wave walks one contiguous range of the level's candidates (equal shares, multiples of
// kLevelStep).  It finds the record of its first candidate with a 64-ary search; per kLevelStep
// candidates it loads the next kLevelStep + 1 record offsets into LDS (coalesced) and each lane
// finds the records of its kLevelIlp candidates there (a global search only past the window, when
// records have under one candidate each).  The kLevelIlp candidates of a lane are tested in
// lock-step so their independent loads are in flight together: the kernel is latency-bound.  Four
// candidates per lane keep the kernel at 76 VGPRs, six waves per SIMD; with the lock-step searches
// that beat eight (four waves per SIMD) and two on the 400K-vertex Chung-Lu workload.
constexpr uint32_t kLevelIlp = 4;
constexpr uint32_t kLevelStep = 64 * kLevelIlp;
static_assert((kLevelStep & (kLevelStep - 1)) == 0, "the window search takes power-of-two steps");

// Write mode stages passing (record, candidate) pairs per wave in LDS and reserves ring space once
// per kOutFlush pairs: an atomic per iteration would put its round trip on every iteration.
constexpr uint32_t kOutFlush = 256;
constexpr uint32_t kOutBuf = kOutFlush + kLevelStep;

__device__ __forceinline__ void flush_records(const uint32_t* rec, uint32_t d0, const uint32_t* sk, const uint32_t* sc, uint32_t cnt, Ring& r,
                                              uint32_t lane) {
    uint32_t wb = 0;
    if (lane == 0) wb = atomicAdd(r.out_count, cnt);
    wb = uniform(__shfl(wb, 0, 64));
    if (static_cast<uint64_t>(wb) + cnt > r.out_cap) {
        if (lane == 0) atomicOr(r.overflow, 1u);
        return;
    }
    // consecutive words of the reserved region: coalesced stores
    const uint32_t rw = d0 + 1;
    uint32_t* o = r.out + static_cast<uint64_t>(wb) * rw;
    for (uint32_t w = lane; w < cnt * rw; w += 64) {
        const uint32_t e = w / rw, q = w - e * rw;
        o[w] = q < d0 ? rec[static_cast<uint64_t>(sk[e]) * d0 + q] : sc[e];
    }
}

template <int kMode>
__device__ __forceinline__ void level_flat_body(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n,
                                                const unsigned long long* off, const uint4* prun, const uint4* runs, const uint8_t* piv,
                                                const unsigned long long* ph, Ring r, const MatchOut& out) {
    // window of record offsets, relative to off[rb] (a window spans fewer than 2^32 candidates)
    __shared__ uint32_t s_win[4][kLevelStep + 1];
    __shared__ uint32_t s_out[4][kMode == kWriteAtEnd ? 2 * kOutBuf : 1];
    const uint32_t lane = lane_id();
    uint32_t* win = s_win[threadIdx.x >> 6];
    uint32_t* const out_k = s_out[threadIdx.x >> 6];
    uint32_t* const out_c = out_k + (kMode == kWriteAtEnd ? kOutBuf : 0);
    uint32_t staged = 0;  // pairs in the wave's LDS buffer
    const uint32_t e0 = d.boff[d0], ne = d.boff[d0 + 1] - e0;
    const uint32_t* filter = d.filter[d0];
    const uint32_t qv = d.order[d0];
    const unsigned long long total = off[n];
    const unsigned long long waves = static_cast<unsigned long long>(gridDim.x) * (blockDim.x >> 6);
    const unsigned long long gw = static_cast<unsigned long long>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const unsigned long long per = ((total + waves - 1) / waves + kLevelStep - 1) / kLevelStep * kLevelStep;
    const unsigned long long w0 = gw * per, w1 = min(total, w0 + per);
    unsigned long long cnt = 0, hsum = 0, bytes = 0;
    uint32_t rb = w0 < total ? find_record(off, n, w0, lane) : 0u;  // off[rb] <= w0
    for (unsigned long long t0 = w0; t0 < w1; t0 += kLevelStep) {
        // a level that overflows is redone by the host with fewer records: stop early (polled every
        // 16 iterations; a load per iteration would sit on every iteration's latency chain)
        if (kMode == kWriteAtEnd && ((t0 - w0) & (16ull * kLevelStep - 1)) == 0 && t0 != w0 &&
            uniform(__hip_atomic_load(r.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
            break;
        const unsigned long long wbase = off[rb];
        auto rel = [&](uint32_t i) {
            return static_cast<uint32_t>(min(off[min(rb + i, n)] - wbase, 0xFFFFFFFFull));
        };
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++) win[lane + 64 * j] = rel(lane + 64 * j);
        if (lane == 0) win[kLevelStep] = rel(kLevelStep);
        wave_sync();
        unsigned long long t[kLevelIlp];
        uint32_t k[kLevelIlp], key[kLevelIlp], c[kLevelIlp];
        bool ok[kLevelIlp];
        // record of each candidate: last i < kLevelStep with win[i] <= t - off[rb] (win[0] <= t0 <= t),
        // a fixed-step search run for all kLevelIlp candidates at once so their LDS reads overlap
        uint32_t tr[kLevelIlp], pos[kLevelIlp];
        bool past = false;
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++) {
            t[j] = t0 + 64 * j + lane;
            ok[j] = t[j] < w1;
            const unsigned long long x = ok[j] ? t[j] - wbase : 0ull;  // < 2^32: inside record rb or the window
            tr[j] = static_cast<uint32_t>(min(x, 0xFFFFFFFFull));
            past = past || (ok[j] && win[kLevelStep] <= x);
            pos[j] = 0;
        }
#pragma unroll
        for (uint32_t st = kLevelStep / 2; st; st >>= 1) {
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++)
                if (win[pos[j] + st] <= tr[j]) pos[j] += st;
        }
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++) k[j] = rb + pos[j];
        if (past) {  // records with fewer than one candidate each: last k in [rb + kLevelStep, n) with off[k] <= t
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++) {
                if (!ok[j] || win[kLevelStep] > t[j] - wbase) continue;
                uint32_t lo = rb + kLevelStep, hi = n;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (off[mid] <= t[j]) lo = mid;
                    else hi = mid;
                }
                k[j] = lo;
            }
        }
        // the record's generating run (prun) and its entry | view << 4 | dir << 6 (piv): independent loads
        RunView rv[kLevelIlp];
        unsigned long long base[kLevelIlp];
        uint32_t pv[kLevelIlp];
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++) {
            rv[j] = RunView{0, 0, 0, 0};
            base[j] = 0;
            pv[j] = 0;
            if (!ok[j]) continue;
            const uint4 x = prun[k[j]];
            rv[j] = RunView{x.x, x.y, x.z, x.w};
            pv[j] = piv[k[j]];
            base[j] = k[j] - rb < kLevelStep ? wbase + win[k[j] - rb] : off[k[j]];
        }
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++) {
            const uint32_t view = (pv[j] >> 4) & 3u;
            key[j] = ok[j] ? run_elem(g, phase_delta(g, pv[j] >> 6, d.phase), rv[j], run_base_len(rv[j], view),
                                      static_cast<uint32_t>(t[j] - base[j]))
                           : 0u;
            c[j] = key[j] & g.id_mask;
        }
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++) ok[j] = ok[j] && bit_test(filter, c[j]);
#pragma unroll
        for (uint32_t j = 0; j < kLevelIlp; j++)
            ok[j] = ok[j] && run_elem_live(phase_delta(g, pv[j] >> 6, d.phase), rv[j], (pv[j] >> 4) & 3u, key[j]);
        for (uint32_t im = d.inj[d0]; im; im &= im - 1) {
            const uint32_t q = static_cast<uint32_t>(__ffs(im)) - 1;
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++) ok[j] = ok[j] && rec[static_cast<uint64_t>(k[j]) * d0 + q] != c[j];
        }
        // probe the other entries' precomputed runs (each with its own slot bits); the searches of
        // the kLevelIlp candidates advance in lock-step so their loads are in flight together
        for (uint32_t e = 0; e < ne; e++) {
            const BackEntry be = d.be[e0 + e];
            const uint32_t hi_bits = be.slot << g.id_bits;
            // runs as {lo, hi} / {dlo, dhi} halves: the delta half is loaded only for views that use it
            const uint2* rh = reinterpret_cast<const uint2*>(runs);
            uint32_t pk[kLevelIlp];
            bool act[kLevelIlp], inb[kLevelIlp], ind[kLevelIlp];
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++) {
                act[j] = ok[j] && (pv[j] & 15u) != e;
                pk[j] = hi_bits | c[j];
                inb[j] = false;
                ind[j] = false;
            }
            if (be.view != kViewDelta) {
                uint32_t lo[kLevelIlp], hi[kLevelIlp];
#pragma unroll
                for (uint32_t j = 0; j < kLevelIlp; j++) {
                    const uint2 x = act[j] ? rh[2 * (static_cast<uint64_t>(k[j]) * ne + e)] : make_uint2(0, 0);
                    lo[j] = x.x;
                    hi[j] = x.y;
                }
                contains_many<kLevelIlp>(g.arena, lo, hi, pk, act, inb);
            }
            if (be.view != kViewBase) {
                const uint32_t* dl = phase_delta(g, be.dir, d.phase);
                uint32_t lo[kLevelIlp], hi[kLevelIlp];
                bool need[kLevelIlp];
#pragma unroll
                for (uint32_t j = 0; j < kLevelIlp; j++) {
                    need[j] = act[j] && (be.view == kViewDelta || (be.view == kViewBasePlusDelta ? !inb[j] : inb[j]));
                    const uint2 x = need[j] ? rh[2 * (static_cast<uint64_t>(k[j]) * ne + e) + 1] : make_uint2(0, 0);
                    lo[j] = x.x;
                    hi[j] = x.y;
                }
                contains_many<kLevelIlp>(dl, lo, hi, pk, need, ind);
            }
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++) {
                if (!act[j]) continue;
                switch (be.view) {
                    case kViewBase: ok[j] = inb[j]; break;
                    case kViewBasePlusDelta: ok[j] = inb[j] || ind[j]; break;
                    case kViewDelta: ok[j] = ind[j]; break;
                    default: ok[j] = inb[j] && !ind[j]; break;
                }
            }
        }
        if (kMode == kCountAtEnd) {
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++)
                if (ok[j]) {
                    cnt += 1;
                    hsum += embedding_hash(ph[k[j]] + pair_hash(qv, c[j]));
                }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++) {
                const uint64_t m = __ballot(ok[j]);
                if (ok[j]) {
                    const uint32_t idx = staged + lanes_below(m);
                    out_k[idx] = k[j];
                    out_c[idx] = c[j];
                }
                staged += static_cast<uint32_t>(__popcll(m));
            }
            if (staged >= kOutFlush || t0 + kLevelStep >= w1) {
                wave_sync();
                if (staged) flush_records(rec, d0, out_k, out_c, staged, r, lane);
                staged = 0;
            }
        }
        if (out.bytes) {
#pragma unroll
            for (uint32_t j = 0; j < kLevelIlp; j++) bytes += ok[j] ? 4 : 0;
        }
        rb = uniform(__shfl(k[kLevelIlp - 1], 63, 64));  // highest record of this step (all lanes valid unless last)
        wave_sync();                                    // the window is rewritten next iteration
    }
    if (kMode == kCountAtEnd) {
        // one pair of device atomics per block (they serialise on one address across the grid)
        __shared__ unsigned long long s_cnt[4], s_hash[4];
        cnt = wave_sum_u64(cnt);
        hsum = wave_sum_u64(hsum);
        if (lane == 0) { s_cnt[threadIdx.x >> 6] = cnt; s_hash[threadIdx.x >> 6] = hsum; }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long c = 0, h = 0;
            for (uint32_t w = 0; w < (blockDim.x >> 6); w++) { c += s_cnt[w]; h += s_hash[w]; }
            if (c) { atomicAdd(out.count, c); atomicAdd(out.hash, h); }
        }
    }
    if (out.bytes) {
        bytes = wave_sum_u64(bytes);
        if (lane == 0 && bytes) atomicAdd(out.bytes, bytes);
    }
}

template <int kMode>
__global__ void __launch_bounds__(256) k_level_flat(DevGraph g, StepDesc d, uint32_t d0, const uint32_t* rec, uint32_t n,
                                                    const unsigned long long* off, const uint4* prun, const uint4* runs,
                                                    const uint8_t* piv, const unsigned long long* ph, Ring r, MatchOut out) {
    level_flat_body<kMode>(g, d, d0, rec, n, off, prun, runs, piv, ph, r, out);
}

// Level k of a device chain: reads its records where chain_at says and writes the records of the
// next depth right after them ({count, overflow} in ctr[2k], ctr[2k + 1]).
template <int kMode>
__global__ void __launch_bounds__(256) k_level_flat_chain(DevGraph g, StepDesc d, LevelChain c, uint32_t k, const unsigned long long* off,
                                                          const uint4* prun, const uint4* runs, const uint8_t* piv,
                                                          const unsigned long long* ph, MatchOut out) {
    uint64_t o;
    uint32_t n;
    if (!chain_at(c, k, o, n)) return;
    const uint32_t d0 = c.depth0 + k;
    const uint64_t out_off = o + static_cast<uint64_t>(n) * d0;
    const uint64_t room = out_off < c.ring_words ? (c.ring_words - out_off) / (d0 + 1) : 0ull;
    const Ring r{c.ring + o, nullptr, n, c.ring + out_off, c.ctr + 2 * k, static_cast<uint32_t>(min(room, 0xFFFFFFFFull)),
                 c.ctr + 2 * k + 1};
    level_flat_body<kMode>(g, d, d0, c.ring + o, n, off, prun, runs, piv, ph, r, out);
}

// Level k of many steps' chains (grid.y = job): the kernels above with the job's descriptor, chain
// and scratch read from the job table (uniform per block: scalar loads).
__global__ void __launch_bounds__(kLenThreads) k_level_len_chain_multi(DevGraph g, const ChainJob* jobs, uint32_t k) {
    const ChainJob& J = jobs[blockIdx.y];
    if (k > J.L) return;
    if (k == 0 && blockIdx.x == 0 && threadIdx.x < 2 * kMaxQV) J.c.ctr[threadIdx.x] = 0;
    uint64_t off;
    uint32_t n;
    if (!chain_at(J.c, k, off, n)) return;
    const bool counting = k == J.L;
    level_len_tiles(g, J.d, J.c.depth0 + k, J.c.ring + off, n, J.ls.off, J.ls.prun, J.ls.runs, J.ls.piv, counting ? J.ls.ph : nullptr,
                    J.ls.tile_sum, counting ? J.out.bytes : J.write_bytes);
}

__global__ void __launch_bounds__(kLenThreads) k_level_scan_chain_multi(const ChainJob* jobs, uint32_t k) {
    const ChainJob& J = jobs[blockIdx.y];
    if (k > J.L) return;
    uint64_t o;
    uint32_t n;
    if (!chain_at(J.c, k, o, n)) return;
    level_scan_tiles(J.ls.off, n, J.ls.tile_sum);
}

// At least six waves per SIMD.  The writing form otherwise takes 86 VGPRs and five waves; bounded it
// takes 80 and spills one 8-byte value per lane, stored once before its candidate loop and reloaded
// only when staged records are flushed to the ring.  The kernel is latency-bound and shares the GPU
// with the other teams' launches, so the extra resident wave measured faster.
template <int kMode>
__global__ void __launch_bounds__(256, 6) k_level_flat_chain_multi(DevGraph g, const ChainJob* jobs, uint32_t k) {
    const ChainJob& J = jobs[blockIdx.y];
    if (k > J.L || (k == J.L) != (kMode == kCountAtEnd)) return;
    uint64_t o;
    uint32_t n;
    if (!chain_at(J.c, k, o, n)) return;
    const uint32_t d0 = J.c.depth0 + k;
    const uint64_t out_off = o + static_cast<uint64_t>(n) * d0;
    const uint64_t room = out_off < J.c.ring_words ? (J.c.ring_words - out_off) / (d0 + 1) : 0ull;
    const Ring r{J.c.ring + o, nullptr, n, J.c.ring + out_off, J.c.ctr + 2 * k, static_cast<uint32_t>(min(room, 0xFFFFFFFFull)),
                 J.c.ctr + 2 * k + 1};
    const MatchOut out = kMode == kCountAtEnd ? J.out : MatchOut{nullptr, nullptr, J.write_bytes};
    level_flat_body<kMode>(g, J.d, d0, J.c.ring + o, n, J.ls.off, J.ls.prun, J.ls.runs, J.ls.piv, J.ls.ph, r, out);
}

__global__ void __launch_bounds__(256) k_root_chunks(DevGraph g, StepDesc d, uint32_t d0, const uint32_t* rec, const uint32_t* n_rec,
                                                     uint32_t max_rec, uint32_t max_chunks, unsigned long long* chunks) {
    const uint32_t n = min(*n_rec, max_rec);
    const BackEntry be = d.be[d.pivot[d0]];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t v = rec[static_cast<uint64_t>(i) * d0 + be.pos];
        const RunView rv = make_run(g, d.phase, be.view, be.dir, v, be.slot);
        chunks[i] = min(max_chunks, (run_seq_len(rv, be.view) + 63u) / 64u);
    }
}

__global__ void k_init_from_pairs(const uint32_t* pairs, const uint32_t* n_pairs, uint32_t max_pairs, int swap, uint32_t* out,
                                  uint32_t* out_count) {
    const uint32_t n = min(*n_pairs, max_pairs);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t a = pairs[2 * i], b = pairs[2 * i + 1];
        out[2 * i] = swap ? b : a;
        out[2 * i + 1] = swap ? a : b;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *out_count = n;
}

// Pairs (v, x) with v in the start list, x in the view run of v along a->b and x in filter(b).
// The runs of a wave's G start vertices are flattened (a wave prefix of their lengths, one element
// per lane per round, its vertex found by a search of the prefix in LDS), and the wave reserves
// its survivors' output once per round: most start vertices have a run of a few elements.
__device__ __forceinline__ void init_set_body(const DevGraph& g, const InitJob& J) {
    __shared__ uint32_t s_pre[4][65];
    __shared__ RunView s_run[4][64];
    __shared__ uint32_t s_nb[4][64], s_v[4][64];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wv, waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t n = min(*J.len, J.max_len);
    const uint32_t* dl = phase_delta(g, J.dir, J.phase);
    const uint32_t G = items_per_wave(n);
    for (uint32_t base = gw * G; base < n; base += waves * G) {
        const uint32_t i = base + lane;
        uint32_t tot = 0, nb = 0, v = 0;
        RunView r{0, 0, 0, 0};
        if (lane < G && i < n) {
            v = J.list[i];
            r = make_run(g, J.phase, J.view, J.dir, v, J.slot);
            nb = run_base_len(r, J.view);
            tot = run_seq_len(r, J.view);
        }
        uint32_t T;
        const uint32_t pre = wave_prefix(tot, T);
        s_pre[wv][lane] = pre;
        s_run[wv][lane] = r;
        s_nb[wv][lane] = nb;
        s_v[wv][lane] = v;
        if (lane == 63) s_pre[wv][64] = T;
        wave_sync();
        for (uint32_t e0 = 0; e0 < T; e0 += 64) {  // wave-uniform trip count
            const uint32_t e = e0 + lane;
            bool ok = false;
            uint32_t sv = 0, x = 0;
            if (e < T) {
                const uint32_t k = owner_of(s_pre[wv], e);
                const RunView rr = s_run[wv][k];
                sv = s_v[wv][k];
                const uint32_t key = run_elem(g, dl, rr, s_nb[wv][k], e - s_pre[wv][k]);
                x = key & g.id_mask;
                ok = x != sv && bit_test(J.filter_b, x) && run_elem_live(dl, rr, J.view, key);
            }
            const uint64_t m = __ballot(ok);
            if (!m) continue;
            uint32_t o = 0;
            if (lane == 0) o = atomicAdd(J.out_count, static_cast<uint32_t>(__popcll(m)));
            o = __shfl(o, 0, 64);
            if (static_cast<uint64_t>(o) + __popcll(m) > J.out_cap) {
                if (lane == 0) atomicOr(J.overflow, 1u);
            } else if (ok) {
                const uint64_t p = o + lanes_below(m);
                J.out[2 * p] = sv;
                J.out[2 * p + 1] = x;
            }
        }
        wave_sync();  // the LDS rows are rewritten by the next items
    }
}

__global__ void __launch_bounds__(256) k_init_from_set(DevGraph g, InitJob J) { init_set_body(g, J); }

__global__ void __launch_bounds__(256) k_init_from_set_multi(DevGraph g, const InitJob* jobs) { init_set_body(g, jobs[blockIdx.y]); }

// ---- Cartesian product -------------------------------------------------------------------------
// The trailing leaves d0 .. n-1 of a record each hang off one matched vertex.  A segment is one
// (record, leaf) pair, segment i = record * leaves + leaf.  k_cp_len sizes every segment's view run
// and the record's prefix hash; after a scan, k_cp_filter tests every run element (load-balanced
// over the grid, as k_level_flat) against the leaf's candidate bitmap, deletion liveness and the
// matched prefix, and appends survivors to the segment's list (order within a list is free: the
// enumeration takes every combination once); k_cp_products multiplies the list lengths per record.
__global__ void __launch_bounds__(256) k_cp_len(DevGraph g, StepDesc d, uint32_t d0, const uint32_t* rec, uint32_t n, CpScratch c) {
    const uint32_t nl = d.n - d0;
    const uint64_t nseg = static_cast<uint64_t>(n) * nl;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < nseg; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t r = static_cast<uint32_t>(i / nl), kk = static_cast<uint32_t>(i % nl);
        const uint32_t* f = rec + static_cast<uint64_t>(r) * d0;
        const BackEntry be = d.be[d.boff[d0 + kk]];
        const RunView rv = make_run(g, d.phase, be.view, be.dir, f[be.pos], be.slot);
        c.seg_off[i] = run_seq_len(rv, be.view);
        c.seg_run[i] = make_uint4(rv.lo, rv.hi, rv.dlo, rv.dhi);
        c.seg_cnt[i] = 0;
        if (kk == 0) {
            unsigned long long h = 0;
            for (uint32_t j = 0; j < d0; j++) h += pair_hash(d.order[j], f[j]);
            c.rec_hash[r] = h;
        }
    }
}

__global__ void __launch_bounds__(256) k_cp_filter(DevGraph g, StepDesc d, uint32_t d0, const uint32_t* rec, uint32_t nseg, CpScratch c) {
    __shared__ unsigned long long s_win[4][65];
    const uint32_t lane = lane_id();
    unsigned long long* win = s_win[threadIdx.x >> 6];
    const uint32_t nl = d.n - d0;
    const unsigned long long* off = c.seg_off;
    const unsigned long long total = off[nseg];
    if (total > c.lists_cap) {  // lists do not fit: the host counts these records another way
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(c.err, kErrScratchFull);
        return;
    }
    const unsigned long long waves = static_cast<unsigned long long>(gridDim.x) * (blockDim.x >> 6);
    const unsigned long long gw = static_cast<unsigned long long>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const unsigned long long per = ((total + waves - 1) / waves + 63) & ~63ull;
    const unsigned long long w0 = gw * per, w1 = min(total, w0 + per);
    uint32_t rb = w0 < total ? find_record(off, nseg, w0, lane) : 0u;  // off[rb] <= w0
    for (unsigned long long t0 = w0; t0 < w1; t0 += 64) {
        win[lane] = off[min(rb + lane, nseg)];
        if (lane == 0) win[64] = off[min(rb + 64, nseg)];
        wave_sync();
        const unsigned long long t = t0 + lane;
        uint32_t i = rb, x = 0;
        unsigned long long base = 0;
        bool ok = false;
        if (t < w1) {
            if (win[64] <= t) {
                uint32_t lo = rb + 64, hi = nseg;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (off[mid] <= t) lo = mid;
                    else hi = mid;
                }
                i = lo;
            } else {
                uint32_t lo = 0, hi = 64;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (win[mid] <= t) lo = mid;
                    else hi = mid;
                }
                i = rb + lo;
            }
            const uint32_t r = i / nl, kk = i % nl;
            const BackEntry be = d.be[d.boff[d0 + kk]];
            const uint32_t* dl = phase_delta(g, be.dir, d.phase);
            const uint4 x4 = c.seg_run[i];
            const RunView rv{x4.x, x4.y, x4.z, x4.w};
            base = i - rb < 64 ? win[i - rb] : off[i];
            const uint32_t key = run_elem(g, dl, rv, run_base_len(rv, be.view), static_cast<uint32_t>(t - base));
            x = key & g.id_mask;
            ok = bit_test(d.filter[d0 + kk], x) && run_elem_live(dl, rv, be.view, key);
            const uint32_t* f = rec + static_cast<uint64_t>(r) * d0;
            for (uint32_t im = d.inj[d0 + kk] & ((1u << d0) - 1u); im && ok; im &= im - 1) ok = f[__ffs(im) - 1] != x;
        }
        // One counter reservation per segment per wave: the lanes of a segment are consecutive
        // (i never decreases along the lanes), so each group's first lane reserves for the group.
        const uint64_t pm = __ballot(ok);
        const uint64_t starts = __ballot(lane == 0 || __shfl_up(i, 1, 64) != i);
        const uint64_t lt = (1ull << lane) - 1ull;
        const uint32_t leader = 63u - static_cast<uint32_t>(__clzll(starts & (lt | (1ull << lane))));
        const uint64_t after = starts & ~(lt | (1ull << lane));  // group starts above this lane
        const uint64_t grp_end = after ? (after & (0ull - after)) : 0ull;  // lowest later start (0: lane 63 ends)
        const uint64_t grp = (grp_end ? grp_end - 1ull : ~0ull) & ~((1ull << leader) - 1ull);
        uint32_t wb = 0;
        if (lane == leader) {
            const uint32_t cnt = static_cast<uint32_t>(__popcll(pm & grp));
            if (cnt) wb = atomicAdd(&c.seg_cnt[i], cnt);
        }
        wb = __shfl(wb, static_cast<int>(leader), 64);
        if (ok) c.lists[base + wb + static_cast<uint32_t>(__popcll(pm & lt & grp))] = x;
        rb = uniform(__shfl(i, 63, 64));
        wave_sync();
    }
}

// Thread per record: the product of its leaf list lengths (0 when a list is empty).
__global__ void __launch_bounds__(256) k_cp_products(StepDesc d, uint32_t d0, uint32_t n, CpScratch c) {
    const uint32_t nl = d.n - d0;
    unsigned long long mx = 0;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        unsigned long long prod = 1;
        bool overflow = false;
        for (uint32_t kk = 0; kk < nl; kk++) {
            const uint32_t sz = c.seg_cnt[static_cast<uint64_t>(r) * nl + kk];
            if (sz == 0) { prod = 0; break; }
            if (prod > (~0ull >> 1) / sz) overflow = true;
            else prod *= sz;
        }
        if (overflow) {
            atomicOr(c.err, kErrProductOverflow);
            prod = 0;
        }
        c.prod[r] = prod;
        mx = max(mx, prod);
    }
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, static_cast<unsigned long long>(__shfl_xor(mx, o, 64)));
    if (lane_id() == 0 && mx) atomicMax(c.max_prod, mx);
}

// Thread per flattened combination t: record r owns [off[r], off[r+1]); digits of t - off[r] in the
// mixed radix of the leaf list sizes select one candidate per leaf; leaves must be distinct.
__global__ void __launch_bounds__(256) k_cp_enumerate(StepDesc d, uint32_t d0, uint32_t n, CpScratch c, MatchOut out) {
    __shared__ unsigned long long sc[256], sh[256];
    const unsigned long long T = c.prod[n];
    const uint32_t nl = d.n - d0;
    unsigned long long cnt = 0, hsum = 0;
    for (unsigned long long t = blockIdx.x * static_cast<unsigned long long>(blockDim.x) + threadIdx.x; t < T;
         t += static_cast<unsigned long long>(gridDim.x) * blockDim.x) {
        // last record with off[r] <= t
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (c.prod[mid] <= t) lo = mid; else hi = mid;
        }
        const uint32_t r = lo;
        unsigned long long local = t - c.prod[r];
        uint32_t x[kMaxQV];
        unsigned long long h = c.rec_hash[r];
        bool ok = true;
        for (uint32_t k = 0; k < nl; k++) {
            const uint64_t i = static_cast<uint64_t>(r) * nl + k;
            const uint32_t sz = c.seg_cnt[i];
            const uint32_t idx = static_cast<uint32_t>(local % sz);
            local /= sz;
            x[k] = c.lists[c.seg_off[i] + idx];
            for (uint32_t j = 0; j < k; j++) ok = ok && (!(d.inj[d0 + k] >> (d0 + j) & 1u) || x[j] != x[k]);
            h += pair_hash(d.order[d0 + k], x[k]);
        }
        if (ok) { cnt += 1; hsum += embedding_hash(h); }
    }
    sc[threadIdx.x] = cnt;
    sh[threadIdx.x] = hsum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sc[threadIdx.x] += sc[threadIdx.x + o]; sh[threadIdx.x] += sh[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0 && sc[0]) { atomicAdd(out.count, sc[0]); atomicAdd(out.hash, sh[0]); }
}

__global__ void __launch_bounds__(256) k_count_records(StepDesc d, const uint32_t* rec, const uint32_t* n_rec, uint32_t max_rec,
                                                       MatchOut out) {
    __shared__ unsigned long long sc[256], sh[256];
    const uint32_t n = min(*n_rec, max_rec);
    unsigned long long cnt = 0, hsum = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long h = 0;
        for (uint32_t j = 0; j < d.n; j++) h += pair_hash(d.order[j], rec[static_cast<uint64_t>(i) * d.n + j]);
        cnt += 1;
        hsum += embedding_hash(h);
    }
    sc[threadIdx.x] = cnt;
    sh[threadIdx.x] = hsum;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sc[threadIdx.x] += sc[threadIdx.x + o]; sh[threadIdx.x] += sh[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0 && sc[0]) { atomicAdd(out.count, sc[0]); atomicAdd(out.hash, sh[0]); }
}

__global__ void __launch_bounds__(256) k_bitmap_to_list(const uint32_t* bitmap, uint32_t dv, uint32_t* list, uint32_t* len) {
    const uint32_t words = (dv + 31) / 32;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += gridDim.x * blockDim.x) {
        uint32_t b = bitmap[w];
        if (!b) continue;
        const uint32_t p = atomicAdd(len, static_cast<uint32_t>(__popc(b)));
        uint32_t k = 0;
        while (b) {
            const uint32_t bit = static_cast<uint32_t>(__ffs(b)) - 1;
            list[p + k++] = w * 32 + bit;
            b &= b - 1;
        }
    }
}

int grid_for(uint64_t n, int per_block, int cap) {
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(static_cast<uint64_t>(cap), (n + per_block - 1) / per_block)));
}

}  // namespace

void init_from_pairs(const uint32_t* pairs, const uint32_t* n_pairs, uint32_t max_pairs, bool swap, uint32_t* out,
                     uint32_t* out_count, hipStream_t s) {
    k_init_from_pairs<<<grid_for(max_pairs, 256, 4096), 256, 0, s>>>(pairs, n_pairs, max_pairs, swap ? 1 : 0, out, out_count);
}

void init_from_set(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot, const uint32_t* list,
                   const uint32_t* len, uint32_t max_len, const uint32_t* filter_b, Ring r, hipStream_t s) {
    const InitJob J{phase, view, dir, slot, list, len, filter_b, r.out, r.out_count, r.overflow, r.out_cap, max_len};
    k_init_from_set<<<grid_for(max_len, 4 * 64, 512), 256, 0, s>>>(g, J);
}

void init_from_set_multi(const DevGraph& g, const InitJob* jobs, uint32_t njobs, hipStream_t s) {
    if (njobs) k_init_from_set_multi<<<dim3(256, njobs), 256, 0, s>>>(g, jobs);
}

void count_records(const StepDesc& d, const uint32_t* rec, const uint32_t* n_rec, uint32_t max_rec, MatchOut out, hipStream_t s) {
    k_count_records<<<grid_for(max_rec, 256, 4096), 256, 0, s>>>(d, rec, n_rec, max_rec, out);
}

void bitmap_to_list(const uint32_t* bitmap, uint32_t dv, uint32_t* list, uint32_t* len, hipStream_t s) {
    k_bitmap_to_list<<<grid_for((dv + 31) / 32, 256, 4096), 256, 0, s>>>(bitmap, dv, list, len);
}

void extend(const DevGraph& g, const StepDesc& d, uint32_t d0, uint32_t end_depth, int mode, Ring r, MatchOut out,
            const unsigned long long* split, unsigned long long* work_counter, int schedule, int blocks, Donation dn, hipStream_t s) {
    // Persistent grid: `blocks` (2048 by default: enough waves to fill 256 CUs several times over),
    // never more than records (a split launch has at least as many items as records with candidates;
    // a donating launch keeps the full grid, its waves share the heavy records' ranges).
    blocks = std::min<int>(blocks, static_cast<int>(kMaxDfsBlocks));
    const int grid = split || schedule == 2 ? blocks : grid_for(r.in_max, kWaves, blocks);
    const uint32_t rows = end_depth > d0 + 1 ? end_depth - 1 - d0 : 0u;
    const size_t lds = static_cast<size_t>(kWaves) * wave_lds_bytes(rows);
    if (schedule == 2) fill_u32(reinterpret_cast<uint32_t*>(dn.range), 0u, 2ull * grid * kWaves, s);  // every word: nothing to take
    if (mode == kCountAtEnd)
        k_extend<kCountAtEnd><<<grid, kWaves * 64, lds, s>>>(g, d, d0, end_depth, r, out, split, work_counter, schedule, dn.range, dn.rec);
    else
        k_extend<kWriteAtEnd><<<grid, kWaves * 64, lds, s>>>(g, d, d0, end_depth, r, out, split, work_counter, schedule, dn.range, dn.rec);
}

void level_prepare(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n, const LevelScratch& ls,
                   bool with_hash, uint32_t* zero, unsigned long long* bytes, hipStream_t s) {
    unsigned long long* ph = with_hash ? ls.ph : nullptr;
    if (n < kSmallLevel) {
        k_level_len_small<<<1, kSmallThreads, 0, s>>>(g, d, d0, rec, n, ls.off, ls.prun, ls.runs, ls.piv, ph, zero, bytes);
        return;
    }
    const uint32_t tiles = (n + kLenTile - 1) / kLenTile;
    k_level_len<<<tiles, kLenThreads, 0, s>>>(g, d, d0, rec, n, ls.off, ls.prun, ls.runs, ls.piv, ph, ls.tile_sum, zero, bytes);
    k_level_scan<<<tiles, kLenThreads, 0, s>>>(ls.off, n, ls.tile_sum);
}

void level_flat(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n, int mode, const LevelScratch& ls,
                Ring r, MatchOut out, hipStream_t s) {
    // Waves take equal contiguous shares of the level's candidates (their total is known on the
    // device only): the grid follows the record count, so small levels do not dispatch thousands
    // of idle waves.
    const int grid = static_cast<int>(std::min<uint32_t>(2048, std::max<uint32_t>(16, n / 2)));
    if (mode == kCountAtEnd)
        k_level_flat<kCountAtEnd><<<grid, 256, 0, s>>>(g, d, d0, rec, n, ls.off, ls.prun, ls.runs, ls.piv, ls.ph, r, out);
    else
        k_level_flat<kWriteAtEnd><<<grid, 256, 0, s>>>(g, d, d0, rec, n, ls.off, ls.prun, ls.runs, ls.piv, ls.ph, r, out);
}

void level_prepare_chain(const DevGraph& g, const StepDesc& d, const LevelChain& c, uint32_t k, const LevelScratch& ls, bool with_hash,
                         unsigned long long* bytes, hipStream_t s) {
    k_level_len_chain<<<kChainLenGrid, kLenThreads, 0, s>>>(g, d, c, k, ls.off, ls.prun, ls.runs, ls.piv, with_hash ? ls.ph : nullptr,
                                                           ls.tile_sum, bytes);
    k_level_scan_chain<<<kChainLenGrid, kLenThreads, 0, s>>>(c, k, ls.off, ls.tile_sum);
}

void level_flat_chain(const DevGraph& g, const StepDesc& d, const LevelChain& c, uint32_t k, int mode, const LevelScratch& ls,
                      MatchOut out, hipStream_t s) {
    // Level sizes are device-side, so the grid is fixed: four four-wave blocks per CU, short of a
    // full residency (six per CU at four candidates per lane) so the other lanes' kernels keep
    // finding room (measured on a 400K-vertex Chung-Lu graph with 65,536-edge batches: 1024 blocks
    // beat 768 and 1536 at four candidates per lane, as 512-768 beat 256, 1024 and 2048 at eight)
    if (mode == kCountAtEnd)
        k_level_flat_chain<kCountAtEnd><<<kChainFlatGrid, 256, 0, s>>>(g, d, c, k, ls.off, ls.prun, ls.runs, ls.piv, ls.ph, out);
    else
        k_level_flat_chain<kWriteAtEnd><<<kChainFlatGrid, 256, 0, s>>>(g, d, c, k, ls.off, ls.prun, ls.runs, ls.piv, ls.ph, out);
}

void level_prepare_chain_multi(const DevGraph& g, const ChainJob* jobs, uint32_t njobs, uint32_t k, int blocks, hipStream_t s) {
    if (!njobs) return;
    const dim3 grid(static_cast<uint32_t>(std::max(16, blocks / 2)), njobs);
    k_level_len_chain_multi<<<grid, kLenThreads, 0, s>>>(g, jobs, k);
    k_level_scan_chain_multi<<<grid, kLenThreads, 0, s>>>(jobs, k);
}

void level_flat_chain_multi(const DevGraph& g, const ChainJob* jobs, uint32_t njobs, uint32_t k, int mode, int blocks, hipStream_t s) {
    if (!njobs) return;
    const dim3 grid(static_cast<uint32_t>(std::max(16, blocks)), njobs);
    if (mode == kCountAtEnd) k_level_flat_chain_multi<kCountAtEnd><<<grid, 256, 0, s>>>(g, jobs, k);
    else k_level_flat_chain_multi<kWriteAtEnd><<<grid, 256, 0, s>>>(g, jobs, k);
}

uint32_t level_max_records(const StepDesc& d, uint32_t d0, const LevelScratch& ls) {
    const uint32_t ne = d.boff[d0 + 1] - d.boff[d0];
    return static_cast<uint32_t>(std::min<uint64_t>(ls.cap, ne > 1 ? ls.runs_cap / ne : ls.cap));
}

void root_chunks(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, const uint32_t* n_rec, uint32_t max_rec,
                 uint32_t max_chunks, unsigned long long* chunks, unsigned long long* scan_tmp, hipStream_t s) {
    k_root_chunks<<<grid_for(max_rec, 256, 4096), 256, 0, s>>>(g, d, d0, rec, n_rec, max_rec, max_chunks, chunks);
    scan_exclusive<unsigned long long>(chunks, chunks, n_rec, max_rec, true, scan_tmp, s);
}

void cp_prepare(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n, CpScratch c, hipStream_t s) {
    const uint32_t nseg = n * (d.n - d0);
    k_cp_len<<<grid_for(nseg, 256, 4096), 256, 0, s>>>(g, d, d0, rec, n, c);
    scan_exclusive<unsigned long long>(c.seg_off, c.seg_off, nullptr, nseg, true, c.scan_tmp, s);
    k_cp_filter<<<static_cast<int>(std::min<uint32_t>(2048, std::max<uint32_t>(16, nseg / 2))), 256, 0, s>>>(g, d, d0, rec, nseg, c);
    k_cp_products<<<grid_for(n, 256, 4096), 256, 0, s>>>(d, d0, n, c);
    scan_exclusive<unsigned long long>(c.prod, c.prod, nullptr, n, true, c.scan_tmp, s);
}

void cp_enumerate(const StepDesc& d, uint32_t d0, uint32_t n, CpScratch c, MatchOut out, hipStream_t s) {
    k_cp_enumerate<<<static_cast<int>(std::min<uint32_t>(4096, std::max<uint32_t>(64, n))), 256, 0, s>>>(d, d0, n, c, out);
}

}  // namespace csm
