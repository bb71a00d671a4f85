// Global candidate bitmaps (NLF), relation summary and the per-step local index.
//
// Global index.  The reference keeps, per query relation, a materialised copy of the data arcs whose
// endpoints are NLF candidates (UpdateGlobalIndex, match_gpu.cu:611-834).  Here the data graph is
// shared by all queries and the global index is only the candidate bitmap of every query vertex:
// a relation's global-index arcs are the relation run of the adjacency filtered by the two bitmaps.
// The NLF test is the reference's (indexing.cu:303-338): for every neighbour group (edge label,
// neighbour label) of u, v needs at least as many neighbours in the group as u has.
//
// Local index.  For batch step i the reference rebuilds candidate bitmaps along the indexing order
// (BuildLocalIndex_v2, match_gpu.cu:1098-1343, kernels indexing.cu:645-1436): the anchors give
// L(u0), L(u1); a later vertex u keeps the candidates that have a neighbour in L(b) for every
// earlier neighbour b.  Here it is computed sparsely: expand from one earlier neighbour's list and
// semi-join each new candidate against the other earlier neighbours' bitmaps in the same pass.
#include "kernels.h"
#include "view.h"

namespace csm {
namespace {

__device__ __forceinline__ bool nlf_pass(const DevGraph& g, const NlfEntry& e, const NlfGroup* groups, uint32_t v, int mode) {
    for (uint32_t k = e.g0; k < e.g1; k++) {
        const NlfGroup gr = groups[k];
        uint32_t lo, hi;
        base_range(g, gr.dir, v, gr.slot, lo, hi);
        uint32_t deg = hi - lo;
        if (mode == 0) {
            const uint4 s = g.delta[gr.dir].seg[v];
            if (s.y) {
                uint32_t dlo, dhi;
                slot_range(g.delta[gr.dir].ins, s.x, s.x + s.y, gr.slot, g.id_bits, dlo, dhi);
                deg += dhi - dlo;
            }
        }
        if (deg < gr.count) return false;
    }
    return true;
}

// One wave per 64 consecutive vertices: every bitmap word is rewritten.
__global__ void __launch_bounds__(256) k_nlf_full(DevGraph g, const NlfEntry* entries, uint32_t n_entries, const NlfGroup* groups) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t nw = (g.dv + 63) / 64;
    const uint32_t words = (g.dv + 31) / 32;
    for (uint32_t w = blockIdx.x * wpb + (threadIdx.x >> 6); w < nw; w += gridDim.x * wpb) {
        const uint32_t v = w * 64 + lane;
        const uint32_t vr = v < g.dv ? g.vrank[v] : kNoVertex;
        for (uint32_t e = 0; e < n_entries; e++) {
            const NlfEntry en = entries[e];
            const bool pass = v < g.dv && vr == en.vrank && en.vrank != kNoVertex && nlf_pass(g, en, groups, v, 1);
            const uint64_t m = __ballot(pass);
            if (lane == 0) en.bitmap[2 * w] = static_cast<uint32_t>(m);
            if (lane == 1 && 2 * w + 1 < words) en.bitmap[2 * w + 1] = static_cast<uint32_t>(m >> 32);
        }
    }
}

__global__ void k_nlf_touched(DevGraph g, const NlfEntry* entries, uint32_t n_entries, const NlfGroup* groups,
                              const uint32_t* touched, const uint32_t* n_touched, int mode) {
    const uint32_t T = *n_touched;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = touched[i];
        const uint32_t vr = g.vrank[v];
        for (uint32_t e = 0; e < n_entries; e++) {
            const NlfEntry en = entries[e];
            if (en.vrank != vr || vr == kNoVertex) continue;
            const bool pass = nlf_pass(g, en, groups, v, mode);
            const uint32_t bit = 1u << (v & 31);
            if (pass) atomicOr(&en.bitmap[v >> 5], bit);
            else if (mode == 1) atomicAnd(&en.bitmap[v >> 5], ~bit);
        }
    }
}

__global__ void __launch_bounds__(256) k_relation_summary(DevGraph g, uint32_t dir, uint32_t slot, uint32_t vrank_u,
                                                          const uint32_t* cand_u, const uint32_t* cand_w, unsigned long long* out) {
    __shared__ unsigned long long sc[256], ss[256];
    unsigned long long card = 0, srcs = 0;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < g.dv; v += gridDim.x * blockDim.x) {
        if (g.vrank[v] != vrank_u || !bit_test(cand_u, v)) continue;
        uint32_t lo, hi;
        base_range(g, dir, v, slot, lo, hi);
        uint32_t c = 0;
        for (uint32_t k = lo; k < hi; k++) c += bit_test(cand_w, g.arena[k] & g.id_mask) ? 1u : 0u;
        card += c;
        srcs += c ? 1u : 0u;
    }
    sc[threadIdx.x] = card;
    ss[threadIdx.x] = srcs;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) { sc[threadIdx.x] += sc[threadIdx.x + o]; ss[threadIdx.x] += ss[threadIdx.x + o]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { atomicAdd(&out[0], sc[0]); atomicAdd(&out[1], ss[0]); }
}

__device__ __forceinline__ void vset_add(VSet s, uint32_t v) {
    const uint32_t bit = 1u << (v & 31);
    const uint32_t old = atomicOr(&s.bitmap[v >> 5], bit);
    if (!(old & bit)) {
        const uint32_t p = atomicAdd(s.len, 1u);
        if (p < s.cap) s.list[p] = v;
    }
}

__global__ void k_local_anchors(DevGraph g, uint32_t phase, uint32_t dir, uint32_t slot, uint32_t vr0, const uint32_t* cand0,
                                const uint32_t* cand1, const uint32_t* touched, const uint32_t* n_touched, uint32_t* pairs,
                                uint32_t* n_pairs, uint32_t max_pairs, VSet L0, VSet L1) {
    const uint32_t T = *n_touched;
    const uint32_t* dl = phase_delta(g, dir, phase);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = touched[i];
        if (g.vrank[v] != vr0 || !bit_test(cand0, v)) continue;
        const uint4 s = g.delta[dir].seg[v];
        const uint32_t off = phase ? s.z : s.x, len = phase ? s.w : s.y;
        if (!len) continue;
        uint32_t lo, hi;
        slot_range(dl, off, off + len, slot, g.id_bits, lo, hi);
        bool any = false;
        for (uint32_t k = lo; k < hi; k++) {
            const uint32_t x = dl[k] & g.id_mask;
            if (!bit_test(cand1, x)) continue;
            const uint32_t p = atomicAdd(n_pairs, 1u);
            if (p < max_pairs) { pairs[2 * p] = v; pairs[2 * p + 1] = x; }
            vset_add(L1, x);
            any = true;
        }
        if (any) vset_add(L0, v);
    }
}

// Work split of the local-index kernels.  The frontier (or queue) size is only known on the
// device, so each launch picks it there: G items per wave, the smallest power of two that covers
// the items with the grid's waves, each item scanned by a sub-group of W = 64 / G lanes.  Small
// frontiers thus get a whole wave per item (short latency), large ones one lane per item (64
// items' loads in flight per wave).  A run longer than kRoundsPerItem * W keys is left to the whole
// wave afterwards, so a hub's run never serialises one sub-group.
constexpr uint32_t kRoundsPerItem = 4;

struct SubWave {
    uint32_t G, W, sub, sl;  // items per wave, lanes per item, this lane's item slot and lane in it
    uint64_t mask;           // lanes of this lane's sub-group
};

__device__ __forceinline__ SubWave sub_wave(uint32_t items) {
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    uint32_t G = 1;
    while (G < 64 && static_cast<uint64_t>(G) * waves < items) G <<= 1;
    SubWave w;
    w.G = G;
    w.W = 64 / G;
    const uint32_t lane = lane_id();
    w.sub = lane / w.W;
    w.sl = lane % w.W;
    w.mask = (w.W == 64 ? ~0ull : ((1ull << w.W) - 1)) << (w.sub * w.W);
    return w;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o, 64)));
    return x;
}

__device__ __forceinline__ RunView shfl_run(const RunView& r, uint32_t src) {
    return RunView{static_cast<uint32_t>(__shfl(static_cast<int>(r.lo), src, 64)), static_cast<uint32_t>(__shfl(static_cast<int>(r.hi), src, 64)),
                   static_cast<uint32_t>(__shfl(static_cast<int>(r.dlo), src, 64)), static_cast<uint32_t>(__shfl(static_cast<int>(r.dhi), src, 64))};
}

// Expands the frontier along (view, dir, slot) into cand_u: every element y of a frontier vertex's
// run that is a candidate of u and live in the view is offered to `bitmap` (atomic test-and-set),
// and `emit(fresh, y)` runs once per round for the whole wave (fresh: this lane set y's bit first).
// Returns this lane's algorithmic bytes (items read, runs located, keys scanned).
template <class Emit>
__device__ __forceinline__ uint64_t expand_frontier(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                                                    const uint32_t* frontier, uint32_t F, const uint32_t* cand_u, uint32_t* bitmap,
                                                    Emit emit) {
    uint64_t bytes = 0;
    const uint32_t lane = lane_id();
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const SubWave sw = sub_wave(F);
    const uint32_t* dl = phase_delta(g, dir, phase);
    auto offer = [&](uint32_t key, const RunView& r, bool& fresh, uint32_t& y) {
        y = key & g.id_mask;
        if (bit_test(cand_u, y) && run_elem_live(dl, r, view, key)) {
            const uint32_t bit = 1u << (y & 31);
            fresh = !(atomicOr(&bitmap[y >> 5], bit) & bit);  // else offered by another visitor
        }
    };
    for (uint32_t base = gw * sw.G; base < F; base += waves * sw.G) {
        const uint32_t i = base + sw.sub;
        RunView r{0, 0, 0, 0};
        uint32_t nb = 0, tot = 0;
        if (i < F) {
            r = make_run(g, phase, view, dir, frontier[i], slot);
            nb = run_base_len(r, view);
            tot = run_seq_len(r, view);
            if (sw.sl == 0) bytes += 12 + 4ull * tot;
        }
        const bool own = tot <= kRoundsPerItem * sw.W;
        const uint32_t rounds = wave_max_u32(own ? (tot + sw.W - 1) / sw.W : 0u);
        for (uint32_t t = 0; t < rounds; t++) {
            bool fresh = false;
            uint32_t y = 0;
            const uint32_t e = t * sw.W + sw.sl;
            if (own && e < tot) offer(run_elem(g, dl, r, nb, e), r, fresh, y);
            emit(fresh, y);
        }
        for (uint64_t longs = __ballot(!own && sw.sl == 0); longs; longs &= longs - 1) {
            const uint32_t src = static_cast<uint32_t>(__ffsll(static_cast<long long>(longs)) - 1);
            const RunView lr = shfl_run(r, src);
            const uint32_t lnb = static_cast<uint32_t>(__shfl(static_cast<int>(nb), src, 64));
            const uint32_t ltot = static_cast<uint32_t>(__shfl(static_cast<int>(tot), src, 64));
            for (uint32_t t0 = 0; t0 < ltot; t0 += 64) {  // wave-uniform trip count
                bool fresh = false;
                uint32_t y = 0;
                if (t0 + lane < ltot) offer(run_elem(g, dl, lr, lnb, t0 + lane), lr, fresh, y);
                emit(fresh, y);
            }
        }
    }
    return bytes;
}

// Adds the wave's byte counts to *out (once per wave).
__device__ __forceinline__ void add_bytes(unsigned long long* out, uint64_t bytes) {
    if (!out) return;
    bytes = wave_sum_u64(bytes);
    if (lane_id() == 0 && bytes) atomicAdd(out, static_cast<unsigned long long>(bytes));
}

// Appends the wave's flagged vertices to list[*len ...] with one reservation (entries past cap are
// dropped; the caller sizes cap for every vertex).
__device__ __forceinline__ void append_fresh(bool fresh, uint32_t y, uint32_t* list, uint32_t* len, uint32_t cap) {
    const uint64_t m = __ballot(fresh);
    if (!m) return;
    uint32_t p = 0;
    if (lane_id() == 0) p = atomicAdd(len, static_cast<uint32_t>(__popcll(m)));
    p = static_cast<uint32_t>(__shfl(static_cast<int>(p), 0, 64)) + lanes_below(m);
    if (fresh && p < cap) list[p] = y;
}

__global__ void __launch_bounds__(256) k_local_expand(DevGraph g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                                                      const uint32_t* frontier, const uint32_t* frontier_len, uint32_t max_frontier,
                                                      const uint32_t* cand_u, VSet out, unsigned long long* nbytes) {
    const uint32_t F = min(*frontier_len, max_frontier);
    uint64_t written = 0;
    const uint64_t bytes = expand_frontier(g, phase, view, dir, slot, frontier, F, cand_u, out.bitmap, [&](bool fresh, uint32_t y) {
        written += fresh ? 4u : 0u;
        append_fresh(fresh, y, out.list, out.len, out.cap);
    });
    add_bytes(nbytes, bytes + written);
}

// Expansion with the semi-joins in a second pass: k_local_queue expands the frontier along
// (view, dir, slot) into cand_u and queues every new candidate y once (its first visitor sets its
// bit in out.bitmap); k_local_check keeps the queued y that have a neighbour in every filter's
// bitmap.  Survivors join out.list; the bits of the others are cleared.
__global__ void __launch_bounds__(256) k_local_queue(DevGraph g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                                                     const uint32_t* frontier, const uint32_t* frontier_len, uint32_t max_frontier,
                                                     const uint32_t* cand_u, uint32_t* bitmap, uint32_t* queue, uint32_t* queue_len,
                                                     unsigned long long* nbytes) {
    const uint32_t F = min(*frontier_len, max_frontier);
    uint64_t written = 0;
    const uint64_t bytes = expand_frontier(g, phase, view, dir, slot, frontier, F, cand_u, bitmap, [&](bool fresh, uint32_t y) {
        written += fresh ? 4u : 0u;
        append_fresh(fresh, y, queue, queue_len, 0xFFFFFFFFu);
    });
    add_bytes(nbytes, bytes + written);
}

// Each queued vertex gets a sub-group (the same split as the expansion), which scans its run of
// every filter relation W keys a round for a member of the filter's bitmap, stopping at the first
// hit; runs longer than kRoundsPerItem * W keys are scanned afterwards by the whole wave.
__global__ void __launch_bounds__(256) k_local_check(DevGraph g, uint32_t phase, const uint32_t* queue, const uint32_t* queue_len,
                                                     uint32_t max_len, LocalFilters fl, VSet out, unsigned long long* nbytes) {
    const uint32_t lane = lane_id();
    uint64_t bytes = 0;  // queued vertex read, runs located, keys scanned, survivor written
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t n = min(*queue_len, max_len);
    const SubWave sw = sub_wave(n);
    for (uint32_t base = gw * sw.G; base < n; base += waves * sw.G) {
        const uint32_t i = base + sw.sub;
        const bool active = i < n;
        const uint32_t y = active ? queue[i] : 0u;
        bool pass = active;
        uint32_t pending = 0;  // filters whose runs are left to the whole wave
        for (uint32_t k = 0; k < fl.n; k++) {
            const LocalFilter f = fl.f[k];
            const uint32_t* fdl = phase_delta(g, f.dir, phase);
            RunView fr{0, 0, 0, 0};
            uint32_t fnb = 0, ftot = 0;
            if (pass) {
                fr = make_run(g, phase, f.view, f.dir, y, f.slot);
                fnb = run_base_len(fr, f.view);
                ftot = run_seq_len(fr, f.view);
                if (sw.sl == 0) bytes += 8;
            }
            const bool own = pass && ftot <= kRoundsPerItem * sw.W;
            if (pass && !own) pending |= 1u << k;
            const uint32_t rounds = wave_max_u32(own ? (ftot + sw.W - 1) / sw.W : 0u);
            bool hit = false;
            for (uint32_t t = 0; t < rounds; t++) {
                const uint32_t e = t * sw.W + sw.sl;
                bool h = false;
                if (own && !hit && e < ftot) {
                    const uint32_t fk = run_elem(g, fdl, fr, fnb, e);
                    h = bit_test(f.bitmap, fk & g.id_mask) && run_elem_live(fdl, fr, f.view, fk);
                    bytes += 4;
                }
                const uint64_t m = __ballot(h);
                hit = hit || (m & sw.mask) != 0;
                if (__ballot(own && !hit && (t + 1) * sw.W < ftot) == 0) break;  // every sub-group decided
            }
            if (own) pass = hit;
        }
        for (uint64_t todo = __ballot(pass && pending && sw.sl == 0); todo; todo &= todo - 1) {
            const uint32_t src = static_cast<uint32_t>(__ffsll(static_cast<long long>(todo)) - 1);
            const uint32_t ys = static_cast<uint32_t>(__shfl(static_cast<int>(y), src, 64));
            const uint32_t pm = static_cast<uint32_t>(__shfl(static_cast<int>(pending), src, 64));
            bool ok = true;
            for (uint32_t k = 0; k < fl.n && ok; k++) {
                if (!(pm >> k & 1u)) continue;
                const LocalFilter f = fl.f[k];
                const uint32_t* fdl = phase_delta(g, f.dir, phase);
                const RunView fr = make_run(g, phase, f.view, f.dir, ys, f.slot);
                const uint32_t fnb = run_base_len(fr, f.view), ftot = run_seq_len(fr, f.view);
                bool hit = false;
                for (uint32_t t0 = 0; t0 < ftot && !hit; t0 += 64) {
                    bool h = false;
                    if (t0 + lane < ftot) {
                        const uint32_t fk = run_elem(g, fdl, fr, fnb, t0 + lane);
                        h = bit_test(f.bitmap, fk & g.id_mask) && run_elem_live(fdl, fr, f.view, fk);
                        bytes += 4;
                    }
                    hit = __ballot(h) != 0;
                }
                ok = hit;
            }
            if ((sw.mask >> src) & 1ull) pass = ok;  // the whole sub-group of the pending vertex
        }
        append_fresh(pass && sw.sl == 0, y, out.list, out.len, out.cap);
        if (active && !pass && sw.sl == 0) atomicAnd(&out.bitmap[y >> 5], ~(1u << (y & 31)));
        if (active && sw.sl == 0) bytes += pass ? 8 : 4;
    }
    add_bytes(nbytes, bytes);
}

// Clears the bitmaps of every set in `mask` (whole words of its listed members).
__global__ void __launch_bounds__(256) k_clear_sets(VSetTable t, uint32_t mask, uint32_t max_len) {
    for (uint32_t u = 0; u < kMaxQV; u++) {
        if (!(mask >> u & 1u)) continue;
        const VSet v = t.s[u];
        const uint32_t n = min(*v.len, max_len);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v.bitmap[v.list[i] >> 5] = 0u;
    }
}

__global__ void k_store_words(uint32_t* p, Words w) {
    if (threadIdx.x < w.n) p[threadIdx.x] = w.v[threadIdx.x];
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x) p[i] = v;
}

__global__ void k_copy_u32(const uint32_t* src, uint32_t* dst, uint64_t n) {
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

int grid_for(uint64_t n, int threads, int cap) {
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(static_cast<uint64_t>(cap), (n + threads - 1) / threads)));
}

}  // namespace

void nlf_full(const DevGraph& g, const NlfEntry* entries, uint32_t n_entries, const NlfGroup* groups, hipStream_t s) {
    const uint32_t nw = (g.dv + 63) / 64;
    k_nlf_full<<<grid_for(nw, 4, 16384), 256, 0, s>>>(g, entries, n_entries, groups);
}

void nlf_touched(const DevGraph& g, const NlfEntry* entries, uint32_t n_entries, const NlfGroup* groups, const uint32_t* touched,
                 const uint32_t* n_touched, uint32_t max_touched, int mode, hipStream_t s) {
    k_nlf_touched<<<grid_for(max_touched, 256, 4096), 256, 0, s>>>(g, entries, n_entries, groups, touched, n_touched, mode);
}

void relation_summary(const DevGraph& g, uint32_t dir, uint32_t slot, uint32_t vrank_u, const uint32_t* cand_u,
                      const uint32_t* cand_w, unsigned long long* out, hipStream_t s) {
    k_relation_summary<<<grid_for(g.dv, 256, 2048), 256, 0, s>>>(g, dir, slot, vrank_u, cand_u, cand_w, out);
}

void local_anchors(const DevGraph& g, uint32_t phase, uint32_t dir, uint32_t slot, uint32_t vr0, const uint32_t* cand0,
                   const uint32_t* cand1, const uint32_t* touched, const uint32_t* n_touched, uint32_t max_touched, uint32_t* pairs,
                   uint32_t* n_pairs, uint32_t max_pairs, VSet L0, VSet L1, hipStream_t s) {
    k_local_anchors<<<grid_for(max_touched, 256, 4096), 256, 0, s>>>(g, phase, dir, slot, vr0, cand0, cand1, touched, n_touched,
                                                                      pairs, n_pairs, max_pairs, L0, L1);
}

void local_expand(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot, const uint32_t* frontier,
                  const uint32_t* frontier_len, uint32_t max_frontier, const uint32_t* cand_u, VSet out, int blocks,
                  unsigned long long* bytes, hipStream_t s) {
    k_local_expand<<<grid_for(max_frontier, 4, blocks), 256, 0, s>>>(g, phase, view, dir, slot, frontier, frontier_len, max_frontier,
                                                                     cand_u, out, bytes);
}

void local_queue(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot, const uint32_t* frontier,
                 const uint32_t* frontier_len, uint32_t max_frontier, const uint32_t* cand_u, VSet out, uint32_t* queue,
                 uint32_t* queue_len, int blocks, unsigned long long* bytes, hipStream_t s) {
    k_local_queue<<<grid_for(max_frontier, 4, blocks), 256, 0, s>>>(g, phase, view, dir, slot, frontier, frontier_len, max_frontier, cand_u,
                                                                out.bitmap, queue, queue_len, bytes);
}

void local_check(const DevGraph& g, uint32_t phase, const LocalFilters& fl, VSet out, const uint32_t* queue,
                 const uint32_t* queue_len, int blocks, unsigned long long* bytes, hipStream_t s) {
    k_local_check<<<blocks, 256, 0, s>>>(g, phase, queue, queue_len, g.dv, fl, out, bytes);
}

void clear_sets(const VSetTable& t, uint32_t mask, uint32_t max_len, hipStream_t s) {
    if (mask) k_clear_sets<<<grid_for(max_len, 256, 256), 256, 0, s>>>(t, mask, max_len);
}

void store_words(uint32_t* p, const Words& w, hipStream_t s) { k_store_words<<<1, 64, 0, s>>>(p, w); }

void copy_u32(const uint32_t* src, uint32_t* dst, uint64_t n, hipStream_t s) {
    k_copy_u32<<<grid_for(n, 256, 8192), 256, 0, s>>>(src, dst, n);
}

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    k_fill_u32<<<grid_for(n, 256, 8192), 256, 0, s>>>(p, v, n);
}

}  // namespace csm
