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
                                const uint32_t* cand1, const uint32_t* by_label, const uint32_t* label_off, uint32_t* pairs,
                                uint32_t* n_pairs, uint32_t max_pairs, VSet L0, VSet L1) {
    if (vr0 == kNoVertex) return;
    const uint32_t first = label_off[vr0], T = label_off[vr0 + 1] - first;
    const uint32_t* touched = by_label + first;
    const uint32_t* dl = phase_delta(g, dir, phase);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = touched[i];
        if (!bit_test(cand0, v)) continue;
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

// Work split of the local-index kernels.  The frontier (or queue) size is only known on the device,
// so each launch picks it there: G items per wave, the smallest power of two that covers the items
// with the grid's waves (a whole wave per item for small frontiers, up to one lane per item for
// large ones).  The runs of a wave's G items are then flattened: a wave-wide prefix of their lengths
// lets each lane take one element of one run per round (64 elements a round, whatever the mix of
// short runs and hub runs), finding its item by a binary search of the prefix in LDS.
struct FlatLds {
    uint32_t pre[4][65];   // exclusive prefix of the items' run lengths, [64] = total
    RunView run[4][64];
    uint32_t nb[4][64];    // base-run length of each item's run
    uint32_t job[4][64];   // job of each item
    uint32_t flag[4][64];  // check: the item's run holds a member of the filter bitmap
    const uint32_t* bm[4][64];    // expansion: the job's output bitmap; check: the filter's bitmap
    const uint32_t* cand[4][64];  // expansion: the job's candidate bitmap
    uint32_t vd[4][64];    // view | dir << 8 | phase << 16 of the item's relation
};

// Run elements a lane tests per round in the flattened loops: their loads are issued together, so
// a wave holding a long (hub) run makes four times the progress per round trip.
constexpr uint32_t kFlatIlp = 4;

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t x) {
    for (int o = 32; o > 0; o >>= 1) x = max(x, static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), o, 64)));
    return x;
}


struct Rel4 {
    uint32_t phase, view, dir, slot;
};

// Expands `total` frontier items: item(i, j, x) gives item i's job and frontier vertex, rel(j) the
// job's expansion relation, target(j, cand, bitmap) its candidate bitmap and output bitmap.  Every
// element y of an item's run that is a candidate and live in the view is offered to the output
// bitmap (atomic test-and-set); emit(fresh, y, j) is called by the whole wave once per element
// group.  Returns this lane's algorithmic bytes.
template <class Item, class Rel, class Target, class Emit>
__device__ __forceinline__ uint64_t expand_flat(const DevGraph& g, FlatLds& lds, uint32_t total, Item item, Rel rel, Target target,
                                                Emit emit) {
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wv, waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t G = items_per_wave(total);
    uint64_t bytes = 0;
    for (uint32_t base = gw * G; base < total; base += waves * G) {
        const uint32_t i = base + lane;
        uint32_t tot = 0, j = 0, nb = 0, vd = 0;
        const uint32_t* cand = nullptr;
        uint32_t* bm = nullptr;
        RunView r{0, 0, 0, 0};
        if (lane < G && i < total) {
            uint32_t x;
            item(i, j, x);
            const Rel4 R = rel(j);
            r = make_run(g, R.phase, R.view, R.dir, x, R.slot);
            nb = run_base_len(r, R.view);
            tot = run_seq_len(r, R.view);
            vd = R.view | R.dir << 8 | R.phase << 16;
            target(j, cand, bm);
            bytes += 12 + 4ull * tot;
        }
        uint32_t T;
        const uint32_t pre = wave_prefix(tot, T);
        lds.pre[wv][lane] = pre;
        lds.run[wv][lane] = r;
        lds.nb[wv][lane] = nb;
        lds.job[wv][lane] = j;
        lds.vd[wv][lane] = vd;
        lds.cand[wv][lane] = cand;
        lds.bm[wv][lane] = bm;
        if (lane == 63) lds.pre[wv][64] = T;
        wave_sync();
        for (uint32_t e0 = 0; e0 < T; e0 += 64 * kFlatIlp) {  // wave-uniform trip count
            uint32_t y[kFlatIlp], jj[kFlatIlp], key[kFlatIlp], kk[kFlatIlp];
            bool ok[kFlatIlp];
#pragma unroll
            for (uint32_t u = 0; u < kFlatIlp; u++) {  // the run elements: independent loads
                const uint32_t e = e0 + 64 * u + lane;
                ok[u] = e < T;
                kk[u] = ok[u] ? owner_of(lds.pre[wv], e) : 0u;
                jj[u] = lds.job[wv][kk[u]];
                const uint32_t v = lds.vd[wv][kk[u]];
                key[u] = ok[u] ? run_elem(g, phase_delta(g, (v >> 8) & 0xFFu, v >> 16), lds.run[wv][kk[u]], lds.nb[wv][kk[u]],
                                          e - lds.pre[wv][kk[u]])
                               : 0u;
                y[u] = key[u] & g.id_mask;
            }
#pragma unroll
            for (uint32_t u = 0; u < kFlatIlp; u++) ok[u] = ok[u] && bit_test(lds.cand[wv][kk[u]], y[u]);
#pragma unroll
            for (uint32_t u = 0; u < kFlatIlp; u++) {
                const uint32_t v = lds.vd[wv][kk[u]];
                ok[u] = ok[u] && run_elem_live(phase_delta(g, (v >> 8) & 0xFFu, v >> 16), lds.run[wv][kk[u]], v & 0xFFu, key[u]);
            }
            bool fresh[kFlatIlp];
#pragma unroll
            for (uint32_t u = 0; u < kFlatIlp; u++) {
                const uint32_t bit = 1u << (y[u] & 31);
                // else offered by another visitor
                fresh[u] = ok[u] && !(atomicOr(const_cast<uint32_t*>(lds.bm[wv][kk[u]]) + (y[u] >> 5), bit) & bit);
            }
#pragma unroll
            for (uint32_t u = 0; u < kFlatIlp; u++) {
                if (e0 + 64 * u >= T) break;  // wave-uniform
                emit(fresh[u], y[u], jj[u]);
                bytes += fresh[u] ? 4u : 0u;
            }
        }
        wave_sync();  // the LDS rows are rewritten by the next items
    }
    return bytes;
}

// Keeps the `total` queued items whose run of every filter relation holds a member of the
// filter's bitmap: item(i, j, y, nf) gives item i's job, vertex and number of filters,
// filter(j, k) job j's filter k, phase_of(j) its phase; keep(pass, y, j) is called by the whole wave
// once per item group (pass only on the item's first lane).  The filters' runs are flattened like
// the expansion's, one filter at a time; an item stops being scanned at its first hit.
template <class Item, class Filter, class Phase, class Keep>
__device__ __forceinline__ uint64_t check_flat(const DevGraph& g, FlatLds& lds, uint32_t total, Item item, Filter filter,
                                               Phase phase_of, Keep keep) {
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    const uint32_t gw = blockIdx.x * (blockDim.x >> 6) + wv, waves = gridDim.x * (blockDim.x >> 6);
    const uint32_t G = items_per_wave(total);
    uint64_t bytes = 0;
    for (uint32_t base = gw * G; base < total; base += waves * G) {
        const uint32_t i = base + lane;
        const bool active = lane < G && i < total;
        uint32_t j = 0, y = 0, nf = 0;
        if (active) item(i, j, y, nf);
        bool pass = active;
        const uint32_t maxf = wave_max_u32(nf);
        for (uint32_t k = 0; k < maxf; k++) {
            const bool mine = pass && k < nf;
            uint32_t tot = 0, nb = 0, vd = 0;
            RunView r{0, 0, 0, 0};
            LocalFilter f{0, 0, 0, nullptr};
            if (mine) {
                f = filter(j, k);
                const uint32_t ph = phase_of(j);
                r = make_run(g, ph, f.view, f.dir, y, f.slot);
                nb = run_base_len(r, f.view);
                tot = run_seq_len(r, f.view);
                vd = f.view | f.dir << 8 | ph << 16;
                bytes += 8;
            }
            uint32_t T;
            const uint32_t pre = wave_prefix(tot, T);
            lds.pre[wv][lane] = pre;
            lds.run[wv][lane] = r;
            lds.nb[wv][lane] = nb;
            lds.bm[wv][lane] = f.bitmap;
            lds.vd[wv][lane] = vd;
            lds.flag[wv][lane] = 0;
            if (lane == 63) lds.pre[wv][64] = T;
            wave_sync();
            for (uint32_t e0 = 0; e0 < T; e0 += 64 * kFlatIlp) {  // wave-uniform trip count
                uint32_t o[kFlatIlp], fk[kFlatIlp];
                bool act[kFlatIlp];
#pragma unroll
                for (uint32_t u = 0; u < kFlatIlp; u++) {  // the run elements: independent loads
                    const uint32_t e = e0 + 64 * u + lane;
                    o[u] = e < T ? owner_of(lds.pre[wv], e) : 0u;
                    act[u] = e < T && !lds.flag[wv][o[u]];  // an item is done at its first hit
                    const uint32_t v = lds.vd[wv][o[u]];
                    fk[u] = act[u] ? run_elem(g, phase_delta(g, (v >> 8) & 0xFFu, v >> 16), lds.run[wv][o[u]], lds.nb[wv][o[u]],
                                              e - lds.pre[wv][o[u]])
                                   : 0u;
                    bytes += act[u] ? 4u : 0u;
                }
                bool hit[kFlatIlp];
#pragma unroll
                for (uint32_t u = 0; u < kFlatIlp; u++) hit[u] = act[u] && bit_test(lds.bm[wv][o[u]], fk[u] & g.id_mask);
#pragma unroll
                for (uint32_t u = 0; u < kFlatIlp; u++) {
                    const uint32_t v = lds.vd[wv][o[u]];
                    if (hit[u] && run_elem_live(phase_delta(g, (v >> 8) & 0xFFu, v >> 16), lds.run[wv][o[u]], v & 0xFFu, fk[u]))
                        lds.flag[wv][o[u]] = 1u;
                }
                wave_sync();
            }
            if (mine) pass = lds.flag[wv][lane] != 0;
            wave_sync();
        }
        keep(active, pass, y, j);
        if (active) bytes += pass ? 8 : 4;
    }
    return bytes;
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

// Adds the wave's byte counts to *out (once per wave).
__device__ __forceinline__ void add_bytes(unsigned long long* out, uint64_t bytes) {
    if (!out) return;
    bytes = wave_sum_u64(bytes);
    if (lane_id() == 0 && bytes) atomicAdd(out, static_cast<unsigned long long>(bytes));
}

__global__ void __launch_bounds__(256) k_local_expand(DevGraph g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                                                      const uint32_t* frontier, const uint32_t* frontier_len, uint32_t max_frontier,
                                                      const uint32_t* cand_u, VSet out, unsigned long long* nbytes) {
    __shared__ FlatLds lds;
    const uint32_t F = min(*frontier_len, max_frontier);
    const uint64_t bytes = expand_flat(
        g, lds, F, [&](uint32_t i, uint32_t& j, uint32_t& x) { j = 0; x = frontier[i]; },
        [&](uint32_t) { return Rel4{phase, view, dir, slot}; },
        [&](uint32_t, const uint32_t*& c, uint32_t*& bm) { c = cand_u; bm = out.bitmap; },
        [&](bool fresh, uint32_t y, uint32_t) { append_fresh(fresh, y, out.list, out.len, out.cap); });
    add_bytes(nbytes, bytes);
}

// Expansion with the semi-joins in a second pass: k_local_queue expands the frontier along
// (view, dir, slot) into cand_u and queues every new candidate y once (its first visitor sets its
// bit in out.bitmap); k_local_check keeps the queued y that have a neighbour in every filter's
// bitmap.  Survivors join out.list; the bits of the others are cleared.
__global__ void __launch_bounds__(256) k_local_queue(DevGraph g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                                                     const uint32_t* frontier, const uint32_t* frontier_len, uint32_t max_frontier,
                                                     const uint32_t* cand_u, uint32_t* bitmap, uint32_t* queue, uint32_t* queue_len,
                                                     unsigned long long* nbytes) {
    __shared__ FlatLds lds;
    const uint32_t F = min(*frontier_len, max_frontier);
    const uint64_t bytes = expand_flat(
        g, lds, F, [&](uint32_t i, uint32_t& j, uint32_t& x) { j = 0; x = frontier[i]; },
        [&](uint32_t) { return Rel4{phase, view, dir, slot}; },
        [&](uint32_t, const uint32_t*& c, uint32_t*& bm) { c = cand_u; bm = bitmap; },
        [&](bool fresh, uint32_t y, uint32_t) { append_fresh(fresh, y, queue, queue_len, 0xFFFFFFFFu); });
    add_bytes(nbytes, bytes);
}

__global__ void __launch_bounds__(256) k_local_check(DevGraph g, uint32_t phase, const uint32_t* queue, const uint32_t* queue_len,
                                                     uint32_t max_len, LocalFilters fl, VSet out, unsigned long long* nbytes) {
    __shared__ FlatLds lds;
    const uint32_t n = min(*queue_len, max_len);
    const uint64_t bytes = check_flat(
        g, lds, n, [&](uint32_t i, uint32_t& j, uint32_t& y, uint32_t& nf) { j = 0; y = queue[i]; nf = fl.n; },
        [&](uint32_t, uint32_t k) { return fl.f[k]; }, [&](uint32_t) { return phase; },
        [&](bool active, bool pass, uint32_t y, uint32_t) {
            append_fresh(active && pass, y, out.list, out.len, out.cap);
            if (active && !pass) atomicAnd(&out.bitmap[y >> 5], ~(1u << (y & 31)));
        });
    add_bytes(nbytes, bytes);
}

// ---- many steps at once --------------------------------------------------------------------

// One block row per job; the body is k_local_anchors'.
__global__ void k_local_anchors_multi(DevGraph g, const AnchorJob* jobs, const uint32_t* by_label0, const uint32_t* label_off0,
                                      const uint32_t* by_label1, const uint32_t* label_off1) {
    const AnchorJob J = jobs[blockIdx.y];
    if (J.vr0 == kNoVertex) return;
    const uint32_t* label_off = J.dir ? label_off1 : label_off0;
    const uint32_t first = label_off[J.vr0], T = label_off[J.vr0 + 1] - first;
    const uint32_t* touched = (J.dir ? by_label1 : by_label0) + first;
    const uint32_t* dl = phase_delta(g, J.dir, J.phase);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = touched[i];
        if (!bit_test(J.cand0, v)) continue;
        const uint4 sg = g.delta[J.dir].seg[v];
        const uint32_t off = J.phase ? sg.z : sg.x, len = J.phase ? sg.w : sg.y;
        if (!len) continue;
        uint32_t lo, hi;
        slot_range(dl, off, off + len, J.slot, g.id_bits, lo, hi);
        bool any = false;
        for (uint32_t k = lo; k < hi; k++) {
            const uint32_t x = dl[k] & g.id_mask;
            if (!bit_test(J.cand1, x)) continue;
            const uint32_t p = atomicAdd(J.n_pairs, 1u);
            if (p < J.max_pairs) { J.pairs[2 * p] = v; J.pairs[2 * p + 1] = x; }
            vset_add(J.L1, x);
            any = true;
        }
        if (any) vset_add(J.L0, v);
    }
}

// Every block of a multi-job launch computes off[0..njobs] in LDS, the exclusive prefix of the jobs'
// frontier (kind 0) or queue (kind 1) lengths: the lengths are independent loads, and wave 0 scans
// them 64 at a time.
template <int kKind>
__device__ __forceinline__ void job_offsets(const LocalJob* jobs, uint32_t njobs, uint32_t max_len, uint32_t* off) {
    for (uint32_t t = threadIdx.x; t < njobs; t += blockDim.x)
        off[t + 1] = min(kKind == 0 ? *jobs[t].frontier_len : *jobs[t].queue_len, max_len);
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t carry = 0;
        for (uint32_t c = 0; c < njobs; c += 64) {
            const uint32_t t = c + threadIdx.x;
            uint32_t total;
            const uint32_t x = t < njobs ? off[t + 1] : 0u;
            const uint32_t ex = wave_prefix(x, total);
            if (t < njobs) off[t + 1] = carry + ex + x;
            carry += total;
        }
        if (threadIdx.x == 0) off[0] = 0;
    }
    __syncthreads();
}

// The job whose items [off[j], off[j + 1]) hold item i (off in LDS, njobs >= 1).
__device__ __forceinline__ uint32_t job_of(const uint32_t* off, uint32_t njobs, uint32_t i) {
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Appends the flagged lanes' vertices to their jobs' queues (to_queue) or output lists: one
// reservation when every flagged lane belongs to the same job, else one atomic per lane.
__device__ __forceinline__ void append_jobs(bool flag, uint32_t y, uint32_t j, const LocalJob* jobs, bool to_queue) {
    const uint64_t m = __ballot(flag);
    if (!m) return;
    const uint32_t first = static_cast<uint32_t>(__ffsll(static_cast<long long>(m)) - 1);
    const uint32_t j0 = static_cast<uint32_t>(__shfl(static_cast<int>(j), first, 64));
    if (__ballot(flag && j != j0) == 0) {
        const LocalJob& J = jobs[j0];
        uint32_t* list = to_queue ? J.queue : J.out.list;
        uint32_t* len = to_queue ? J.queue_len : J.out.len;
        const uint32_t cap = to_queue ? 0xFFFFFFFFu : J.out.cap;
        uint32_t p = 0;
        if (lane_id() == first) p = atomicAdd(len, static_cast<uint32_t>(__popcll(m)));
        p = static_cast<uint32_t>(__shfl(static_cast<int>(p), first, 64)) + lanes_below(m);
        if (flag && p < cap) list[p] = y;
    } else if (flag) {
        const LocalJob& J = jobs[j];
        if (to_queue) {
            J.queue[atomicAdd(J.queue_len, 1u)] = y;
        } else {
            const uint32_t p = atomicAdd(J.out.len, 1u);
            if (p < J.out.cap) J.out.list[p] = y;
        }
    }
}

// Launch bounds of the two multi-job local-index kernels: eight and seven waves per SIMD, no spills
// (64 and 72 VGPRs against 66 and 76 unbounded); more resident waves hide more of their dependent
// loads when the teams' launches share the GPU, which measured faster.
__global__ void __launch_bounds__(256, 8) k_local_queue_multi(DevGraph g, const LocalJob* jobs, uint32_t njobs, unsigned long long* nbytes) {
    __shared__ FlatLds lds;
    __shared__ uint32_t s_off[kMaxLocalJobs + 1];
    job_offsets<0>(jobs, njobs, g.dv, s_off);
    const uint64_t bytes = expand_flat(
        g, lds, s_off[njobs],
        [&](uint32_t i, uint32_t& j, uint32_t& x) {
            j = job_of(s_off, njobs, i);
            x = jobs[j].frontier[i - s_off[j]];
        },
        [&](uint32_t j) { return Rel4{jobs[j].phase, jobs[j].view, jobs[j].dir, jobs[j].slot}; },
        [&](uint32_t j, const uint32_t*& c, uint32_t*& bm) { c = jobs[j].cand_u; bm = jobs[j].out.bitmap; },
        [&](bool fresh, uint32_t y, uint32_t j) { append_jobs(fresh, y, j, jobs, true); });
    add_bytes(nbytes, bytes);
}

__global__ void __launch_bounds__(256, 7) k_local_check_multi(DevGraph g, const LocalJob* jobs, uint32_t njobs, unsigned long long* nbytes) {
    __shared__ FlatLds lds;
    __shared__ uint32_t s_off[kMaxLocalJobs + 1];
    job_offsets<1>(jobs, njobs, g.dv, s_off);
    const uint64_t bytes = check_flat(
        g, lds, s_off[njobs],
        [&](uint32_t i, uint32_t& j, uint32_t& y, uint32_t& nf) {
            j = job_of(s_off, njobs, i);
            y = jobs[j].queue[i - s_off[j]];
            nf = jobs[j].fl.n;
        },
        [&](uint32_t j, uint32_t k) { return jobs[j].fl.f[k]; }, [&](uint32_t j) { return jobs[j].phase; },
        [&](bool active, bool pass, uint32_t y, uint32_t j) {
            append_jobs(active && pass, y, j, jobs, false);
            if (active && !pass) atomicAnd(&jobs[j].out.bitmap[y >> 5], ~(1u << (y & 31)));
        });
    add_bytes(nbytes, bytes);
}

__global__ void __launch_bounds__(256) k_clear_sets_multi(const VSet* sets, uint32_t max_len) {
    const VSet v = sets[blockIdx.y];
    const uint32_t n = min(*v.len, max_len);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v.bitmap[v.list[i] >> 5] = 0u;
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
                   const uint32_t* cand1, const uint32_t* by_label, const uint32_t* label_off, uint32_t max_touched, uint32_t* pairs,
                   uint32_t* n_pairs, uint32_t max_pairs, VSet L0, VSet L1, hipStream_t s) {
    k_local_anchors<<<grid_for(max_touched, 256, 1024), 256, 0, s>>>(g, phase, dir, slot, vr0, cand0, cand1, by_label, label_off,
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

void local_anchors_multi(const DevGraph& g, const AnchorJob* jobs, uint32_t njobs, const uint32_t* by_label0, const uint32_t* label_off0,
                         const uint32_t* by_label1, const uint32_t* label_off1, uint32_t max_touched, hipStream_t s) {
    if (!njobs) return;
    const dim3 grid(static_cast<uint32_t>(grid_for(max_touched, 256, static_cast<int>(std::max<uint32_t>(16, 4096 / njobs)))), njobs);
    k_local_anchors_multi<<<grid, 256, 0, s>>>(g, jobs, by_label0, label_off0, by_label1, label_off1);
}

void local_queue_multi(const DevGraph& g, const LocalJob* jobs, uint32_t njobs, int blocks, unsigned long long* bytes, hipStream_t s) {
    if (!njobs) return;
    k_local_queue_multi<<<blocks, 256, 0, s>>>(g, jobs, njobs, bytes);
}

void local_check_multi(const DevGraph& g, const LocalJob* jobs, uint32_t njobs, int blocks, unsigned long long* bytes, hipStream_t s) {
    if (!njobs) return;
    k_local_check_multi<<<blocks, 256, 0, s>>>(g, jobs, njobs, bytes);
}

void clear_sets_multi(const VSet* sets, uint32_t nsets, uint32_t max_len, hipStream_t s) {
    if (nsets) k_clear_sets_multi<<<dim3(64, nsets), 256, 0, s>>>(sets, max_len);
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
