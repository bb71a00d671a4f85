// Relation views of one vertex's neighbour run during a batch step (device side).
#pragma once
#include "dev_common.h"

namespace csm {

// The neighbours of v along one query relation (label slot, direction) as the current step sees
// them: a base run a[lo, hi) and a delta run dl[dlo, dhi) of the phase's batch arcs.
struct RunView {
    uint32_t lo, hi, dlo, dhi;
};

__device__ __forceinline__ const uint32_t* phase_delta(const DevGraph& g, uint32_t dir, uint32_t phase) {
    return phase ? g.delta[dir].del : g.delta[dir].ins;
}

// The arena run of slot `slot` in v's direction-dir list: two adjacent directory words when the
// slot directory is on, else two binary searches over the whole list.
__device__ __forceinline__ void base_range(const DevGraph& g, uint32_t dir, uint32_t v, uint32_t slot, uint32_t& lo, uint32_t& hi) {
    if (g.sdir[dir]) {
        const uint32_t* row = g.sdir[dir] + static_cast<uint64_t>(v) * (g.nslots + 1);
        lo = row[slot];
        hi = row[slot + 1];
    } else {
        const uint2 a = g.dir[dir].adj[v];
        slot_range(g.arena, a.x, a.x + a.y, slot, g.id_bits, lo, hi);
    }
}

__device__ __forceinline__ RunView make_run(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t v, uint32_t slot) {
    RunView r{0, 0, 0, 0};
    if (view != kViewDelta) base_range(g, dir, v, slot, r.lo, r.hi);
    if (view != kViewBase) {
        const uint4 s = g.delta[dir].seg[v];
        const uint32_t off = phase ? s.z : s.x;
        const uint32_t len = phase ? s.w : s.y;
        if (len) slot_range(phase_delta(g, dir, phase), off, off + len, slot, g.id_bits, r.dlo, r.dhi);
    }
    return r;
}

// Number of elements in the iteration sequence of a view (BaseMinus iterates the base run and
// skips deleted keys).
__device__ __forceinline__ uint32_t run_base_len(const RunView& r, uint32_t view) { return view == kViewDelta ? 0u : r.hi - r.lo; }
__device__ __forceinline__ uint32_t run_seq_len(const RunView& r, uint32_t view) {
    return run_base_len(r, view) + ((view == kViewBasePlusDelta || view == kViewDelta) ? r.dhi - r.dlo : 0u);
}
__device__ __forceinline__ uint32_t run_elem(const DevGraph& g, const uint32_t* dl, const RunView& r, uint32_t nb, uint32_t t) {
    return t < nb ? g.arena[r.lo + t] : dl[r.dlo + (t - nb)];
}
// Whether element key (from the iteration sequence) is really in the view: only BaseMinus can
// reject, when the base key is one of the batch's deletions.
__device__ __forceinline__ bool run_elem_live(const uint32_t* dl, const RunView& r, uint32_t view, uint32_t key) {
    return view != kViewBaseMinusDelta || !contains_u32(dl, r.dlo, r.dhi, key);
}
__device__ __forceinline__ bool run_contains(const DevGraph& g, const uint32_t* dl, const RunView& r, uint32_t view, uint32_t key) {
    switch (view) {
        case kViewBase: return contains_u32(g.arena, r.lo, r.hi, key);
        case kViewBasePlusDelta: return contains_u32(g.arena, r.lo, r.hi, key) || contains_u32(dl, r.dlo, r.dhi, key);
        case kViewDelta: return contains_u32(dl, r.dlo, r.dhi, key);
        default: return contains_u32(g.arena, r.lo, r.hi, key) && !contains_u32(dl, r.dlo, r.dhi, key);
    }
}

// Whether `key` (slot and id of a neighbour) is in the view of b's direction-dir list for the
// phase: binary searches of b's slot run and of its short delta segment.
__device__ __forceinline__ bool view_contains(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot, uint32_t b,
                                              uint32_t key) {
    bool base = false;
    if (view != kViewDelta) {
        uint32_t lo, hi;
        base_range(g, dir, b, slot, lo, hi);
        base = contains_u32(g.arena, lo, hi, key);
        if (view == kViewBase || (view == kViewBasePlusDelta && base) || (view == kViewBaseMinusDelta && !base)) return base;
    }
    const uint4 s = g.delta[dir].seg[b];
    const uint32_t off = phase ? s.z : s.x, len = phase ? s.w : s.y;
    bool in_delta = false;
    if (len) {
        const uint32_t* dl = phase_delta(g, dir, phase);
        uint32_t lo, hi;
        slot_range(dl, off, off + len, slot, g.id_bits, lo, hi);
        in_delta = contains_u32(dl, lo, hi, key);
    }
    return view == kViewBaseMinusDelta ? !in_delta : in_delta;  // Minus: base is true here
}

}  // namespace csm
