// Device primitives and batched maintenance of the labelled adjacency.
//
//   scan_exclusive   device-wide exclusive prefix sum (replaces the reference's CUB
//                    DeviceScan::ExclusiveSum calls, match_gpu.cu:739, 784, ...).
//   batch build      the batch's arcs as a per-vertex sorted delta (replaces BuildTries /
//                    BuildTriesFromEdgeList, match_gpu.cu:302-322, 2296-2344: CUB radix sort + RLE +
//                    scan into a CSR).  Counting sort by source into per-vertex segments, then a
//                    wave-level sort of each segment.
//   apply deletions  delete-compact of every touched list (extension; the reference is
//                    insertion-only).
//   apply insertions merge-insert of every touched list, relocating to a larger arena block when
//                    the capacity is exceeded (addTriesToGraph, indexing.cu:76-263; capacity rule
//                    roundCapacities, indexing.cu:46-59, computed with integer arithmetic).
#include "kernels.h"
#include "view.h"

namespace csm {
namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ T block_exclusive_scan(T v, T* sh, T& total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
        T x = t >= o ? sh[t - o] : T(0);
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[kScanThreads - 1];
    const T incl = sh[t];
    __syncthreads();
    return incl - v;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_tile_sums(const T* in, const uint32_t* n_dev, uint32_t n_max, T* sums) {
    __shared__ T sh[kScanThreads];
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n && blockIdx.x != 0) { if (threadIdx.x == 0) sums[blockIdx.x] = T(0); return; }
    T acc = 0;
    for (int k = 0; k < kScanItems; k++) {
        const uint32_t i = base + k * kScanThreads + threadIdx.x;
        if (i < n) acc += in[i];
    }
    T total;
    block_exclusive_scan<T>(acc, sh, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

template <typename T>
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const T* in, T* out, const uint32_t* n_dev, uint32_t n_max,
                                                              const T* tile_offsets, bool write_total) {
    __shared__ T sh[kScanThreads];
    __shared__ T carry;
    const uint32_t n = n_dev ? min(*n_dev, n_max) : n_max;
    const uint32_t base = blockIdx.x * kScanTile;
    if (base > n) return;
    if (threadIdx.x == 0) carry = tile_offsets ? tile_offsets[blockIdx.x] : T(0);
    __syncthreads();
    for (int k = 0; k < kScanItems; k++) {
        const uint32_t i = base + k * kScanThreads + threadIdx.x;
        const T v = i < n ? in[i] : T(0);
        T total;
        const T ex = block_exclusive_scan<T>(v, sh, total);
        if (i < n) out[i] = carry + ex;
        if (write_total && i == n) out[n] = carry + ex;  // ex == sum of in[0..n) here
        __syncthreads();
        if (threadIdx.x == 0) carry += total;
        __syncthreads();
    }
}

}  // namespace

template <typename T>
void scan_exclusive(const T* in, T* out, const uint32_t* n_dev, uint32_t n_max, bool write_total, T* scratch, hipStream_t s) {
    // scratch must hold scratch_words_for_scan(n_max) elements of T.
    const uint32_t tiles = (n_max + 1 + kScanTile - 1) / kScanTile;
    if (tiles <= 1) {
        k_scan_tiles<T><<<1, kScanThreads, 0, s>>>(in, out, n_dev, n_max, nullptr, write_total);
        return;
    }
    T* sums = scratch;
    T* sums_scanned = scratch + tiles;
    k_scan_tile_sums<T><<<tiles, kScanThreads, 0, s>>>(in, n_dev, n_max, sums);
    scan_exclusive<T>(sums, sums_scanned, nullptr, tiles, false, scratch + 2 * tiles, s);
    k_scan_tiles<T><<<tiles, kScanThreads, 0, s>>>(in, out, n_dev, n_max, sums_scanned, write_total);
}

uint64_t scan_scratch_elems(uint32_t n_max) {
    uint64_t total = 0;
    uint64_t n = n_max;
    while (true) {
        const uint64_t tiles = (n + 1 + kScanTile - 1) / kScanTile;
        if (tiles <= 1) break;
        total += 2 * tiles;
        n = tiles;
    }
    return total + 16;
}

template void scan_exclusive<uint32_t>(const uint32_t*, uint32_t*, const uint32_t*, uint32_t, bool, uint32_t*, hipStream_t);
template void scan_exclusive<unsigned long long>(const unsigned long long*, unsigned long long*, const uint32_t*, uint32_t, bool,
                                                 unsigned long long*, hipStream_t);

// ---------------------------------------------------------------------------------------------
// Batch build.
// Arcs of update edge k: undirected -> (src, key_to_dst) and (dst, key_to_src) both in direction 0;
// directed -> (src, key_to_dst) in direction 0 and (dst, key_to_src) in direction 1.
namespace {

__device__ __forceinline__ bool arc_of(const BatchInput& b, uint32_t k, int dirsel, int which, uint32_t& v, uint32_t& key) {
    // which: 0 = source-side arc, 1 = destination-side arc
    if (dirsel == 1 && which == 1) return false;  // directed, out direction: only source arcs
    if (dirsel == 2 && which == 0) return false;  // directed, in direction: only destination arcs
    if (which == 0) { v = b.src[k]; key = b.key_fwd[k]; }
    else { v = b.dst[k]; key = b.key_bwd[k]; }
    return true;
}

__global__ void k_batch_count(BatchInput b, uint32_t lo, uint32_t n, int dirsel, BatchScratch w) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < 2 * n; i += stride) {
        const uint32_t k = lo + (i >> 1);
        uint32_t v, key;
        if (!arc_of(b, k, dirsel, i & 1, v, key)) continue;
        uint32_t* cnt = b.op[k] ? w.cnt_del : w.cnt_ins;
        atomicAdd(&cnt[v], 1u);
        if (atomicExch(&w.tflag[v], 1u) == 0u) {
            const uint32_t p = atomicAdd(w.n_touched, 1u);
            w.touched[p] = v;
        }
    }
}

__global__ void k_batch_gather_counts(BatchScratch w) {
    const uint32_t T = *w.n_touched;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = w.touched[i];
        w.tmp_ins[i] = w.cnt_ins[v];
        w.tmp_del[i] = w.cnt_del[v];
    }
}

__global__ void k_batch_write_segments(BatchScratch w, DevDelta d) {
    const uint32_t T = *w.n_touched;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = w.touched[i];
        d.seg[v] = make_uint4(w.off_ins[i], w.cnt_ins[v], w.off_del[i], w.cnt_del[v]);
        w.cnt_ins[v] = 0;
        w.cnt_del[v] = 0;
    }
}

__global__ void k_batch_scatter(BatchInput b, uint32_t lo, uint32_t n, int dirsel, BatchScratch w, DevDelta d, uint32_t* err) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = tid; i < 2 * n; i += stride) {
        const uint32_t k = lo + (i >> 1);
        uint32_t v, key;
        if (!arc_of(b, k, dirsel, i & 1, v, key)) continue;
        const uint4 s = d.seg[v];
        if (b.op[k]) {
            const uint32_t p = atomicAdd(&w.cnt_del[v], 1u);
            if (p < s.w) d.del[s.z + p] = key; else atomicOr(err, kErrPrecondition);
        } else {
            const uint32_t p = atomicAdd(&w.cnt_ins[v], 1u);
            if (p < s.y) d.ins[s.x + p] = key; else atomicOr(err, kErrPrecondition);
        }
    }
}

// Sorting network with every comparator ascending ("mirror" first step per stage), so elements
// past n behave as +infinity padding and are never touched: valid for any n <= N = 2^k.
__device__ void wave_sort_small(uint32_t* a, uint32_t n) {
    // n <= 64: one element per lane, exchanges through shuffles.
    const uint32_t lane = lane_id();
    uint32_t v = lane < n ? a[lane] : 0xFFFFFFFFu;
    for (uint32_t k = 2; k <= 64; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t partner = (j == (k >> 1)) ? (lane ^ (k - 1)) : (lane ^ j);
            const uint32_t pv = __shfl(v, static_cast<int>(partner), 64);
            v = lane < partner ? min(v, pv) : max(v, pv);
        }
    }
    if (lane < n) a[lane] = v;
}

__device__ void wave_sort_lds(uint32_t* g, uint32_t n, uint32_t* sh) {
    // n <= kSortLdsWave: staged through this wave's LDS slice.
    const uint32_t lane = lane_id();
    for (uint32_t i = lane; i < n; i += 64) sh[i] = g[i];
    wave_sync();
    uint32_t N = 1;
    while (N < n) N <<= 1;
    for (uint32_t k = 2; k <= N; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < N; i += 64) {
                const uint32_t p = (j == (k >> 1)) ? (i ^ (k - 1)) : (i ^ j);
                if (p > i && p < n) {
                    const uint32_t x = sh[i], y = sh[p];
                    if (y < x) { sh[i] = y; sh[p] = x; }
                }
            }
            wave_sync();
        }
    }
    for (uint32_t i = lane; i < n; i += 64) g[i] = sh[i];
    wave_sync();
}

constexpr uint32_t kSortWaves = 4;
constexpr uint32_t kSortLdsWave = 2048;

__global__ void __launch_bounds__(kSortWaves * 64) k_batch_segsort(BatchScratch w, DevDelta d) {
    __shared__ uint32_t sh[kSortWaves][kSortLdsWave];
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t T = *w.n_touched;
    for (uint32_t i = blockIdx.x * kSortWaves + wid; i < T; i += gridDim.x * kSortWaves) {
        const uint32_t v = w.touched[i];
        const uint4 s = d.seg[v];
        for (int which = 0; which < 2; which++) {
            uint32_t* base = which == 0 ? d.ins + s.x : d.del + s.z;
            const uint32_t len = which == 0 ? s.y : s.w;
            if (len <= 1) continue;
            if (len <= 64) wave_sort_small(base, len);
            else if (len <= kSortLdsWave) wave_sort_lds(base, len, sh[wid]);
            else if (lane_id() == 0) {
                const uint32_t p = atomicAdd(w.n_big, 1u);
                if (p < kMaxBigSegments) { w.big_ptr[p] = static_cast<uint32_t>(base - (which == 0 ? d.ins : d.del)); w.big_len[p] = len | (which ? 0x80000000u : 0u); }
            }
        }
    }
}

// Segments longer than one wave's LDS slice: one block each, bitonic network in global memory.
__global__ void __launch_bounds__(1024) k_batch_segsort_big(BatchScratch w, DevDelta d, uint32_t* err) {
    const uint32_t nb = min(*w.n_big, kMaxBigSegments);
    if (*w.n_big > kMaxBigSegments && threadIdx.x == 0 && blockIdx.x == 0) atomicOr(err, kErrSegmentTooLong);
    for (uint32_t s = blockIdx.x; s < nb; s += gridDim.x) {
        const uint32_t raw = w.big_len[s];
        const uint32_t n = raw & 0x7FFFFFFFu;
        uint32_t* a = ((raw >> 31) ? d.del : d.ins) + w.big_ptr[s];
        uint32_t N = 1;
        while (N < n) N <<= 1;
        for (uint32_t k = 2; k <= N; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                for (uint32_t i = threadIdx.x; i < N; i += blockDim.x) {
                    const uint32_t p = (j == (k >> 1)) ? (i ^ (k - 1)) : (i ^ j);
                    if (p > i && p < n) {
                        const uint32_t x = a[i], y = a[p];
                        if (y < x) { a[i] = y; a[p] = x; }
                    }
                }
                __syncthreads();
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Apply the batch to the base adjacency of one direction.
__global__ void __launch_bounds__(256) k_apply_deletions(DevGraph g, int dir, BatchScratch w) {
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t T = *w.n_touched;
    DevAdj A = g.dir[dir];
    DevDelta D = g.delta[dir];
    for (uint32_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < T; i += gridDim.x * wpb) {
        const uint32_t v = w.touched[i];
        const uint4 s = D.seg[v];
        if (s.w == 0) continue;
        const uint2 a = A.adj[v];
        uint32_t out = a.x, removed = 0;
        for (uint32_t c = a.x; c < a.x + a.y; c += 64) {
            const uint32_t idx = c + lane;
            const bool valid = idx < a.x + a.y;
            const uint32_t k = valid ? g.arena[idx] : 0u;
            const bool del = valid && contains_u32(D.del, s.z, s.z + s.w, k);
            const bool keep = valid && !del;
            const uint64_t km = __ballot(keep);
            const uint64_t dm = __ballot(del);
            wave_sync();
            if (keep) g.arena[out + lanes_below(km)] = k;
            out += static_cast<uint32_t>(__popcll(km));
            removed += static_cast<uint32_t>(__popcll(dm));
            wave_sync();
        }
        if (lane == 0) {
            A.adj[v].y = a.y - removed;
            if (removed != s.w) atomicOr(g.err, kErrPrecondition);
        }
    }
}

__global__ void __launch_bounds__(256) k_apply_insertions(DevGraph g, int dir, BatchScratch w, unsigned long long* nbytes) {
    const uint32_t lane = lane_id();
    uint64_t bytes = 0;  // lane 0 of each wave: 8 * (old + inserted) + 16 per list, 8 per inserted key
    const uint32_t wpb = blockDim.x >> 6;
    const uint32_t T = *w.n_touched;
    DevAdj A = g.dir[dir];
    DevDelta D = g.delta[dir];
    for (uint32_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < T; i += gridDim.x * wpb) {
        const uint32_t v = w.touched[i];
        const uint4 s = D.seg[v];
        if (s.y == 0) continue;
        const uint2 a = A.adj[v];
        const uint32_t cap = A.cap[v];
        const uint32_t nd = a.y + s.y;
        if (lane == 0) bytes += 8ull * nd + 16 + 8ull * s.y;
        const uint32_t* ins = D.ins + s.x;
        uint32_t dst = a.x;
        uint32_t newcap = cap;
        bool relocate = nd > cap;
        if (relocate) {
            uint32_t p2 = 1;
            while (p2 < nd) p2 <<= 1;
            newcap = max(kMinNbrSize, 2u * p2);
            unsigned long long base = ~0ull;
            if (lane == 0) {
                const uint32_t c = 31u - __clz(newcap);
                const uint2 cl = g.fl_cls[c];
                if (cl.y) {  // a block freed by an earlier batch
                    const int k = atomicSub(&g.fl_top[c], 1) - 1;
                    if (k >= 0) base = g.fl_blk[cl.x + k];
                }
                if (base == ~0ull) {
                    base = atomicAdd(g.arena_top, static_cast<unsigned long long>(newcap));
                    if (base + newcap > g.arena_words) { atomicOr(g.err, kErrArenaFull); base = ~0ull; }
                }
            }
            base = __shfl(base, 0, 64);
            if (base == ~0ull) continue;
            dst = static_cast<uint32_t>(base);
        }
        // Insert keys: final position = j + #old keys below, computed before any old key moves;
        // an equal old key violates the "inserted edges are new" precondition.
        for (uint32_t j = lane; j < s.y; j += 64) {
            const uint32_t b = ins[j];
            const uint32_t p = lower_bound_u32(g.arena, a.x, a.x + a.y, b);
            if (p < a.x + a.y && g.arena[p] == b) atomicOr(g.err, kErrPrecondition);
            w.ins_pos[s.x + j] = j + (p - a.x);
        }
        wave_sync();
        if (relocate) {
            for (uint32_t k = lane; k < a.y; k += 64) {
                const uint32_t x = g.arena[a.x + k];
                g.arena[dst + k + lower_bound_u32(ins, 0, s.y, x)] = x;
            }
        } else {
            // In place, back to front: every old key moves right by the number of inserted keys
            // below it; a chunk is read completely before any of it is written, and the keys not
            // yet read all sit left of the chunk, below every destination of the chunk.
            const uint32_t first = lower_bound_u32(g.arena, a.x, a.x + a.y, ins[0]) - a.x;
            uint32_t end = a.y;
            while (end > first) {
                const uint32_t start = end - min(64u, end - first);
                const uint32_t k = start + lane;
                const bool valid = k < end;
                const uint32_t x = valid ? g.arena[a.x + k] : 0u;
                const uint32_t pos = valid ? k + lower_bound_u32(ins, 0, s.y, x) : 0u;
                wave_sync();
                if (valid) g.arena[dst + pos] = x;
                wave_sync();
                end = start;
            }
        }
        for (uint32_t j = lane; j < s.y; j += 64) g.arena[dst + w.ins_pos[s.x + j]] = ins[j];
        wave_sync();
        if (lane == 0) {
            A.adj[v] = make_uint2(dst, nd);
            if (relocate) {
                A.cap[v] = newcap;
                const uint32_t c = 31u - __clz(cap);  // the old block (capacities are powers of two)
                const uint2 cl = g.fl_cls[c];
                const uint32_t j = cl.y ? atomicAdd(&g.fl_pend_n[c], 1u) : 0xFFFFFFFFu;
                if (j < cl.y) g.fl_pend[cl.x + j] = a.x;  // else the block stays unused
            }
        }
    }
    if (nbytes && lane == 0 && bytes) atomicAdd(nbytes, static_cast<unsigned long long>(bytes));
}

// After a batch: the blocks its relocations freed become available to the next batches.
__global__ void k_free_merge(DevGraph g) {
    for (uint32_t c = blockIdx.y; c < kFreeClasses; c += gridDim.y) {
        const uint2 cl = g.fl_cls[c];
        if (!cl.y) continue;
        const uint32_t t = static_cast<uint32_t>(max(g.fl_top[c], 0));
        const uint32_t n = min(g.fl_pend_n[c], cl.y - t);
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
            g.fl_blk[cl.x + t + i] = g.fl_pend[cl.x + i];
    }
}
__global__ void k_free_merge_counts(DevGraph g) {
    const uint32_t c = threadIdx.x;
    if (c >= kFreeClasses) return;
    const uint2 cl = g.fl_cls[c];
    const uint32_t t = static_cast<uint32_t>(max(g.fl_top[c], 0));
    g.fl_top[c] = static_cast<int>(t + min(g.fl_pend_n[c], cl.y - t));
    g.fl_pend_n[c] = 0;
}

// Admission of a built batch before anything changes, one wave per touched vertex: the list space
// its insertions will take under k_apply_insertions' relocation rule (adm[0]) and whether the batch
// breaks a precondition of the update semantics (adm[1]): an inserted arc that is already in the
// list (this also catches a delete and re-insert of one arc in the same batch), a deleted arc that
// is not in the list, or the same arc twice in the batch (equal neighbouring keys in the sorted
// segments: an undirected edge given as both (a, b) and (b, a) lands in a's segment twice).
__global__ void __launch_bounds__(256) k_admit(DevGraph g, int dir, BatchScratch w, unsigned long long* adm) {
    __shared__ uint32_t cls_n[kFreeClasses];  // relocations of this block by size class
    if (threadIdx.x < kFreeClasses) cls_n[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t T = *w.n_touched;
    const DevAdj A = g.dir[dir];
    const DevDelta D = g.delta[dir];
    const uint32_t lane = lane_id();
    const uint32_t wpb = blockDim.x >> 6;
    bool bad = false;
    for (uint32_t i = blockIdx.x * wpb + (threadIdx.x >> 6); i < T; i += gridDim.x * wpb) {
        const uint32_t v = w.touched[i];
        const uint4 s = D.seg[v];
        const uint2 a = A.adj[v];
        for (uint32_t j = lane; j < s.y; j += 64) {
            const uint32_t k = D.ins[s.x + j];
            if (j > 0 && D.ins[s.x + j - 1] == k) bad = true;
            if (contains_u32(g.arena, a.x, a.x + a.y, k)) bad = true;
        }
        for (uint32_t j = lane; j < s.w; j += 64) {
            const uint32_t k = D.del[s.z + j];
            if (j > 0 && D.del[s.z + j - 1] == k) bad = true;
            if (!contains_u32(g.arena, a.x, a.x + a.y, k)) bad = true;
        }
        if (lane == 0 && s.w <= a.y) {
            const uint32_t nd = a.y - s.w + s.y;
            if (s.y && nd > A.cap[v]) {
                uint32_t p2 = 1;
                while (p2 < nd) p2 <<= 1;
                atomicAdd(&cls_n[31u - __clz(max(kMinNbrSize, 2u * p2))], 1u);
            }
        }
    }
    const bool any_bad = __ballot(bad) != 0;
    if (lane == 0 && any_bad) atomicOr(reinterpret_cast<uint32_t*>(&adm[1]), 1u);
    __syncthreads();
    if (threadIdx.x < kFreeClasses && cls_n[threadIdx.x]) atomicAdd(&adm[6 + threadIdx.x], static_cast<unsigned long long>(cls_n[threadIdx.x]));
}

// adm[2..5] = {need, bad, arena top, error bits} for the host, need = the words of the relocations
// that the free blocks of their classes cannot serve; adm[0..1] and the class counts adm[6..37]
// cleared for the next batch.
__global__ void k_admit_finish(DevGraph g, unsigned long long* adm) {
    if (threadIdx.x != 0) return;
    unsigned long long need = 0;
    for (uint32_t c = 0; c < kFreeClasses; c++) {
        const unsigned long long n = adm[6 + c];
        const unsigned long long avail = g.fl_cls[c].y ? static_cast<unsigned long long>(max(g.fl_top[c], 0)) : 0ull;
        if (n > avail) need += (n - avail) << c;
        adm[6 + c] = 0;
    }
    adm[2] = need;
    adm[3] = adm[1];
    adm[4] = *g.arena_top;
    adm[5] = *g.err;
    adm[0] = 0;
    adm[1] = 0;
}

__global__ void k_batch_cleanup(BatchScratch w, DevDelta d) {
    const uint32_t T = *w.n_touched;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = w.touched[i];
        d.seg[v] = make_uint4(0, 0, 0, 0);
        w.tflag[v] = 0;
        w.cnt_ins[v] = 0;
        w.cnt_del[v] = 0;
    }
}

// Thread per (vertex, slot): row[s] = lower bound of key (s << id_bits) in the vertex's list;
// row[nslots] = end of the list.
__global__ void __launch_bounds__(256) k_slot_dir(DevGraph g, int dir, const uint32_t* list, const uint32_t* n_list, uint32_t max_n) {
    const uint32_t n = list ? min(*n_list, max_n) : g.dv;
    const uint32_t w = g.nslots + 1;
    const uint64_t total = static_cast<uint64_t>(n) * w;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; t < total;
         t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const uint32_t i = static_cast<uint32_t>(t / w), sl = static_cast<uint32_t>(t % w);
        const uint32_t v = list ? list[i] : i;
        const uint2 a = g.dir[dir].adj[v];
        const uint32_t key0 = sl == 0 ? 0u : (sl << g.id_bits);  // sl < nslots: fits the key width
        const uint32_t r = sl == g.nslots ? a.x + a.y : lower_bound_u32(g.arena, a.x, a.x + a.y, key0);
        g.sdir[dir][static_cast<uint64_t>(v) * w + sl] = r;
    }
}

// Label buckets of the touched vertices: a histogram, its scan, then a scatter.  With few labels
// every block counts in LDS and reserves its run of each bucket with one device atomic per label
// (a device atomic per vertex would serialise on the few bucket counters); with more labels than
// the LDS table holds, per-vertex device atomics.
constexpr uint32_t kBucketLds = 4096;

__global__ void __launch_bounds__(256) k_bucket_count(DevGraph g, BatchScratch w) {
    __shared__ uint32_t h[kBucketLds];
    const uint32_t T = *w.n_touched;
    const bool lds = g.nvl <= kBucketLds;
    if (lds) {
        for (uint32_t r = threadIdx.x; r < g.nvl; r += blockDim.x) h[r] = 0;
        __syncthreads();
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t r = g.vrank[w.touched[i]];
        if (r >= g.nvl) continue;
        if (lds) atomicAdd(&h[r], 1u);
        else atomicAdd(&w.label_cur[r], 1u);
    }
    if (lds) {
        __syncthreads();
        for (uint32_t r = threadIdx.x; r < g.nvl; r += blockDim.x)
            if (h[r]) atomicAdd(&w.label_cur[r], h[r]);
    }
}

// Same walk over the touched vertices as k_bucket_count (same grid): each block counts again in
// LDS, reserves its runs from the bucket cursors, then places its vertices in them.
__global__ void __launch_bounds__(256) k_bucket_scatter(DevGraph g, BatchScratch w) {
    __shared__ uint32_t h[kBucketLds];
    const uint32_t T = *w.n_touched;
    if (g.nvl > kBucketLds) {
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
            const uint32_t v = w.touched[i];
            const uint32_t r = g.vrank[v];
            if (r < g.nvl) w.by_label[atomicAdd(&w.label_cur[r], 1u)] = v;
        }
        return;
    }
    for (uint32_t r = threadIdx.x; r < g.nvl; r += blockDim.x) h[r] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t r = g.vrank[w.touched[i]];
        if (r < g.nvl) atomicAdd(&h[r], 1u);
    }
    __syncthreads();
    for (uint32_t r = threadIdx.x; r < g.nvl; r += blockDim.x)
        if (h[r]) h[r] = atomicAdd(&w.label_cur[r], h[r]);  // the block's first position in bucket r
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < T; i += gridDim.x * blockDim.x) {
        const uint32_t v = w.touched[i];
        const uint32_t r = g.vrank[v];
        if (r < g.nvl) w.by_label[atomicAdd(&h[r], 1u)] = v;
    }
}

__global__ void k_reset_counters(BatchScratch w) {
    if (threadIdx.x == 0 && blockIdx.x == 0) { *w.n_touched = 0; *w.n_big = 0; }
}

}  // namespace

// Builds the delta of batch edges [lo, lo + n) for one direction.  n_touched ends up on the
// device; max_arcs bounds every grid.
void batch_build(const DevGraph& g, int dir, int dirsel, const BatchInput& b, uint32_t lo, uint32_t n,
                 const BatchScratch& w, hipStream_t s) {
    const uint32_t max_arcs = 2 * n;
    const int grid = static_cast<int>(std::min<uint32_t>(4096, (max_arcs + 255) / 256 + 1));
    k_reset_counters<<<1, 64, 0, s>>>(w);
    k_batch_count<<<grid, 256, 0, s>>>(b, lo, n, dirsel, w);
    k_batch_gather_counts<<<grid, 256, 0, s>>>(w);
    scan_exclusive<uint32_t>(w.tmp_ins, w.off_ins, w.n_touched, max_arcs, false, w.scan_tmp, s);
    scan_exclusive<uint32_t>(w.tmp_del, w.off_del, w.n_touched, max_arcs, false, w.scan_tmp, s);
    k_batch_write_segments<<<grid, 256, 0, s>>>(w, g.delta[dir]);
    k_batch_scatter<<<grid, 256, 0, s>>>(b, lo, n, dirsel, w, g.delta[dir], g.err);
    const int sgrid = static_cast<int>(std::min<uint32_t>(2048, (max_arcs + kSortWaves - 1) / kSortWaves + 1));
    k_batch_segsort<<<sgrid, kSortWaves * 64, 0, s>>>(w, g.delta[dir]);
    k_batch_segsort_big<<<64, 1024, 0, s>>>(w, g.delta[dir], g.err);
}

void batch_bucket(const DevGraph& g, uint32_t max_touched, const BatchScratch& w, hipStream_t s) {
    const int grid = static_cast<int>(std::min<uint32_t>(2048, (max_touched + 255) / 256 + 1));
    fill_u32(w.label_cur, 0u, g.nvl + 1, s);
    k_bucket_count<<<grid, 256, 0, s>>>(g, w);
    scan_exclusive<uint32_t>(w.label_cur, w.label_off, nullptr, g.nvl, true, w.label_tmp, s);
    copy_u32(w.label_off, w.label_cur, g.nvl, s);  // the scatter's cursors
    k_bucket_scatter<<<grid, 256, 0, s>>>(g, w);
}

void slot_dir_update(const DevGraph& g, int dir, const uint32_t* list, const uint32_t* n_list, uint32_t max_n, hipStream_t s) {
    if (!g.sdir[dir]) return;
    const uint64_t total = static_cast<uint64_t>(list ? max_n : g.dv) * (g.nslots + 1);
    const int grid = static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>(8192, (total + 255) / 256)));
    k_slot_dir<<<grid, 256, 0, s>>>(g, dir, list, n_list, max_n);
}

void batch_apply_deletions(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, hipStream_t s) {
    const int grid = static_cast<int>(std::min<uint32_t>(8192, (max_touched + 3) / 4 + 1));
    k_apply_deletions<<<grid, 256, 0, s>>>(g, dir, w);
}

void batch_apply_insertions(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, unsigned long long* bytes,
                            hipStream_t s) {
    const int grid = static_cast<int>(std::min<uint32_t>(8192, (max_touched + 3) / 4 + 1));
    k_apply_insertions<<<grid, 256, 0, s>>>(g, dir, w, bytes);
}

void batch_admit(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, unsigned long long* adm, hipStream_t s) {
    const int grid = static_cast<int>(std::min<uint32_t>(8192, (max_touched + 3) / 4 + 1));
    k_admit<<<grid, 256, 0, s>>>(g, dir, w, adm);
}

void batch_admit_finish(const DevGraph& g, unsigned long long* adm, hipStream_t s) { k_admit_finish<<<1, 64, 0, s>>>(g, adm); }

void free_blocks_merge(const DevGraph& g, hipStream_t s) {
    k_free_merge<<<dim3(64, kFreeClasses), 256, 0, s>>>(g);
    k_free_merge_counts<<<1, 64, 0, s>>>(g);
}

void batch_cleanup(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, hipStream_t s) {
    const int grid = static_cast<int>(std::min<uint32_t>(4096, (max_touched + 255) / 256 + 1));
    k_batch_cleanup<<<grid, 256, 0, s>>>(w, g.delta[dir]);
}

}  // namespace csm
