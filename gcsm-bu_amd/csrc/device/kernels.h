// Launch interface of the HIP kernels (graph.hip, index.hip, enum.hip).  Host code passes plain
// device pointers; sizes that only the device knows stay in device memory and every grid is sized
// for the host-known maximum, so no launch waits for a device-to-host copy.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dev_common.h"

namespace csm {

constexpr uint32_t kMaxBigSegments = 4096;

// The whole update stream, resident in HBM: edge k is src[k] -- dst[k] with precomputed keys.
struct BatchInput {
    const uint32_t* src;
    const uint32_t* dst;
    const uint32_t* key_fwd;  // key of dst in src's list
    const uint32_t* key_bwd;  // key of src in dst's list
    const uint8_t* op;        // 0 insert, 1 delete
};

struct BatchScratch {
    uint32_t* cnt_ins;   // [dv]
    uint32_t* cnt_del;   // [dv]
    uint32_t* tflag;     // [dv]
    uint32_t* touched;   // [max_arcs]
    uint32_t* n_touched; // [1]
    uint32_t* tmp_ins;   // [max_arcs + 1]
    uint32_t* tmp_del;
    uint32_t* off_ins;
    uint32_t* off_del;
    uint32_t* ins_pos;   // [max_arcs]
    uint32_t* scan_tmp;  // scan_scratch_elems(max_arcs)
    uint32_t* n_big;     // [1]
    uint32_t* big_ptr;   // [kMaxBigSegments]
    uint32_t* big_len;
    // the touched vertices grouped by label rank (batch_bucket): rank r's at by_label[label_off[r] ..
    // label_off[r + 1])
    uint32_t* by_label;  // [max_arcs]
    uint32_t* label_off; // [nvl + 1]
    uint32_t* label_cur; // [nvl + 1]
    uint32_t* label_tmp; // scan_scratch_elems(nvl + 1)
};

template <typename T>
void scan_exclusive(const T* in, T* out, const uint32_t* n_dev, uint32_t n_max, bool write_total, T* scratch, hipStream_t s);
uint64_t scan_scratch_elems(uint32_t n_max);

void batch_build(const DevGraph& g, int dir, int dirsel, const BatchInput& b, uint32_t lo, uint32_t n,
                 const BatchScratch& w, hipStream_t s);
// Groups the built batch's touched vertices by label rank (w.by_label, w.label_off).
void batch_bucket(const DevGraph& g, uint32_t max_touched, const BatchScratch& w, hipStream_t s);
// Admission of the built batch (per direction, then finish once): adm[6 + c] += the insertions'
// relocations into blocks of size class c, adm[1] |= a precondition violation; finish moves
// {need, bad, arena top, error bits} to adm[2..5], need = words the free blocks cannot serve, and
// clears adm[0..1] and adm[6..37] (all zero before the first batch).
void batch_admit(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, unsigned long long* adm, hipStream_t s);
void batch_admit_finish(const DevGraph& g, unsigned long long* adm, hipStream_t s);
// The blocks freed by a batch's relocations become available (after its insertions are applied).
void free_blocks_merge(const DevGraph& g, hipStream_t s);
// Recomputes label-slot directory rows of direction dir: every vertex when list is null, else
// the vertices list[0, min(*n_list, max_n)).
void slot_dir_update(const DevGraph& g, int dir, const uint32_t* list, const uint32_t* n_list, uint32_t max_n, hipStream_t s);
void batch_apply_deletions(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, hipStream_t s);
// bytes (or null): += algorithmic bytes, 8 * (old + inserted keys) + 16 per touched list plus 8 per
// inserted key (read old and delta, write merged; list header; batch arrays).
void batch_apply_insertions(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, unsigned long long* bytes,
                            hipStream_t s);
void batch_cleanup(const DevGraph& g, int dir, uint32_t max_touched, const BatchScratch& w, hipStream_t s);

// ---- global candidates (NLF), summary ---------------------------------------------------------
// Per query vertex: required label rank and the neighbour-label-frequency groups it must satisfy
// (graph.cpp:295-330 NLF; indexing.cu:303-338 test).  One entry per (query, query vertex).
struct NlfGroup {
    uint32_t slot;
    uint32_t count;
    uint32_t dir;
};
struct NlfEntry {
    uint32_t vrank;     // kNoVertex: label absent from the data graph -> no candidates
    uint32_t g0, g1;    // groups [g0, g1)
    uint32_t* bitmap;   // candidate bitmap of this query vertex
};

void nlf_full(const DevGraph& g, const NlfEntry* entries, uint32_t n_entries, const NlfGroup* groups, hipStream_t s);
// mode 0: upper bound with the pending insertions (only sets bits); mode 1: exact on the base
// adjacency (sets and clears bits).  Runs over the touched vertices of direction `dir`.
void nlf_touched(const DevGraph& g, const NlfEntry* entries, uint32_t n_entries, const NlfGroup* groups,
                 const uint32_t* touched, const uint32_t* n_touched, uint32_t max_touched, int mode, hipStream_t s);
// Cardinality and number of non-empty sources of the relation u->w restricted to candidates
// (statisticIndex, indexing.cu:265-301).  out[0] += card, out[1] += #sources.
void relation_summary(const DevGraph& g, uint32_t dir, uint32_t slot, uint32_t vrank_u, const uint32_t* cand_u,
                      const uint32_t* cand_w, unsigned long long* out, hipStream_t s);

// ---- local index ------------------------------------------------------------------------------
// A vertex set built for one batch step: bitmap + compact list + device length.
struct VSet {
    uint32_t* bitmap;
    uint32_t* list;
    uint32_t* len;
    uint32_t cap;
};

// Anchors of the step's query edge u0->u1: batch arcs (v, x) of the phase with v in cand(u0) of
// label rank vr0 and x in cand(u1).  Writes pairs (v, x), and fills L(u0) = {v}, L(u1) = {x}.
// by_label / label_off: the touched vertices of direction dir grouped by label rank (batch_bucket);
// only rank vr0's group is read.
void local_anchors(const DevGraph& g, uint32_t phase, uint32_t dir, uint32_t slot, uint32_t vr0,
                   const uint32_t* cand0, const uint32_t* cand1, const uint32_t* by_label, const uint32_t* label_off,
                   uint32_t max_touched, uint32_t* pairs, uint32_t* n_pairs, uint32_t max_pairs, VSet L0, VSet L1,
                   hipStream_t s);
// L(u) seed: every vertex reached from the frontier through the relation view (b -> u) that is a
// candidate of u; the output set's bitmap (cleared beforehand) deduplicates the list.
// `blocks`: grid of the local-index kernels (their frontier sizes are device-side).
// bytes (or null): += algorithmic bytes of the sparse sweep: 4 per item read, 8 per run located
// (two directory words), 4 per key scanned, 4 per vertex written; bitmap probes count 0.
void local_expand(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                  const uint32_t* frontier, const uint32_t* frontier_len, uint32_t max_frontier,
                  const uint32_t* cand_u, VSet out, int blocks, unsigned long long* bytes, hipStream_t s);
// Keeps the vertices y of `in` with a neighbour in L(b) along the view (u -> b).
// Local index of one query vertex u in two launches: expand from the frontier along
// (view, dir, slot), keep the candidates of cand_u that have a neighbour in every filter's bitmap.
struct LocalFilter {
    uint32_t view, dir, slot;
    const uint32_t* bitmap;
};
struct LocalFilters {
    uint32_t n;
    LocalFilter f[kMaxQV - 1];
};
// queue (capacity dv) and *queue_len (zero on entry) hold the new candidates between the passes:
// local_queue fills them, local_check filters them into out.
void local_queue(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot, const uint32_t* frontier,
                 const uint32_t* frontier_len, uint32_t max_frontier, const uint32_t* cand_u, VSet out, uint32_t* queue,
                 uint32_t* queue_len, int blocks, unsigned long long* bytes, hipStream_t s);
void local_check(const DevGraph& g, uint32_t phase, const LocalFilters& fl, VSet out, const uint32_t* queue,
                 const uint32_t* queue_len, int blocks, unsigned long long* bytes, hipStream_t s);
struct VSetTable {
    VSet s[kMaxQV];
};

// ---- local index of many steps at once --------------------------------------------------------
// The local index of every step of a group (query, query edge) is built in lockstep: one launch
// computes the anchors of all of them, then, for each indexing-order position, one launch expands
// and one checks the local sets of every step that builds a vertex at that position.  Each launch
// treats its jobs' items as one space (device-side lengths, prefix-summed by the launch itself).
struct AnchorJob {
    uint32_t phase, dir, slot, vr0;
    const uint32_t* cand0;
    const uint32_t* cand1;
    uint32_t* pairs;
    uint32_t* n_pairs;
    uint32_t max_pairs;
    VSet L0, L1;
};
struct LocalJob {
    uint32_t phase, view, dir, slot;   // expansion relation: the run of each frontier vertex
    const uint32_t* frontier;          // L(b0) of an earlier position
    const uint32_t* frontier_len;
    const uint32_t* cand_u;            // global candidates of the vertex built
    VSet out;                          // its local set (bitmap clear, length zero on entry)
    uint32_t* queue;                   // dv words of queue space and its length (zero on entry)
    uint32_t* queue_len;
    LocalFilters fl;                   // the other built neighbours (none: every queued vertex is kept)
};
constexpr uint32_t kMaxLocalJobs = 256;
void local_anchors_multi(const DevGraph& g, const AnchorJob* jobs, uint32_t njobs, const uint32_t* by_label0, const uint32_t* label_off0,
                         const uint32_t* by_label1, const uint32_t* label_off1, uint32_t max_touched, hipStream_t s);
// bytes: as local_queue / local_check.
void local_queue_multi(const DevGraph& g, const LocalJob* jobs, uint32_t njobs, int blocks, unsigned long long* bytes, hipStream_t s);
void local_check_multi(const DevGraph& g, const LocalJob* jobs, uint32_t njobs, int blocks, unsigned long long* bytes, hipStream_t s);
// Clears the bitmaps of sets[0, nsets) (whole words of their listed members).
void clear_sets_multi(const VSet* sets, uint32_t nsets, uint32_t max_len, hipStream_t s);
// Clears the bitmaps of the sets in mask (their lengths stay; the caller zeroes them).
void clear_sets(const VSetTable& t, uint32_t mask, uint32_t max_len, hipStream_t s);
// Stores w.v[0 .. w.n) at p (one tiny launch instead of several fills).
struct Words {
    uint32_t n;
    uint32_t v[15];
};
void store_words(uint32_t* p, const Words& w, hipStream_t s);

// ---- enumeration ------------------------------------------------------------------------------
struct MatchOut {
    unsigned long long* count;  // [1]
    unsigned long long* hash;   // [1]
    unsigned long long* bytes;  // [1] algorithmic bytes (lists read + candidates written), or null
};

struct Ring {
    const uint32_t* in;         // records of in_depth words
    const uint32_t* in_count;   // device count
    uint32_t in_max;            // host bound on the count
    uint32_t* out;              // output records (write modes)
    uint32_t* out_count;
    uint32_t out_cap;           // records
    uint32_t* overflow;         // set when the output does not fit
};

// Initial partial results of depth 2 from the anchor pairs (optionally swapped).
void init_from_pairs(const uint32_t* pairs, const uint32_t* n_pairs, uint32_t max_pairs, bool swap,
                     uint32_t* out, uint32_t* out_count, hipStream_t s);
// Initial partial results of depth 2 from a start vertex set: (v, x) with x in the view a->b of v
// and in filter_b, written as pairs at out[0 ..] (count at *out_count, *overflow set when more
// than out_cap do not fit).
struct InitJob {
    uint32_t phase, view, dir, slot;
    const uint32_t* list;
    const uint32_t* len;
    const uint32_t* filter_b;
    uint32_t* out;
    uint32_t* out_count;
    uint32_t* overflow;
    uint32_t out_cap;
    uint32_t max_len;
};
void init_from_set(const DevGraph& g, uint32_t phase, uint32_t view, uint32_t dir, uint32_t slot,
                   const uint32_t* list, const uint32_t* len, uint32_t max_len, const uint32_t* filter_b,
                   Ring r, hipStream_t s);
// The same for njobs steps at once (jobs in device memory, grid.y = job).
void init_from_set_multi(const DevGraph& g, const InitJob* jobs, uint32_t njobs, hipStream_t s);
// Counts and hashes records that are already complete embeddings (two-vertex queries).
void count_records(const StepDesc& d, const uint32_t* rec, const uint32_t* n_rec, uint32_t max_rec, MatchOut out, hipStream_t s);
// Compacts the set bits of a bitmap over [0, dv) into a list.
void bitmap_to_list(const uint32_t* bitmap, uint32_t dv, uint32_t* list, uint32_t* len, hipStream_t s);

enum ExtendMode : int { kCountAtEnd = 0, kWriteAtEnd = 1 };
// Range donation of the depth-first kernel (schedule 2): one range word and one record index per
// wave of the launch (kMaxDfsBlocks * 4 waves at most); extend clears the words before the launch.
constexpr uint32_t kMaxDfsBlocks = 16384;
struct Donation {
    unsigned long long* range;  // [kMaxDfsBlocks * 4]
    uint32_t* rec;              // [kMaxDfsBlocks * 4]
};
// DFS from the records at depth d0 down to end_depth: counts/hashes complete embeddings when
// end_depth == n, or writes the partial results of depth end_depth to the ring.
// schedule: 0 static (item t by wave t mod W), 1 dynamic (global work counter, zeroed by the
// caller), 2 dynamic with range donation: a wave out of items takes the upper half of the
// first-depth candidate range another wave is still exploring (dn).  Without `split` an item is one record.  With `split` (the exclusive scan of
// root_chunks, total at [n]) an item is one of a record's k chunks: its depth-d0 candidates cut
// into k equal pieces of whole 64-candidate rounds, so a heavy record spreads over k waves.
// blocks: the persistent grid (fewer when there are fewer items).
void extend(const DevGraph& g, const StepDesc& d, uint32_t d0, uint32_t end_depth, int mode, Ring r,
            MatchOut out, const unsigned long long* split, unsigned long long* work_counter, int schedule, int blocks, Donation dn,
            hipStream_t s);
// One level from the n records of depth d0 at rec (n known on the host), flattened over all its
// candidates: level_prepare writes, per record, the view runs of the depth's backward entries, the
// entry with the shortest view sequence (it generates the candidates; the others are probed), the
// scanned lengths of those sequences to off[0..n] (off[n] = total candidates) and, with_hash, the
// record's prefix hash; level_flat then tests every candidate, load-balanced over the grid, and
// counts/hashes complete embeddings (d0 + 1 == n) or writes records of depth d0 + 1.
// n <= level_max_records(d, d0, ls).
constexpr uint32_t kLevelTile = 512;  // records per tile of the multi-block level preparation
struct LevelScratch {
    unsigned long long* off;       // [cap + 1] scanned candidate counts
    uint4* prun;                   // [cap] generating view run of each record
    uint4* runs;                   // [runs_cap] view runs of every backward entry (record-major; depths with one entry skip it)
    uint8_t* piv;                  // [cap] generating entry | view << 4 | dir << 6
    unsigned long long* ph;        // [cap] prefix hash of each record (counting levels)
    unsigned long long* tile_sum;  // [cap / kLevelTile + 2] candidate totals of the preparation's tiles
    uint64_t cap;
    uint64_t runs_cap;
};
// zero (nullable): two device words cleared before the level runs (a writing level's output count
// and overflow flag).
void level_prepare(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n, const LevelScratch& ls,
                   bool with_hash, uint32_t* zero, unsigned long long* bytes, hipStream_t s);
void level_flat(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n, int mode, const LevelScratch& ls,
                Ring r, MatchOut out, hipStream_t s);
// Most records one level call takes at depth d0 (the per-entry runs must fit).
uint32_t level_max_records(const StepDesc& d, uint32_t d0, const LevelScratch& ls);

// The levels of one step chained on the device, with no host round trip between them: level k
// extends the records of depth depth0 + k.  Its input follows from the host-known start and the
// counters of the levels before it (level k writes its records right after its input, with
// {count, overflow} in ctr[2k], ctr[2k + 1]; level 0's preparation clears ctr).  A level whose
// input is empty, holds more than cap[k] records or at least stop_at[k], or that follows an
// overflow, does not run, and neither does any later one: the host reads ctr and resumes there.
struct LevelChain {
    uint32_t* ring;
    uint64_t ring_words;
    uint32_t* ctr;                 // [2 * kMaxQV]
    uint64_t off0;                 // ring word offset of level 0's input records
    uint32_t n0;                   // their count
    uint32_t depth0;
    uint32_t cap[kMaxQV];
    uint32_t stop_at[kMaxQV];
};
void level_prepare_chain(const DevGraph& g, const StepDesc& d, const LevelChain& c, uint32_t k, const LevelScratch& ls, bool with_hash,
                         unsigned long long* bytes, hipStream_t s);
void level_flat_chain(const DevGraph& g, const StepDesc& d, const LevelChain& c, uint32_t k, int mode, const LevelScratch& ls,
                      MatchOut out, hipStream_t s);
// The level chains of several steps run together: level k of every job in one launch per kernel
// (grid.y = job), each job with its own descriptor, ring, counters and level scratch.  Job j runs
// levels 0 .. L (writing levels 0 .. L - 1, counting level L); at launch k it does nothing when
// k > L or when chain_at says its chain stopped.
struct ChainJob {
    StepDesc d;
    LevelChain c;
    LevelScratch ls;
    MatchOut out;                    // the counting level's count, hash and bytes
    unsigned long long* write_bytes; // the writing levels' algorithmic bytes (nullable)
    uint32_t L;
    uint32_t pad;
};
// Level k of the njobs jobs at `jobs` (device memory): preparation (run lengths and scan), then
// the flattened candidate kernel of the jobs that write at k (mode kWriteAtEnd) or count at k.
// blocks: of the candidate kernel per job (the preparation takes half as many).
void level_prepare_chain_multi(const DevGraph& g, const ChainJob* jobs, uint32_t njobs, uint32_t k, int blocks, hipStream_t s);
void level_flat_chain_multi(const DevGraph& g, const ChainJob* jobs, uint32_t njobs, uint32_t k, int mode, int blocks, hipStream_t s);
// chunks[i] = ceil(|pivot view run of record i at depth d0| / 64), then scanned in place (total at [n]).
// Per record, its chunk count k = min(max_chunks, ceil(depth-d0 candidates / 64)), scanned.
void root_chunks(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, const uint32_t* n_rec,
                 uint32_t max_rec, uint32_t max_chunks, unsigned long long* chunks, unsigned long long* scan_tmp, hipStream_t s);

// Cartesian-product flattening of trailing tree leaves (cartesian_product.cu:193-284).  A segment
// is one (record, leaf) pair: segment i = record * leaves + leaf.
struct CpScratch {
    unsigned long long* seg_off;   // [max_segments + 1] scanned view-run lengths of the segments
    uint4* seg_run;                // [max_segments] view run of each segment
    uint32_t* seg_cnt;             // [max_segments] filtered list length of each segment
    unsigned long long* prod;      // [max_records + 1] list-length products, scanned in place into offsets
    unsigned long long* rec_hash;  // [max_records] pair-hash sum of the record's prefix
    unsigned long long* scan_tmp;
    uint32_t* lists;               // segment i's list starts at lists[seg_off[i]]
    unsigned long long lists_cap;
    unsigned long long* max_prod;  // [1]
    uint32_t* err;                 // kErrProductOverflow / kErrScratchFull of this scratch's last preparation
};
// n records of depth d0 (host count), n * (d.n - d0) <= max_segments; max_prod and err zeroed by
// the caller.  prod[n] ends up as the number of combinations.
void cp_prepare(const DevGraph& g, const StepDesc& d, uint32_t d0, const uint32_t* rec, uint32_t n, CpScratch c, hipStream_t s);
void cp_enumerate(const StepDesc& d, uint32_t d0, uint32_t n, CpScratch c, MatchOut out, hipStream_t s);

void fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// Plain copy of n words (calibration of the HBM traffic counters).
void copy_u32(const uint32_t* src, uint32_t* dst, uint64_t n, hipStream_t s);

// Words of the library's static partial-result ring (scratch.hip): 256 MiB on every device used.
constexpr uint64_t kStaticRingWords = 64ull << 20;
// Device address of that ring on `device`, which must be the current device; resolved once per
// device and cached.
hipError_t static_ring(int device, uint32_t** ptr);

}  // namespace csm
