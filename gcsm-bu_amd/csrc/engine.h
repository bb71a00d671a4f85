// Host orchestration of the matching path on one GPU.
//
// Mirrors the reference driver loop (TechniquePermutationProcess, technique_permutation.h:29-225) and
// the MatchGPU class (graph/match_gpu.h:80-162), redesigned around one query-independent device
// graph: load -> offline index (candidate bitmaps) -> summary -> plan -> per batch: build the batch
// delta, deletion steps, apply deletions, insertion steps, apply insertions.  Each step is one query
// edge of one query: local index, then one of the four search strategies.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/csm.h"
#include "common/errors.h"
#include "device/kernels.h"
#include "host/graph_io.h"
#include "host/plan.h"
#include "host/query.h"

namespace csm {

struct Technique {
    MatchingOrder order = MatchingOrder::kGcsm;
    SearchStrategy strategy = SearchStrategy::kDynamicBfsDfs;
    bool local_index = true;
    bool cartesian_product = true;
    bool heavy_vertex_avoidance = true;
    bool relation_switch = true;
};

struct EngineOptions {
    int device = 0;
    uint64_t pool_bytes = 0;          // HBM budget for the device pool (0: no cap)
    uint64_t ring_bytes = 1ull << 30; // partial-result ring of each lane (a stack of levels)
    uint64_t cp_list_bytes = 512ull << 20;  // per lane
    uint32_t cp_max_records = 1u << 22;
    uint32_t max_batch_edges = 0;     // 0: the stream's batch size
    uint32_t semi_bfs_threshold = CSM_DEFAULT_SEMI_BFS_THRESHOLD;
    bool directed = false;
    // Launch settings that change scheduling only, never results (Tunables):
    uint32_t waves_per_anchor = 1024; // most waves one record of the counting DFS is split over (1: no split)
    uint32_t stack_depth = kMaxQV;    // most LDS stack rows of a counting DFS launch; deeper searches run levels first
    uint32_t dfs_blocks = 2048;       // blocks of the persistent DFS grid
    uint32_t level_blocks = 384;      // blocks per step of a team's multi-step level kernels
    bool slot_directory = true;       // per-vertex label-slot offsets (when they fit kMaxSlotDirectoryBytes)
    bool chain_levels = true;         // semi-BFS levels of a step chained on the device (one host wait per step)
    // A team's chained levels launched as a captured, replayed HIP graph.  Off by default: measured
    // slower than direct launches on every workload (C3 34.4-37.3 against 33.7-34.1 ms per batch,
    // C4 85-91 against 81-82 ms, C2 19.7-20.8 against 18.8-18.9 ms).
    bool graph_chains = false;
    uint32_t lanes = 32;              // concurrent matching steps, each with its own scratch
    uint32_t team = 0;                // lanes per team (0: a quarter of the lanes, four teams): a team's steps share a
                                      // stream, host waits and multi-step launches
    uint64_t arena_slack_words = 1u << 20;  // adjacency arena room beyond the growth rule
};

constexpr uint64_t kMaxSlotDirectoryBytes = 4ull << 30;
constexpr uint32_t kSplitMaxRecords = 1u << 14;  // root splitting only below this many records
constexpr uint32_t kMaxLanes = 64;
// Per-entry view runs a lane's level scratch holds, in records' worth of cp_max_records: levels
// whose depth has more backward entries than this take fewer records per call.
constexpr uint64_t kLevelRunsPerRecord = 3;
constexpr uint64_t kLiveBatchesOfRoom = 16;  // arena room for in-memory batches, in max-size batches
// Per lane, zeroed at every step: 16 set lengths, list counters (+16, +17), one local-index queue
// counter per query vertex (+kQueueCounters), then the step counters (+kStepCounters).
constexpr uint32_t kQueueCounters = 18;
constexpr uint32_t kStepCounters = 48;
constexpr uint32_t kLaneWords = 64;
constexpr uint32_t kSlotWords = 64;
constexpr uint64_t kGroupBudgetBytes = 6ull << 30;  // device memory of the local-index slots
constexpr uint32_t kMaxGroup = 64;
constexpr uint32_t kGroupMinSteps = 256;  // standing query edges from which steps' local indexes are built together

// The launch settings a session may change between batches; they never change results.  The
// reference fixes all of them at compile time (config.h:47 MIN_NUM_RESULTS_TO_GPU, the block and
// warp counts of match_gpu.cu's launches).
struct Tunables {
    uint32_t semi_bfs_threshold = 0;  // records at a Cartesian-product depth at which the levels stop expanding
    uint32_t waves_per_anchor = 0;    // most waves one record's first-depth candidates are split over (1..65535)
    uint32_t stack_depth = 0;         // most LDS stack rows (depths) of a counting DFS launch (1..kMaxQV)
    uint32_t dfs_blocks = 0;          // blocks of the persistent DFS grid (64..16384)
    uint32_t local_blocks = 0;        // blocks of the local-index kernels (16..2048)
    uint32_t team = 0;                // lanes per team (1: every lane runs its steps alone; up to the lanes)
    uint32_t level_blocks = 0;        // blocks per step of a team's level kernels (64..4096)
};

struct QueryResult {
    uint64_t pos = 0, neg = 0, pos_hash = 0, neg_hash = 0;
};

struct BatchInfo {
    uint64_t edges = 0, inserts = 0, deletes = 0;
    double ms_total = 0, ms_index = 0, ms_match = 0;
    uint32_t empty_steps = 0, steps = 0;
};

// Bump allocator over one device allocation (the reference's MemPool, utils/mem_pool.h:8-37, sized
// at build time from the workload instead of fixed compile-time constants).
class DevicePool {
  public:
    ~DevicePool();
    std::string reserve(uint64_t bytes);
    template <typename T>
    T* take(uint64_t count) { return static_cast<T*>(take_bytes(count * sizeof(T))); }
    void* take_bytes(uint64_t bytes);
    uint64_t used() const { return used_; }
    uint64_t capacity() const { return cap_; }
    bool failed() const { return failed_; }

  private:
    char* base_ = nullptr;
    uint64_t cap_ = 0, used_ = 0;
    bool failed_ = false;
};

class Engine {
  public:
    Engine() = default;
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    std::string init(const EngineOptions& o);
    std::string set_graph(const DataGraph& g);
    std::string set_stream(const UpdateStream& s, uint32_t batch_size);
    std::string add_query(const Query& q, uint32_t& qid);
    std::string build();                                  // upload + offline index + summaries
    std::string plan(uint32_t qid, const Technique& t);   // plans need the summaries
    std::string run_batch(uint64_t lo, uint64_t n, std::vector<QueryResult>& out, BatchInfo& info);
    // run_batch with every lane's partial-result ring taken from the library's static device ring
    // (an equal slice per lane) instead of the session's pool; allocates nothing.  Batches over
    // static_ring_max_edges() edges are refused before any device work.
    std::string run_batch_static(uint64_t lo, uint64_t n, std::vector<QueryResult>& out, BatchInfo& info);
    uint64_t static_ring_max_edges() const;
    // In-memory batch (serving): the edges are copied to the device scratch region after the
    // resident stream and applied like a stream batch.  Rejected before any device work when the
    // adjacency arena cannot hold the lists it grows (admission).
    std::string run_edges(const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op, uint64_t n,
                          std::vector<QueryResult>& out, BatchInfo& info);
    std::string synchronize();
    // Serving loop (continuous batching of in-memory batches): requests run in submission order;
    // while one runs, the next ones are validated, keyed and copied to the device on a copy stream
    // into the other of two staging slots.  Other batch calls are refused while it runs.
    std::string serve_start();
    std::string serve_submit(const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op, uint64_t n,
                             uint64_t& ticket);
    std::string serve_wait(uint64_t ticket, std::vector<QueryResult>& out, BatchInfo& info, double& latency_ms);
    std::string serve_stop();
    // Device copy of the graph state (adjacency arena in use, list headers, label-slot directory,
    // candidate bitmaps, arena bump pointer) taken between batches, and its restoration: the session
    // then continues as if the batches after the snapshot had never been applied.
    std::string snapshot();
    std::string restore();
    // Launch settings (fields of 0 keep their current value); tune() times candidate settings on
    // the stream edges [first, first + (warmup + timed) * batch) replayed from a snapshot taken on
    // entry (it replaces any earlier snapshot), checks that every candidate's results equal the
    // first one's, leaves the graph as it found it and keeps the fastest settings.  Without
    // candidates it runs a coordinate search from the current settings over built-in value lists
    // (each setting in turn, the others held at the best so far); trail gets every evaluation.
    Tunables tunables() const;
    std::string set_tunables(const Tunables& t);
    struct TuneEval {
        Tunables t;
        double ms_per_batch = 0;
    };
    std::string tune(uint64_t first, uint64_t batch, uint32_t warmup, uint32_t timed, const std::vector<Tunables>& cands,
                     std::vector<TuneEval>& trail, uint32_t& best);

    uint32_t num_queries() const { return static_cast<uint32_t>(qs_.size()); }
    uint64_t num_batches() const;
    uint32_t batch_size() const { return batch_size_; }
    const Query& query(uint32_t q) const { return qs_[q]->q; }
    const Plan& plan_of(uint32_t q) const { return qs_[q]->plan; }
    bool planned(uint32_t q) const { return qs_[q]->planned; }
    const std::vector<uint32_t>& cardinalities(uint32_t q) const { return qs_[q]->card; }
    const std::vector<float>& avg_degrees(uint32_t q) const { return qs_[q]->avg; }
    uint64_t pool_used() const { return pool_.used(); }
    uint64_t pool_capacity() const { return pool_.capacity(); }
    const DataGraph& graph() const { return host_g_; }
    const UpdateStream& stream() const { return stream_; }
    uint32_t dv() const { return dv_; }
    // Read-back helpers for parity tests.
    std::string read_candidates(uint32_t qid, uint32_t u, std::vector<uint32_t>& bitmap);
    std::string read_adjacency(uint32_t v, uint32_t dir, std::vector<uint32_t>& ids, std::vector<uint32_t>& labels);
    uint32_t id_bits() const { return id_bits_; }
    uint64_t arena_used_words();  // words of the adjacency arena below the device bump pointer
    uint64_t arena_free_words();  // of those, the words in free blocks (reused by later relocations)
    uint64_t arena_words() const { return h_arena_words_; }
    uint32_t last_error_bits() const { return last_err_; }

    // Kernel families timed with HIP events on the engine's stream when profiling is enabled.
    enum KernelFamily : uint32_t { kKExtendCount = 0, kKExtendWrite, kKCpEnumerate, kKLocalIndex, kKBatchBuild, kKApply, kKCpPrepare,
                                   kKAnchors, kKLevelCount, kKLevelWrite, kKLevelPrepare, kKLocalCheck, kKMergeInsert, kKFamilies };
    struct KernelStat {
        uint64_t launches = 0;
        double ms = 0;
        uint64_t bytes = 0;
    };
    std::string profile_enable(uint32_t mask);
    std::string profile_read(uint32_t family, KernelStat& st);
    std::string calibration_copy(uint64_t words, double& ms);

  private:
    struct RelInfo {
        uint32_t slot = 0, dir = 0, step = 0;
        bool valid = false;
    };
    struct QState {
        Query q;
        Plan plan;
        Technique tech;
        bool planned = false;
        bool matchable = true;  // every label and edge label of the query occurs in the data
        std::vector<RelInfo> rel;           // per relation id
        std::vector<uint32_t> vrank;        // per query vertex
        uint32_t* cand = nullptr;           // n bitmaps of `words` words
        std::vector<uint32_t> card;
        std::vector<float> avg;
        uint32_t entry0 = 0, entry1 = 0;    // NLF entries [entry0, entry1)
    };

    // HIP-event timing of kernel families on one stream.
    struct Prof {
        std::vector<hipEvent_t> pool;
        size_t used = 0;
        std::vector<std::pair<uint32_t, std::pair<hipEvent_t, hipEvent_t>>> pending;
        hipEvent_t open[kKFamilies]{};
    };
    // One lane of concurrent matching steps: its own stream, local index sets, anchor pairs,
    // partial-result ring, counters, Cartesian-product scratch and pinned read-back buffer.
    struct Lane {
        hipStream_t s = nullptr;
        VSet L[kMaxQV]{};
        uint16_t L_used = 0;  // query vertices whose local set holds data
        uint32_t* tmp_list = nullptr;
        uint32_t* tmp_len = nullptr;
        uint32_t* pairs = nullptr;
        uint32_t* n_pairs = nullptr;
        uint32_t* ring = nullptr;
        uint64_t ring_words = 0;
        uint32_t* counters = nullptr;  // [0]: input count, [1]: output count, [2]: overflow, [3]: finishing count, [8..9]: u64 work counter
        CpScratch cp{};
        LevelScratch level{};          // offsets, hashes and scan scratch shared with cp
        uint32_t* h_pinned = nullptr;
        uint32_t* chain_ctr = nullptr; // [2 * kMaxQV] per-level {count, overflow} of a device chain
        hipEvent_t done = nullptr;     // end of the lane's work in a phase
        Prof prof;
        std::string trace_line;
        int sched = 1;                 // DFS schedule of the current step: 0 static, 1 dynamic, 2 dynamic + range donation
        Donation dn{};                 // range words of the depth-first kernel's waves (schedule 2)
        unsigned long long* work_counter() { return reinterpret_cast<unsigned long long*>(counters + 8); }
    };

    // Local index of many steps at once: each slot holds one step's local sets, anchor pairs and
    // expansion queue; slots_ holds two groups of group_ slots, so the main thread can build the
    // next group's local index while the lanes search the steps of the current one.
    struct Slot {
        VSet L[kMaxQV]{};
        uint32_t* pairs = nullptr;
        uint32_t* n_pairs = nullptr;
        uint32_t* queue = nullptr;
        uint32_t* counters = nullptr;  // [kSlotWords]: set lengths [0, 16), pairs [16], queue lengths [kQueueCounters, +16)
        uint16_t used = 0;             // sets holding data (cleared before the slot is reused)
    };
    struct GroupStep {
        uint32_t q = 0, i = 0;  // query, query edge
        int slot = -1;          // -1: the lane builds the local index itself (run_step)
        uint32_t built = 0, np = 0;
    };
    std::string stage_group(uint32_t ph, std::vector<GroupStep>& steps);
    std::vector<Slot> slots_;
    uint32_t group_ = 0;  // steps per group; 0: every step builds its own local index on its lane
    AnchorJob* d_anchor_jobs_ = nullptr;
    LocalJob* d_local_jobs_ = nullptr;   // kMaxQV regions of kMaxLocalJobs jobs
    VSet* d_clear_sets_ = nullptr;
    uint8_t* h_jobs_ = nullptr;          // pinned staging of the job arrays, same layout, then counter read-back
    uint32_t* slot_ctr_ = nullptr;       // every slot's counters, kSlotWords each, contiguous

    std::string run_resident(uint64_t lo, uint64_t n, const uint32_t* hsrc, const uint32_t* hdst, const uint8_t* hop,
                             std::vector<QueryResult>& out, BatchInfo& info);
    std::string run_phase(uint32_t ph);
    std::string run_step(Lane& w, QState& qs, uint32_t i, uint32_t ph);
    // the stages of run_step, so that a team can run each stage for all its steps before one wait
    void step_anchors(Lane& w, QState& qs, uint32_t i, uint32_t ph);
    uint32_t step_local(Lane& w, QState& qs, uint32_t i, uint32_t ph);
    std::string enumerate_step(Lane& w, QState& qs, uint32_t i, uint32_t ph, const VSet* L, uint32_t built, const uint32_t* pairs,
                               const uint32_t* n_pairs, uint32_t np);
    std::string enumerate_begin(Lane& w, QState& qs, uint32_t i, uint32_t ph, const VSet* L, uint32_t built, const uint32_t* pairs,
                                const uint32_t* n_pairs, uint32_t np, StepDesc& d, bool& pending, uint32_t& count, InitJob* defer = nullptr);
    static std::string initial_count(const uint32_t* counters, uint32_t& count);
    std::string enumerate_rest(Lane& w, QState& qs, uint32_t i, uint32_t ph, const StepDesc& d, uint32_t count);
    std::string search_chained(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t count, uint32_t res_slot);
    uint32_t chain_build(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t count, LevelChain& c);
    MatchOut chain_out(const StepDesc& d, uint32_t res_slot);
    std::string chain_finish(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t count, uint32_t res_slot, const LevelChain& c,
                             uint32_t L, const uint32_t* ctr);
    uint32_t slot_of(const QState& qs) const;
    static int sched_of(SearchStrategy st);
    // Teams: lanes [t * team_, (t + 1) * team_) run up to team_ steps at a time on the first lane's
    // stream from one host thread, each stage for all of them before one wait, and their level
    // chains as one chain of multi-job launches (ChainJob tables at d_chain_jobs_ + first lane).
    std::string run_team_steps(size_t first_lane, std::vector<GroupStep>& items, uint32_t ph);
    std::string team_local(size_t first_lane, std::vector<GroupStep>& items, uint32_t ph, std::vector<uint32_t>& np,
                           std::vector<uint32_t>& built);
    uint32_t team_ = 1;
    // Captured level chains of the teams (graph_chains), by launch shape.
    struct ChainGraphKey {
        size_t first_lane;
        uint32_t njobs, wmask, cmask, blocks;
        bool operator<(const ChainGraphKey& o) const {
            return std::tie(first_lane, njobs, wmask, cmask, blocks) < std::tie(o.first_lane, o.njobs, o.wmask, o.cmask, o.blocks);
        }
    };
    std::map<ChainGraphKey, hipGraphExec_t> chain_graphs_;
    std::mutex chain_graphs_mu_;
    // The teams' multi-step jobs: per lane one anchor job, one initial-results job, kMaxQV local
    // jobs and kMaxQV sets to clear; a team uses the regions of its lanes (device arrays and their
    // pinned staging, same layout).
    uint32_t* lens_all_ = nullptr;       // every lane's kLaneWords counters, lane after lane
    AnchorJob* d_tajobs_ = nullptr;
    LocalJob* d_tljobs_ = nullptr;
    VSet* d_tclear_ = nullptr;
    AnchorJob* h_tajobs_ = nullptr;
    LocalJob* h_tljobs_ = nullptr;
    VSet* h_tclear_ = nullptr;
    uint32_t* h_tlens_ = nullptr;
    uint32_t* chain_ctr_all_ = nullptr;  // every lane's 2 * kMaxQV chain counters, lane after lane
    uint32_t* h_tctr_ = nullptr;         // their pinned read-back, same layout
    InitJob* d_tinit_ = nullptr;         // one initial-results job per lane, and its pinned staging
    InitJob* h_tinit_ = nullptr;
    ChainJob* d_chain_jobs_ = nullptr;   // one per lane
    ChainJob* h_chain_jobs_ = nullptr;   // pinned staging, one per lane
    std::string search(Lane& w, QState& qs, const StepDesc& d, uint32_t i, const uint32_t* rec, uint32_t count, uint32_t depth,
                       uint64_t top, uint32_t target, uint32_t res_slot);
    std::string extend_to(Lane& w, const StepDesc& d, const uint32_t* rec, uint32_t count, uint32_t depth, uint32_t end, uint32_t* out,
                          uint64_t cap, bool& ok, uint32_t& n_out);
    // any_depth: count depth-first whatever the stack depth (a single record that no level holds)
    std::string finish_records(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t depth, uint32_t count, const uint32_t* rec,
                               uint32_t res_slot, bool any_depth = false);
    uint32_t view_of(const QState& qs, uint32_t rel, uint32_t i, uint32_t ph) const;
    uint32_t key_of_edge(uint32_t nbr, uint32_t elabel_raw) const;
    std::string check_err(const char* where);
    static uint32_t read_u32(hipStream_t st, uint32_t* pinned, const uint32_t* p);

    EngineOptions opt_;
    hipStream_t s_ = nullptr;
    DevicePool pool_;
    bool built_ = false;

    // host copies
    DataGraph host_g_;
    UpdateStream stream_;
    uint32_t batch_size_ = 0;
    std::vector<std::unique_ptr<QState>> qs_;

    // label ranks
    std::vector<std::pair<uint32_t, uint32_t>> vlabel_rank_, elabel_rank_;  // sorted (raw, rank)
    uint32_t nvl_ = 0, nel_ = 0, id_bits_ = 1, dv_ = 0, words_ = 0;
    int ndirs_ = 1;

    // device state
    DevGraph g_{};
    BatchInput bin_{};
    uint64_t stream_cap_ = 0;       // edges the device stream buffers hold
    uint32_t max_batch_ = 0;        // edges per batch the scratch is sized for
    BatchScratch bs_[2]{};
    NlfEntry* d_entries_ = nullptr;
    NlfGroup* d_groups_ = nullptr;
    uint32_t n_entries_ = 0;
    std::vector<NlfEntry> h_entries_;
    std::vector<NlfGroup> h_groups_;
    std::vector<Lane> lanes_;
    std::vector<hipStream_t> lane_streams_;   // created at init, before the main stream
    std::vector<double> step_cost_[2];        // per phase, per (query, edge): host seconds in the last batch
    // host mirrors of list degrees and capacities per direction, and of the arena bump pointer
    uint64_t h_arena_words_ = 0;
    unsigned long long* adm_ = nullptr;  // [8] batch admission words (batch_admit)
    uint64_t ring_words_ = 0;                 // words of each lane's ring
    uint64_t ring_words_exact_ = 0;           // CSM_RING_WORDS: ring size without the 1M-word minimum
    hipEvent_t phase_start_ = nullptr;
    unsigned long long* results_ = nullptr;   // per query: pos, pos_hash, neg, neg_hash
    unsigned long long* summary_ = nullptr;   // [2]
    uint32_t* h_pinned_ = nullptr;            // small pinned read-back buffer (main stream)
    uint32_t* static_ring_ = nullptr;         // this device's address of the static ring (scratch.hip)
    uint64_t free_entries_ = 0;               // entries of the free-block lists (all classes)
    std::vector<uint2> h_free_cls_;           // {first entry, capacity} of each size class
    int local_grid_ = 128;                    // blocks of the local-index kernels (set at build)
    uint32_t last_err_ = 0;
    std::string failed_;  // set when a batch failed after changing the graph: every later batch is refused
    // serving loop state (serve.cpp)
    struct ServeReq {
        uint64_t id = 0, n = 0;
        std::vector<uint32_t> src, dst, elabel;
        std::vector<uint8_t> op;
        int slot = -1;
        bool done = false;
        std::string err;
        std::vector<QueryResult> out;
        BatchInfo info;
        std::chrono::steady_clock::time_point t_submit, t_done;
    };
    struct Server {
        std::mutex mu;
        std::condition_variable cv;
        std::deque<std::shared_ptr<ServeReq>> incoming, staged;
        std::map<uint64_t, std::shared_ptr<ServeReq>> open;  // submitted and not yet waited for
        bool slot_busy[2] = {false, false};
        bool stop = false;
        uint64_t next_id = 1;
        hipStream_t copy = nullptr;
        hipEvent_t ready[2] = {nullptr, nullptr};
        uint8_t* pinned[2] = {nullptr, nullptr};  // staging: src, dst, key_fwd, key_bwd (u32 each), op
        std::thread stager, runner;
    };
    std::unique_ptr<Server> srv_;
    void serve_stage_loop();
    void serve_run_loop();
    std::string stage_keys(const ServeReq& r, uint32_t* kf, uint32_t* kb) const;
    bool knows_elabel(uint32_t raw) const;
    // snapshot(): one device allocation outside the pool holding the graph state
    char* snap_ = nullptr;
    uint64_t snap_bytes_ = 0, snap_top_ = 0;

    // profiling
    void prof_begin(Prof& P, hipStream_t st, uint32_t family);
    void prof_end(Prof& P, hipStream_t st, uint32_t family);
    std::string prof_collect();
    uint32_t prof_mask_ = 0;
    KernelStat kstat_[kKFamilies];
    Prof prof_main_;
    unsigned long long* prof_bytes_ = nullptr;  // device counters per family
    unsigned long long* bytes_ptr(uint32_t family) { return (prof_mask_ >> family & 1u) ? prof_bytes_ + family : nullptr; }
    // Per-step trace (environment variable CSM_TRACE = output path): one JSON line per step.
    FILE* trace_ = nullptr;
    std::mutex trace_mu_;
};

}  // namespace csm
