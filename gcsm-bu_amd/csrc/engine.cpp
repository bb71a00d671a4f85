#include "engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <numeric>
#include <thread>

namespace csm {

#define CSM_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return std::string("HIP error: ") + hipGetErrorString(e_) + " at " #call; \
    } while (0)

// ---- pool -------------------------------------------------------------------------------------
DevicePool::~DevicePool() {
    if (base_) (void)hipFree(base_);
}

std::string DevicePool::reserve(uint64_t bytes) {
    if (base_) return tagged(Code::kState, "device pool already reserved");
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&base_), bytes);
    if (e != hipSuccess) {
        base_ = nullptr;
        return tagged(Code::kMemory, std::string("device pool allocation of ") + std::to_string(bytes >> 20) + " MiB failed: " + hipGetErrorString(e));
    }
    cap_ = bytes;
    used_ = 0;
    return "";
}

void* DevicePool::take_bytes(uint64_t bytes) {
    const uint64_t aligned = (used_ + 255) & ~255ull;
    if (aligned + bytes > cap_) {
        failed_ = true;
        return nullptr;
    }
    used_ = aligned + bytes;
    return base_ + aligned;
}

// ---- engine -----------------------------------------------------------------------------------
Engine::~Engine() {
    (void)serve_stop();
    if (s_) (void)hipStreamSynchronize(s_);
    for (Lane& w : lanes_)
        if (w.s) (void)hipStreamSynchronize(w.s);
    for (hipEvent_t e : prof_main_.pool) (void)hipEventDestroy(e);
    for (Lane& w : lanes_) {
        for (hipEvent_t e : w.prof.pool) (void)hipEventDestroy(e);
        if (w.done) (void)hipEventDestroy(w.done);
        if (w.h_pinned) (void)hipHostFree(w.h_pinned);
    }
    for (auto& kv : chain_graphs_) (void)hipGraphExecDestroy(kv.second);
    for (hipStream_t st : lane_streams_) (void)hipStreamDestroy(st);
    if (phase_start_) (void)hipEventDestroy(phase_start_);
    if (snap_) (void)hipFree(snap_);
    if (trace_) std::fclose(trace_);
    if (h_pinned_) (void)hipHostFree(h_pinned_);
    if (h_jobs_) (void)hipHostFree(h_jobs_);
    if (h_chain_jobs_) (void)hipHostFree(h_chain_jobs_);
    if (h_tajobs_) (void)hipHostFree(h_tajobs_);
    if (s_) (void)hipStreamDestroy(s_);
}

std::string Engine::init(const EngineOptions& o) {
    opt_ = o;
    int n = 0;
    CSM_HIP(hipGetDeviceCount(&n));
    if (n == 0) return tagged(Code::kDevice, "no HIP device available");
    if (o.device < 0 || o.device >= n) return tagged(Code::kArgument, "device id out of range");
    CSM_HIP(hipSetDevice(o.device));
    if (const char* t = std::getenv("CSM_LANES")) opt_.lanes = static_cast<uint32_t>(std::max(1, std::atoi(t)));
    // The lanes' streams first, then the main stream: the runtime gives streams hardware queues in
    // creation order (new queues while it has fewer than its maximum, then the least used one), so
    // the first lanes (the teams' streams) get queues of their own also in a fresh process.
    lane_streams_.resize(std::max<uint32_t>(1, std::min<uint32_t>(opt_.lanes, kMaxLanes)));
    for (hipStream_t& st : lane_streams_) CSM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    CSM_HIP(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    CSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_pinned_), 4096, hipHostMallocDefault));
    CSM_HIP(hipEventCreateWithFlags(&phase_start_, hipEventDisableTiming));
    CSM_HIP(static_ring(o.device, &static_ring_));
    if (const char* t = std::getenv("CSM_TRACE")) trace_ = std::fopen(t, "a");
    if (const char* t = std::getenv("CSM_TEAM")) opt_.team = static_cast<uint32_t>(std::max(1, std::atoi(t)));
    // A/B switches of the scheduling and layout aids (all on by default)
    if (const char* t = std::getenv("CSM_SPLIT_ROOTS")) if (std::atoi(t) == 0) opt_.waves_per_anchor = 1;
    if (const char* t = std::getenv("CSM_SLOT_DIRECTORY")) opt_.slot_directory = std::atoi(t) != 0;
    if (const char* t = std::getenv("CSM_CHAIN_LEVELS")) opt_.chain_levels = std::atoi(t) != 0;
    if (const char* t = std::getenv("CSM_GRAPH_CHAINS")) opt_.graph_chains = std::atoi(t) != 0;
    if (const char* t = std::getenv("CSM_ARENA_SLACK_WORDS")) opt_.arena_slack_words = std::strtoull(t, nullptr, 10);
    if (const char* t = std::getenv("CSM_CP_MAX_RECORDS")) opt_.cp_max_records = static_cast<uint32_t>(std::max(1024L, std::atol(t)));
    // test hook: an exact (possibly tiny) partial-result ring per lane, in words
    if (const char* t = std::getenv("CSM_RING_WORDS")) ring_words_exact_ = std::max<uint64_t>(8, std::strtoull(t, nullptr, 10)) & ~1ull;
    return "";
}

std::string Engine::set_graph(const DataGraph& g) {
    if (built_) return tagged(Code::kState, "graph already uploaded");
    host_g_ = g;
    return "";
}

std::string Engine::set_stream(const UpdateStream& s, uint32_t batch_size) {
    if (built_) return tagged(Code::kState, "stream must be set before the index is built");
    if (batch_size == 0) return tagged(Code::kArgument, "batch size must be positive");
    stream_ = s;
    batch_size_ = batch_size;
    return "";
}

std::string Engine::add_query(const Query& q, uint32_t& qid) {
    if (built_) return tagged(Code::kState, "queries must be added before the index is built");
    auto st = std::make_unique<QState>();
    st->q = q;
    qid = static_cast<uint32_t>(qs_.size());
    qs_.push_back(std::move(st));
    return "";
}

uint64_t Engine::num_batches() const {
    if (!batch_size_) return 0;
    return (stream_.size() + batch_size_ - 1) / batch_size_;
}

namespace {
uint32_t rank_of(const std::vector<std::pair<uint32_t, uint32_t>>& t, uint32_t raw) {
    auto it = std::lower_bound(t.begin(), t.end(), std::make_pair(raw, 0u));
    return (it != t.end() && it->first == raw) ? it->second : kNoVertex;
}
std::vector<std::pair<uint32_t, uint32_t>> make_ranks(std::vector<uint32_t> v) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
    std::vector<std::pair<uint32_t, uint32_t>> t;
    for (uint32_t i = 0; i < v.size(); i++) t.emplace_back(v[i], i);
    return t;
}
uint32_t cap_for(uint32_t deg) {  // max(8, 2^(ceil(log2 deg) + 1)), deg 0 -> 8 (indexing.cu:46-59)
    uint32_t p = 1;
    while (p < deg) p <<= 1;
    return std::max<uint32_t>(kMinNbrSize, deg == 0 ? 0u : 2u * p);
}
}  // namespace

bool Engine::knows_elabel(uint32_t raw) const { return rank_of(elabel_rank_, raw) != kNoVertex; }

uint32_t Engine::key_of_edge(uint32_t nbr, uint32_t elabel_raw) const {
    const uint32_t er = rank_of(elabel_rank_, elabel_raw);
    const uint32_t slot = er * nvl_ + rank_of(vlabel_rank_, host_g_.vlabel[nbr]);
    return (slot << id_bits_) | nbr;
}

uint32_t Engine::read_u32(hipStream_t st, uint32_t* pinned, const uint32_t* p) {
    (void)hipMemcpyAsync(pinned, p, 4, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    return pinned[0];
}

std::string Engine::build() {
    if (built_) return tagged(Code::kState, "index already built");
    if (host_g_.dv == 0) return tagged(Code::kState, "no data graph loaded");
    if (qs_.empty()) return tagged(Code::kState, "no query loaded");
    host_g_.directed = opt_.directed;
    ndirs_ = opt_.directed ? 2 : 1;
    dv_ = host_g_.dv;
    words_ = (dv_ + 31) / 32;

    // label ranks over the graph, the stream and the queries
    std::vector<uint32_t> vl, el;
    for (uint32_t v = 0; v < dv_; v++) if (host_g_.vlabel[v] != kNoVertex) vl.push_back(host_g_.vlabel[v]);
    el = host_g_.elabel;
    el.insert(el.end(), stream_.elabel.begin(), stream_.elabel.end());
    for (auto& qs : qs_) {
        for (uint32_t x : qs->q.vlabel_raw) vl.push_back(x);
        for (uint32_t x : qs->q.elabel_raw) el.push_back(x);
    }
    vlabel_rank_ = make_ranks(vl);
    elabel_rank_ = make_ranks(el);
    nvl_ = std::max<uint32_t>(1, static_cast<uint32_t>(vlabel_rank_.size()));
    nel_ = std::max<uint32_t>(1, static_cast<uint32_t>(elabel_rank_.size()));
    id_bits_ = 1;
    while ((1ull << id_bits_) < dv_) id_bits_++;
    const uint64_t nslots = static_cast<uint64_t>(nvl_) * nel_;
    if ((nslots << id_bits_) > (1ull << 32))
        return tagged(Code::kLimit, "label ranks times vertex ids exceed 32-bit adjacency keys (too many labels for this vertex count)");

    max_batch_ = opt_.max_batch_edges ? opt_.max_batch_edges : std::max<uint32_t>(1, std::min<uint64_t>(batch_size_, stream_.size()));
    if (batch_size_ > 0 && batch_size_ < 0xFFFFFFFFu) max_batch_ = std::max<uint32_t>(max_batch_, std::min<uint64_t>(batch_size_, stream_.size()));
    max_batch_ = std::max<uint32_t>(max_batch_, 1);
    // The local-index kernels run beside the other teams' kernels on fixed grids (their frontier
    // sizes are device-side; each launch splits the frontier over its waves itself).  Measured with
    // 32 lanes in teams of 8: 256 four-wave blocks with 65,536-edge batches on a 400K-vertex graph
    // (18.6 ms per batch; 128: 19.0; also 64 standing queries with 64 lanes: 74.6 against 75.7 ms),
    // 512 with 1,048,576-edge batches on a 4M-vertex graph (34.1 ms; 256 and 1024: 35.3, 128: 38.2).
    local_grid_ = max_batch_ >= (1u << 18) ? 512 : 256;
    if (const char* t = std::getenv("CSM_LOCAL_GRID")) local_grid_ = std::max(1, std::atoi(t));

    // ---- host CSR of the initial graph ----
    std::vector<std::vector<uint32_t>> deg(ndirs_, std::vector<uint32_t>(dv_, 0));
    const size_t ne = host_g_.src.size();
    for (size_t k = 0; k < ne; k++) {
        deg[0][host_g_.src[k]]++;
        deg[opt_.directed ? 1 : 0][host_g_.dst[k]]++;
    }
    std::vector<std::vector<uint64_t>> off(ndirs_, std::vector<uint64_t>(dv_ + 1, 0));
    uint64_t total_cap = 0;
    for (int d = 0; d < ndirs_; d++)
        for (uint32_t v = 0; v < dv_; v++) {
            off[d][v] = total_cap;
            total_cap += cap_for(deg[d][v]);
        }
    uint64_t stream_arcs = 2ull * stream_.size();
    // growth room: three times the initial lists, eight words per streamed arc, and the same for
    // in-memory batches when the caller sized the scratch for them (max_batch_edges)
    const uint64_t live_arcs = opt_.max_batch_edges ? 2ull * opt_.max_batch_edges * kLiveBatchesOfRoom : 0;
    const uint64_t arena_words = 3 * total_cap + 8 * (stream_arcs + live_arcs) + opt_.arena_slack_words;
    if (arena_words >= (1ull << 32)) return tagged(Code::kLimit, "adjacency arena exceeds 2^32 words");

    // ---- sizes of everything else, then one reservation ----
    const uint64_t max_arcs = 2ull * max_batch_;
    ring_words_ = ring_words_exact_ ? ring_words_exact_ : std::max<uint64_t>(opt_.ring_bytes / 4, 1u << 20) & ~1ull;
    const uint64_t cp_rec = std::max<uint32_t>(1024, opt_.cp_max_records);
    const uint64_t cp_list_words = std::max<uint64_t>(opt_.cp_list_bytes / 4, 1u << 20);
    uint64_t need = 0;
    auto add = [&](uint64_t bytes) { need += (bytes + 255) & ~255ull; };
    add(arena_words * 4);
    add(dv_ * 4ull);                                            // vrank
    for (int d = 0; d < ndirs_; d++) {
        add(dv_ * 8ull); add(dv_ * 4ull);                       // adj, cap
        add(dv_ * 16ull); add(max_arcs * 4); add(max_arcs * 4); // delta seg, ins, del
        add(dv_ * 12ull);                                       // cnt_ins, cnt_del, tflag
        add(max_arcs * 4); add(16); add(4 * (max_arcs + 1) * 4); add(max_arcs * 4);
        add(scan_scratch_elems(static_cast<uint32_t>(max_arcs)) * 4); add(16); add(2 * kMaxBigSegments * 4);
        add(max_arcs * 4); add(2 * (nvl_ + 1) * 4ull); add(scan_scratch_elems(nvl_ + 1) * 4);  // label buckets
    }
    add(stream_.size() * 17ull + 2 * max_batch_ * 17ull);      // stream + two live-batch slots
    const uint32_t n_lanes = std::max<uint32_t>(1, std::min<uint32_t>(opt_.lanes, kMaxLanes));
    for (uint32_t l = 0; l < n_lanes; l++) {  // per lane: mirrors the takes of alloc below
        add(kMaxQV * (words_ * 4ull + dv_ * 4ull + 512) + 128);  // local sets
        add(dv_ * 4ull + 16);                                      // tmp list
        add(max_arcs * 8 + 16);                                    // anchor pairs
        add(ring_words_ * 4); add(256);
        add((cp_rec + 1) * 8); add(cp_rec * 8); add(scan_scratch_elems(static_cast<uint32_t>(cp_rec)) * 8);
        add((cp_rec + 1) * 8); add(cp_rec * 16); add(cp_rec * 4);  // CP segments (runs shared with levels)
        add(kLevelRunsPerRecord * cp_rec * 16);                    // level: per-entry runs
        add(2 * kMaxQV * 4);                                       // level chain counters
        add((cp_rec / kLevelTile + 2) * 8);                        // level tile totals
        add(cp_list_words * 4); add(64);
        add(64 << 10);                                             // alignment slack of the takes
        add(sizeof(ChainJob) + 256);                               // the lane's level-chain job
        add(kMaxDfsBlocks * 4ull * 12 + 512);                      // range donation words
        add(kLaneWords * 4 + sizeof(AnchorJob) + sizeof(InitJob) + kMaxQV * (sizeof(LocalJob) + sizeof(VSet)) + 2 * kMaxQV * 4 + 1024);  // team jobs
    }
    // local-index slots: two groups of group_ steps (each: its sets, anchor pairs, queue, counters)
    uint32_t max_qn = 2;
    for (auto& qs : qs_) max_qn = std::max(max_qn, qs->q.n);
    const uint64_t slot_bytes = max_qn * (words_ * 4ull + dv_ * 4ull + 512) + 2 * max_arcs * 4 + dv_ * 4ull + kSlotWords * 4 + (16 << 10);
    {
        // Measured with eight lanes: building the local index of many steps together pays when a
        // phase has many steps (64 standing queries, ~500 steps: 107.9 against 120.6 ms per batch
        // with groups of 64), not with six queries (4M-vertex graph: 61.4 against 60.2 ms with
        // groups of 15; 400K-vertex graph: 25.0 against 21.6 ms with groups of 64).
        uint32_t nsteps = 0;
        for (auto& qs : qs_) nsteps += qs->q.m;
        uint64_t want = nsteps >= kGroupMinSteps ? kMaxGroup : 0;
        if (const char* t = std::getenv("CSM_GROUP_STEPS")) want = static_cast<uint64_t>(std::max(0L, std::atol(t)));
        if (opt_.team > 1) want = 0;  // teams build their steps' local indexes themselves
        uint64_t budget = kGroupBudgetBytes;
        if (opt_.pool_bytes) budget = std::min<uint64_t>(budget, opt_.pool_bytes / 8);
        group_ = static_cast<uint32_t>(std::min<uint64_t>(want, std::min<uint64_t>(kMaxGroup, budget / (2 * slot_bytes))));
    }
    if (group_) {
        add(2ull * group_ * slot_bytes);
        add(kMaxGroup * sizeof(AnchorJob)); add(kMaxQV * kMaxLocalJobs * sizeof(LocalJob));
        add(2ull * kMaxGroup * kMaxQV * sizeof(VSet));
    }
    for (auto& qs : qs_) add(static_cast<uint64_t>(qs->q.n) * words_ * 4);
    add(qs_.size() * 32 + 64);
    add(kKFamilies * 8);
    add(64 * 8);                                                // batch admission words
    // free arena blocks: per size class 2^c (c >= 3) at most min(2^20, arena / 2^c) entries, twice
    uint64_t free_entries = 0;
    for (uint32_t c = 3; c < kFreeClasses; c++) free_entries += std::min<uint64_t>(1u << 20, arena_words >> c);
    add(free_entries * 4); add(free_entries * 4); add(kFreeClasses * 16);
    add(kMaxQV * qs_.size() * sizeof(NlfEntry) + kMaxRel * qs_.size() * sizeof(NlfGroup) + 512);
    const uint64_t sdir_bytes = (nslots + 1) * dv_ * 4ull;
    const bool use_sdir = opt_.slot_directory && sdir_bytes * ndirs_ <= kMaxSlotDirectoryBytes;
    if (use_sdir)
        for (int d = 0; d < ndirs_; d++) add(sdir_bytes);
    need += 1 << 20;
    if (opt_.pool_bytes && need > opt_.pool_bytes)
        return tagged(Code::kMemory, "device pool budget too small: need " + std::to_string(need >> 20) + " MiB, budget " + std::to_string(opt_.pool_bytes >> 20) + " MiB");
    std::string err = pool_.reserve(need);
    if (!err.empty()) return err;

    // ---- graph ----
    g_.dv = dv_;
    g_.id_bits = id_bits_;
    g_.id_mask = id_bits_ >= 32 ? 0xFFFFFFFFu : ((1u << id_bits_) - 1u);
    g_.nvl = nvl_;
    g_.nslots = static_cast<uint32_t>(nslots);
    g_.sdir[0] = g_.sdir[1] = nullptr;
    uint32_t* d_vrank = pool_.take<uint32_t>(dv_);
    g_.vrank = d_vrank;
    g_.arena = pool_.take<uint32_t>(arena_words);
    g_.arena_words = arena_words;
    g_.err = nullptr;
    {
        std::vector<uint32_t> vr(dv_);
        for (uint32_t v = 0; v < dv_; v++) vr[v] = host_g_.vlabel[v] == kNoVertex ? kNoVertex : rank_of(vlabel_rank_, host_g_.vlabel[v]);
        CSM_HIP(hipMemcpyAsync(d_vrank, vr.data(), dv_ * 4ull, hipMemcpyHostToDevice, s_));
        h_arena_words_ = arena_words;
        std::vector<uint32_t> arena(total_cap, 0u);
        std::vector<std::vector<uint64_t>> pos = off;
        auto key = [&](uint32_t nbr, uint32_t el) {
            return ((rank_of(elabel_rank_, el) * nvl_ + vr[nbr]) << id_bits_) | nbr;
        };
        for (size_t k = 0; k < ne; k++) {
            const uint32_t a = host_g_.src[k], b = host_g_.dst[k], l = host_g_.elabel[k];
            arena[pos[0][a]++] = key(b, l);
            arena[pos[opt_.directed ? 1 : 0][b]++] = key(a, l);
        }
        const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++)
            th.emplace_back([&, t]() {
                for (int d = 0; d < ndirs_; d++)
                    for (uint32_t v = t; v < dv_; v += nth) {
                        auto* b = arena.data() + off[d][v];
                        std::sort(b, b + deg[d][v]);
                    }
            });
        for (auto& x : th) x.join();
        for (int d = 0; d < ndirs_; d++)
            for (uint32_t v = 0; v < dv_; v++)
                if (deg[d][v] > 1)
                    for (uint64_t i = off[d][v] + 1; i < off[d][v] + deg[d][v]; i++)
                        if (arena[i] == arena[i - 1]) return tagged(Code::kFormat, "duplicate edge in the data graph");
        CSM_HIP(hipMemcpyAsync(g_.arena, arena.data(), total_cap * 4, hipMemcpyHostToDevice, s_));
        for (int d = 0; d < ndirs_; d++) {
            g_.dir[d].adj = pool_.take<uint2>(dv_);
            g_.dir[d].cap = pool_.take<uint32_t>(dv_);
            std::vector<uint2> adj(dv_);
            std::vector<uint32_t> cap(dv_);
            for (uint32_t v = 0; v < dv_; v++) {
                adj[v] = make_uint2(static_cast<uint32_t>(off[d][v]), deg[d][v]);
                cap[v] = cap_for(deg[d][v]);
            }
            CSM_HIP(hipMemcpyAsync(g_.dir[d].adj, adj.data(), dv_ * 8ull, hipMemcpyHostToDevice, s_));
            CSM_HIP(hipMemcpyAsync(g_.dir[d].cap, cap.data(), dv_ * 4ull, hipMemcpyHostToDevice, s_));
            g_.delta[d].seg = pool_.take<uint4>(dv_);
            g_.delta[d].ins = pool_.take<uint32_t>(max_arcs);
            g_.delta[d].del = pool_.take<uint32_t>(max_arcs);
            CSM_HIP(hipMemsetAsync(g_.delta[d].seg, 0, dv_ * 16ull, s_));
            BatchScratch& w = bs_[d];
            w.cnt_ins = pool_.take<uint32_t>(dv_);
            w.cnt_del = pool_.take<uint32_t>(dv_);
            w.tflag = pool_.take<uint32_t>(dv_);
            CSM_HIP(hipMemsetAsync(w.cnt_ins, 0, dv_ * 4ull, s_));
            CSM_HIP(hipMemsetAsync(w.cnt_del, 0, dv_ * 4ull, s_));
            CSM_HIP(hipMemsetAsync(w.tflag, 0, dv_ * 4ull, s_));
            w.touched = pool_.take<uint32_t>(max_arcs);
            w.n_touched = pool_.take<uint32_t>(4);
            w.n_big = w.n_touched + 1;
            CSM_HIP(hipMemsetAsync(w.n_touched, 0, 16, s_));
            w.tmp_ins = pool_.take<uint32_t>(max_arcs + 1);
            w.tmp_del = pool_.take<uint32_t>(max_arcs + 1);
            w.off_ins = pool_.take<uint32_t>(max_arcs + 1);
            w.off_del = pool_.take<uint32_t>(max_arcs + 1);
            w.ins_pos = pool_.take<uint32_t>(max_arcs);
            w.scan_tmp = pool_.take<uint32_t>(scan_scratch_elems(static_cast<uint32_t>(max_arcs)));
            w.big_ptr = pool_.take<uint32_t>(kMaxBigSegments);
            w.big_len = pool_.take<uint32_t>(kMaxBigSegments);
            w.by_label = pool_.take<uint32_t>(max_arcs);
            w.label_off = pool_.take<uint32_t>(nvl_ + 1);
            w.label_cur = pool_.take<uint32_t>(nvl_ + 1);
            w.label_tmp = pool_.take<uint32_t>(scan_scratch_elems(nvl_ + 1));
            CSM_HIP(hipMemsetAsync(w.label_off, 0, (nvl_ + 1) * 4ull, s_));
        }
        if (use_sdir)
            for (int d = 0; d < ndirs_; d++) {
                g_.sdir[d] = pool_.take<uint32_t>((nslots + 1) * dv_);
                slot_dir_update(g_, d, nullptr, nullptr, 0, s_);
            }
        if (ndirs_ == 1) { g_.dir[1] = g_.dir[0]; g_.delta[1] = g_.delta[0]; bs_[1] = bs_[0]; g_.sdir[1] = g_.sdir[0]; }
        unsigned long long* top = pool_.take<unsigned long long>(4);
        g_.arena_top = top;
        uint32_t* errp = reinterpret_cast<uint32_t*>(top + 2);
        g_.err = errp;
        h_pinned_[0] = static_cast<uint32_t>(total_cap);
        h_pinned_[1] = static_cast<uint32_t>(total_cap >> 32);
        h_pinned_[2] = 0; h_pinned_[3] = 0; h_pinned_[4] = 0; h_pinned_[5] = 0;
        CSM_HIP(hipMemcpyAsync(top, h_pinned_, 32, hipMemcpyHostToDevice, s_));
        CSM_HIP(hipStreamSynchronize(s_));
    }

    // ---- stream ----
    {
        stream_cap_ = stream_.size();
        uint32_t* sb = pool_.take<uint32_t>(4 * (stream_cap_ + 2ull * max_batch_));
        uint8_t* ob = pool_.take<uint8_t>(stream_cap_ + 2ull * max_batch_);
        if (!sb || !ob) return tagged(Code::kMemory, "device pool exhausted (stream)");
        const uint64_t S = stream_cap_ + 2ull * max_batch_;
        bin_.src = sb;
        bin_.dst = sb + S;
        bin_.key_fwd = sb + 2 * S;
        bin_.key_bwd = sb + 3 * S;
        bin_.op = ob;
        if (stream_cap_) {
            std::vector<uint32_t> kf(stream_cap_), kb(stream_cap_);
            for (uint64_t k = 0; k < stream_cap_; k++) {
                const uint32_t a = stream_.src[k], b = stream_.dst[k], l = stream_.elabel[k];
                kf[k] = ((rank_of(elabel_rank_, l) * nvl_ + rank_of(vlabel_rank_, host_g_.vlabel[b])) << id_bits_) | b;
                kb[k] = ((rank_of(elabel_rank_, l) * nvl_ + rank_of(vlabel_rank_, host_g_.vlabel[a])) << id_bits_) | a;
            }
            CSM_HIP(hipMemcpyAsync(sb, stream_.src.data(), stream_cap_ * 4, hipMemcpyHostToDevice, s_));
            CSM_HIP(hipMemcpyAsync(sb + S, stream_.dst.data(), stream_cap_ * 4, hipMemcpyHostToDevice, s_));
            CSM_HIP(hipMemcpyAsync(sb + 2 * S, kf.data(), stream_cap_ * 4, hipMemcpyHostToDevice, s_));
            CSM_HIP(hipMemcpyAsync(sb + 3 * S, kb.data(), stream_cap_ * 4, hipMemcpyHostToDevice, s_));
            CSM_HIP(hipMemcpyAsync(ob, stream_.op.data(), stream_cap_, hipMemcpyHostToDevice, s_));
            CSM_HIP(hipStreamSynchronize(s_));
        }
    }

    // ---- runtime buffers: one set per lane ----
    lanes_.assign(n_lanes, Lane{});
    // default: four teams, one per hardware queue of the runtime's default four
    team_ = std::max<uint32_t>(1, std::min<uint32_t>(opt_.team ? opt_.team : (n_lanes + 3) / 4, n_lanes));
    lens_all_ = pool_.take<uint32_t>(static_cast<uint64_t>(kLaneWords) * n_lanes);
    chain_ctr_all_ = pool_.take<uint32_t>(2ull * kMaxQV * n_lanes);
    d_tajobs_ = pool_.take<AnchorJob>(n_lanes);
    d_tinit_ = pool_.take<InitJob>(n_lanes);
    d_tljobs_ = pool_.take<LocalJob>(static_cast<uint64_t>(kMaxQV) * n_lanes);
    d_tclear_ = pool_.take<VSet>(static_cast<uint64_t>(kMaxQV) * n_lanes);
    if (!lens_all_ || !chain_ctr_all_ || !d_tajobs_ || !d_tinit_ || !d_tljobs_ || !d_tclear_) return tagged(Code::kMemory, "device pool exhausted (lane counters)");
    {
        const uint64_t hb =
            n_lanes * (sizeof(AnchorJob) + kMaxQV * sizeof(LocalJob) + kMaxQV * sizeof(VSet) + kLaneWords * 4ull + 2 * kMaxQV * 4ull + sizeof(InitJob)) +
            1024;
        uint8_t* h = nullptr;
        CSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h), hb, hipHostMallocDefault));
        h_tajobs_ = reinterpret_cast<AnchorJob*>(h);
        h_tljobs_ = reinterpret_cast<LocalJob*>(h_tajobs_ + n_lanes);
        h_tclear_ = reinterpret_cast<VSet*>(h_tljobs_ + static_cast<uint64_t>(kMaxQV) * n_lanes);
        h_tlens_ = reinterpret_cast<uint32_t*>(h_tclear_ + static_cast<uint64_t>(kMaxQV) * n_lanes);
        h_tctr_ = h_tlens_ + static_cast<uint64_t>(kLaneWords) * n_lanes;
        h_tinit_ = reinterpret_cast<InitJob*>(h_tctr_ + 2ull * kMaxQV * n_lanes);
    }
    for (Lane& w : lanes_) {
        w.s = lane_streams_[static_cast<size_t>(&w - lanes_.data())];
        CSM_HIP(hipEventCreateWithFlags(&w.done, hipEventDisableTiming));
        CSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&w.h_pinned), 4096, hipHostMallocDefault));
        for (uint32_t u = 0; u < kMaxQV; u++) {
            w.L[u].bitmap = pool_.take<uint32_t>(words_);
            w.L[u].list = pool_.take<uint32_t>(dv_);
            w.L[u].cap = dv_;
            if (!w.L[u].bitmap || !w.L[u].list) return tagged(Code::kMemory, "device pool exhausted (local sets)");
            CSM_HIP(hipMemsetAsync(w.L[u].bitmap, 0, words_ * 4ull, s_));
        }
        uint32_t* lens = lens_all_ + static_cast<uint64_t>(&w - lanes_.data()) * kLaneWords;  // set lengths, list and step counters
        for (uint32_t u = 0; u < kMaxQV; u++) w.L[u].len = lens + u;
        w.tmp_len = lens + kMaxQV;
        w.n_pairs = lens + kMaxQV + 1;
        w.counters = lens + kStepCounters;
        w.tmp_list = pool_.take<uint32_t>(dv_);
        w.pairs = pool_.take<uint32_t>(2 * max_arcs);
        w.ring = pool_.take<uint32_t>(ring_words_);
        w.ring_words = ring_words_;
        w.cp.prod = pool_.take<unsigned long long>(cp_rec + 1);
        w.cp.rec_hash = pool_.take<unsigned long long>(cp_rec);
        w.cp.scan_tmp = pool_.take<unsigned long long>(scan_scratch_elems(static_cast<uint32_t>(cp_rec)));
        w.cp.seg_off = pool_.take<unsigned long long>(cp_rec + 1);
        w.cp.seg_run = pool_.take<uint4>(cp_rec);
        w.cp.seg_cnt = pool_.take<uint32_t>(cp_rec);
        w.cp.lists = pool_.take<uint32_t>(cp_list_words);
        w.cp.lists_cap = cp_list_words;
        unsigned long long* cpc = pool_.take<unsigned long long>(4);
        if (!cpc) return tagged(Code::kMemory, "device pool exhausted (lane scratch)");
        w.cp.max_prod = cpc;
        w.cp.err = reinterpret_cast<uint32_t*>(cpc + 1);
        // level scratch: offsets, generating runs, entry tags and hashes shared with the CP scratch
        uint4* level_runs = pool_.take<uint4>(kLevelRunsPerRecord * cp_rec);
        unsigned long long* tile_sum = pool_.take<unsigned long long>(cp_rec / kLevelTile + 2);
        w.chain_ctr = chain_ctr_all_ + static_cast<uint64_t>(&w - lanes_.data()) * 2 * kMaxQV;
        w.dn.range = pool_.take<unsigned long long>(kMaxDfsBlocks * 4ull);
        w.dn.rec = pool_.take<uint32_t>(kMaxDfsBlocks * 4ull);
        if (!w.dn.range || !w.dn.rec) return tagged(Code::kMemory, "device pool exhausted (lane scratch)");
        if (!level_runs || !tile_sum || !w.chain_ctr) return tagged(Code::kMemory, "device pool exhausted (lane scratch)");
        w.level = LevelScratch{w.cp.prod, w.cp.seg_run, level_runs, reinterpret_cast<uint8_t*>(w.cp.seg_cnt), w.cp.rec_hash,
                               tile_sum, cp_rec, kLevelRunsPerRecord * cp_rec};
    }
    opt_.cp_max_records = static_cast<uint32_t>(cp_rec);
    slots_.assign(2ull * group_, Slot{});
    slot_ctr_ = group_ ? pool_.take<uint32_t>(2ull * group_ * kSlotWords) : nullptr;
    if (group_ && !slot_ctr_) return tagged(Code::kMemory, "device pool exhausted (local-index slots)");
    if (group_) CSM_HIP(hipMemsetAsync(slot_ctr_, 0, 2ull * group_ * kSlotWords * 4, s_));
    for (size_t k = 0; k < slots_.size(); k++) {
        Slot& sl = slots_[k];
        for (uint32_t u = 0; u < max_qn; u++) {
            sl.L[u].bitmap = pool_.take<uint32_t>(words_);
            sl.L[u].list = pool_.take<uint32_t>(dv_);
            sl.L[u].cap = dv_;
            if (!sl.L[u].bitmap || !sl.L[u].list) return tagged(Code::kMemory, "device pool exhausted (local-index slots)");
            CSM_HIP(hipMemsetAsync(sl.L[u].bitmap, 0, words_ * 4ull, s_));
        }
        sl.counters = slot_ctr_ + k * kSlotWords;
        sl.pairs = pool_.take<uint32_t>(2 * max_arcs);
        sl.queue = pool_.take<uint32_t>(dv_);
        if (!sl.pairs || !sl.queue) return tagged(Code::kMemory, "device pool exhausted (local-index slots)");
        for (uint32_t u = 0; u < max_qn; u++) sl.L[u].len = sl.counters + u;
        sl.n_pairs = sl.counters + kMaxQV;
    }
    if (group_) {
        d_anchor_jobs_ = pool_.take<AnchorJob>(kMaxGroup);
        d_local_jobs_ = pool_.take<LocalJob>(kMaxQV * kMaxLocalJobs);
        d_clear_sets_ = pool_.take<VSet>(2ull * kMaxGroup * kMaxQV);
        if (!d_anchor_jobs_ || !d_local_jobs_ || !d_clear_sets_) return tagged(Code::kMemory, "device pool exhausted (local-index jobs)");
        const uint64_t hbytes = kMaxGroup * sizeof(AnchorJob) + kMaxQV * kMaxLocalJobs * sizeof(LocalJob) +
                                2ull * kMaxGroup * kMaxQV * sizeof(VSet) + 2ull * kMaxGroup * kSlotWords * 4 + 4096;
        CSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_jobs_), hbytes, hipHostMallocDefault));
    }
    d_chain_jobs_ = pool_.take<ChainJob>(lanes_.size());
    if (!d_chain_jobs_) return tagged(Code::kMemory, "device pool exhausted (level-chain jobs)");
    CSM_HIP(hipHostMalloc(reinterpret_cast<void**>(&h_chain_jobs_), lanes_.size() * sizeof(ChainJob), hipHostMallocDefault));
    results_ = pool_.take<unsigned long long>(qs_.size() * 4 + 4);
    prof_bytes_ = pool_.take<unsigned long long>(kKFamilies);
    CSM_HIP(hipMemsetAsync(prof_bytes_, 0, kKFamilies * 8, s_));
    summary_ = pool_.take<unsigned long long>(4);
    adm_ = pool_.take<unsigned long long>(64);
    if (!adm_) return tagged(Code::kMemory, "device pool exhausted (admission words)");
    CSM_HIP(hipMemsetAsync(adm_, 0, 64 * 8, s_));
    {
        std::vector<uint2> cls(kFreeClasses, make_uint2(0, 0));
        uint64_t n = 0;
        for (uint32_t c = 3; c < kFreeClasses; c++) {
            const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(1u << 20, g_.arena_words >> c));
            cls[c] = make_uint2(static_cast<uint32_t>(n), k);
            n += k;
        }
        free_entries_ = n;
        g_.fl_blk = pool_.take<uint32_t>(n);
        g_.fl_pend = pool_.take<uint32_t>(n);
        g_.fl_top = pool_.take<int>(kFreeClasses);
        g_.fl_pend_n = pool_.take<uint32_t>(kFreeClasses);
        uint2* d_cls = pool_.take<uint2>(kFreeClasses);
        if (!g_.fl_blk || !g_.fl_pend || !d_cls) return tagged(Code::kMemory, "device pool exhausted (free arena blocks)");
        g_.fl_cls = d_cls;
        h_free_cls_ = cls;
        CSM_HIP(hipMemcpyAsync(d_cls, cls.data(), kFreeClasses * sizeof(uint2), hipMemcpyHostToDevice, s_));
        CSM_HIP(hipMemsetAsync(g_.fl_top, 0, kFreeClasses * 4, s_));
        CSM_HIP(hipMemsetAsync(g_.fl_pend_n, 0, kFreeClasses * 4, s_));
        CSM_HIP(hipStreamSynchronize(s_));
    }

    // ---- queries: relation slots, NLF programs, candidate bitmaps ----
    for (auto& qsp : qs_) {
        QState& qs = *qsp;
        const Query& q = qs.q;
        qs.vrank.assign(q.n, kNoVertex);
        for (uint32_t u = 0; u < q.n; u++) {
            qs.vrank[u] = rank_of(vlabel_rank_, q.vlabel_raw[u]);
            if (qs.vrank[u] == kNoVertex) qs.matchable = false;
        }
        qs.rel.assign(2 * q.m, RelInfo{});
        for (uint32_t r = 0; r < 2 * q.m; r++) {
            const auto ep = q.rel_endpoints(r);
            uint32_t el_local = 0;
            for (const auto& e : q.nbrs[ep.first]) if (e.first == ep.second) el_local = e.second;
            const uint32_t er = rank_of(elabel_rank_, q.elabel_raw[el_local]);
            RelInfo& ri = qs.rel[r];
            ri.step = q.rel_step(r);
            ri.dir = opt_.directed ? (q.rel_out[r] ? 0u : 1u) : 0u;
            ri.valid = er != kNoVertex && qs.vrank[ep.second] != kNoVertex;
            ri.slot = ri.valid ? er * nvl_ + qs.vrank[ep.second] : 0u;
            if (!ri.valid) qs.matchable = false;
        }
        qs.cand = pool_.take<uint32_t>(static_cast<uint64_t>(q.n) * words_);
        if (!qs.cand) return tagged(Code::kMemory, "device pool exhausted (candidates)");
        CSM_HIP(hipMemsetAsync(qs.cand, 0, static_cast<uint64_t>(q.n) * words_ * 4, s_));
        qs.entry0 = static_cast<uint32_t>(h_entries_.size());
        for (uint32_t u = 0; u < q.n; u++) {
            std::map<std::pair<uint32_t, uint32_t>, uint32_t> grp;  // (dir, slot) -> count
            for (uint32_t r = q.qv_offs[u]; r < q.qv_offs[u + 1]; r++)
                if (qs.rel[r].valid) grp[{qs.rel[r].dir, qs.rel[r].slot}]++;
            NlfEntry e{};
            e.vrank = qs.matchable ? qs.vrank[u] : kNoVertex;
            e.g0 = static_cast<uint32_t>(h_groups_.size());
            for (const auto& kv : grp) h_groups_.push_back(NlfGroup{kv.first.second, kv.second, kv.first.first});
            e.g1 = static_cast<uint32_t>(h_groups_.size());
            e.bitmap = qs.cand + static_cast<uint64_t>(u) * words_;
            h_entries_.push_back(e);
        }
        qs.entry1 = static_cast<uint32_t>(h_entries_.size());
    }
    n_entries_ = static_cast<uint32_t>(h_entries_.size());
    d_entries_ = pool_.take<NlfEntry>(std::max<size_t>(1, h_entries_.size()));
    d_groups_ = pool_.take<NlfGroup>(std::max<size_t>(1, h_groups_.size()));
    if (pool_.failed()) return tagged(Code::kMemory, "device pool exhausted");
    CSM_HIP(hipMemcpyAsync(d_entries_, h_entries_.data(), h_entries_.size() * sizeof(NlfEntry), hipMemcpyHostToDevice, s_));
    if (!h_groups_.empty())
        CSM_HIP(hipMemcpyAsync(d_groups_, h_groups_.data(), h_groups_.size() * sizeof(NlfGroup), hipMemcpyHostToDevice, s_));

    // ---- offline index: candidate bitmaps on G_0, then per-relation summaries ----
    nlf_full(g_, d_entries_, n_entries_, d_groups_, s_);
    for (auto& qsp : qs_) {
        QState& qs = *qsp;
        const Query& q = qs.q;
        qs.card.assign(2 * q.m, 0);
        qs.avg.assign(2 * q.m, 0.0f);
        for (uint32_t r = 0; r < 2 * q.m; r++) {
            const auto ep = q.rel_endpoints(r);
            uint64_t c = 0, ns = 0;
            if (qs.rel[r].valid && qs.matchable) {
                CSM_HIP(hipMemsetAsync(summary_, 0, 16, s_));
                relation_summary(g_, qs.rel[r].dir, qs.rel[r].slot, qs.vrank[ep.first], qs.cand + static_cast<uint64_t>(ep.first) * words_,
                                 qs.cand + static_cast<uint64_t>(ep.second) * words_, summary_, s_);
                CSM_HIP(hipMemcpyAsync(h_pinned_, summary_, 16, hipMemcpyDeviceToHost, s_));
                CSM_HIP(hipStreamSynchronize(s_));
                unsigned long long v[2];
                std::memcpy(v, h_pinned_, 16);
                c = v[0];
                ns = v[1];
            }
            // match_gpu.cu:606-607: 32-bit sums, float division (0/0 gives NaN)
            const uint32_t c32 = static_cast<uint32_t>(c), n32 = static_cast<uint32_t>(ns);
            qs.card[r] = c32;
            qs.avg[r] = static_cast<float>(c32) / static_cast<float>(n32);
        }
    }
    CSM_HIP(hipStreamSynchronize(s_));
    built_ = true;
    return check_err("build");
}

std::string Engine::plan(uint32_t qid, const Technique& t) {
    if (!built_) return tagged(Code::kState, "build the index before planning");
    if (qid >= qs_.size()) return tagged(Code::kArgument, "query id out of range");
    QState& qs = *qs_[qid];
    qs.tech = t;
    std::string err = make_plan(qs.q, t.order, qs.card.data(), qs.avg.data(), qs.plan);
    if (!err.empty()) return err;
    qs.planned = true;
    return "";
}

uint64_t Engine::arena_used_words() {
    if (!built_) return 0;
    unsigned long long top = 0;
    if (hipMemcpyAsync(h_pinned_ + 64, g_.arena_top, 8, hipMemcpyDeviceToHost, s_) != hipSuccess || hipStreamSynchronize(s_) != hipSuccess)
        return 0;
    std::memcpy(&top, h_pinned_ + 64, 8);
    return top;
}

uint64_t Engine::arena_free_words() {
    if (!built_) return 0;
    if (hipMemcpyAsync(h_pinned_ + 128, g_.fl_top, kFreeClasses * 4, hipMemcpyDeviceToHost, s_) != hipSuccess ||
        hipStreamSynchronize(s_) != hipSuccess)
        return 0;
    uint64_t words = 0;
    for (uint32_t c = 0; c < kFreeClasses; c++) {
        const int t = static_cast<int>(h_pinned_[128 + c]);
        if (t > 0) words += static_cast<uint64_t>(t) << c;
    }
    return words;
}

std::string Engine::check_err(const char* where) {
    uint32_t e = read_u32(s_, h_pinned_, g_.err);
    last_err_ = e;
    if (!e) return "";
    (void)hipMemsetAsync(g_.err, 0, 4, s_);
    std::string msg = std::string(where) + ": device error";
    if (e & kErrArenaFull) msg += " [adjacency arena full]";
    if (e & kErrPrecondition) msg += " [batch precondition violated: inserted edge exists or deleted edge missing]";
    if (e & kErrSegmentTooLong) msg += " [too many batch arcs on one vertex]";
    if (e & kErrProductOverflow) msg += " [Cartesian product exceeds 2^63]";
    if (e & kErrScratchFull) msg += " [Cartesian-product scratch full]";
    const Code c = (e & kErrPrecondition) ? Code::kPrecondition
                   : (e & kErrArenaFull) ? Code::kMemory
                   : (e & kErrSegmentTooLong) ? Code::kLimit
                                              : Code::kDevice;
    return tagged(c, msg);
}

uint32_t Engine::view_of(const QState& qs, uint32_t rel, uint32_t i, uint32_t ph) const {
    const uint32_t s = qs.rel[rel].step;
    if (s < i) return ph ? kViewBaseMinusDelta : kViewBasePlusDelta;
    if (s == i) return kViewDelta;
    return kViewBase;
}

// First stage of a step: clear the lane's local sets of its previous step, zero their lengths and
// the step counters, and enqueue the anchor pairs of query edge i (their count at *w.n_pairs).
void Engine::step_anchors(Lane& w, QState& qs, uint32_t i, uint32_t ph) {
    const Query& q = qs.q;
    const uint32_t u0 = q.qe_list[i].first, u1 = q.qe_list[i].second;
    const uint32_t max_arcs = 2 * max_batch_;
    auto cand = [&](uint32_t u) { return qs.cand + static_cast<uint64_t>(u) * words_; };
    if (w.L_used) {
        VSetTable t;
        for (uint32_t u = 0; u < kMaxQV; u++) t.s[u] = w.L[u];
        clear_sets(t, w.L_used, dv_, w.s);
    }
    w.L_used = 0;
    fill_u32(w.L[0].len, 0u, kLaneWords, w.s);
    const RelInfo& R = qs.rel[q.rel(u0, u1)];
    prof_begin(w.prof, w.s, kKAnchors);
    local_anchors(g_, ph, R.dir, R.slot, qs.vrank[u0], cand(u0), cand(u1), bs_[R.dir].by_label, bs_[R.dir].label_off, max_arcs, w.pairs,
                  w.n_pairs, max_arcs, w.L[u0], w.L[u1], w.s);
    prof_end(w.prof, w.s, kKAnchors);
    w.L_used |= static_cast<uint16_t>((1u << u0) | (1u << u1));
}

// Second stage (local index on): enqueue the local sets of the step's indexing order; returns the
// query vertices whose set is built (their lengths at w.L[u].len).
uint32_t Engine::step_local(Lane& w, QState& qs, uint32_t i, uint32_t ph) {
    const Query& q = qs.q;
    const Plan& p = qs.plan;
    const uint32_t u0 = q.qe_list[i].first, u1 = q.qe_list[i].second;
    auto cand = [&](uint32_t u) { return qs.cand + static_cast<uint64_t>(u) * words_; };
    uint32_t built = (1u << u0) | (1u << u1);
    if (!qs.tech.local_index) return built;
    const OrderPerEdge& io = p.indexing[i];
    for (uint32_t pos = 2; pos < q.n; pos++) {
        const uint32_t u = io.vs[pos];
        if (!(p.rebuild_B[i] >> u & 1u)) continue;
        std::vector<uint32_t> bsv;
        for (uint32_t k = io.bni_offs[pos]; k < io.bni_offs[pos + 1]; k++) {
            const uint32_t b = io.vs[io.bni[k]];
            if (built >> b & 1u) bsv.push_back(b);
        }
        if (bsv.empty()) continue;
        size_t first = 0;
        if (qs.tech.heavy_vertex_avoidance) {
            for (size_t k = 1; k < bsv.size(); k++)
                if (qs.avg[q.rel(bsv[k], u)] < qs.avg[q.rel(bsv[first], u)]) first = k;
        }
        const uint32_t b0 = bsv[first];
        const uint32_t rb = q.rel(b0, u);
        prof_begin(w.prof, w.s, kKLocalIndex);
        if (bsv.size() == 1) {
            local_expand(g_, ph, view_of(qs, rb, i, ph), qs.rel[rb].dir, qs.rel[rb].slot, w.L[b0].list, w.L[b0].len, dv_, cand(u), w.L[u],
                         local_grid_, bytes_ptr(kKLocalIndex), w.s);
            prof_end(w.prof, w.s, kKLocalIndex);
        } else {
            // semi-joins with the other built neighbours, fused into the expansion
            LocalFilters fl{};
            for (size_t k = 0; k < bsv.size(); k++) {
                if (k == first) continue;
                const uint32_t rub = q.rel(u, bsv[k]);
                fl.f[fl.n++] = LocalFilter{view_of(qs, rub, i, ph), qs.rel[rub].dir, qs.rel[rub].slot, w.L[bsv[k]].bitmap};
            }
            uint32_t* queue_len = w.L[0].len + kQueueCounters + u;
            local_queue(g_, ph, view_of(qs, rb, i, ph), qs.rel[rb].dir, qs.rel[rb].slot, w.L[b0].list, w.L[b0].len, dv_, cand(u), w.L[u],
                        w.tmp_list, queue_len, local_grid_, bytes_ptr(kKLocalIndex), w.s);
            prof_end(w.prof, w.s, kKLocalIndex);
            prof_begin(w.prof, w.s, kKLocalCheck);
            local_check(g_, ph, fl, w.L[u], w.tmp_list, queue_len, local_grid_, bytes_ptr(kKLocalCheck), w.s);
            prof_end(w.prof, w.s, kKLocalCheck);
        }
        w.L_used |= static_cast<uint16_t>(1u << u);
        built |= 1u << u;
    }
    return built;
}

std::string Engine::run_step(Lane& w, QState& qs, uint32_t i, uint32_t ph) {
    step_anchors(w, qs, i, ph);
    const uint32_t np = read_u32(w.s, w.h_pinned, w.n_pairs);
    if (np == 0) return "empty";
    const uint32_t built = step_local(w, qs, i, ph);
    if (qs.tech.local_index) {
        CSM_HIP(hipMemcpyAsync(w.h_pinned, w.L[0].len, kMaxQV * 4, hipMemcpyDeviceToHost, w.s));
        CSM_HIP(hipStreamSynchronize(w.s));
        for (uint32_t u = 0; u < qs.q.n; u++)
            if ((built >> u & 1u) && w.h_pinned[u] == 0) return "empty";
    }
    return enumerate_step(w, qs, i, ph, w.L, built, w.pairs, w.n_pairs, np);
}

// The search of one step once its local index exists: L[u] for the query vertices in `built`,
// the np anchor pairs of the step's query edge in pairs / *n_pairs.
std::string Engine::enumerate_step(Lane& w, QState& qs, uint32_t i, uint32_t ph, const VSet* L, uint32_t built, const uint32_t* pairs,
                                   const uint32_t* n_pairs, uint32_t np) {
    StepDesc d{};
    bool pending = false;
    uint32_t count = 0;
    std::string err = enumerate_begin(w, qs, i, ph, L, built, pairs, n_pairs, np, d, pending, count);
    if (!err.empty()) return err;
    if (pending) {
        CSM_HIP(hipStreamSynchronize(w.s));
        err = initial_count(w.h_pinned, count);
        if (!err.empty()) return err;
    }
    return enumerate_rest(w, qs, i, ph, d, count);
}

// The initial partial results of a step started by init_from_set: their count (or the overflow)
// from a host copy of the lane's step counters, once the stream has passed the copy.
std::string Engine::initial_count(const uint32_t* counters, uint32_t& count) {
    if (counters[2]) return tagged(Code::kMemory, "partial-result ring too small for the initial partial results");
    count = counters[0];
    return "";
}

// Step descriptor and initial partial results (depth 2, at the bottom of the lane's ring).  With
// the anchor pairs as the first two vertices their count is np; otherwise init_from_set runs and
// `pending` says that its count arrives in the lane's pinned buffer (initial_count, after a wait).
std::string Engine::enumerate_begin(Lane& w, QState& qs, uint32_t i, uint32_t ph, const VSet* L, uint32_t built, const uint32_t* pairs,
                                    const uint32_t* n_pairs, uint32_t np, StepDesc& d, bool& pending, uint32_t& count, InitJob* defer) {
    const Query& q = qs.q;
    w.sched = sched_of(qs.tech.strategy);
    const Plan& p = qs.plan;
    const uint32_t u0 = q.qe_list[i].first, u1 = q.qe_list[i].second;
    const uint32_t max_arcs = 2 * max_batch_;
    auto cand = [&](uint32_t u) { return qs.cand + static_cast<uint64_t>(u) * words_; };
    fill_u32(w.counters, 0u, kLaneWords - kStepCounters, w.s);

    // ---- step descriptor ----
    const OrderPerEdge& o = p.orders[i];
    d = StepDesc{};
    pending = false;
    count = 0;
    d.n = q.n;
    d.phase = ph;
    uint32_t ne = 0;
    for (uint32_t k = 0; k < q.n; k++) {
        const uint32_t u = o.vs[k];
        d.order[k] = u;
        d.vrank[k] = qs.vrank[u];
        d.boff[k] = static_cast<uint8_t>(ne);
        d.pivot[k] = static_cast<uint8_t>(ne);
        d.filter[k] = (built >> u & 1u) ? L[u].bitmap : cand(u);
        if (k == 0) continue;
        bool have_delta = false;
        for (uint32_t kk = o.bni_offs[k]; kk < o.bni_offs[k + 1]; kk++) {
            const uint32_t pos = o.bni[kk];
            const uint32_t r = q.rel(o.vs[pos], u);
            BackEntry e{};
            e.slot = qs.rel[r].slot;
            e.pos = static_cast<uint8_t>(pos);
            e.view = static_cast<uint8_t>(view_of(qs, r, i, ph));
            e.dir = static_cast<uint8_t>(qs.rel[r].dir);
            if (e.view == kViewDelta && !have_delta) { d.pivot[k] = static_cast<uint8_t>(ne); have_delta = true; }
            d.be[ne++] = e;
        }
    }
    d.boff[q.n] = static_cast<uint8_t>(ne);
    for (uint32_t k = 0; k < q.n; k++) {
        uint32_t back = 0;
        for (uint32_t e = d.boff[k]; e < d.boff[k + 1]; e++) back |= 1u << d.be[e].pos;
        for (uint32_t j = 0; j < k; j++)
            if (d.vrank[j] == d.vrank[k] && !(back >> j & 1u)) d.inj[k] |= static_cast<uint16_t>(1u << j);
    }

    // ---- initial partial results (depth 2) at the bottom of the lane's ring ----
    uint32_t* ring = w.ring;
    const uint32_t a = o.vs[0], b = o.vs[1];
    if ((a == u0 && b == u1) || (a == u1 && b == u0)) {
        if (2ull * np > w.ring_words) return tagged(Code::kMemory, "partial-result ring too small for the batch's anchors");
        init_from_pairs(pairs, n_pairs, max_arcs, a == u1, ring, w.counters, w.s);
        count = np;
    } else {
        const uint32_t rab = q.rel(a, b);
        const uint32_t* list;
        const uint32_t* len;
        if (built >> a & 1u) { list = L[a].list; len = L[a].len; }
        else {
            fill_u32(w.tmp_len, 0u, 1, w.s);
            bitmap_to_list(cand(a), dv_, w.tmp_list, w.tmp_len, w.s);
            list = w.tmp_list;
            len = w.tmp_len;
        }
        const uint32_t cap = static_cast<uint32_t>(std::min<uint64_t>(w.ring_words / 2, 0xFFFFFFFFull));
        if (defer) {  // the caller launches it with other steps' (init_from_set_multi) and reads the count
            *defer = InitJob{ph, view_of(qs, rab, i, ph), qs.rel[rab].dir, qs.rel[rab].slot, list, len, d.filter[1], ring, w.counters,
                             w.counters + 2, cap, dv_};
        } else {
            Ring r{nullptr, nullptr, 0, ring, w.counters, cap, w.counters + 2};
            init_from_set(g_, ph, view_of(qs, rab, i, ph), qs.rel[rab].dir, qs.rel[rab].slot, list, len, dv_, d.filter[1], r, w.s);
            CSM_HIP(hipMemcpyAsync(w.h_pinned, w.counters, 12, hipMemcpyDeviceToHost, w.s));
        }
        pending = true;
    }
    if (trace_) w.trace_line += ", \"anchors\": " + std::to_string(np);
    return "";
}

// The depth-first kernel's schedule of a strategy: static for DFS without work stealing (match_gpu.cu
// 1721-1914 with a fixed partition), range donation for DFS with work stealing, dynamic otherwise.
int Engine::sched_of(SearchStrategy st) {
    return st == SearchStrategy::kDfsWithoutWorkStealing ? 0 : st == SearchStrategy::kDfsWithWorkStealing ? 2 : 1;
}

uint32_t Engine::slot_of(const QState& qs) const {
    return static_cast<uint32_t>(std::find_if(qs_.begin(), qs_.end(), [&](const std::unique_ptr<QState>& x) { return x.get() == &qs; }) -
                                 qs_.begin());
}

// The search of a step from its `count` initial partial results (enumerate_begin).
std::string Engine::enumerate_rest(Lane& w, QState& qs, uint32_t i, uint32_t ph, const StepDesc& d, uint32_t count) {
    const Query& q = qs.q;
    const Plan& p = qs.plan;
    uint32_t* ring = w.ring;
    if (trace_) w.trace_line += ", \"initial\": " + std::to_string(count);
    if (count == 0) return "";
    const uint32_t res_slot = slot_of(qs);
    if (q.n == 2) {
        count_records(d, ring, w.counters, count, MatchOut{results_ + 4 * res_slot + 2 * ph, results_ + 4 * res_slot + 2 * ph + 1, nullptr}, w.s);
        return "";
    }

    const uint8_t* cpi = p.cp_info[i];
    auto tail_from = [&](uint32_t from) {
        uint32_t nd = from;
        while (nd < q.n && cpi[nd] != kCpTreeProduct && cpi[nd] != kCpTreeSingle) nd++;
        return nd;
    };
    const uint64_t top = 2ull * count;  // first free word of the ring
    switch (qs.tech.strategy) {
        case SearchStrategy::kDynamicBfsDfs:  // match_gpu.cu:1916-2127
            if (opt_.chain_levels && !trace_) return search_chained(w, qs, d, i, count, res_slot);
            return search(w, qs, d, i, ring, count, 2, top, q.n, res_slot);
        case SearchStrategy::kBfs: {  // match_gpu.cu:2129-2294: breadth-first to the Cartesian-product tail
            const uint32_t target = qs.tech.cartesian_product ? std::min(tail_from(2), q.n - 1) : q.n - 1;
            return search(w, qs, d, i, ring, count, 2, top, target, res_slot);
        }
        default: {  // DFS, with or without work stealing (match_gpu.cu:1721-1914)
            const uint32_t t = tail_from(2);
            if (qs.tech.cartesian_product && t < q.n && t > 2 && cpi[t] == kCpTreeProduct) {
                // depth-first down to the tail, written for the Cartesian product
                bool ok;
                uint32_t n_out;
                uint32_t* out = ring + top;
                std::string err = extend_to(w, d, ring, count, 2, t, out, (w.ring_words - top) / t, ok, n_out);
                if (!err.empty()) return err;
                if (ok) return n_out ? finish_records(w, qs, d, i, t, n_out, out, res_slot) : "";
            }
            return finish_records(w, qs, d, i, 2, count, ring, res_slot);
        }
    }
}

// Writes the records of depth `end` that extend the `count` records of depth `depth` at `rec` to
// `out` (room for `cap` records): one level with the flattened kernels, several with the
// depth-first kernel.  ok = false when they do not fit.
std::string Engine::extend_to(Lane& w, const StepDesc& d, const uint32_t* rec, uint32_t count, uint32_t depth, uint32_t end, uint32_t* out,
                              uint64_t cap, bool& ok, uint32_t& n_out) {
    cap = std::min<uint64_t>(cap, 0xFFFFFFFFull);
    if (end == depth + 1) {
        // [1]: output count, [2]: overflow, cleared by the first chunk's preparation
        Ring r{rec, w.counters, count, out, w.counters + 1, static_cast<uint32_t>(cap), w.counters + 2};
        const uint32_t chunk = level_max_records(d, depth, w.level);
        for (uint32_t s0 = 0; s0 < count; s0 += chunk) {
            const uint32_t nc = std::min(chunk, count - s0);
            const uint32_t* rc = rec + static_cast<uint64_t>(s0) * depth;
            prof_begin(w.prof, w.s, kKLevelPrepare);
            level_prepare(g_, d, depth, rc, nc, w.level, false, s0 == 0 ? w.counters + 1 : nullptr, bytes_ptr(kKLevelWrite), w.s);
            prof_end(w.prof, w.s, kKLevelPrepare);
            if (trace_) {  // candidates of the level (the flat kernel's work), for the trace only
                CSM_HIP(hipMemcpyAsync(w.h_pinned + 12, w.level.off + nc, 8, hipMemcpyDeviceToHost, w.s));
                CSM_HIP(hipStreamSynchronize(w.s));
                uint64_t t;
                std::memcpy(&t, w.h_pinned + 12, 8);
                w.trace_line += ", \"cands_" + std::to_string(end) + "\": " + std::to_string(t);
            }
            prof_begin(w.prof, w.s, kKLevelWrite);
            level_flat(g_, d, depth, rc, nc, kWriteAtEnd, w.level, r, MatchOut{nullptr, nullptr, bytes_ptr(kKLevelWrite)}, w.s);
            prof_end(w.prof, w.s, kKLevelWrite);
        }
        CSM_HIP(hipMemcpyAsync(w.h_pinned, w.counters, 12, hipMemcpyDeviceToHost, w.s));
        CSM_HIP(hipStreamSynchronize(w.s));
        ok = w.h_pinned[2] == 0;
        n_out = w.h_pinned[1];
    } else {
        // [0]: input count, [1]: output count, [2]: overflow, [3]: finishing count, [8..9]: work counter
        store_words(w.counters, Words{10, {count, 0, 0, 0, 0, 0, 0, 0, 0, 0}}, w.s);
        Ring r{rec, w.counters, count, out, w.counters + 1, static_cast<uint32_t>(cap), w.counters + 2};
        prof_begin(w.prof, w.s, kKExtendWrite);
        extend(g_, d, depth, end, kWriteAtEnd, r, MatchOut{nullptr, nullptr, bytes_ptr(kKExtendWrite)}, nullptr, w.work_counter(), w.sched,
               static_cast<int>(opt_.dfs_blocks), w.dn, w.s);
        prof_end(w.prof, w.s, kKExtendWrite);
        CSM_HIP(hipMemcpyAsync(w.h_pinned, w.counters, 12, hipMemcpyDeviceToHost, w.s));
        CSM_HIP(hipStreamSynchronize(w.s));
        ok = w.h_pinned[2] == 0;
        n_out = w.h_pinned[1];
    }
    if (trace_) w.trace_line += ", \"level_" + std::to_string(end) + "\": " + (ok ? std::to_string(n_out) : std::string("\"overflow\""));
    return "";
}

// The semi-BFS search of `search` from the initial records (depth 2 at the bottom of the ring)
// with its levels chained on the device: every writing level and the final counting level are
// enqueued at once, each reading its size from the counters of the levels before it, and the host
// waits once.  The chain runs exactly the levels `search` would run while they need no host
// decision (a level larger than one call, a Cartesian-product depth with enough records for the
// semi-BFS threshold, a Cartesian-product leaf, a ring overflow); at the first level that does,
// the chain stops and `search` goes on from that level's records.
std::string Engine::search_chained(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t count, uint32_t res_slot) {
    LevelChain c{};
    const uint32_t L = chain_build(w, qs, d, i, count, c);
    for (uint32_t k = 0; k < L; k++) {
        prof_begin(w.prof, w.s, kKLevelPrepare);
        level_prepare_chain(g_, d, c, k, w.level, false, bytes_ptr(kKLevelWrite), w.s);
        prof_end(w.prof, w.s, kKLevelPrepare);
        prof_begin(w.prof, w.s, kKLevelWrite);
        level_flat_chain(g_, d, c, k, kWriteAtEnd, w.level, MatchOut{nullptr, nullptr, bytes_ptr(kKLevelWrite)}, w.s);
        prof_end(w.prof, w.s, kKLevelWrite);
    }
    const MatchOut out = chain_out(d, res_slot);
    prof_begin(w.prof, w.s, kKLevelPrepare);
    level_prepare_chain(g_, d, c, L, w.level, true, out.bytes, w.s);
    prof_end(w.prof, w.s, kKLevelPrepare);
    prof_begin(w.prof, w.s, kKLevelCount);
    level_flat_chain(g_, d, c, L, kCountAtEnd, w.level, out, w.s);
    prof_end(w.prof, w.s, kKLevelCount);
    if (L) CSM_HIP(hipMemcpyAsync(w.h_pinned, w.chain_ctr, 8ull * L, hipMemcpyDeviceToHost, w.s));
    CSM_HIP(hipStreamSynchronize(w.s));
    return chain_finish(w, qs, d, i, count, res_slot, c, L, w.h_pinned);
}

// The chain of a step from its `count` initial records (depth 2 at the bottom of the lane's ring):
// writing levels 0 .. L - 1, counting level L.
uint32_t Engine::chain_build(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t count, LevelChain& c) {
    const uint32_t depth0 = 2, L = d.n - 1 - depth0;
    const uint8_t* cpi = qs.plan.cp_info[i];
    c = LevelChain{};
    c.ring = w.ring;
    c.ring_words = w.ring_words;
    c.ctr = w.chain_ctr;
    c.off0 = 0;
    c.n0 = count;
    c.depth0 = depth0;
    for (uint32_t k = 0; k <= L; k++) {
        const uint32_t depth = depth0 + k;
        const bool cp_depth = qs.tech.cartesian_product && cpi[depth] == kCpTreeProduct;
        c.cap[k] = level_max_records(d, depth, w.level);
        c.stop_at[k] = !cp_depth ? 0xFFFFFFFFu : k < L ? opt_.semi_bfs_threshold : 0u;
    }
    return L;
}

// The counting level's outputs: the step's query's count and hash of this phase.
MatchOut Engine::chain_out(const StepDesc& d, uint32_t res_slot) {
    return MatchOut{results_ + 4 * res_slot + 2 * d.phase, results_ + 4 * res_slot + 2 * d.phase + 1, bytes_ptr(kKLevelCount)};
}

// After the chain ran and its per-level {count, overflow} words reached host memory at `ctr`:
// replay the chain's decisions (chain_at) to find where it stopped, and go on from there with
// `search`.
std::string Engine::chain_finish(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t count, uint32_t res_slot, const LevelChain& c,
                                 uint32_t L, const uint32_t* ctr) {
    const uint32_t depth0 = c.depth0;
    uint64_t off = 0;
    uint32_t n = count;
    for (uint32_t k = 0;; k++) {
        const uint32_t depth = depth0 + k;
        if (n == 0) return "";
        if (n > c.cap[k] || n >= c.stop_at[k] || (k < L && ctr[2 * k + 1]))
            return search(w, qs, d, i, w.ring + off, n, depth, off + static_cast<uint64_t>(n) * depth, d.n, res_slot);
        if (k == L) return "";  // the leaf level counted
        off += static_cast<uint64_t>(n) * depth;
        n = ctr[2 * k];
    }
}

// Breadth-first levels kept as a stack in the lane's ring: the `count` records of depth `depth` at
// `rec` are extended one level into the free words above `top`, and so on, until the strategy's
// finishing depth (the leaf level, `target`, or a Cartesian-product depth once the level holds at
// least the semi-BFS threshold of records).  A level that does not fit is redone for each half of
// its input records in turn (one record that does not fit is counted depth-first), so the ring
// size bounds memory, never the search.
std::string Engine::search(Lane& w, QState& qs, const StepDesc& d, uint32_t i, const uint32_t* rec, uint32_t count, uint32_t depth,
                           uint64_t top, uint32_t target, uint32_t res_slot) {
    const uint8_t* cpi = qs.plan.cp_info[i];
    while (count) {
        const bool leaf = depth + 1 >= d.n;
        const bool cp_here = qs.tech.cartesian_product && cpi[depth] == kCpTreeProduct && count >= opt_.semi_bfs_threshold;
        if (leaf || depth >= target || (target == d.n && cp_here)) return finish_records(w, qs, d, i, depth, count, rec, res_slot);
        uint32_t* out = w.ring + top;
        bool ok;
        uint32_t n_out;
        std::string err = extend_to(w, d, rec, count, depth, depth + 1, out, (w.ring_words - top) / (depth + 1), ok, n_out);
        if (!err.empty()) return err;
        if (!ok) {
            if (count == 1) return finish_records(w, qs, d, i, depth, count, rec, res_slot, true);
            const uint32_t h = count / 2;
            err = search(w, qs, d, i, rec, h, depth, top, target, res_slot);
            if (!err.empty()) return err;
            rec += static_cast<uint64_t>(h) * depth;
            count -= h;
            continue;
        }
        rec = out;
        count = n_out;
        top += static_cast<uint64_t>(n_out) * (depth + 1);
        depth += 1;
    }
    return "";
}

std::string Engine::finish_records(Lane& w, QState& qs, const StepDesc& d, uint32_t i, uint32_t depth, uint32_t count, const uint32_t* rec,
                                   uint32_t res_slot, bool any_depth) {
    const uint32_t ph = d.phase;
    MatchOut out{results_ + 4 * res_slot + 2 * ph, results_ + 4 * res_slot + 2 * ph + 1, nullptr};
    uint32_t* rec_count = w.counters + 3;
    const uint64_t n_seg = static_cast<uint64_t>(count) * (d.n - depth);
    bool use_cp = qs.tech.cartesian_product && depth < d.n && qs.plan.cp_info[i][depth] == kCpTreeProduct && n_seg <= opt_.cp_max_records;
    if (use_cp) {
        // match_gpu.cu:2046-2085: flatten unless the products are uniformly large
        CSM_HIP(hipMemsetAsync(w.cp.max_prod, 0, 16, w.s));  // max_prod, err
        prof_begin(w.prof, w.s, kKCpPrepare);
        cp_prepare(g_, d, depth, rec, count, w.cp, w.s);
        prof_end(w.prof, w.s, kKCpPrepare);
        unsigned long long v[2];
        CSM_HIP(hipMemcpyAsync(w.h_pinned, w.cp.max_prod, 8, hipMemcpyDeviceToHost, w.s));
        CSM_HIP(hipMemcpyAsync(w.h_pinned + 2, w.cp.prod + count, 8, hipMemcpyDeviceToHost, w.s));
        CSM_HIP(hipMemcpyAsync(w.h_pinned + 4, w.cp.err, 4, hipMemcpyDeviceToHost, w.s));
        CSM_HIP(hipStreamSynchronize(w.s));
        std::memcpy(&v[0], w.h_pinned, 8);
        std::memcpy(&v[1], w.h_pinned + 2, 8);
        if (w.h_pinned[4] & (kErrScratchFull | kErrProductOverflow)) {
            use_cp = false;  // fall back to counting: same embeddings
        } else {
            const float avg = static_cast<float>(v[1]) / static_cast<float>(count);
            const float ratio = static_cast<float>(v[0]) / avg;
            if (ratio < 50.f && avg >= static_cast<float>(kSemiBfsThreshold)) use_cp = false;
        }
        if (trace_) w.trace_line += ", \"cp_total\": " + std::to_string(v[1]) + ", \"cp_max\": " + std::to_string(v[0]) + ", \"cp_used\": " + std::to_string(use_cp ? 1 : 0);
        if (use_cp) {
            prof_begin(w.prof, w.s, kKCpEnumerate);
            cp_enumerate(d, depth, count, w.cp, out, w.s);
            prof_end(w.prof, w.s, kKCpEnumerate);
            return "";
        }
    }
    Ring r{rec, rec_count, count, nullptr, nullptr, 0, nullptr};
    if (trace_) w.trace_line += ", \"count_from\": " + std::to_string(depth) + ", \"records\": " + std::to_string(count);
    if (depth + 1 == d.n && w.sched != 0) {
        out.bytes = bytes_ptr(kKLevelCount);
        const uint32_t chunk = level_max_records(d, depth, w.level);
        for (uint32_t s0 = 0; s0 < count; s0 += chunk) {
            const uint32_t nc = std::min(chunk, count - s0);
            const uint32_t* rc = rec + static_cast<uint64_t>(s0) * depth;
            prof_begin(w.prof, w.s, kKLevelPrepare);
            level_prepare(g_, d, depth, rc, nc, w.level, true, nullptr, out.bytes, w.s);
            prof_end(w.prof, w.s, kKLevelPrepare);
            if (trace_) {
                CSM_HIP(hipMemcpyAsync(w.h_pinned + 12, w.level.off + nc, 8, hipMemcpyDeviceToHost, w.s));
                CSM_HIP(hipStreamSynchronize(w.s));
                uint64_t t;
                std::memcpy(&t, w.h_pinned + 12, 8);
                w.trace_line += ", \"cands_leaf\": " + std::to_string(t);
            }
            prof_begin(w.prof, w.s, kKLevelCount);
            level_flat(g_, d, depth, rc, nc, kCountAtEnd, w.level, r, out, w.s);
            prof_end(w.prof, w.s, kKLevelCount);
        }
        return "";
    }
    // Stack depth: a counting DFS needs one LDS row per depth below its first; with more than the
    // setting allows, breadth-first levels run first (records stay on the lane's ring: the words
    // above `rec`'s records are free, as every caller leaves its records at the top).
    const uint32_t rows = d.n - 1 - depth;
    if (!any_depth && rows > opt_.stack_depth && rec >= w.ring && rec < w.ring + w.ring_words) {
        const uint64_t top = static_cast<uint64_t>(rec - w.ring) + static_cast<uint64_t>(count) * depth;
        return search(w, qs, d, i, rec, count, depth, top, d.n - 1 - opt_.stack_depth, res_slot);
    }
    store_words(rec_count, Words{7, {count, 0, 0, 0, 0, 0, 0}}, w.s);  // [3]: count, [8..9]: work counter
    out.bytes = bytes_ptr(kKExtendCount);
    // Root splitting (dynamic schedule, few records: a handful of heavy records would otherwise
    // serialise on single waves).  The CP offsets buffer is free here.
    const bool split = opt_.waves_per_anchor > 1 && w.sched != 0 && count <= kSplitMaxRecords;
    prof_begin(w.prof, w.s, kKExtendCount);
    if (split) root_chunks(g_, d, depth, rec, rec_count, count, opt_.waves_per_anchor, w.cp.prod, w.cp.scan_tmp, w.s);
    extend(g_, d, depth, d.n, kCountAtEnd, r, out, split ? w.cp.prod : nullptr, w.work_counter(), w.sched,
           static_cast<int>(opt_.dfs_blocks), w.dn, w.s);
    prof_end(w.prof, w.s, kKExtendCount);
    return "";
}

// The local index of a group of steps (each with its slot) in lockstep on the main stream: clear
// the slots, compute every step's anchors in one launch, then for each indexing-order position one
// expansion and one check launch over every step that builds a vertex there (the same rule as
// run_step: heavy-vertex avoidance picks the expansion side, the other built neighbours filter).
// Steps whose anchors or local sets come out empty get np = 0.
std::string Engine::stage_group(uint32_t ph, std::vector<GroupStep>& steps) {
    const uint32_t n = static_cast<uint32_t>(steps.size());
    if (!n) return "";
    const uint32_t max_arcs = 2 * max_batch_;
    uint8_t* h_anchor = h_jobs_;
    uint8_t* h_local = h_anchor + kMaxGroup * sizeof(AnchorJob);
    uint8_t* h_clear = h_local + kMaxQV * kMaxLocalJobs * sizeof(LocalJob);
    uint32_t* h_ctr = reinterpret_cast<uint32_t*>(h_clear + 2ull * kMaxGroup * kMaxQV * sizeof(VSet));
    const int first_slot = steps[0].slot;
    for (uint32_t k = 0; k < n; k++)
        if (steps[k].slot != first_slot + static_cast<int>(k)) return tagged(Code::kState, "group slots must be consecutive");
    uint32_t* ctr = slots_[first_slot].counters;

    // clear what the slots' previous steps left, then zero their counters
    VSet* cs = reinterpret_cast<VSet*>(h_clear);
    uint32_t nclear = 0;
    for (auto& st : steps) {
        Slot& sl = slots_[st.slot];
        for (uint32_t u = 0; u < kMaxQV; u++)
            if (sl.used >> u & 1u) cs[nclear++] = sl.L[u];
        sl.used = 0;
    }
    if (nclear) {
        CSM_HIP(hipMemcpyAsync(d_clear_sets_, cs, nclear * sizeof(VSet), hipMemcpyHostToDevice, s_));
        clear_sets_multi(d_clear_sets_, nclear, dv_, s_);
    }
    fill_u32(ctr, 0u, static_cast<uint64_t>(n) * kSlotWords, s_);

    // anchors of every step
    AnchorJob* aj = reinterpret_cast<AnchorJob*>(h_anchor);
    for (uint32_t k = 0; k < n; k++) {
        QState& qs = *qs_[steps[k].q];
        const Query& q = qs.q;
        const uint32_t i = steps[k].i;
        const uint32_t u0 = q.qe_list[i].first, u1 = q.qe_list[i].second;
        const RelInfo& R = qs.rel[q.rel(u0, u1)];
        Slot& sl = slots_[steps[k].slot];
        aj[k] = AnchorJob{ph, R.dir, R.slot, qs.vrank[u0], qs.cand + static_cast<uint64_t>(u0) * words_,
                          qs.cand + static_cast<uint64_t>(u1) * words_, sl.pairs, sl.n_pairs, max_arcs, sl.L[u0], sl.L[u1]};
        sl.used = static_cast<uint16_t>((1u << u0) | (1u << u1));
        steps[k].built = (1u << u0) | (1u << u1);
    }
    CSM_HIP(hipMemcpyAsync(d_anchor_jobs_, aj, n * sizeof(AnchorJob), hipMemcpyHostToDevice, s_));
    prof_begin(prof_main_, s_, kKAnchors);
    local_anchors_multi(g_, d_anchor_jobs_, n, bs_[0].by_label, bs_[0].label_off, bs_[1].by_label, bs_[1].label_off, max_arcs, s_);
    prof_end(prof_main_, s_, kKAnchors);
    CSM_HIP(hipMemcpyAsync(h_ctr, ctr, static_cast<uint64_t>(n) * kSlotWords * 4, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    for (uint32_t k = 0; k < n; k++) steps[k].np = h_ctr[k * kSlotWords + kMaxQV];

    // local sets, one indexing-order position at a time
    for (uint32_t pos = 2; pos < kMaxQV; pos++) {
        LocalJob* lj = reinterpret_cast<LocalJob*>(h_local) + static_cast<uint64_t>(pos) * kMaxLocalJobs;
        uint32_t nj = 0;
        for (uint32_t k = 0; k < n; k++) {
            GroupStep& st = steps[k];
            QState& qs = *qs_[st.q];
            const Query& q = qs.q;
            if (st.np == 0 || pos >= q.n) continue;
            const Plan& p = qs.plan;
            const uint32_t i = st.i;
            const OrderPerEdge& io = p.indexing[i];
            const uint32_t u = io.vs[pos];
            if (!(p.rebuild_B[i] >> u & 1u)) continue;
            std::vector<uint32_t> bsv;
            for (uint32_t kk = io.bni_offs[pos]; kk < io.bni_offs[pos + 1]; kk++) {
                const uint32_t b = io.vs[io.bni[kk]];
                if (st.built >> b & 1u) bsv.push_back(b);
            }
            if (bsv.empty()) continue;
            size_t first = 0;
            if (qs.tech.heavy_vertex_avoidance) {
                for (size_t kk = 1; kk < bsv.size(); kk++)
                    if (qs.avg[q.rel(bsv[kk], u)] < qs.avg[q.rel(bsv[first], u)]) first = kk;
            }
            const uint32_t b0 = bsv[first];
            const uint32_t rb = q.rel(b0, u);
            Slot& sl = slots_[st.slot];
            LocalJob& J = lj[nj++];
            J = LocalJob{};
            J.phase = ph;
            J.view = view_of(qs, rb, i, ph);
            J.dir = qs.rel[rb].dir;
            J.slot = qs.rel[rb].slot;
            J.frontier = sl.L[b0].list;
            J.frontier_len = sl.L[b0].len;
            J.cand_u = qs.cand + static_cast<uint64_t>(u) * words_;
            J.out = sl.L[u];
            J.queue = sl.queue;
            J.queue_len = sl.counters + kQueueCounters + u;
            for (size_t kk = 0; kk < bsv.size(); kk++) {
                if (kk == first) continue;
                const uint32_t rub = q.rel(u, bsv[kk]);
                J.fl.f[J.fl.n++] = LocalFilter{view_of(qs, rub, i, ph), qs.rel[rub].dir, qs.rel[rub].slot, sl.L[bsv[kk]].bitmap};
            }
            sl.used |= static_cast<uint16_t>(1u << u);
            st.built |= 1u << u;
        }
        if (!nj) continue;
        LocalJob* dj = d_local_jobs_ + static_cast<uint64_t>(pos) * kMaxLocalJobs;
        CSM_HIP(hipMemcpyAsync(dj, lj, nj * sizeof(LocalJob), hipMemcpyHostToDevice, s_));
        prof_begin(prof_main_, s_, kKLocalIndex);
        local_queue_multi(g_, dj, nj, local_grid_, bytes_ptr(kKLocalIndex), s_);
        prof_end(prof_main_, s_, kKLocalIndex);
        prof_begin(prof_main_, s_, kKLocalCheck);
        local_check_multi(g_, dj, nj, local_grid_, bytes_ptr(kKLocalCheck), s_);
        prof_end(prof_main_, s_, kKLocalCheck);
    }
    CSM_HIP(hipMemcpyAsync(h_ctr, ctr, static_cast<uint64_t>(n) * kSlotWords * 4, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    for (uint32_t k = 0; k < n; k++) {
        GroupStep& st = steps[k];
        if (st.np == 0) continue;
        for (uint32_t u = 0; u < qs_[st.q]->q.n; u++)
            if ((st.built >> u & 1u) && h_ctr[k * kSlotWords + u] == 0) { st.np = 0; break; }
    }
    return check_err("local index");
}

// The local indexes of a team's steps (each on its lane's sets) with the multi-step kernels: the
// anchors of every step in one launch, then per indexing-order position one expansion and one
// check launch over every step that builds a vertex there (the rule of step_local); one wait after
// the anchors and one after the sets.  np[k] = 0 for a step without anchors or with an empty set.
std::string Engine::team_local(size_t first_lane, std::vector<GroupStep>& items, uint32_t ph, std::vector<uint32_t>& np,
                               std::vector<uint32_t>& built) {
    const size_t n = items.size();
    Lane& lead = lanes_[first_lane];
    hipStream_t s = lead.s;
    const uint32_t max_arcs = 2 * max_batch_;
    auto lane = [&](size_t k) -> Lane& { return lanes_[first_lane + k]; };
    // clear the sets the lanes' previous steps left, then zero the lanes' counters
    VSet* hc = h_tclear_ + first_lane * kMaxQV;
    VSet* dc = d_tclear_ + first_lane * kMaxQV;
    uint32_t nclear = 0;
    for (size_t k = 0; k < n; k++) {
        Lane& w = lane(k);
        for (uint32_t u = 0; u < kMaxQV; u++)
            if (w.L_used >> u & 1u) hc[nclear++] = w.L[u];
        w.L_used = 0;
    }
    if (nclear) {
        CSM_HIP(hipMemcpyAsync(dc, hc, nclear * sizeof(VSet), hipMemcpyHostToDevice, s));
        clear_sets_multi(dc, nclear, dv_, s);
    }
    uint32_t* lens = lens_all_ + first_lane * kLaneWords;
    uint32_t* hl = h_tlens_ + first_lane * kLaneWords;
    fill_u32(lens, 0u, n * kLaneWords, s);
    // anchors of every step
    AnchorJob* ha = h_tajobs_ + first_lane;
    for (size_t k = 0; k < n; k++) {
        Lane& w = lane(k);
        QState& qs = *qs_[items[k].q];
        const Query& q = qs.q;
        const uint32_t u0 = q.qe_list[items[k].i].first, u1 = q.qe_list[items[k].i].second;
        const RelInfo& R = qs.rel[q.rel(u0, u1)];
        ha[k] = AnchorJob{ph, R.dir, R.slot, qs.vrank[u0], qs.cand + static_cast<uint64_t>(u0) * words_,
                          qs.cand + static_cast<uint64_t>(u1) * words_, w.pairs, w.n_pairs, max_arcs, w.L[u0], w.L[u1]};
        w.L_used = static_cast<uint16_t>((1u << u0) | (1u << u1));
        built[k] = (1u << u0) | (1u << u1);
    }
    CSM_HIP(hipMemcpyAsync(d_tajobs_ + first_lane, ha, n * sizeof(AnchorJob), hipMemcpyHostToDevice, s));
    prof_begin(lead.prof, s, kKAnchors);
    local_anchors_multi(g_, d_tajobs_ + first_lane, static_cast<uint32_t>(n), bs_[0].by_label, bs_[0].label_off, bs_[1].by_label,
                        bs_[1].label_off, max_arcs, s);
    prof_end(lead.prof, s, kKAnchors);
    CSM_HIP(hipMemcpyAsync(hl, lens, n * kLaneWords * 4ull, hipMemcpyDeviceToHost, s));
    CSM_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < n; k++) np[k] = hl[k * kLaneWords + kMaxQV + 1];  // Lane::n_pairs
    // local sets, one indexing-order position at a time: every position's jobs first (they depend
    // only on the plans), copied to the device at once, then the launches
    uint32_t njobs[kMaxQV] = {};
    uint32_t last = 0;
    for (uint32_t pos = 2; pos < kMaxQV; pos++) {
        LocalJob* lj = h_tljobs_ + first_lane * kMaxQV + pos * n;
        uint32_t& nj = njobs[pos];
        for (size_t k = 0; k < n; k++) {
            QState& qs = *qs_[items[k].q];
            const Query& q = qs.q;
            if (!np[k] || !qs.tech.local_index || pos >= q.n) continue;
            const Plan& p = qs.plan;
            const uint32_t i = items[k].i;
            const OrderPerEdge& io = p.indexing[i];
            const uint32_t u = io.vs[pos];
            if (!(p.rebuild_B[i] >> u & 1u)) continue;
            std::vector<uint32_t> bsv;
            for (uint32_t kk = io.bni_offs[pos]; kk < io.bni_offs[pos + 1]; kk++) {
                const uint32_t b = io.vs[io.bni[kk]];
                if (built[k] >> b & 1u) bsv.push_back(b);
            }
            if (bsv.empty()) continue;
            size_t first = 0;
            if (qs.tech.heavy_vertex_avoidance) {
                for (size_t kk = 1; kk < bsv.size(); kk++)
                    if (qs.avg[q.rel(bsv[kk], u)] < qs.avg[q.rel(bsv[first], u)]) first = kk;
            }
            const uint32_t b0 = bsv[first];
            const uint32_t rb = q.rel(b0, u);
            Lane& w = lane(k);
            LocalJob& J = lj[nj++];
            J = LocalJob{};
            J.phase = ph;
            J.view = view_of(qs, rb, i, ph);
            J.dir = qs.rel[rb].dir;
            J.slot = qs.rel[rb].slot;
            J.frontier = w.L[b0].list;
            J.frontier_len = w.L[b0].len;
            J.cand_u = qs.cand + static_cast<uint64_t>(u) * words_;
            J.out = w.L[u];
            J.queue = w.tmp_list;
            J.queue_len = w.L[0].len + kQueueCounters + u;
            for (size_t kk = 0; kk < bsv.size(); kk++) {
                if (kk == first) continue;
                const uint32_t rub = q.rel(u, bsv[kk]);
                J.fl.f[J.fl.n++] = LocalFilter{view_of(qs, rub, i, ph), qs.rel[rub].dir, qs.rel[rub].slot, w.L[bsv[kk]].bitmap};
            }
            w.L_used |= static_cast<uint16_t>(1u << u);
            built[k] |= 1u << u;
        }
        if (nj) last = pos;
    }
    if (!last) return "";
    CSM_HIP(hipMemcpyAsync(d_tljobs_ + first_lane * kMaxQV + 2 * n, h_tljobs_ + first_lane * kMaxQV + 2 * n,
                           (last - 1) * n * sizeof(LocalJob), hipMemcpyHostToDevice, s));
    for (uint32_t pos = 2; pos <= last; pos++) {
        const uint32_t nj = njobs[pos];
        if (!nj) continue;
        LocalJob* dj = d_tljobs_ + first_lane * kMaxQV + pos * n;
        prof_begin(lead.prof, s, kKLocalIndex);
        local_queue_multi(g_, dj, nj, local_grid_, bytes_ptr(kKLocalIndex), s);
        prof_end(lead.prof, s, kKLocalIndex);
        prof_begin(lead.prof, s, kKLocalCheck);
        local_check_multi(g_, dj, nj, local_grid_, bytes_ptr(kKLocalCheck), s);
        prof_end(lead.prof, s, kKLocalCheck);
    }
    CSM_HIP(hipMemcpyAsync(hl, lens, n * kLaneWords * 4ull, hipMemcpyDeviceToHost, s));
    CSM_HIP(hipStreamSynchronize(s));
    for (size_t k = 0; k < n; k++) {
        if (!np[k] || !qs_[items[k].q]->tech.local_index) continue;
        for (uint32_t u = 0; u < qs_[items[k].q]->q.n; u++)
            if ((built[k] >> u & 1u) && hl[k * kLaneWords + u] == 0) { np[k] = 0; break; }
    }
    return "";
}

std::string Engine::run_team_steps(size_t first_lane, std::vector<GroupStep>& items, uint32_t ph) {
    const size_t n = items.size();
    Lane& lead = lanes_[first_lane];
    hipStream_t s = lead.s;  // every member's stream while the team runs
    auto lane = [&](size_t k) -> Lane& { return lanes_[first_lane + k]; };
    std::vector<uint32_t> np(n, 0), built(n, 0), count(n, 0);
    for (size_t k = 0; k < n; k++) lane(k).sched = sched_of(qs_[items[k].q]->tech.strategy);
    std::string lerr = team_local(first_lane, items, ph, np, built);
    if (!lerr.empty()) return lerr;
    // initial partial results (the steps that start from a vertex set in one launch), one wait
    std::vector<StepDesc> d(n);
    std::vector<char> pend(n, 0);
    bool wait = false;
    InitJob* hi = h_tinit_ + first_lane;
    uint32_t ni = 0;
    for (size_t k = 0; k < n; k++) {
        if (!np[k]) continue;
        Lane& w = lane(k);
        bool p = false;
        std::string err =
            enumerate_begin(w, *qs_[items[k].q], items[k].i, ph, w.L, built[k], w.pairs, w.n_pairs, np[k], d[k], p, count[k], hi + ni);
        if (!err.empty()) return err;
        pend[k] = p;
        if (p) ni++;
        wait = wait || p;
    }
    if (ni) {
        CSM_HIP(hipMemcpyAsync(d_tinit_ + first_lane, hi, ni * sizeof(InitJob), hipMemcpyHostToDevice, s));
        init_from_set_multi(g_, d_tinit_ + first_lane, ni, s);
    }
    uint32_t* hl = h_tlens_ + first_lane * kLaneWords;
    if (wait) {  // every lane's counters in one copy (they are adjacent)
        CSM_HIP(hipMemcpyAsync(hl, lens_all_ + first_lane * kLaneWords, n * kLaneWords * 4ull, hipMemcpyDeviceToHost, s));
        CSM_HIP(hipStreamSynchronize(s));
    }
    for (size_t k = 0; k < n; k++)
        if (np[k] && pend[k]) {
            std::string err = initial_count(hl + k * kLaneWords + kStepCounters, count[k]);
            if (!err.empty()) return err;
        }
    // the level chains of the semi-BFS steps as one chain of multi-job launches, one wait
    ChainJob* hj = h_chain_jobs_ + first_lane;
    std::vector<size_t> job_step;
    uint32_t max_l = 0;
    std::vector<char> chained(n, 0);
    for (size_t k = 0; k < n; k++) {
        QState& qs = *qs_[items[k].q];
        if (!np[k] || !count[k] || qs.q.n <= 2 || qs.tech.strategy != SearchStrategy::kDynamicBfsDfs || !opt_.chain_levels || trace_) continue;
        Lane& w = lane(k);
        ChainJob& J = hj[job_step.size()];
        J = ChainJob{};
        J.d = d[k];
        J.L = chain_build(w, qs, d[k], items[k].i, count[k], J.c);
        J.ls = w.level;
        J.out = chain_out(d[k], slot_of(qs));
        J.write_bytes = bytes_ptr(kKLevelWrite);
        max_l = std::max(max_l, J.L);
        job_step.push_back(k);
        chained[k] = 1;
    }
    const uint32_t nj = static_cast<uint32_t>(job_step.size());
    if (nj) {
        ChainJob* dj = d_chain_jobs_ + first_lane;
        CSM_HIP(hipMemcpyAsync(dj, hj, nj * sizeof(ChainJob), hipMemcpyHostToDevice, s));
        // which levels have jobs that write, and which have jobs that count
        uint32_t wmask = 0, cmask = 0;
        for (uint32_t j = 0; j < nj; j++) {
            wmask |= (1u << hj[j].L) - 1u;
            cmask |= 1u << hj[j].L;
        }
        auto enqueue = [&](bool timed) {
            for (uint32_t lv = 0; lv <= max_l; lv++) {
                if (timed) prof_begin(lead.prof, s, kKLevelPrepare);
                level_prepare_chain_multi(g_, dj, nj, lv, static_cast<int>(opt_.level_blocks), s);
                if (timed) prof_end(lead.prof, s, kKLevelPrepare);
                if (wmask >> lv & 1u) {
                    if (timed) prof_begin(lead.prof, s, kKLevelWrite);
                    level_flat_chain_multi(g_, dj, nj, lv, kWriteAtEnd, static_cast<int>(opt_.level_blocks), s);
                    if (timed) prof_end(lead.prof, s, kKLevelWrite);
                }
                if (cmask >> lv & 1u) {
                    if (timed) prof_begin(lead.prof, s, kKLevelCount);
                    level_flat_chain_multi(g_, dj, nj, lv, kCountAtEnd, static_cast<int>(opt_.level_blocks), s);
                    if (timed) prof_end(lead.prof, s, kKLevelCount);
                }
            }
        };
        if (prof_mask_ || !opt_.graph_chains) {
            enqueue(prof_mask_ != 0);
        } else {
            // The launches depend only on the team, the job count, the write / count levels and the
            // grid (the jobs themselves are read from device memory): one captured graph per such
            // shape, replayed.  The graph holds kernels only (no memset or copy nodes).
            const ChainGraphKey key{first_lane, nj, wmask, cmask, opt_.level_blocks};
            hipGraphExec_t exec = nullptr;
            {
                std::lock_guard<std::mutex> lk(chain_graphs_mu_);
                auto it = chain_graphs_.find(key);
                if (it != chain_graphs_.end()) exec = it->second;
            }
            if (!exec) {
                hipGraph_t graph = nullptr;
                CSM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                enqueue(false);
                CSM_HIP(hipStreamEndCapture(s, &graph));
                const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (ie != hipSuccess) return std::string("HIP error: ") + hipGetErrorString(ie) + " instantiating a level-chain graph";
                std::lock_guard<std::mutex> lk(chain_graphs_mu_);
                chain_graphs_.emplace(key, exec);
            }
            CSM_HIP(hipGraphLaunch(exec, s));
        }
        // every chain's counters in one copy (the lanes' counters are adjacent)
        uint32_t* hc = h_tctr_ + first_lane * 2 * kMaxQV;
        CSM_HIP(hipMemcpyAsync(hc, chain_ctr_all_ + first_lane * 2 * kMaxQV, n * 2 * kMaxQV * 4ull, hipMemcpyDeviceToHost, s));
        CSM_HIP(hipStreamSynchronize(s));
        for (uint32_t j = 0; j < nj; j++) {
            const size_t k = job_step[j];
            QState& qs = *qs_[items[k].q];
            std::string err = chain_finish(lane(k), qs, d[k], items[k].i, count[k], slot_of(qs), hj[j].c, hj[j].L, hc + k * 2 * kMaxQV);
            if (!err.empty()) return err;
        }
    }
    // every other step on its own (other strategies, Cartesian-product-free two-vertex queries, ...)
    for (size_t k = 0; k < n; k++) {
        if (!np[k] || chained[k]) continue;
        std::string err = enumerate_rest(lane(k), *qs_[items[k].q], items[k].i, ph, d[k], count[k]);
        if (!err.empty()) return err;
    }
    return "";
}

std::string Engine::run_phase(uint32_t ph) {
    std::vector<std::pair<uint32_t, uint32_t>> steps;  // (query, edge)
    for (size_t qi = 0; qi < qs_.size(); qi++) {
        const QState& qs = *qs_[qi];
        if (!qs.planned || !qs.matchable) continue;
        for (uint32_t i = 0; i < qs.q.m; i++) steps.push_back({static_cast<uint32_t>(qi), i});
    }
    if (steps.empty()) return "";
    // Longest steps first (host time of the same step in the previous batch): the phase ends when
    // its last step does, so the heavy ones must not start last.
    std::vector<double>& cost = step_cost_[ph];
    auto cost_of = [&](const std::pair<uint32_t, uint32_t>& st) {
        const size_t k = static_cast<size_t>(st.first) * kMaxQE + st.second;
        return k < cost.size() ? cost[k] : 0.0;
    };
    std::stable_sort(steps.begin(), steps.end(), [&](const auto& x, const auto& y) { return cost_of(x) > cost_of(y); });
    cost.resize(qs_.size() * kMaxQE, 0.0);
    CSM_HIP(hipEventRecord(phase_start_, s_));
    const auto phase_t0 = std::chrono::steady_clock::now();
    // Work items for the lanes: steps that build their own local index (slot -1), and steps
    // whose local index a group stage built (their slot).  The main thread stages group g + 1
    // while the lanes search group g; a group reuses the slots of group g - 1 once its steps are done.
    struct Work {
        std::mutex mu;
        std::condition_variable cv;
        std::deque<GroupStep> q;
        bool closed = false;
        uint32_t pending[2] = {0, 0};  // staged steps of each half of the slots not yet finished
    } work;
    std::atomic<bool> failed{false};
    std::vector<std::string> errs(lanes_.size());
    std::vector<GroupStep> grouped;
    {
        std::lock_guard<std::mutex> lk(work.mu);
        for (const auto& st : steps) {
            GroupStep g;
            g.q = st.first;
            g.i = st.second;
            if (group_ && team_ == 1 && qs_[st.first]->tech.local_index) grouped.push_back(g);
            else work.q.push_back(g);
        }
    }
    // A team (team_ > 1) of lanes [li, li + team_): up to team_ steps at a time on the stream of
    // lane li / team_.  The runtime hands hardware queues to streams round-robin in creation order,
    // so the first lanes' streams (created one after another) sit on different queues, while
    // every team_-th lane's stream could share one.
    std::vector<hipStream_t> team_stream;
    for (size_t li = 0; li < lanes_.size(); li += team_) team_stream.push_back(lanes_[li / team_].s);
    auto team_worker = [&](size_t li) {
        Lane& lead = lanes_[li];
        const size_t size = std::min<size_t>(team_, lanes_.size() - li);
        std::vector<hipStream_t> own;
        for (size_t k = 0; k < size; k++) own.push_back(lanes_[li + k].s);
        for (size_t k = 0; k < size; k++) lanes_[li + k].s = team_stream[li / team_];
        auto run = [&]() -> std::string {
            CSM_HIP(hipSetDevice(opt_.device));
            CSM_HIP(hipStreamWaitEvent(lead.s, phase_start_, 0));
            for (;;) {
                std::vector<GroupStep> items;
                {
                    std::unique_lock<std::mutex> lk(work.mu);
                    work.cv.wait(lk, [&]() { return failed.load() || work.closed || !work.q.empty(); });
                    if (failed.load() || work.q.empty()) break;
                    while (items.size() < size && !work.q.empty()) {
                        items.push_back(work.q.front());
                        work.q.pop_front();
                    }
                }
                const auto t0 = std::chrono::steady_clock::now();
                for (size_t k = 0; k < size; k++) lanes_[li + k].trace_line.clear();
                const std::string e = run_team_steps(li, items, ph);
                const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                for (const GroupStep& it : items) cost[static_cast<size_t>(it.q) * kMaxQE + it.i] = sec;
                if (!e.empty() && e != "empty") return e;
            }
            for (size_t k = 0; k < size; k++) CSM_HIP(hipEventRecord(lanes_[li + k].done, lead.s));
            return "";
        };
        errs[li] = run();
        for (size_t k = 0; k < size; k++) lanes_[li + k].s = own[k];
        if (!errs[li].empty()) {
            failed.store(true);
            std::lock_guard<std::mutex> lk(work.mu);
            work.cv.notify_all();
        }
    };
    auto worker = [&](size_t li) {
        Lane& w = lanes_[li];
        auto run = [&]() -> std::string {
            CSM_HIP(hipSetDevice(opt_.device));
            CSM_HIP(hipStreamWaitEvent(w.s, phase_start_, 0));
            for (;;) {
                GroupStep it;
                {
                    std::unique_lock<std::mutex> lk(work.mu);
                    work.cv.wait(lk, [&]() { return failed.load() || work.closed || !work.q.empty(); });
                    if (failed.load() || work.q.empty()) break;
                    it = work.q.front();
                    work.q.pop_front();
                }
                QState& qs = *qs_[it.q];
                const auto t0 = std::chrono::steady_clock::now();
                w.trace_line.clear();
                std::string e;
                if (it.slot < 0) {
                    e = run_step(w, qs, it.i, ph);
                } else {
                    const Slot& sl = slots_[it.slot];
                    e = enumerate_step(w, qs, it.i, ph, sl.L, it.built, sl.pairs, sl.n_pairs, it.np);
                    // the slot is reused by a later group: its readers must be done
                    const hipError_t se = hipStreamSynchronize(w.s);
                    if (e.empty() && se != hipSuccess) e = std::string("HIP error: ") + hipGetErrorString(se) + " finishing a step";
                    std::lock_guard<std::mutex> lk(work.mu);
                    work.pending[static_cast<uint32_t>(it.slot) / std::max(1u, group_)]--;
                    work.cv.notify_all();
                }
                cost[static_cast<size_t>(it.q) * kMaxQE + it.i] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                if (trace_) {
                    (void)hipStreamSynchronize(w.s);
                    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
                    std::lock_guard<std::mutex> lock(trace_mu_);
                    const double at = std::chrono::duration<double, std::milli>(t0 - phase_t0).count();
                    std::fprintf(trace_, "{\"query\": %u, \"step\": %u, \"phase\": %u, \"lane\": %zu, \"slot\": %d, \"at\": %.3f, \"ms\": %.3f, \"empty\": %d%s}\n",
                                 it.q, it.i, ph, li, it.slot, at, ms, e == "empty" ? 1 : 0, w.trace_line.c_str());
                }
                if (e == "empty") continue;
                if (!e.empty()) return e;
            }
            CSM_HIP(hipEventRecord(w.done, w.s));
            return "";
        };
        errs[li] = run();
        if (!errs[li].empty()) {
            failed.store(true);
            std::lock_guard<std::mutex> lk(work.mu);
            work.cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    if (team_ > 1)
        for (size_t li = 0; li < lanes_.size(); li += team_) th.emplace_back(team_worker, li);
    else
        for (size_t li = 0; li < lanes_.size(); li++) th.emplace_back(worker, li);
    std::string stage_err;
    for (size_t g0 = 0, gi = 0; g0 < grouped.size() && !failed.load(); g0 += group_, gi++) {
        const uint32_t half = static_cast<uint32_t>(gi % 2);
        {
            std::unique_lock<std::mutex> lk(work.mu);
            work.cv.wait(lk, [&]() { return failed.load() || work.pending[half] == 0; });
        }
        if (failed.load()) break;
        std::vector<GroupStep> grp(grouped.begin() + g0, grouped.begin() + std::min(grouped.size(), g0 + group_));
        for (size_t k = 0; k < grp.size(); k++) grp[k].slot = static_cast<int>(half * group_ + k);
        const auto ts = std::chrono::steady_clock::now();
        stage_err = stage_group(ph, grp);
        if (trace_) {
            std::lock_guard<std::mutex> lock(trace_mu_);
            std::fprintf(trace_, "{\"stage\": %zu, \"phase\": %u, \"steps\": %zu, \"at\": %.3f, \"ms\": %.3f}\n", gi, ph, grp.size(),
                         std::chrono::duration<double, std::milli>(ts - phase_t0).count(),
                         std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count());
        }
        if (!stage_err.empty()) {
            failed.store(true);
            break;
        }
        std::lock_guard<std::mutex> lk(work.mu);
        for (const GroupStep& st : grp) {
            if (st.np == 0) continue;  // no anchors, or an empty local set: nothing to search
            work.q.push_back(st);
            work.pending[half]++;
        }
        work.cv.notify_all();
    }
    {
        std::lock_guard<std::mutex> lk(work.mu);
        work.closed = true;
        work.cv.notify_all();
    }
    for (auto& t : th) t.join();
    if (trace_) std::fflush(trace_);
    if (!stage_err.empty()) {
        for (Lane& w : lanes_) (void)hipStreamSynchronize(w.s);
        return stage_err;
    }
    for (size_t li = 0; li < lanes_.size(); li++)
        if (!errs[li].empty()) {
            for (Lane& w : lanes_) (void)hipStreamSynchronize(w.s);
            return errs[li];
        }
    for (Lane& w : lanes_) CSM_HIP(hipStreamWaitEvent(s_, w.done, 0));
    return "";
}

std::string Engine::run_batch(uint64_t lo, uint64_t n, std::vector<QueryResult>& out, BatchInfo& info) {
    if (srv_) return tagged(Code::kState, "the session is serving: batches go through the request loop");
    if (!built_) return tagged(Code::kState, "index not built");
    if (lo + n > stream_.size()) return tagged(Code::kArgument, "batch range beyond the update stream");
    return run_resident(lo, n, stream_.src.data() + lo, stream_.dst.data() + lo, stream_.op.data() + lo, out, info);
}

uint64_t Engine::static_ring_max_edges() const {
    // a step's initial partial results take 2 words per anchor arc, at most 2 arcs per edge: half
    // of a lane's slice holds them, the other half the levels (which split when they overflow)
    const uint64_t lanes = std::max<uint64_t>(1, lanes_.empty() ? opt_.lanes : lanes_.size());
    return (kStaticRingWords / lanes & ~63ull) / 8;
}

std::string Engine::run_batch_static(uint64_t lo, uint64_t n, std::vector<QueryResult>& out, BatchInfo& info) {
    if (srv_) return tagged(Code::kState, "the session is serving: batches go through the request loop");
    if (!built_) return tagged(Code::kState, "index not built");
    if (lo + n > stream_.size()) return tagged(Code::kArgument, "batch range beyond the update stream");
    if (!static_ring_) return tagged(Code::kState, "static ring unavailable on this device");
    const uint64_t max_n = static_ring_max_edges();
    if (n > max_n)
        return tagged(Code::kLimit, "batch of " + std::to_string(n) + " edges exceeds the static ring's limit of " + std::to_string(max_n) +
                                        " edges with " + std::to_string(lanes_.size()) + " lanes");
    const uint64_t slice = kStaticRingWords / lanes_.size() & ~63ull;
    std::vector<std::pair<uint32_t*, uint64_t>> own;
    for (size_t k = 0; k < lanes_.size(); k++) {
        own.emplace_back(lanes_[k].ring, lanes_[k].ring_words);
        lanes_[k].ring = static_ring_ + k * slice;
        lanes_[k].ring_words = slice;
    }
    // run_resident returns with every lane's work finished (each phase joins the lanes, and its
    // failure path synchronises them), so no kernel touches the static ring after this call
    std::string err = run_resident(lo, n, stream_.src.data() + lo, stream_.dst.data() + lo, stream_.op.data() + lo, out, info);
    for (size_t k = 0; k < lanes_.size(); k++) {
        lanes_[k].ring = own[k].first;
        lanes_[k].ring_words = own[k].second;
    }
    return err;
}

// Arena admission: the words the batch's relocated lists will take, replaying the device rule
// (deletions first; a list whose new length exceeds its capacity moves to a fresh block of
// max(8, 2 * next_pow2(length))) on the host mirrors of degrees and capacities.
std::string Engine::run_edges(const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op, uint64_t n,
                              std::vector<QueryResult>& out, BatchInfo& info) {
    if (srv_) return tagged(Code::kState, "the session is serving: batches go through the request loop");
    if (!built_) return tagged(Code::kState, "index not built");
    if (n > max_batch_) return tagged(Code::kLimit, "batch larger than the scratch was sized for (max_batch_edges)");
    std::vector<uint32_t> kf(n), kb(n);
    for (uint64_t k = 0; k < n; k++) {
        const uint32_t a = src[k], b = dst[k];
        if (a >= dv_ || b >= dv_) return tagged(Code::kArgument, "batch edge endpoint out of range");
        if (a == b) return tagged(Code::kArgument, "self-loops are not supported");
        if (host_g_.vlabel[a] == kNoVertex || host_g_.vlabel[b] == kNoVertex) return tagged(Code::kArgument, "batch edge endpoint was never declared");
        if (rank_of(elabel_rank_, elabel[k]) == kNoVertex) return tagged(Code::kArgument, "batch edge label unknown to the session");
        kf[k] = key_of_edge(b, elabel[k]);
        kb[k] = key_of_edge(a, elabel[k]);
    }
    if (n) {  // the scratch region after the resident stream (the engine's own buffers)
        CSM_HIP(hipMemcpyAsync(const_cast<uint32_t*>(bin_.src) + stream_cap_, src, n * 4, hipMemcpyHostToDevice, s_));
        CSM_HIP(hipMemcpyAsync(const_cast<uint32_t*>(bin_.dst) + stream_cap_, dst, n * 4, hipMemcpyHostToDevice, s_));
        CSM_HIP(hipMemcpyAsync(const_cast<uint32_t*>(bin_.key_fwd) + stream_cap_, kf.data(), n * 4, hipMemcpyHostToDevice, s_));
        CSM_HIP(hipMemcpyAsync(const_cast<uint32_t*>(bin_.key_bwd) + stream_cap_, kb.data(), n * 4, hipMemcpyHostToDevice, s_));
        CSM_HIP(hipMemcpyAsync(const_cast<uint8_t*>(bin_.op) + stream_cap_, op, n, hipMemcpyHostToDevice, s_));
    }
    return run_resident(stream_cap_, n, src, dst, op, out, info);
}

// One batch whose edges sit on the device at [lo, lo + n) of the stream buffers (host copies at
// src / dst / op for admission and counting).
std::string Engine::run_resident(uint64_t lo, uint64_t n, const uint32_t* hsrc, const uint32_t* hdst, const uint8_t* hop,
                                 std::vector<QueryResult>& out, BatchInfo& info) {
    if (!failed_.empty()) return failed_;
    if (n > max_batch_) return tagged(Code::kLimit, "batch larger than the scratch was sized for");
    for (auto& qs : qs_) if (!qs->planned) return tagged(Code::kState, "every query must be planned before matching");
    (void)hsrc;
    (void)hdst;
    info = BatchInfo{};
    info.edges = n;
    for (uint64_t k = 0; k < n; k++) (hop[k] ? info.deletes : info.inserts)++;
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t max_arcs = static_cast<uint32_t>(2 * n);
    if (n == 0) { out.assign(qs_.size(), QueryResult{}); return ""; }
    prof_begin(prof_main_, s_, kKBatchBuild);
    for (int d = 0; d < ndirs_; d++) batch_build(g_, d, ndirs_ == 1 ? 0 : d + 1, bin_, static_cast<uint32_t>(lo), static_cast<uint32_t>(n), bs_[d], s_);
    // admission before anything changes: list space of the insertions, deletions of missing arcs
    for (int d = 0; d < ndirs_; d++) batch_admit(g_, d, max_arcs, bs_[d], adm_, s_);
    batch_admit_finish(g_, adm_, s_);
    prof_end(prof_main_, s_, kKBatchBuild);
    CSM_HIP(hipMemcpyAsync(h_pinned_, adm_ + 2, 32, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    unsigned long long a[4];
    std::memcpy(a, h_pinned_, 32);
    // Every failure from here on drops the built batch (delta segments, touched flags, counters),
    // so the next batch starts clean; a failure after the deletions were applied also leaves the
    // session refusing further batches, since the graph then holds half a batch.
    bool graph_changed = false;
    auto fail = [&](const std::string& why) -> std::string {
        for (Lane& w : lanes_) (void)hipStreamSynchronize(w.s);
        for (int d = 0; d < ndirs_; d++) batch_cleanup(g_, d, max_arcs, bs_[d], s_);
        (void)hipStreamSynchronize(s_);
        if (graph_changed) {
            failed_ = tagged(Code::kState, "an earlier batch failed after changing the graph (" + text_of(why) + "); reopen the session");
            return why + " (the session is now unusable: the graph holds part of the batch)";
        }
        return why;
    };
    std::string err = a[3] ? check_err("batch build") : std::string();
    if (!err.empty()) return fail(err);
    if (a[1] || a[2] + a[0] > h_arena_words_) {  // rejected: drop the built batch, nothing else changed
        if (a[1])
            return fail(tagged(Code::kPrecondition, "batch precondition violated: an inserted edge exists, a deleted edge is missing, or an edge occurs twice in the batch"));
        return fail(tagged(Code::kMemory, "adjacency arena budget exhausted: the batch needs " + std::to_string(a[0]) + " more words of list space"));
    }
    for (int d = 0; d < ndirs_; d++) {
        batch_bucket(g_, max_arcs, bs_[d], s_);
        nlf_touched(g_, d_entries_, n_entries_, d_groups_, bs_[d].touched, bs_[d].n_touched, max_arcs, 0, s_);
    }
    fill_u32(reinterpret_cast<uint32_t*>(results_), 0u, qs_.size() * 8, s_);
    const auto t1 = std::chrono::steady_clock::now();
    double match_ms = 0;
    if (info.deletes) {
        const auto a = std::chrono::steady_clock::now();
        err = run_phase(1);
        if (!err.empty()) return fail(err);
        CSM_HIP(hipStreamSynchronize(s_));
        match_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
        prof_begin(prof_main_, s_, kKApply);
        for (int d = 0; d < ndirs_; d++) {
            batch_apply_deletions(g_, d, max_arcs, bs_[d], s_);
            slot_dir_update(g_, d, bs_[d].touched, bs_[d].n_touched, max_arcs, s_);
        }
        prof_end(prof_main_, s_, kKApply);
        graph_changed = true;
    }
    if (info.inserts) {
        CSM_HIP(hipStreamSynchronize(s_));
        const auto a = std::chrono::steady_clock::now();
        err = run_phase(0);
        if (!err.empty()) return fail(err);
        CSM_HIP(hipStreamSynchronize(s_));
        match_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count();
        prof_begin(prof_main_, s_, kKApply);
        for (int d = 0; d < ndirs_; d++) {
            prof_begin(prof_main_, s_, kKMergeInsert);
            batch_apply_insertions(g_, d, max_arcs, bs_[d], bytes_ptr(kKMergeInsert), s_);
            prof_end(prof_main_, s_, kKMergeInsert);
            slot_dir_update(g_, d, bs_[d].touched, bs_[d].n_touched, max_arcs, s_);
        }
        free_blocks_merge(g_, s_);  // blocks the relocations left behind serve the next batches
        prof_end(prof_main_, s_, kKApply);
    }
    for (int d = 0; d < ndirs_; d++)
        nlf_touched(g_, d_entries_, n_entries_, d_groups_, bs_[d].touched, bs_[d].n_touched, max_arcs, 1, s_);
    for (int d = 0; d < ndirs_; d++) batch_cleanup(g_, d, max_arcs, bs_[d], s_);
    std::vector<unsigned long long> res(qs_.size() * 4);
    CSM_HIP(hipMemcpyAsync(res.data(), results_, res.size() * 8, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    err = check_err("batch apply");
    if (!err.empty()) {
        failed_ = tagged(Code::kState, "an earlier batch failed while it was applied to the graph (" + text_of(err) + "); reopen the session");
        return err;
    }
    err = prof_collect();
    if (!err.empty()) return err;
    out.assign(qs_.size(), QueryResult{});
    for (size_t q = 0; q < qs_.size(); q++) {
        out[q].pos = res[4 * q];
        out[q].pos_hash = res[4 * q + 1];
        out[q].neg = res[4 * q + 2];
        out[q].neg_hash = res[4 * q + 3];
    }
    const auto t2 = std::chrono::steady_clock::now();
    info.ms_total = std::chrono::duration<double, std::milli>(t2 - t0).count();
    info.ms_match = match_ms;
    info.ms_index = info.ms_total - match_ms;
    (void)t1;
    return "";
}


void Engine::prof_begin(Prof& P, hipStream_t st, uint32_t family) {
    if (!(prof_mask_ >> family & 1u)) return;
    if (P.used + 2 > P.pool.size()) {
        for (int k = 0; k < 64; k++) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            P.pool.push_back(e);
        }
    }
    hipEvent_t a = P.pool[P.used++];
    (void)hipEventRecord(a, st);
    P.open[family] = a;
}

void Engine::prof_end(Prof& P, hipStream_t st, uint32_t family) {
    if (!(prof_mask_ >> family & 1u) || !P.open[family]) return;
    hipEvent_t b = P.pool[P.used++];
    (void)hipEventRecord(b, st);
    P.pending.push_back({family, {P.open[family], b}});
    P.open[family] = nullptr;
}

std::string Engine::prof_collect() {
    if (!prof_mask_) return "";
    CSM_HIP(hipStreamSynchronize(s_));
    for (Lane& w : lanes_) CSM_HIP(hipStreamSynchronize(w.s));
    auto drain = [&](Prof& P) -> std::string {
        for (const auto& p : P.pending) {
            float ms = 0;
            CSM_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
            kstat_[p.first].launches += 1;
            kstat_[p.first].ms += ms;
        }
        P.pending.clear();
        P.used = 0;
        return "";
    };
    std::string err = drain(prof_main_);
    for (Lane& w : lanes_)
        if (err.empty()) err = drain(w.prof);
    return err;
}

std::string Engine::profile_enable(uint32_t mask) {
    if (!built_) return tagged(Code::kState, "build the index before profiling");
    std::string err = prof_collect();
    if (!err.empty()) return err;
    prof_mask_ = mask;
    for (auto& k : kstat_) k = KernelStat{};
    CSM_HIP(hipMemsetAsync(prof_bytes_, 0, kKFamilies * 8, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    return "";
}

std::string Engine::profile_read(uint32_t family, KernelStat& st) {
    if (family >= kKFamilies) return tagged(Code::kArgument, "kernel family out of range");
    std::string err = prof_collect();
    if (!err.empty()) return err;
    unsigned long long b[kKFamilies];
    CSM_HIP(hipMemcpyAsync(b, prof_bytes_, sizeof b, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    st = kstat_[family];
    st.bytes = b[family];
    return "";
}

std::string Engine::calibration_copy(uint64_t words, double& ms) {
    uint32_t *a = nullptr, *b = nullptr;
    CSM_HIP(hipMalloc(reinterpret_cast<void**>(&a), words * 4));
    CSM_HIP(hipMalloc(reinterpret_cast<void**>(&b), words * 4));
    hipEvent_t e0, e1;
    CSM_HIP(hipEventCreate(&e0));
    CSM_HIP(hipEventCreate(&e1));
    fill_u32(a, 1u, words, s_);
    CSM_HIP(hipEventRecord(e0, s_));
    copy_u32(a, b, words, s_);
    CSM_HIP(hipEventRecord(e1, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    float f = 0;
    CSM_HIP(hipEventElapsedTime(&f, e0, e1));
    ms = f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return "";
}

std::string Engine::synchronize() {
    CSM_HIP(hipStreamSynchronize(s_));
    for (Lane& w : lanes_) CSM_HIP(hipStreamSynchronize(w.s));
    return "";
}

// The regions snapshot() saves, in order: arena words [0, top), per direction the list headers,
// capacities and label-slot directory, then every query's candidate bitmaps.
namespace {
struct Region {
    void* p;
    uint64_t bytes;
};
}  // namespace

std::string Engine::snapshot() {
    if (srv_) return tagged(Code::kState, "the session is serving: batches go through the request loop");
    if (!built_) return tagged(Code::kState, "build the index before taking a snapshot");
    if (!failed_.empty()) return failed_;
    CSM_HIP(hipStreamSynchronize(s_));
    const uint64_t top = arena_used_words();
    std::vector<Region> rs{{g_.arena, top * 4}};
    for (int d = 0; d < ndirs_; d++) {
        rs.push_back({g_.dir[d].adj, dv_ * 8ull});
        rs.push_back({g_.dir[d].cap, dv_ * 4ull});
        if (g_.sdir[d]) rs.push_back({g_.sdir[d], (g_.nslots + 1ull) * dv_ * 4});
    }
    for (auto& qs : qs_) rs.push_back({qs->cand, static_cast<uint64_t>(qs->q.n) * words_ * 4});
    rs.push_back({g_.fl_top, kFreeClasses * 4ull});
    rs.push_back({g_.fl_blk, free_entries_ * 4});
    uint64_t need = 0;
    for (const Region& r : rs) need += (r.bytes + 255) & ~255ull;
    // the arena part may grow between snapshots: keep room for the whole arena
    const uint64_t cap = need - ((top * 4 + 255) & ~255ull) + ((h_arena_words_ * 4 + 255) & ~255ull);
    if (!snap_ || snap_bytes_ < cap) {
        if (snap_) CSM_HIP(hipFree(snap_));
        snap_ = nullptr;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&snap_), cap);
        if (e != hipSuccess) {
            snap_ = nullptr;
            snap_bytes_ = 0;
            return tagged(Code::kMemory, std::string("snapshot allocation of ") + std::to_string(cap >> 20) + " MiB failed: " + hipGetErrorString(e));
        }
        snap_bytes_ = cap;
    }
    uint64_t off = 0;
    for (const Region& r : rs) {
        if (r.bytes) CSM_HIP(hipMemcpyAsync(snap_ + off, r.p, r.bytes, hipMemcpyDeviceToDevice, s_));
        off += (r.bytes + 255) & ~255ull;
    }
    CSM_HIP(hipStreamSynchronize(s_));
    snap_top_ = top;
    return "";
}

std::string Engine::restore() {
    if (srv_) return tagged(Code::kState, "the session is serving: batches go through the request loop");
    if (!snap_) return tagged(Code::kState, "no snapshot taken");
    CSM_HIP(hipStreamSynchronize(s_));
    for (Lane& w : lanes_) CSM_HIP(hipStreamSynchronize(w.s));
    std::vector<Region> rs{{g_.arena, snap_top_ * 4}};
    for (int d = 0; d < ndirs_; d++) {
        rs.push_back({g_.dir[d].adj, dv_ * 8ull});
        rs.push_back({g_.dir[d].cap, dv_ * 4ull});
        if (g_.sdir[d]) rs.push_back({g_.sdir[d], (g_.nslots + 1ull) * dv_ * 4});
    }
    for (auto& qs : qs_) rs.push_back({qs->cand, static_cast<uint64_t>(qs->q.n) * words_ * 4});
    rs.push_back({g_.fl_top, kFreeClasses * 4ull});
    rs.push_back({g_.fl_blk, free_entries_ * 4});
    uint64_t off = 0;
    for (const Region& r : rs) {
        if (r.bytes) CSM_HIP(hipMemcpyAsync(r.p, snap_ + off, r.bytes, hipMemcpyDeviceToDevice, s_));
        off += (r.bytes + 255) & ~255ull;
    }
    const unsigned long long top = snap_top_;
    std::memcpy(h_pinned_ + 64, &top, 8);
    CSM_HIP(hipMemcpyAsync(g_.arena_top, h_pinned_ + 64, 8, hipMemcpyHostToDevice, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    failed_.clear();
    return check_err("restore");
}

std::string Engine::read_candidates(uint32_t qid, uint32_t u, std::vector<uint32_t>& bitmap) {
    if (!built_ || qid >= qs_.size() || u >= qs_[qid]->q.n) return tagged(Code::kArgument, "bad arguments");
    bitmap.resize(words_);
    CSM_HIP(hipMemcpyAsync(bitmap.data(), qs_[qid]->cand + static_cast<uint64_t>(u) * words_, words_ * 4ull, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    return "";
}

std::string Engine::read_adjacency(uint32_t v, uint32_t dir, std::vector<uint32_t>& ids, std::vector<uint32_t>& labels) {
    if (!built_ || v >= dv_ || dir >= static_cast<uint32_t>(ndirs_)) return tagged(Code::kArgument, "bad arguments");
    uint2 a;
    CSM_HIP(hipMemcpyAsync(&a, g_.dir[dir].adj + v, 8, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    std::vector<uint32_t> keys(a.y);
    if (a.y) CSM_HIP(hipMemcpyAsync(keys.data(), g_.arena + a.x, a.y * 4ull, hipMemcpyDeviceToHost, s_));
    CSM_HIP(hipStreamSynchronize(s_));
    ids.resize(a.y);
    labels.resize(a.y);
    for (uint32_t k = 0; k < a.y; k++) {
        ids[k] = keys[k] & g_.id_mask;
        const uint32_t slot = keys[k] >> id_bits_;
        labels[k] = elabel_rank_.empty() ? 0 : elabel_rank_[slot / nvl_].first;  // edge label value
    }
    return "";
}

}  // namespace csm
