// Launch settings of a session and their autotuner.
//
// The reference compiles its tunables in: the semi-BFS switch at MIN_NUM_RESULTS_TO_GPU = 65,536
// partial results (config.h:47, used at match_gpu.cu:1961) and the block and warp counts of its
// enumeration launches (match_gpu.cu:1721-2294).  Here they are session settings that change
// scheduling only, never results, and Engine::tune picks them for a workload by measurement:
// it replays a prefix of the update stream from a snapshot under each candidate setting, times
// the batches after the warm-up ones, and requires every candidate's per-batch results (counts
// and hashes of every query, both phases) to equal the first candidate's.
#include "engine.h"

#include <algorithm>
#include <chrono>

namespace csm {

Tunables Engine::tunables() const {
    Tunables t;
    t.semi_bfs_threshold = opt_.semi_bfs_threshold;
    t.waves_per_anchor = opt_.waves_per_anchor;
    t.stack_depth = opt_.stack_depth;
    t.dfs_blocks = opt_.dfs_blocks;
    t.local_blocks = static_cast<uint32_t>(local_grid_);
    t.team = team_;
    t.level_blocks = opt_.level_blocks;
    return t;
}

std::string Engine::set_tunables(const Tunables& t) {
    if (srv_) return tagged(Code::kState, "the session is serving: settings change between batches only");
    if (t.waves_per_anchor > 65535) return tagged(Code::kArgument, "waves_per_anchor outside 1..65535");
    if (t.stack_depth > kMaxQV) return tagged(Code::kArgument, "stack_depth outside 1..16");
    if (t.dfs_blocks && (t.dfs_blocks < 64 || t.dfs_blocks > 16384)) return tagged(Code::kArgument, "dfs_blocks outside 64..16384");
    if (t.local_blocks && (t.local_blocks < 16 || t.local_blocks > 2048)) return tagged(Code::kArgument, "local_blocks outside 16..2048");
    if (t.team > kMaxLanes) return tagged(Code::kArgument, "team outside 1..64");
    if (t.level_blocks && (t.level_blocks < 64 || t.level_blocks > 4096)) return tagged(Code::kArgument, "level_blocks outside 64..4096");
    if (t.semi_bfs_threshold) opt_.semi_bfs_threshold = t.semi_bfs_threshold;
    if (t.waves_per_anchor) opt_.waves_per_anchor = t.waves_per_anchor;
    if (t.stack_depth) opt_.stack_depth = t.stack_depth;
    if (t.dfs_blocks) opt_.dfs_blocks = t.dfs_blocks;
    if (t.local_blocks) local_grid_ = static_cast<int>(t.local_blocks);
    if (t.team) team_ = std::min<uint32_t>(t.team, std::max<size_t>(1, lanes_.size()));
    if (t.level_blocks) opt_.level_blocks = t.level_blocks;
    return "";
}

namespace {

bool same_results(const std::vector<QueryResult>& a, const std::vector<QueryResult>& b) {
    if (a.size() != b.size()) return false;
    for (size_t k = 0; k < a.size(); k++)
        if (a[k].pos != b[k].pos || a[k].neg != b[k].neg || a[k].pos_hash != b[k].pos_hash || a[k].neg_hash != b[k].neg_hash) return false;
    return true;
}

// Values the coordinate search tries for each setting, in this order of settings.
constexpr uint32_t kSemiBfs[] = {1u << 14, 1u << 16, 1u << 18, 1u << 20};
constexpr uint32_t kLocalBlocks[] = {64, 128, 256, 512};
constexpr uint32_t kWavesPerAnchor[] = {1, 16, 256, 1024};
constexpr uint32_t kStackDepth[] = {2, 4, kMaxQV};
constexpr uint32_t kDfsBlocks[] = {1024, 2048, 4096};
constexpr uint32_t kLevelBlocks[] = {256, 384, 512, 1024};
// A candidate replaces the best so far only when it is faster by this fraction (timing noise).
constexpr double kMinGain = 0.01;

}  // namespace

std::string Engine::tune(uint64_t first, uint64_t batch, uint32_t warmup, uint32_t timed, const std::vector<Tunables>& cands,
                         std::vector<TuneEval>& trail, uint32_t& best) {
    trail.clear();
    best = 0;
    if (srv_) return tagged(Code::kState, "the session is serving: settings change between batches only");
    if (!built_) return tagged(Code::kState, "index not built");
    if (batch == 0) batch = std::min<uint64_t>(batch_size_, stream_.size());
    if (batch == 0 || timed == 0) return tagged(Code::kArgument, "tuning needs a batch size and at least one timed batch");
    const uint64_t nb = static_cast<uint64_t>(warmup) + timed;
    // every batch of the range holds edges; the last one may end early, at the end of the stream
    if (batch > max_batch_ || first >= stream_.size() || (nb - 1) * batch >= stream_.size() - first)
        return tagged(Code::kArgument, "tuning range beyond the update stream or the session's batch size");
    for (const Tunables& c : cands) {  // reject bad candidates before anything runs
        const Tunables keep = tunables();
        std::string err = set_tunables(c);
        set_tunables(keep);
        if (!err.empty()) return err;
    }
    std::string err = snapshot();
    if (!err.empty()) return err;
    const Tunables start = tunables();
    std::vector<std::vector<QueryResult>> want;  // per batch, from the first evaluation

    auto eval = [&](const Tunables& t) -> std::string {
        std::string e = set_tunables(t);
        if (e.empty()) e = restore();
        if (!e.empty()) return e;
        double ms = 0;
        for (uint64_t b = 0; b < nb; b++) {
            std::vector<QueryResult> out;
            BatchInfo info;
            const auto t0 = std::chrono::steady_clock::now();
            const uint64_t lo = first + b * batch;
            e = run_batch(lo, std::min<uint64_t>(batch, stream_.size() - lo), out, info);  // returns after the batch completed
            if (!e.empty()) return e;
            if (b >= warmup) ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            if (want.size() <= b) want.push_back(out);
            else if (!same_results(out, want[b]))
                return tagged(Code::kMismatch, "tuning: two launch settings gave different results on batch " + std::to_string(b) +
                                                   " of the tuning range (a library fault)");
        }
        trail.push_back(TuneEval{tunables(), ms / timed});
        return "";
    };

    if (!cands.empty()) {
        for (const Tunables& c : cands) {
            Tunables t = start;  // fields of 0 keep the session's setting
            if (c.semi_bfs_threshold) t.semi_bfs_threshold = c.semi_bfs_threshold;
            if (c.waves_per_anchor) t.waves_per_anchor = c.waves_per_anchor;
            if (c.stack_depth) t.stack_depth = c.stack_depth;
            if (c.dfs_blocks) t.dfs_blocks = c.dfs_blocks;
            if (c.local_blocks) t.local_blocks = c.local_blocks;
            if (c.team) t.team = c.team;
            if (c.level_blocks) t.level_blocks = c.level_blocks;
            if (!(err = eval(t)).empty()) break;
            if (trail.back().ms_per_batch < trail[best].ms_per_batch) best = static_cast<uint32_t>(trail.size() - 1);
        }
    } else {
        // coordinate search from the current settings
        err = eval(start);
        auto coord = [&](uint32_t Tunables::*field, const uint32_t* vals, size_t nv) {
            const Tunables base = trail[best].t;
            for (size_t k = 0; k < nv && err.empty(); k++) {
                if (vals[k] == base.*field) continue;
                Tunables t = base;
                t.*field = vals[k];
                if (!(err = eval(t)).empty()) return;
                if (trail.back().ms_per_batch < trail[best].ms_per_batch * (1.0 - kMinGain)) best = static_cast<uint32_t>(trail.size() - 1);
            }
        };
        if (err.empty()) coord(&Tunables::semi_bfs_threshold, kSemiBfs, std::size(kSemiBfs));
        if (err.empty()) coord(&Tunables::local_blocks, kLocalBlocks, std::size(kLocalBlocks));
        if (err.empty()) coord(&Tunables::waves_per_anchor, kWavesPerAnchor, std::size(kWavesPerAnchor));
        if (err.empty()) coord(&Tunables::stack_depth, kStackDepth, std::size(kStackDepth));
        if (err.empty()) coord(&Tunables::dfs_blocks, kDfsBlocks, std::size(kDfsBlocks));
        if (err.empty() && lanes_.size() > 1) {  // an eighth, a quarter and half of the lanes: 8, 4 or 2 teams
            const uint32_t nl = static_cast<uint32_t>(lanes_.size());
            const uint32_t teams[] = {std::max(1u, nl / 8), std::max(1u, nl / 4), std::max(1u, nl / 2)};
            coord(&Tunables::team, teams, std::size(teams));
        }
        if (err.empty() && trail[best].t.team > 1) coord(&Tunables::level_blocks, kLevelBlocks, std::size(kLevelBlocks));
    }
    // leave the graph as it was, with the best settings (the starting ones after a failure)
    const std::string back = set_tunables(err.empty() ? trail[best].t : start);
    const std::string res = restore();
    if (!err.empty()) return err;
    return !back.empty() ? back : res;
}

}  // namespace csm
