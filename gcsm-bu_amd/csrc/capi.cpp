// C ABI over the engine (include/csm.h).
#include <algorithm>
#include "../../include/csm.h"

#include <cstring>
#include <string>
#include <vector>

#include "engine.h"

using namespace csm;

struct csm_session {
    Engine engine;
    DataGraph graph;
    bool have_graph = false;
    std::string error;
    EngineOptions opt;
};

namespace {

csm_status fail(csm_session* s, csm_status st, const std::string& msg) {
    if (s) s->error = msg;
    return st;
}

// The status a failing engine or parser call carries in its message tag (untagged: HIP failure).
csm_status classify(csm_session* s, const std::string& err) {
    if (err.empty()) return CSM_OK;
    return fail(s, static_cast<csm_status>(code_of(err) - '0'), text_of(err));
}

}  // namespace

extern "C" {

csm_status csm_session_create(const csm_options* options, csm_session** out) {
    if (!out) return CSM_ERR_ARGUMENT;
    *out = nullptr;
    auto* s = new csm_session();
    EngineOptions o;
    if (options) {
        o.device = options->device;
        o.directed = options->directed != 0;
        o.pool_bytes = options->pool_bytes;
        if (options->ring_bytes) o.ring_bytes = options->ring_bytes;
        if (options->cp_list_bytes) o.cp_list_bytes = options->cp_list_bytes;
        o.max_batch_edges = options->max_batch_edges;
        if (options->semi_bfs_threshold) o.semi_bfs_threshold = options->semi_bfs_threshold;
        if (options->lanes) o.lanes = options->lanes;
    }
    s->opt = o;
    std::string err = s->engine.init(o);
    if (!err.empty()) {
        // keep the session so the caller can read the message
        s->error = err;
        *out = s;
        return CSM_ERR_DEVICE;
    }
    *out = s;
    return CSM_OK;
}

void csm_session_destroy(csm_session* s) { delete s; }

const char* csm_last_error(const csm_session* s) { return s ? s->error.c_str() : "no session"; }

csm_status csm_load_graph_file(csm_session* s, const char* path) {
    if (!s || !path) return CSM_ERR_ARGUMENT;
    std::string err = load_data_graph(path, s->graph);
    if (!err.empty()) return classify(s, err);
    s->have_graph = true;
    return classify(s, s->engine.set_graph(s->graph));
}

csm_status csm_load_query_file(csm_session* s, const char* path, uint32_t* query_id) {
    if (!s || !path) return CSM_ERR_ARGUMENT;
    Query q;
    std::string err = parse_query(path, q);
    if (!err.empty()) return classify(s, err);
    uint32_t id = 0;
    err = s->engine.add_query(q, id);
    if (!err.empty()) return classify(s, err);
    if (query_id) *query_id = id;
    return CSM_OK;
}

csm_status csm_open_update_file(csm_session* s, const char* path, uint32_t batch_size, uint64_t* num_batches) {
    if (!s || !path) return CSM_ERR_ARGUMENT;
    if (!s->have_graph) return fail(s, CSM_ERR_STATE, "load the data graph before the update file");
    UpdateStream us;
    std::string err = load_update_stream(path, s->graph, us);
    if (!err.empty()) return classify(s, err);
    err = s->engine.set_stream(us, batch_size);
    if (!err.empty()) return classify(s, err);
    if (num_batches) *num_batches = s->engine.num_batches();
    return CSM_OK;
}

csm_status csm_build_index(csm_session* s) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.build());
}

csm_status csm_plan(csm_session* s, uint32_t query_id, const csm_technique* t) {
    if (!s) return CSM_ERR_ARGUMENT;
    Technique tech;
    if (t) {
        if (t->matching_order < 0 || t->matching_order > 3 || t->search_strategy < 0 || t->search_strategy > 3)
            return fail(s, CSM_ERR_ARGUMENT, "matching order and search strategy must be 0..3");
        tech.order = static_cast<MatchingOrder>(t->matching_order);
        tech.strategy = static_cast<SearchStrategy>(t->search_strategy);
        tech.local_index = t->local_index != 0;
        tech.cartesian_product = t->cartesian_product != 0;
        tech.heavy_vertex_avoidance = t->heavy_vertex_avoidance != 0;
        tech.relation_switch = t->relation_switch != 0;
    }
    return classify(s, s->engine.plan(query_id, tech));
}

csm_status csm_get_plan_text(csm_session* s, uint32_t query_id, char* buf, size_t cap, size_t* needed) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (query_id >= s->engine.num_queries() || !s->engine.planned(query_id)) return fail(s, CSM_ERR_STATE, "query not planned");
    const std::string t = print_orders(s->engine.query(query_id), s->engine.plan_of(query_id));
    if (needed) *needed = t.size() + 1;
    if (buf && cap) {
        const size_t n = std::min(cap - 1, t.size());
        std::memcpy(buf, t.data(), n);
        buf[n] = 0;
    }
    return CSM_OK;
}

csm_status csm_get_summary(csm_session* s, uint32_t query_id, uint32_t* card, float* avg, uint32_t cap, uint32_t* num_relations) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (query_id >= s->engine.num_queries()) return fail(s, CSM_ERR_ARGUMENT, "query id out of range");
    const auto& c = s->engine.cardinalities(query_id);
    const auto& a = s->engine.avg_degrees(query_id);
    if (c.empty()) return fail(s, CSM_ERR_STATE, "index not built");
    if (num_relations) *num_relations = static_cast<uint32_t>(c.size());
    for (uint32_t r = 0; r < c.size() && r < cap; r++) {
        if (card) card[r] = c[r];
        if (avg) avg[r] = a[r];
    }
    return CSM_OK;
}

csm_status csm_apply_batch(csm_session* s, uint64_t batch_index, csm_query_result* results, uint32_t num_queries, csm_batch_info* info) {
    if (!s) return CSM_ERR_ARGUMENT;
    const uint64_t bs = s->engine.batch_size();
    if (batch_index >= s->engine.num_batches()) return fail(s, CSM_ERR_ARGUMENT, "batch index out of range");
    const uint64_t lo = batch_index * bs;
    const uint64_t n = std::min<uint64_t>(bs, s->engine.stream().size() - lo);
    std::vector<QueryResult> out;
    BatchInfo bi;
    std::string err = s->engine.run_batch(lo, n, out, bi);
    if (!err.empty()) return classify(s, err);
    for (uint32_t q = 0; q < num_queries && q < out.size(); q++) {
        results[q].positive = out[q].pos;
        results[q].negative = out[q].neg;
        results[q].positive_hash = out[q].pos_hash;
        results[q].negative_hash = out[q].neg_hash;
    }
    if (info) {
        info->edges = bi.edges;
        info->inserts = bi.inserts;
        info->deletes = bi.deletes;
        info->ms_total = bi.ms_total;
        info->ms_index = bi.ms_index;
        info->ms_match = bi.ms_match;
    }
    return CSM_OK;
}

namespace {
csm_status copy_results(const std::vector<QueryResult>& out, const BatchInfo& bi, csm_query_result* results, uint32_t num_queries,
                        csm_batch_info* info) {
    for (uint32_t q = 0; q < num_queries && q < out.size(); q++) {
        results[q].positive = out[q].pos;
        results[q].negative = out[q].neg;
        results[q].positive_hash = out[q].pos_hash;
        results[q].negative_hash = out[q].neg_hash;
    }
    if (info) {
        info->edges = bi.edges;
        info->inserts = bi.inserts;
        info->deletes = bi.deletes;
        info->ms_total = bi.ms_total;
        info->ms_index = bi.ms_index;
        info->ms_match = bi.ms_match;
    }
    return CSM_OK;
}
}  // namespace

csm_status csm_apply_range(csm_session* s, uint64_t first, uint64_t n, csm_query_result* results, uint32_t num_queries,
                           csm_batch_info* info) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (num_queries && !results) return fail(s, CSM_ERR_ARGUMENT, "results array missing");
    std::vector<QueryResult> out;
    BatchInfo bi;
    std::string err = s->engine.run_batch(first, n, out, bi);
    if (!err.empty()) return classify(s, err);
    return copy_results(out, bi, results, num_queries, info);
}

csm_status csm_apply_range_static(csm_session* s, uint64_t first, uint64_t n, csm_query_result* results, uint32_t num_queries,
                                  csm_batch_info* info) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (num_queries && !results) return fail(s, CSM_ERR_ARGUMENT, "results array missing");
    std::vector<QueryResult> out;
    BatchInfo bi;
    std::string err = s->engine.run_batch_static(first, n, out, bi);
    if (!err.empty()) return classify(s, err);
    return copy_results(out, bi, results, num_queries, info);
}

csm_status csm_apply_edges(csm_session* s, const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op,
                           uint64_t n, csm_query_result* results, uint32_t num_queries, csm_batch_info* info) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (n && (!src || !dst || !elabel || !op)) return fail(s, CSM_ERR_ARGUMENT, "edge arrays missing");
    if (num_queries && !results) return fail(s, CSM_ERR_ARGUMENT, "results array missing");
    std::vector<QueryResult> out;
    BatchInfo bi;
    std::string err = s->engine.run_edges(src, dst, elabel, op, n, out, bi);
    if (!err.empty()) return classify(s, err);
    return copy_results(out, bi, results, num_queries, info);
}

csm_status csm_serve_start(csm_session* s) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.serve_start());
}

csm_status csm_serve_submit(csm_session* s, const uint32_t* src, const uint32_t* dst, const uint32_t* elabel, const uint8_t* op,
                            uint64_t n, uint64_t* ticket) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (!ticket) return fail(s, CSM_ERR_ARGUMENT, "ticket pointer missing");
    if (n && (!src || !dst || !elabel || !op)) return fail(s, CSM_ERR_ARGUMENT, "edge arrays missing");
    return classify(s, s->engine.serve_submit(src, dst, elabel, op, n, *ticket));
}

csm_status csm_serve_wait(csm_session* s, uint64_t ticket, csm_query_result* results, uint32_t num_queries, csm_batch_info* info,
                          double* latency_ms) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (num_queries && !results) return fail(s, CSM_ERR_ARGUMENT, "results array missing");
    std::vector<QueryResult> out;
    BatchInfo bi;
    double ms = 0;
    std::string err = s->engine.serve_wait(ticket, out, bi, ms);
    if (latency_ms) *latency_ms = ms;
    if (!err.empty()) return classify(s, err);
    return copy_results(out, bi, results, num_queries, info);
}

csm_status csm_serve_stop(csm_session* s) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.serve_stop());
}

csm_status csm_arena_usage(csm_session* s, uint64_t* used_words, uint64_t* capacity_words) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (used_words) {
        const uint64_t top = s->engine.arena_used_words();
        *used_words = top - std::min(top, s->engine.arena_free_words());
    }
    if (capacity_words) *capacity_words = s->engine.arena_words();
    return CSM_OK;
}

csm_status csm_get_relation_edge_counts(csm_session* s, uint32_t query_id, uint64_t batch_index, uint64_t* counts, uint32_t cap,
                                        uint32_t* num_relations) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (query_id >= s->engine.num_queries()) return fail(s, CSM_ERR_ARGUMENT, "query id out of range");
    const Query& q = s->engine.query(query_id);
    std::vector<uint64_t> c;
    if (batch_index == UINT64_MAX) {
        relation_edge_counts(q, s->graph, s->graph.src.data(), s->graph.dst.data(), s->graph.elabel.data(), s->graph.src.size(), c);
    } else {
        const UpdateStream& us = s->engine.stream();
        const uint64_t bs = s->engine.batch_size();
        if (batch_index >= s->engine.num_batches()) return fail(s, CSM_ERR_ARGUMENT, "batch index out of range");
        const uint64_t lo = batch_index * bs, n = std::min<uint64_t>(bs, us.size() - lo);
        relation_edge_counts(q, s->graph, us.src.data() + lo, us.dst.data() + lo, us.elabel.data() + lo, n, c);
    }
    if (num_relations) *num_relations = static_cast<uint32_t>(c.size());
    for (uint32_t r = 0; r < c.size() && r < cap; r++) counts[r] = c[r];
    return CSM_OK;
}

csm_status csm_get_counts(csm_session* s, uint32_t* nv, uint32_t* nq) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (nv) *nv = s->graph.dv;
    if (nq) *nq = s->engine.num_queries();
    return CSM_OK;
}

csm_status csm_pool_usage(csm_session* s, uint64_t* used, uint64_t* capacity) {
    if (!s) return CSM_ERR_ARGUMENT;
    if (used) *used = s->engine.pool_used();
    if (capacity) *capacity = s->engine.pool_capacity();
    return CSM_OK;
}

csm_status csm_read_candidates(csm_session* s, uint32_t query_id, uint32_t u, uint32_t* words, uint32_t cap) {
    if (!s || !words) return CSM_ERR_ARGUMENT;
    std::vector<uint32_t> bm;
    std::string err = s->engine.read_candidates(query_id, u, bm);
    if (!err.empty()) return classify(s, err);
    if (cap < bm.size()) return fail(s, CSM_ERR_ARGUMENT, "buffer too small");
    std::memcpy(words, bm.data(), bm.size() * 4);
    return CSM_OK;
}

csm_status csm_read_adjacency(csm_session* s, uint32_t v, uint32_t dir, uint32_t* ids, uint32_t cap, uint32_t* degree) {
    if (!s) return CSM_ERR_ARGUMENT;
    std::vector<uint32_t> idv, lab;
    std::string err = s->engine.read_adjacency(v, dir, idv, lab);
    if (!err.empty()) return classify(s, err);
    if (degree) *degree = static_cast<uint32_t>(idv.size());
    if (ids)
        for (uint32_t k = 0; k < idv.size() && k < cap; k++) ids[k] = idv[k];
    return CSM_OK;
}

csm_status csm_profile_enable(csm_session* s, uint32_t mask) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.profile_enable(mask));
}

csm_status csm_profile_read(csm_session* s, uint32_t family, uint64_t* launches, double* ms, uint64_t* bytes) {
    if (!s) return CSM_ERR_ARGUMENT;
    Engine::KernelStat st;
    std::string err = s->engine.profile_read(family, st);
    if (!err.empty()) return classify(s, err);
    if (launches) *launches = st.launches;
    if (ms) *ms = st.ms;
    if (bytes) *bytes = st.bytes;
    return CSM_OK;
}

csm_status csm_calibration_copy(csm_session* s, uint64_t words, double* ms) {
    if (!s) return CSM_ERR_ARGUMENT;
    double m = 0;
    std::string err = s->engine.calibration_copy(words, m);
    if (!err.empty()) return classify(s, err);
    if (ms) *ms = m;
    return CSM_OK;
}

csm_status csm_snapshot(csm_session* s) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.snapshot());
}

csm_status csm_restore(csm_session* s) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.restore());
}

csm_status csm_get_tunables(csm_session* s, csm_tunables* out) {
    if (!s || !out) return CSM_ERR_ARGUMENT;
    const Tunables t = s->engine.tunables();
    *out = csm_tunables{t.semi_bfs_threshold, t.waves_per_anchor, t.stack_depth, t.dfs_blocks, t.local_blocks, t.team, t.level_blocks};
    return CSM_OK;
}

csm_status csm_set_tunables(csm_session* s, const csm_tunables* t) {
    if (!s || !t) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.set_tunables(Tunables{t->semi_bfs_threshold, t->waves_per_anchor, t->stack_depth, t->dfs_blocks,
                                                       t->local_blocks, t->team, t->level_blocks}));
}

csm_status csm_tune(csm_session* s, uint64_t first, uint64_t batch_edges, uint32_t warmup, uint32_t batches,
                    const csm_tunables* candidates, uint32_t num_candidates, csm_tunables* evaluated, double* ms_per_batch,
                    uint32_t cap, uint32_t* num_evaluated, uint32_t* best) {
    if (!s || (num_candidates && !candidates)) return CSM_ERR_ARGUMENT;
    std::vector<Tunables> cands;
    for (uint32_t k = 0; k < num_candidates; k++) {
        const csm_tunables& c = candidates[k];
        cands.push_back(Tunables{c.semi_bfs_threshold, c.waves_per_anchor, c.stack_depth, c.dfs_blocks, c.local_blocks, c.team, c.level_blocks});
    }
    std::vector<Engine::TuneEval> trail;
    uint32_t b = 0;
    const csm_status st = classify(s, s->engine.tune(first, batch_edges, warmup, batches, cands, trail, b));
    if (num_evaluated) *num_evaluated = static_cast<uint32_t>(trail.size());
    if (best) *best = b;
    for (size_t k = 0; k < trail.size() && k < cap; k++) {
        const Tunables& t = trail[k].t;
        if (evaluated) evaluated[k] = csm_tunables{t.semi_bfs_threshold, t.waves_per_anchor, t.stack_depth, t.dfs_blocks, t.local_blocks, t.team, t.level_blocks};
        if (ms_per_batch) ms_per_batch[k] = trail[k].ms_per_batch;
    }
    return st;
}

csm_status csm_synchronize(csm_session* s) {
    if (!s) return CSM_ERR_ARGUMENT;
    return classify(s, s->engine.synchronize());
}

}  // extern "C"
