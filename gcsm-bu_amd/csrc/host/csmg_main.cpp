// csmg: command line of the matching path, over the C ABI (include/csm.h).
//
// Flags, defaults and algorithm presets follow main.cpp:162-189 and TechniqueOptionBuilder
// (utils/technique_option.h:57-86); argv lookup follows InputParser (simple_command_parser.h:13-76):
// "--flag value", first occurrence wins, booleans "true"/"True"/"false"/"False".  Output lines keep the
// reference's formats: "<Phase>, time (ms): X(host), Y(kernel)" (cuda_helpers.h:98-102), the
// "Orders:" block (plan.cpp:1200-1291) and "Num Positive Matches: N" (technique_permutation.h:224).
// Extensions: --queries a,b,c (several standing queries), --directed, --pool_mb, --ring_mb, --lanes,
// --semi_bfs_threshold, --json <path> (one JSON record per batch and query), the launch settings
// --waves_per_anchor, --stack_depth, --dfs_blocks, --local_blocks, --team, --level_blocks, and --autotune true (the
// library's autotuner, csm_tune, on the first --autotune_batches batches of the update stream
// before the timed matching; it leaves the graph as it found it).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/csm.h"

namespace {

class InputParser {
  public:
    InputParser(int argc, char** argv) {
        for (int i = 1; i < argc; i++) tokens_.emplace_back(argv[i]);
    }
    std::string get(const std::string& o) const {
        auto it = std::find(tokens_.begin(), tokens_.end(), o);
        if (it != tokens_.end() && ++it != tokens_.end()) return *it;
        return "";
    }
    int32_t get_i32(const std::string& o, int32_t d) const {
        const std::string v = get(o);
        return v.empty() ? d : std::stoi(v);
    }
    uint32_t get_u32(const std::string& o, uint32_t d) const {
        const std::string v = get(o);
        return v.empty() ? d : static_cast<uint32_t>(std::stoul(v));
    }
    bool get_bool(const std::string& o, bool d) const {
        const std::string v = get(o);
        if (v == "true" || v == "True") return true;
        if (v == "false" || v == "False") return false;
        return d;
    }
    std::string cmd() const {
        std::string s = "csmg";
        for (const auto& t : tokens_) s += " " + t;
        return s;
    }

  private:
    std::vector<std::string> tokens_;
};

const char* kRule = "--------------------------------------------------------------------";

void print_info(const std::string& q, const std::string& d, const std::string& u, int dev, uint32_t bs, bool mem, const csm_technique& t,
                uint32_t semi_bfs) {
    static const char* orders[] = {"BFS Order", "Gamma Order", "RapidCSM Order", "RI Order"};
    static const char* strat[] = {"DFS w/ Work Stealing", "DFS w/o Work Stealing", "Dynamic BFS DFS", "BFS"};
    std::cout << kRule << std::endl;
    std::cout << "Command Line:\n\tQuery Graph: " << q << "\n\tData Graph: " << d << "\n\tData Graph Update: " << u
              << "\n\tDevice ID: " << dev << "\n\tBatch Size: " << bs << "\n\tShow GPU Memory?: " << (mem ? "yes" : "no") << '\n';
    std::cout << kRule << std::endl;
    std::cout << "Technique Option:\n\tMatching Order: " << orders[t.matching_order & 3] << "\n\tSearch Strategy: " << strat[t.search_strategy & 3]
              << "\n\tUse Local Index?: " << (t.local_index ? "w/ Local Index" : "w/o Local Index")
              << "\n\tUse Cartesian Product?: " << (t.cartesian_product ? "w/ Cartesian Product" : "w/o Use Cartesian Product")
              << "\n\tUse Heavy Vertex Avoidance?: "
              << (t.heavy_vertex_avoidance ? "w/ Heavy Vertex Avoidance" : "w/o Use Heavy Vertex Avoidance")
              << "\n\tUse Relation Switch?: " << (t.relation_switch ? "w/ Relation Switch" : "w/o Use Relation Switch") << '\n';
    std::cout << kRule << std::endl;
    std::cout << "Macros (Items after the third one are for GPU configuration.):\nMAX_QV_COUNT: 16\nMAX_QE_COUNT: 42\nMIN_NBR_SIZE: 8\n"
              << "WAVE_SIZE: 64\nMIN_NUM_RESULTS_TO_GPU: " << semi_bfs << '\n';
    std::cout << kRule << std::endl;
}

void time_line(const char* name, double host_ms, double kernel_ms) {
    std::cout << name << ", time (ms): " << static_cast<unsigned long>(host_ms) << "(host), " << static_cast<unsigned long>(kernel_ms)
              << "(kernel)\n";
}

int die(csm_session* s, csm_status st) {
    std::cout << "error (" << st << "): " << csm_last_error(s) << std::endl;
    csm_session_destroy(s);
    return 1;
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

int main(int argc, char** argv) {
    InputParser p(argc, argv);
    std::cout << p.cmd() << std::endl;
    const std::string query = p.get("--query");
    const std::string data = p.get("--data");
    const std::string update = p.get("--update");
    const int32_t device = p.get_i32("--device", 0);
    const uint32_t batch_size = p.get_u32("--batch_size", UINT32_MAX);
    const int32_t algorithm = p.get_i32("--algorithm", 0);
    const bool show_mem = p.get_bool("--show_gpu_memory", false);
    const bool print_index = p.get_bool("--print_indexing_time", false);

    csm_technique t{};
    switch (algorithm) {  // technique_option.h:57-86
        case 1: t = csm_technique{2, 2, 1, 1, 1, 1}; break;
        case 2: t = csm_technique{1, 0, 0, 1, 0, 0}; break;
        case 4: t = csm_technique{0, 2, 0, 1, 1, 1}; break;
        case 5: t = csm_technique{1, 2, 0, 1, 1, 1}; break;
        default:
            t.matching_order = p.get_i32("--matching_order", 2);
            t.search_strategy = p.get_i32("--search_strategy", 2);
            t.local_index = p.get_bool("--local_index", true);
            t.cartesian_product = p.get_bool("--cartesian_product", true);
            t.heavy_vertex_avoidance = p.get_bool("--heavy_vertex_avoidance", true);
            t.relation_switch = p.get_bool("--relation_switch", true);
    }
    std::vector<std::string> queries;
    if (!query.empty()) queries.push_back(query);
    {
        std::stringstream ss(p.get("--queries"));
        std::string item;
        while (std::getline(ss, item, ',')) if (!item.empty()) queries.push_back(item);
    }
    uint32_t semi_bfs = p.get_u32("--semi_bfs_threshold", CSM_DEFAULT_SEMI_BFS_THRESHOLD);
    if (semi_bfs == 0) semi_bfs = CSM_DEFAULT_SEMI_BFS_THRESHOLD;  // the library's meaning of 0: print the value in use
    print_info(query, data, update, device, batch_size, show_mem, t, semi_bfs);

    csm_options o{};
    o.device = device;
    o.directed = p.get_bool("--directed", false) ? 1 : 0;
    o.pool_bytes = static_cast<uint64_t>(p.get_u32("--pool_mb", 0)) << 20;
    o.ring_bytes = static_cast<uint64_t>(p.get_u32("--ring_mb", 0)) << 20;
    o.lanes = p.get_u32("--lanes", 0);
    o.semi_bfs_threshold = semi_bfs;
    csm_session* s = nullptr;
    csm_status st = csm_session_create(&o, &s);
    if (st != CSM_OK) return die(s, st);

    std::cout << "----------- Read Graphs from Files ------------\n";
    auto t0 = std::chrono::steady_clock::now();
    std::vector<uint32_t> qids;
    for (const auto& q : queries) {
        uint32_t id;
        if ((st = csm_load_query_file(s, q.c_str(), &id)) != CSM_OK) return die(s, st);
        qids.push_back(id);
    }
    if ((st = csm_load_graph_file(s, data.c_str())) != CSM_OK) return die(s, st);
    uint64_t nb = 0;
    if ((st = csm_open_update_file(s, update.c_str(), batch_size, &nb)) != CSM_OK) return die(s, st);
    uint32_t dv = 0, nq = 0;
    csm_get_counts(s, &dv, &nq);
    std::cout << "#data graph vertices: " << dv << std::endl << "#initial data graph edges: " << std::endl;
    {
        std::vector<uint64_t> c(128);
        uint32_t nr = 0;
        csm_get_relation_edge_counts(s, 0, UINT64_MAX, c.data(), 128, &nr);
        uint64_t tot = 0;
        for (uint32_t r = 0; r < nr; r++) { std::cout << "edge_idx: " << r << ", #edges: " << c[r] << std::endl; tot += c[r]; }
        std::cout << "total #edges: " << tot << std::endl;
    }
    time_line("Read Graphs from Files", ms_since(t0), 0);

    std::cout << "----------- Preprocessing ------------\n";
    t0 = std::chrono::steady_clock::now();
    if ((st = csm_build_index(s)) != CSM_OK) return die(s, st);
    time_line("Build Global Index Offline", ms_since(t0), ms_since(t0));
    t0 = std::chrono::steady_clock::now();
    for (uint32_t id : qids)
        if ((st = csm_plan(s, id, &t)) != CSM_OK) return die(s, st);
    time_line("Build Matching Order", ms_since(t0), 0);
    for (uint32_t id : qids) {
        size_t need = 0;
        csm_get_plan_text(s, id, nullptr, 0, &need);
        std::string txt(need, '\0');
        csm_get_plan_text(s, id, &txt[0], need, &need);
        if (qids.size() > 1) std::cout << "query " << id << " (" << queries[id] << ")\n";
        std::cout << txt.c_str();
    }
    {
        csm_tunables tn{};
        tn.waves_per_anchor = p.get_u32("--waves_per_anchor", 0);
        tn.stack_depth = p.get_u32("--stack_depth", 0);
        tn.dfs_blocks = p.get_u32("--dfs_blocks", 0);
        tn.local_blocks = p.get_u32("--local_blocks", 0);
        tn.team = p.get_u32("--team", 0);
        tn.level_blocks = p.get_u32("--level_blocks", 0);
        if ((st = csm_set_tunables(s, &tn)) != CSM_OK) return die(s, st);
    }
    if (p.get_bool("--autotune", false) && nb > 0) {
        const uint32_t k = static_cast<uint32_t>(std::min<uint64_t>(nb, std::max<uint32_t>(1, p.get_u32("--autotune_batches", 3))));
        const uint32_t warm = k > 1 ? 1 : 0;
        std::vector<csm_tunables> ev(64);
        std::vector<double> ms(64);
        uint32_t n = 0, best = 0;
        t0 = std::chrono::steady_clock::now();
        if ((st = csm_tune(s, 0, 0, warm, k - warm, nullptr, 0, ev.data(), ms.data(), 64, &n, &best)) != CSM_OK) return die(s, st);
        for (uint32_t e = 0; e < n && e < 64; e++)
            std::cout << "Autotune candidate " << e << ": semi_bfs_threshold " << ev[e].semi_bfs_threshold << ", waves_per_anchor "
                      << ev[e].waves_per_anchor << ", stack_depth " << ev[e].stack_depth << ", dfs_blocks " << ev[e].dfs_blocks
                      << ", local_blocks " << ev[e].local_blocks << ", team " << ev[e].team << ", level_blocks " << ev[e].level_blocks << ": " << ms[e] << " ms per batch"
                      << (e == best ? " (best)" : "") << '\n';
        time_line("Autotune", ms_since(t0), 0);
    }
    {
        csm_tunables tn{};
        csm_get_tunables(s, &tn);
        std::cout << "Launch settings: semi_bfs_threshold " << tn.semi_bfs_threshold << ", waves_per_anchor " << tn.waves_per_anchor
                  << ", stack_depth " << tn.stack_depth << ", dfs_blocks " << tn.dfs_blocks << ", local_blocks " << tn.local_blocks
                  << ", team " << tn.team << ", level_blocks " << tn.level_blocks << '\n';
    }
    if (show_mem) {
        uint64_t used = 0, cap = 0;
        csm_pool_usage(s, &used, &cap);
        std::cout << "Device pool: " << (used >> 20) << " MiB used of " << (cap >> 20) << " MiB" << std::endl;
    }

    std::cout << "--------- Incremental Matching --------\n";
    FILE* js = nullptr;
    const std::string json = p.get("--json");
    if (!json.empty()) js = std::fopen(json.c_str(), "w");
    std::vector<unsigned long long> total_pos(nq, 0), total_neg(nq, 0);
    std::vector<csm_query_result> res(nq);
    double kernel_ms = 0;
    t0 = std::chrono::steady_clock::now();
    for (uint64_t b = 0; b < nb; b++) {
        std::cout << std::endl << "Batch #" << b << " --------" << std::endl;
        csm_batch_info info{};
        if ((st = csm_apply_batch(s, b, res.data(), nq, &info)) != CSM_OK) return die(s, st);
        kernel_ms += info.ms_total;
        if (print_index)
            std::cout << "Finish Index Maintenance, time (ms): " << static_cast<unsigned long>(info.ms_index) << "(host), "
                      << static_cast<unsigned long>(info.ms_index) << "(kernel), Matching time (ms): " << static_cast<unsigned long>(info.ms_match)
                      << "(host)\n";
        for (uint32_t q = 0; q < nq; q++) {
            total_pos[q] += res[q].positive;
            total_neg[q] += res[q].negative;
            std::cout << "query " << q << ": +" << res[q].positive << " -" << res[q].negative << std::endl;
            if (js)
                std::fprintf(js,
                             "{\"batch\": %llu, \"query\": %u, \"n_ins\": %llu, \"n_del\": %llu, \"pos\": %llu, \"neg\": %llu, "
                             "\"pos_hash\": \"%016llx\", \"neg_hash\": \"%016llx\", \"t_index\": %.3f, \"t_match\": %.3f, \"t_total\": %.3f}\n",
                             (unsigned long long)b, q, (unsigned long long)info.inserts, (unsigned long long)info.deletes,
                             (unsigned long long)res[q].positive, (unsigned long long)res[q].negative,
                             (unsigned long long)res[q].positive_hash, (unsigned long long)res[q].negative_hash, info.ms_index, info.ms_match,
                             info.ms_total);
        }
    }
    if (js) std::fclose(js);
    time_line("Incremental Matching", ms_since(t0), kernel_ms);
    std::cout << kRule << std::endl;
    unsigned long long all = 0;
    for (uint32_t q = 0; q < nq; q++) all += total_pos[q];
    if (nq > 1)
        for (uint32_t q = 0; q < nq; q++) std::cout << "query " << q << " Num Positive Matches: " << total_pos[q] << ", Num Negative Matches: " << total_neg[q] << '\n';
    std::cout << "Num Positive Matches: " << all << '\n';
    csm_session_destroy(s);
    return 0;
}
