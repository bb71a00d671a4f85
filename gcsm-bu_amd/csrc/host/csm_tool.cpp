// csm_tool: host-side utilities around the matching path.
//   csm_tool plan <query> <summary|-> <matching_order>   query metadata + plan dump + PrintOrders
//   csm_tool load <query> <data> <update> <batch_size>    per-relation edge grouping dump
//   csm_tool meta <query>                                  relation-id lines as SetQueryMeta prints them
//   csm_tool gen  <config> <out_dir> [--seed S] [...]      seeded synthetic workload (see gen.h)
//   csm_tool count <data> <query> [--cap C] [--budget W] [--threads T] [--directed]
//                                                          the generator's exact, budgeted embedding count
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

#include "gen.h"
#include "gen_count.h"
#include "graph_io.h"
#include "plan.h"
#include "query.h"

using namespace csm;

static int cmd_plan(int argc, char** argv) {
    if (argc != 5) return 2;
    Query q;
    std::string err = parse_query(argv[2], q);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::vector<uint32_t> card(2 * q.m, 0u);
    std::vector<float> avg(2 * q.m);
    for (uint32_t r = 0; r < 2 * q.m; r++) avg[r] = card[r] / (float)0u;
    if (std::string(argv[3]) != "-") {
        std::ifstream sf(argv[3]);
        uint32_t r, c, bits;
        while (sf >> r >> c >> std::hex >> bits >> std::dec)
            if (r < 2 * q.m) { card[r] = c; std::memcpy(&avg[r], &bits, 4); }
    }
    Plan p;
    err = make_plan(q, static_cast<MatchingOrder>(std::atoi(argv[4])), card.data(), avg.data(), p);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::cout << dump_query_meta(q) << dump_plan(q, p) << "PrintOrders:\n" << print_orders(q, p);
    return 0;
}

static int cmd_meta(int argc, char** argv) {
    if (argc != 3) return 2;
    Query q;
    std::string err = parse_query(argv[2], q);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::cout << query_meta_lines(q);
    return 0;
}

static int cmd_load(int argc, char** argv) {
    if (argc != 6) return 2;
    Query q;
    std::string err = parse_query(argv[2], q);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    DataGraph g;
    err = load_data_graph(argv[3], g);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    UpdateStream s;
    err = load_update_stream(argv[4], g, s);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    std::cout << dump_query_meta(q) << dump_reference_grouping(q, g, s, (uint32_t)std::stoul(argv[5]));
    return 0;
}

static int cmd_count(int argc, char** argv) {
    if (argc < 4) return 2;
    uint64_t cap = ~0ull, budget = ~0ull;
    unsigned threads = std::max(1u, std::thread::hardware_concurrency());
    bool directed = false;
    for (int i = 4; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--directed") directed = true;
        else if (i + 1 < argc && a == "--cap") cap = std::stoull(argv[++i]);
        else if (i + 1 < argc && a == "--budget") budget = std::stoull(argv[++i]);
        else if (i + 1 < argc && a == "--threads") threads = static_cast<unsigned>(std::stoul(argv[++i]));
        else return 2;
    }
    DataGraph g;
    std::string err = load_data_graph(argv[2], g);
    if (!err.empty()) { std::fprintf(stderr, "%s\n", err.c_str()); return 1; }
    // query file: v <id> <label>, e <a> <b> <label> with raw labels
    std::ifstream qf(argv[3]);
    if (!qf) { std::fprintf(stderr, "cannot read %s\n", argv[3]); return 1; }
    CountQuery q;
    std::string t;
    while (qf >> t) {
        if (t == "t") { std::string a, b; qf >> a >> b; }
        else if (t == "v") { uint32_t id, l; qf >> id >> l; if (id >= q.vlab.size()) q.vlab.resize(id + 1); q.vlab[id] = l; }
        else if (t == "e") { CountEdge e; qf >> e.a >> e.b >> e.l; q.edges.push_back(e); }
    }
    q.n = static_cast<uint32_t>(q.vlab.size());
    std::vector<uint32_t> vlab(g.dv);
    for (uint32_t v = 0; v < g.dv; v++) vlab[v] = g.vlabel[v] == kNoVertex ? 0xFFFFu : g.vlabel[v];
    uint32_t nel = 1;
    for (uint32_t l : g.elabel) nel = std::max(nel, l + 1);
    for (const auto& e : q.edges) nel = std::max(nel, e.l + 1);
    std::vector<CountEdge> es(g.src.size());
    for (size_t k = 0; k < es.size(); k++) es[k] = CountEdge{g.src[k], g.dst[k], g.elabel[k]};
    CountGraph cg;
    cg.build(g.dv, vlab, nel, directed, es, threads);
    const auto r = count_embeddings(cg, q, cap, budget, threads);
    std::printf("within %d count %llu work %llu\n", r.within ? 1 : 0, (unsigned long long)r.count, (unsigned long long)r.work);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: csm_tool plan|load|gen ...\n");
        return 2;
    }
    const std::string c = argv[1];
    int rc = 2;
    if (c == "plan") rc = cmd_plan(argc, argv);
    else if (c == "load") rc = cmd_load(argc, argv);
    else if (c == "meta") rc = cmd_meta(argc, argv);
    else if (c == "gen") rc = gen_main(argc - 1, argv + 1);
    else if (c == "count") rc = cmd_count(argc, argv);
    if (rc == 2) std::fprintf(stderr, "bad arguments\n");
    return rc;
}
