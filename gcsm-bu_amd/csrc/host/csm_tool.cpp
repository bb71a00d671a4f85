// csm_tool: host-side utilities around the matching path.
//   csm_tool plan <query> <summary|-> <matching_order>   query metadata + plan dump + PrintOrders
//   csm_tool load <query> <data> <update> <batch_size>    per-relation edge grouping dump
//   csm_tool meta <query>                                  relation-id lines as SetQueryMeta prints them
//   csm_tool gen  <config> <out_dir> [--seed S] [...]      seeded synthetic workload (see gen.h)
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "gen.h"
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
    if (rc == 2) std::fprintf(stderr, "bad arguments\n");
    return rc;
}
