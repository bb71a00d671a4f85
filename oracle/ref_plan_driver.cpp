// Oracle-side driver around the reference's own host code (graph/graph.cpp, graph/plan.cpp).
//
// TEST INFRASTRUCTURE ONLY.  This file is linked against the reference's host sources where they
// lie (see oracle/Makefile.ref); the resulting binary oracle/_ref/ref_plan is used to write golden
// fixtures for the query parser (graph.cpp:191-357), the per-relation edge grouping of the data and
// update files (graph.cpp:359-563) and the planner (plan.cpp:62-76, 185-193, 195-798, 1548-1841,
// 1843-2146, PrintOrders 1200-1291).  Nothing in the product links or runs it.
//
// Usage:
//   ref_plan plan  <query> <summary|-> <matching_order 0..3>
//   ref_plan load  <query> <data> <update> <batch_size>
//
// The summary file holds one line per relation "r card avg_bits_hex" (the float bits of
// avg_degrees[r]); "-" means all-zero cardinalities with NaN averages, which is what the
// reference computes for an empty global index (match_gpu.cu:607).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "graph/graph.h"
#include "graph/plan.h"
#include "utils/constants.h"
#include "utils/technique_option.h"

// QueryGraph / DataGraphManager / GraphFileIoManager declare MatchCPU a friend; the driver uses that
// name to read their private members for the dump.
class MatchCPU {
  public:
    static void DumpQuery(const QueryGraph& q, std::ostream& os) {
        os << "qv " << q.vcount_ << " qe " << q.ecount_ << "\n";
        os << "vlabels:";
        for (uint32_t u = 0; u < q.vcount_; u++) os << ' ' << q.vlabels_[u];
        os << "\nnbrs:";
        for (uint32_t u = 0; u < q.vcount_; u++) {
            os << " [";
            for (size_t k = 0; k < q.nbrs_[u].size(); k++) {
                if (k) os << ' ';
                os << q.nbrs_[u][k].first << ':' << q.nbrs_[u][k].second;
            }
            os << ']';
        }
        os << "\nqv_offs:";
        for (uint32_t u = 0; u <= q.vcount_; u++) os << ' ' << (uint32_t)q.qv_offs_[u];
        os << "\nqv_nbrs:";
        for (uint32_t r = 0; r < 2 * q.ecount_; r++) os << ' ' << (uint32_t)q.qv_nbrs_[r];
        os << "\neidx:";
        for (uint32_t u = 0; u < q.vcount_; u++)
            for (uint32_t w = 0; w < q.vcount_; w++) os << ' ' << (uint32_t)q.eidx_[u * q.vcount_ + w];
        os << "\nnlf:";
        for (uint32_t r = 0; r < 2 * q.ecount_; r++) os << ' ' << (uint32_t)q.NLF_[r];
        os << "\nfirst_nl:";
        for (uint32_t r = 0; r < 2 * q.ecount_; r++) os << ' ' << (uint32_t)q.first_NL_[r];
        os << "\nlast_nl:";
        for (uint32_t r = 0; r < 2 * q.ecount_; r++) os << ' ' << (uint32_t)q.last_NL_[r];
        os << "\nqe_list:";
        for (uint32_t i = 0; i < q.ecount_; i++)
            os << ' ' << q.qe_list_[i].first << ',' << q.qe_list_[i].second << ','
               << q.qe_eidx_[i].first << ',' << q.qe_eidx_[i].second << ','
               << std::get<0>(q.qe_labels_[i]) << '/' << std::get<1>(q.qe_labels_[i]) << '/'
               << std::get<2>(q.qe_labels_[i]);
        os << "\n";
    }
    static uint32_t vcount(const QueryGraph& q) { return q.vcount_; }
    static uint32_t ecount(const QueryGraph& q) { return q.ecount_; }
    static uint32_t dv_count(const DataGraphManager& d) { return d.vcount_; }
};

namespace {

// FNV-1a over the pushed (src, dst) pairs, in push order.
uint64_t HashList(const EdgeList& l) {
    uint64_t h = 1469598103934665603ull;
    for (size_t k = 0; k < l.first.size(); k++) {
        uint32_t pair[2] = {l.first[k], l.second[k]};
        const unsigned char* p = reinterpret_cast<const unsigned char*>(pair);
        for (int b = 0; b < 8; b++) { h ^= p[b]; h *= 1099511628211ull; }
    }
    return h;
}

void DumpBatch(const EdgeBatch& b, uint32_t nrel, std::ostream& os) {
    for (uint32_t r = 0; r < nrel; r++) {
        os << ' ' << r << ':' << b[r].first.size() << ':' << std::hex << HashList(b[r]) << std::dec;
    }
    os << "\n";
}

void DumpPlan(const PlanManager& pm, uint32_t n, uint32_t m, std::ostream& os) {
    for (uint32_t i = 0; i < m; i++) {
        os << "edge " << i << "\n  indexing vs:";
        for (uint32_t j = 0; j < n; j++) os << ' ' << (uint32_t)pm.indexing_orders_[i].vs_[j];
        os << "\n  indexing bni_offs:";
        for (uint32_t j = 0; j <= n; j++) os << ' ' << (uint32_t)pm.indexing_orders_[i].bni_offs_[j];
        os << "\n  indexing bni:";
        for (uint32_t k = 0; k < pm.indexing_orders_[i].bni_offs_[n]; k++) os << ' ' << (uint32_t)pm.indexing_orders_[i].bni_[k];
        os << "\n  order vs:";
        for (uint32_t j = 0; j < n; j++) os << ' ' << (uint32_t)pm.orders_[i].vs_[j];
        os << "\n  order bni_offs:";
        for (uint32_t j = 0; j <= n; j++) os << ' ' << (uint32_t)pm.orders_[i].bni_offs_[j];
        os << "\n  order bni:";
        for (uint32_t k = 0; k < pm.orders_[i].bni_offs_[n]; k++) os << ' ' << (uint32_t)pm.orders_[i].bni_[k];
        os << "\n  rebuild_R:";
        for (uint32_t u = 0; u < n; u++) os << (pm.rebuild_R_flags_[i][u] ? '1' : '0');
        os << "\n  rebuild_B:";
        for (uint32_t u = 0; u < n; u++) os << (pm.rebuild_B_flags_[i][u] ? '1' : '0');
        os << "\n  cp:";
        for (uint32_t j = 0; j < n; j++) os << ' ' << (uint32_t)pm.cartesian_product_info_[i][j];
        os << "\n";
    }
    std::vector<uint8_t> d = pm.getNonTailLeafDepths();
    os << "non_tail_leaf:";
    for (uint32_t i = 0; i < m; i++) os << ' ' << (i < d.size() ? (uint32_t)d[i] : 0u);
    os << "\n";
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: ref_plan plan|load ...\n");
        return 2;
    }
    const std::string mode = argv[1];
    std::ostringstream out;
    std::streambuf* saved = std::cout.rdbuf();
    std::ostringstream sink;
    std::cout.rdbuf(sink.rdbuf());  // the reference prints progress; keep it out of the dump

    QueryGraph qg;
    if (mode == "plan" && argc == 5) {
        GraphFileIoManager io(argv[2], qg);
        io.SetQueryMeta();
        const uint32_t n = MatchCPU::vcount(qg), m = MatchCPU::ecount(qg);
        std::vector<uint32_t> card(2 * m, 0u);
        std::vector<float> avg(2 * m, 0.0f);
        for (uint32_t r = 0; r < 2 * m; r++) avg[r] = card[r] / (float)0u;
        if (std::string(argv[3]) != "-") {
            std::ifstream sf(argv[3]);
            uint32_t r, c, bits;
            while (sf >> r >> c >> std::hex >> bits >> std::dec) {
                if (r < 2 * m) { card[r] = c; std::memcpy(&avg[r], &bits, 4); }
            }
        }
        const int order_id = std::atoi(argv[4]);
        TechniqueOption opt(static_cast<MatchingOrder>(order_id), SearchStrategy::kDynamicBfsDfs,
                            LocalIndexOption::kUseLocalIndex, true, true, true);
        PlanManager pm(qg);
        pm.GenerateIndexingOrders();
        pm.GenerateMatchingOrders(opt, card.data(), avg.data());
        MatchCPU::DumpQuery(qg, out);
        DumpPlan(pm, n, m, out);
        std::ostringstream po;
        std::cout.rdbuf(po.rdbuf());
        pm.PrintOrders();
        std::cout.rdbuf(sink.rdbuf());
        out << "PrintOrders:\n" << po.str();
    } else if (mode == "load" && argc == 6) {
        GraphFileIoManager io(argv[2], qg);
        io.SetQueryMeta();
        DataGraphManager dg;
        io.LoadInitial(argv[3], dg);
        io.LoadUpdate(argv[4], dg, (uint32_t)std::stoul(argv[5]));
        const uint32_t m = MatchCPU::ecount(qg);
        MatchCPU::DumpQuery(qg, out);
        out << "dv " << MatchCPU::dv_count(dg) << "\n";
        out << "initial:";
        DumpBatch(dg.initial_edges_, 2 * m, out);
        out << "batches " << dg.updated_edges_.size() << "\n";
        for (size_t b = 0; b < dg.updated_edges_.size(); b++) {
            out << "batch " << b << ":";
            DumpBatch(dg.updated_edges_[b], 2 * m, out);
        }
    } else {
        std::cout.rdbuf(saved);
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    std::cout.rdbuf(saved);
    std::cout << out.str();
    return 0;
}
