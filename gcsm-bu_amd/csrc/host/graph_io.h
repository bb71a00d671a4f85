// Data graph and update stream files.
//
// File format and parse semantics follow GraphFileIoManager::LoadInitial / LoadUpdate
// (graph/graph.cpp:359-468, 482-563): whitespace tokens, "t" skips two tokens, "v id label",
// "e a b label" (undirected).  Update files hold only edge lines; batch k is edge lines
// [k*B, (k+1)*B) counting every edge line.  A "v" line in an update file is an error, as in the
// reference (graph.cpp:502-506).  Extension: "-e a b label" deletes an edge.
//
// The graph is kept query-independent (one labelled graph shared by every standing query); the
// reference's per-query-relation grouping (graph.cpp:435-459) is derived from it on demand by
// dump_reference_grouping / relation_edge_counts.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "query.h"

namespace csm {

struct DataGraph {
    uint32_t dv = 0;                  // max declared vertex id + 1 (graph.cpp:464 DV_COUNT)
    std::vector<uint32_t> vlabel;     // label value per vertex, kNoVertex if undeclared
    std::vector<uint32_t> src, dst, elabel;  // edges in file order
    bool directed = false;            // extension: arcs src->dst only
};

struct UpdateStream {
    std::vector<uint32_t> src, dst, elabel;
    std::vector<uint8_t> op;  // 0 = insert, 1 = delete
    size_t size() const { return src.size(); }
};

std::string load_data_graph(const std::string& path, DataGraph& g);
std::string load_update_stream(const std::string& path, const DataGraph& g, UpdateStream& s);

// Per-relation counts of the reference's EdgeBatch grouping for one query.
void relation_edge_counts(const Query& q, const DataGraph& g, const uint32_t* src, const uint32_t* dst,
                          const uint32_t* elabel, size_t n, std::vector<uint64_t>& counts);

// Dump identical in layout to the oracle driver's "load" mode: dv, per-relation sizes and FNV-1a
// hashes of the initial grouping and of every batch.
std::string dump_reference_grouping(const Query& q, const DataGraph& g, const UpdateStream& s, uint32_t batch_size);

}  // namespace csm
