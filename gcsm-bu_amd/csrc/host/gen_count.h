// Exact, budgeted embedding counts for the workload generator's query selection.
//
// The generator keeps a candidate query only if its number of embeddings in the union of every
// graph the stream passes through (initial graph plus all insertions) is at most a cap: every
// batch's positive and negative counts are embeddings of some intermediate graph, all of which are
// subgraphs of that union, so the cap bounds every batch's work.  The count is exact (a
// depth-first search with the trailing tree leaves counted by inclusion-exclusion over their
// candidate sets instead of enumerated), and it stops as soon as the count passes the cap or the
// search passes a work budget; either outcome is independent of the number of threads.
#pragma once

#include <cstdint>
#include <vector>

namespace csm {

struct CountEdge { uint32_t a, b, l; };

// Labelled adjacency: per vertex, neighbours sorted by (neighbour label, edge label, id), so the
// neighbours along one (vertex label, edge label) pair form one sorted run.
struct CountGraph {
    uint32_t n = 0, nel = 1;
    bool directed = false;
    std::vector<uint32_t> vlab;
    std::vector<uint64_t> off[2];  // [0]: out-arcs (all arcs when undirected), [1]: in-arcs (directed only)
    std::vector<uint64_t> key[2];  // (slot << 32) | neighbour, slot = vlab[neighbour] * nel + edge label
    std::vector<std::vector<uint32_t>> by_label;  // vertices of each label
    void build(uint32_t n, const std::vector<uint32_t>& vlab, uint32_t nel, bool directed, const std::vector<CountEdge>& edges,
               unsigned threads);
};

struct CountQuery {
    uint32_t n = 0;
    std::vector<uint32_t> vlab;    // raw data labels of the query vertices
    std::vector<CountEdge> edges;  // query vertex ids; directed: arc a -> b
};

struct CountResult {
    bool within = false;  // count <= cap and work <= budget
    uint64_t count = 0;   // exact when within
    uint64_t work = 0;    // list entries scanned and searches made (a lower bound when not within)
};

CountResult count_embeddings(const CountGraph& g, const CountQuery& q, uint64_t cap, uint64_t work_budget, unsigned threads);

}  // namespace csm
