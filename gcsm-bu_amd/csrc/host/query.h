// Query graph: parsing and metadata.
//
// Mirrors the reference's GraphFileIoManager(query_path) constructor (graph/graph.cpp:191-269) and
// SetQueryMeta (graph/graph.cpp:271-357); the data layout follows QueryGraph (graph/graph.h:29-186).
// Differences are limited to error handling: where the reference calls exit(), these functions
// return an error string.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../common/csm_common.h"

namespace csm {

struct Query {
    uint32_t n = 0;  // vertices  (QV_COUNT)
    uint32_t m = 0;  // undirected edges (QE_COUNT)
    std::string path;

    // Labels remapped to first-seen dense ids, as graph.cpp:221-225 / 245-249 do.
    std::vector<uint32_t> vlabel;        // remapped vertex label of u
    std::vector<uint32_t> vlabel_raw;    // label value written in the file
    std::vector<uint32_t> elabel_raw;    // remapped edge label -> file value
    // nbrs[u] = (neighbour, remapped edge label), sorted ascending (graph.cpp:251-258).
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> nbrs;

    std::array<uint8_t, kMaxQV + 1> qv_offs{};   // CSR of the query (graph.cpp:283-293)
    std::array<uint8_t, kMaxRel> qv_nbrs{};
    std::array<uint8_t, kMaxQV * kMaxQV> eidx{}; // eidx[u*n+w] = relation id (UINT8_MAX if none)
    std::array<uint8_t, kMaxRel> nlf{};          // NLF count at the group's first relation
    std::array<uint8_t, kMaxRel> first_nl{};
    std::array<uint8_t, kMaxRel> last_nl{};
    // rel_out[r] = 1 when the file lists relation r's edge as source -> destination (u -> w);
    // only directed data graphs look at it.
    std::array<uint8_t, kMaxRel> rel_out{};
    std::array<uint16_t, kMaxQV> out_mask{};   // out_mask[a] bit b: the file has "e a b".
    // Step order of the undirected query edges: u ascending, then neighbour ascending, u <= w.
    std::array<std::pair<uint32_t, uint32_t>, kMaxQE> qe_list{};
    std::array<std::pair<uint32_t, uint32_t>, kMaxQE> qe_eidx{};  // (eidx[u][w], eidx[w][u])
    std::array<std::tuple<uint32_t, uint32_t, uint32_t>, kMaxQE> qe_labels{};
    std::array<std::tuple<uint32_t, uint32_t, uint32_t>, kMaxQE> qe_rlabels{};

    uint32_t rel(uint32_t u, uint32_t w) const { return eidx[u * n + w]; }
    bool adjacent(uint32_t u, uint32_t w) const {
        for (const auto& p : nbrs[u]) if (p.first == w) return true;
        return false;
    }
    uint32_t degree(uint32_t u) const { return static_cast<uint32_t>(nbrs[u].size()); }
    // Relation id -> (source query vertex, destination query vertex).
    std::pair<uint32_t, uint32_t> rel_endpoints(uint32_t r) const;
    // Step index (position in qe_list) of the undirected edge that relation r belongs to.
    uint32_t rel_step(uint32_t r) const;
};

// Parse a query file ("t", "v id label", "e a b label").  Returns "" on success.
std::string parse_query(const std::string& path, Query& q);
std::string parse_query_text(const std::string& text, Query& q);

// Writes the metadata block in the same layout as the oracle's reference driver dump.
std::string dump_query_meta(const Query& q);
// The relation-id lines SetQueryMeta prints ("[u,w]: r " per neighbour, one line per vertex;
// graph.cpp:349-356).
std::string query_meta_lines(const Query& q);

}  // namespace csm
