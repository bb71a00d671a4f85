// Per-query-edge indexing and matching orders, relation-switch flags and Cartesian-product info.
//
// Restates the live planner of the reference (graph/plan.cpp):
//   BuildIndexingOrder            plan.cpp:1548-1602   (P1)
//   GenerateMatchingOrders        plan.cpp:2117-2146   (P3 dispatch)
//     kIndexingOrder              plan.cpp:2120-2132
//     GenerateGammaMatchingOrders plan.cpp:1843-2115
//     GenerateMatchingOrders_v2   plan.cpp:386-782     (GCSM order, the default)
//     GenerateMatchingOrders_v1   plan.cpp:195-384     (RI order)
//   GenerateRebuildVFlagsWithOrder_v2 plan.cpp:1704-1783 (P4, relation switch)
//   GenerateCartesianProductInfo  plan.cpp:1785-1841   (P5)
//   ComputeNumQueryVerticesNonTailLeaf plan.cpp:62-76
//   PrintOrders                   plan.cpp:1200-1291
// OrderPerEdge keeps the layout of plan.h:13-21.
#pragma once
#include <array>
#include <cstdint>
#include <string>

#include "query.h"

namespace csm {

enum class MatchingOrder : int32_t { kIndexing = 0, kGamma = 1, kGcsm = 2, kRi = 3 };
enum class SearchStrategy : int32_t { kDfsWithWorkStealing = 0, kDfsWithoutWorkStealing = 1, kDynamicBfsDfs = 2, kBfs = 3 };

enum CpType : uint8_t { kCpNone = 0, kCpTreeSingle = 1, kCpNonTreeSingle = 2, kCpTreeProduct = 3, kCpNonTreeProduct = 4 };

struct OrderPerEdge {
    uint8_t vs[kMaxQV];
    uint8_t bni_offs[kMaxQV + 1];
    uint8_t bni[kMaxQE];  // positions (in vs) of backward neighbours
};

struct Plan {
    MatchingOrder order_kind = MatchingOrder::kGcsm;
    OrderPerEdge indexing[kMaxQE];
    OrderPerEdge orders[kMaxQE];
    uint16_t rebuild_R[kMaxQE];  // bit u: rebuild the local relation of query vertex u
    uint16_t rebuild_B[kMaxQE];  // bit u: rebuild the local candidate bitmap of u
    uint8_t cp_info[kMaxQE][kMaxQV];
    uint8_t non_tail_leaf[kMaxQE];  // as the reference reports it (all 0 for the RI order)
};

// cardinalities/avg_degrees are indexed by relation id (2*m entries), as produced by the
// global-index summary (match_gpu.cu:589-609).  Returns "" on success.
std::string make_plan(const Query& q, MatchingOrder kind, const uint32_t* cardinalities,
                      const float* avg_degrees, Plan& plan);

// First depth whose Cartesian-product type is TreeProduct or TreeSingle, or n if none: where
// trailing tree leaves start in orders[i].  Valid for every order kind.
uint32_t tail_leaf_depth(const Query& q, const Plan& p, uint32_t i);

std::string print_orders(const Query& q, const Plan& p);      // plan.cpp:1200-1291 text
std::string dump_plan(const Query& q, const Plan& p);          // raw arrays, oracle-driver layout

}  // namespace csm
