#include "plan.h"

#include "../common/errors.h"

#include <algorithm>
#include <cstring>
#include <limits>
#include <sstream>
#include <utility>
#include <vector>

namespace csm {
namespace {

// libstdc++'s std::sort on at most 16 elements is exactly __insertion_sort (the introsort loop only
// runs above 16 elements).  The planner sorts at most 16 items, and the reference compares floats
// that may be NaN, so the same insertion procedure is reproduced here rather than relying on a
// stable sort: with an incomparable NaN the two can differ.
template <class T, class Cmp>
void insertion_sort(T* first, T* last, Cmp comp) {
    if (first == last) return;
    for (T* i = first + 1; i != last; ++i) {
        T val = *i;
        if (comp(val, *first)) {
            std::move_backward(first, i, i + 1);
            *first = val;
        } else {
            T* hole = i;
            T* prev = i - 1;
            while (comp(val, *prev)) {
                *hole = *prev;
                hole = prev;
                --prev;
            }
            *hole = val;
        }
    }
}

bool is_adjacent_sorted(const Query& q, uint32_t u, uint32_t w) {
    // plan.cpp uses lower_bound on (w, 0) over the (neighbour, label)-sorted list.
    auto it = std::lower_bound(q.nbrs[u].begin(), q.nbrs[u].end(), std::make_pair(w, 0u));
    return it != q.nbrs[u].end() && it->first == w;
}

// plan.cpp:1548-1602
void build_indexing_order(const Query& q, uint32_t u0, uint32_t u1, OrderPerEdge& o) {
    std::vector<uint32_t> order{u0, u1};
    std::vector<uint32_t> pre{u0, u1}, cur;
    uint32_t visited = (1u << u0) | (1u << u1);
    while (!pre.empty()) {
        for (uint32_t v : pre)
            for (const auto& e : q.nbrs[v])
                if (!(visited >> e.first & 1u)) {
                    cur.push_back(e.first);
                    visited |= 1u << e.first;
                    order.push_back(e.first);
                }
        pre.clear();
        std::swap(pre, cur);
    }
    std::memset(&o, 0, sizeof(o));
    o.vs[0] = static_cast<uint8_t>(order[0]);
    o.bni_offs[0] = 0;
    o.bni[0] = 1;  // the first vertex lists the second as its "backward" neighbour
    uint32_t cum = 1;
    o.bni_offs[1] = 1;
    for (uint32_t j = 1; j < q.n; j++) {
        const uint32_t u = order[j];
        o.vs[j] = static_cast<uint8_t>(u);
        for (uint32_t k = 0; k < j; k++)
            if (is_adjacent_sorted(q, u, o.vs[k])) o.bni[cum++] = static_cast<uint8_t>(k);
        o.bni_offs[j + 1] = static_cast<uint8_t>(cum);
    }
}

// plan.cpp:1785-1841
void cartesian_product_info(const Query& q, const std::vector<uint32_t>& order, uint8_t* info) {
    const uint32_t n = q.n;
    uint8_t type;
    uint32_t v = order[n - 1];
    if (q.degree(v) == 1) { info[n - 1] = kCpTreeSingle; type = kCpTreeProduct; }
    else { info[n - 1] = kCpNonTreeSingle; type = kCpNonTreeProduct; }
    for (uint32_t i = 1; i + 1 < n; i++) {
        const uint32_t pos = n - 1 - i;
        if (type == kCpNone) { info[pos] = kCpNone; continue; }
        v = order[pos];
        bool no_forward = true;
        for (uint32_t j = 0; j < i; j++)
            if (is_adjacent_sorted(q, v, order[n - 1 - j])) { no_forward = false; break; }
        if (no_forward) {
            if (q.degree(v) == 1 && type == kCpTreeProduct) info[pos] = kCpTreeProduct;
            else { info[pos] = kCpNonTreeProduct; type = kCpNonTreeProduct; }
        } else {
            info[pos] = kCpNone;
            type = kCpNone;
        }
    }
}

// plan.cpp:62-76
void non_tail_leaf_depths(const Query& q, Plan& p) {
    for (uint32_t i = 0; i < q.m; i++) p.non_tail_leaf[i] = static_cast<uint8_t>(tail_leaf_depth(q, p, i));
}

// plan.cpp:1704-1783 (relation switch enabled, which is what GenerateMatchingOrders always passes,
// plan.cpp:2134-2138).
void rebuild_flags_v2(const Query& q, uint32_t index, const std::vector<uint32_t>& order, Plan& p) {
    const OrderPerEdge& io = p.indexing[index];
    auto before = [&](uint32_t a, uint32_t b) {  // does a appear before b in the matching order?
        for (uint32_t j = 0; j < q.n; j++) {
            if (order[j] == b) return false;
            if (order[j] == a) return true;
        }
        return false;
    };
    uint16_t R = p.rebuild_R[index], B = p.rebuild_B[index];
    for (uint32_t i = 2; i < q.n; i++) {
        const uint32_t u = io.vs[i];
        if (io.bni_offs[i + 1] - io.bni_offs[i] == 1) {
            const uint32_t ub = io.vs[io.bni[io.bni_offs[i]]];
            if (before(ub, u)) R &= ~(1u << u); else R |= 1u << u;
        } else {
            R |= 1u << u;
        }
    }
    std::vector<uint32_t> deg(q.n);
    for (uint32_t u = 0; u < q.n; u++) deg[u] = q.degree(u);
    for (uint32_t i = q.n - 1; i >= 2 && i < q.n; i--) {
        const uint32_t u = io.vs[i];
        if (deg[u] == 1 && io.bni_offs[i + 1] - io.bni_offs[i] == 1) {
            const uint32_t ub = io.vs[io.bni[io.bni_offs[i]]];
            if (before(ub, u)) {
                B &= ~(1u << u);
                deg[u]--;
                deg[ub]--;
            } else {
                B |= 1u << u;
            }
        } else {
            B |= 1u << u;
        }
    }
    p.rebuild_R[index] = R;
    p.rebuild_B[index] = B;
}

// plan.cpp:1606-1663: flag every query edge on a simple path from a start vertex to u0 or u1.
void rebuild_edge_flags(const Query& q, uint32_t u0, uint32_t u1, std::initializer_list<uint32_t> starts,
                        std::vector<char>& flags) {
    const uint32_t n = q.n;
    flags[u0 * n + u1] = 1;
    flags[u1 * n + u0] = 1;
    std::vector<uint32_t> path(n, 0xFF);
    for (uint32_t s : starts) {
        uint32_t visited = 1u << s;
        path[0] = s;
        std::vector<uint32_t> ci(n, 0);
        uint32_t depth = 1;
        while (true) {
            while (ci[depth] < q.nbrs[path[depth - 1]].size()) {
                const uint32_t nv = q.nbrs[path[depth - 1]][ci[depth]].first;
                if (visited >> nv & 1u) { ci[depth]++; continue; }
                if (nv == u0 || nv == u1) {
                    path[depth] = nv;
                    for (uint32_t i = 0; i < depth; i++) {
                        flags[path[i] * n + path[i + 1]] = 1;
                        flags[path[i + 1] * n + path[i]] = 1;
                    }
                    path[depth] = 0xFF;
                    ci[depth]++;
                    continue;
                }
                visited |= 1u << nv;
                path[depth] = nv;
                depth++;
            }
            if (ci[depth] >= q.nbrs[path[depth - 1]].size()) {
                ci[depth] = 0;
                depth--;
                if (depth == 0) break;
                visited &= ~(1u << path[depth]);
                ci[depth]++;
            }
        }
    }
}

// plan.cpp:1665-1701: only its consistency check matters (R/B flags are not written by it).
std::string rebuild_vertex_flags_check(const Query& q, uint32_t index, const Plan& p, const std::vector<char>& flags) {
    const OrderPerEdge& io = p.indexing[index];
    for (uint32_t i = 1; i < q.n; i++) {
        const uint32_t u = io.vs[i];
        uint32_t sum = 0;
        for (uint32_t k = io.bni_offs[i]; k < io.bni_offs[i + 1]; k++)
            sum += flags[io.vs[io.bni[k]] * q.n + u] ? 1u : 0u;
        const uint32_t cnt = io.bni_offs[i + 1] - io.bni_offs[i];
        if (sum != 0 && sum != cnt) return tagged(Code::kState, "Error in order generation!");
    }
    return "";
}

// Fills vs/bni/bni_offs of orders[index] from a vertex order; bni of each vertex sorted by the
// average degree of (backward neighbour -> vertex) ascending (plan.cpp:750-778).
void fill_order_meta(const Query& q, uint32_t index, const std::vector<uint32_t>& order, const float* avg, Plan& p) {
    OrderPerEdge& o = p.orders[index];
    std::memset(&o, 0, sizeof(o));
    uint32_t cum = 0;
    o.bni_offs[0] = 0;
    for (uint32_t j = 0; j < q.n; j++) {
        const uint32_t u = order[j];
        o.vs[j] = static_cast<uint8_t>(u);
        for (uint32_t k = j - 1; k < q.n; k--)
            if (is_adjacent_sorted(q, u, o.vs[k])) o.bni[cum++] = static_cast<uint8_t>(k);
        o.bni_offs[j + 1] = static_cast<uint8_t>(cum);
        insertion_sort(o.bni + o.bni_offs[j], o.bni + o.bni_offs[j + 1], [&](uint8_t a, uint8_t b) {
            return avg[q.rel(o.vs[a], u)] < avg[q.rel(o.vs[b], u)];
        });
    }
}

// Greedy extension shared by the GCSM and Gamma orders (plan.cpp:536-744 == 1866-2074).
std::string extend_order(const Query& q, std::vector<uint32_t>& order, uint32_t& visited, const float* avg) {
    const uint32_t n = q.n;
    auto vis = [&](uint32_t v) { return (visited >> v & 1u) != 0; };
    for (uint32_t step = 2; step < n; step++) {
        std::vector<uint32_t> sel, fur;
        {
            // 0. extendable vertices with their minimum backward average degree, stored as uint32
            // (the pair's second member), compared against new floats after conversion to float.
            std::vector<std::pair<uint8_t, uint32_t>> cand;
            for (uint32_t u = 0; u < n; u++) {
                if (vis(u)) continue;
                for (const auto& e : q.nbrs[u]) {
                    const uint32_t bn = e.first;
                    if (!vis(bn)) continue;
                    const float a = avg[q.rel(bn, u)];
                    if (cand.empty() || cand.back().first != u) {
                        cand.emplace_back(static_cast<uint8_t>(u), f2u_x86(a));
                    } else if (a < static_cast<float>(cand.back().second)) {
                        cand.back().second = f2u_x86(a);
                    }
                }
            }
            if (cand.empty()) return tagged(Code::kState, "planner: no extendable query vertex (query not connected)");
            insertion_sort(cand.data(), cand.data() + cand.size(),
                           [](const std::pair<uint8_t, uint32_t>& a, const std::pair<uint8_t, uint32_t>& b) {
                               return a.second < b.second;
                           });
            size_t new_end = cand.size();
            for (size_t j = 0; j + 1 < cand.size(); j++) {
                if (cand[j + 1].second >= cand[j].second * 1024u) {  // (2 << 9), uint32 wrap-around
                    new_end = j + 1;
                    break;
                }
            }
            for (size_t j = 0; j < new_end; j++) sel.push_back(cand[j].first);
            std::sort(sel.begin(), sel.end());
        }
        // 1. most backward neighbours
        if (sel.size() > 1) {
            uint32_t best = 0;
            for (uint32_t u : sel) {
                uint32_t c = 0;
                for (const auto& e : q.nbrs[u]) c += vis(e.first) ? 1u : 0u;
                if (c > best) { best = c; fur.clear(); fur.push_back(u); }
                else if (c == best) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        // 2. most distinct visited vertices adjacent to u's unvisited neighbours
        if (sel.size() > 1) {
            uint32_t best = 0;
            for (uint32_t u : sel) {
                uint32_t c = 0, seen = 0;
                for (const auto& e : q.nbrs[u]) {
                    if (vis(e.first)) continue;
                    for (const auto& f : q.nbrs[e.first])
                        if (vis(f.first) && !(seen >> f.first & 1u)) { seen |= 1u << f.first; c++; }
                }
                if (c > best) { best = c; fur.clear(); fur.push_back(u); }
                else if (c == best) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        // 3. most dangling neighbours (unvisited, with no visited neighbour)
        if (sel.size() > 1) {
            uint32_t best = 0;
            for (uint32_t u : sel) {
                uint32_t c = 0;
                for (const auto& e : q.nbrs[u]) {
                    if (vis(e.first)) continue;
                    bool dangling = true;
                    for (const auto& f : q.nbrs[e.first])
                        if (vis(f.first)) { dangling = false; break; }
                    if (dangling) c++;
                }
                if (c > best) { best = c; fur.clear(); fur.push_back(u); }
                else if (c == best) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        // 4. smallest minimum backward average degree (float; std::min keeps the left operand
        // unless the right one compares smaller, so NaN never wins)
        if (sel.size() > 1) {
            float best = std::numeric_limits<float>::max();
            for (uint32_t u : sel) {
                float c = std::numeric_limits<float>::max();
                for (const auto& e : q.nbrs[u])
                    if (vis(e.first)) c = std::min(c, avg[q.rel(e.first, u)]);
                if (c < best) { best = c; fur.clear(); fur.push_back(u); }
                else if (c == best) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        if (sel.empty()) return tagged(Code::kState, "planner: empty candidate set (reference behaviour undefined)");
        order.push_back(sel[0]);
        visited |= 1u << sel[0];
    }
    return "";
}

// plan.cpp:399-534: starting edge of the GCSM order.
std::string gcsm_start_edge(const Query& q, uint32_t index, const uint32_t* card, const float* avg,
                            std::pair<uint32_t, uint32_t>& out) {
    const auto [u0, u1] = q.qe_list[index];
    std::vector<std::pair<uint32_t, uint32_t>> sel, fur;
    uint32_t best = 0;
    for (uint32_t i = 0; i < q.m; i++) {
        const auto e = q.qe_list[i];
        const uint32_t s = q.degree(e.first) + q.degree(e.second);
        if (s > best) { best = s; sel.clear(); sel.push_back(e); }
        else if (s == best) sel.push_back(e);
    }
    if (sel.size() > 1) {
        uint32_t bt = 0;
        for (const auto& e : sel) {
            uint32_t t = 0;
            for (uint32_t w = 0; w < q.n; w++) {
                bool f0 = false, f1 = false;
                for (const auto& x : q.nbrs[w]) {
                    if (x.first == e.first) f0 = true;
                    else if (x.first == e.second) f1 = true;
                }
                if (f0 && f1) t++;
            }
            if (t > bt) { bt = t; fur.clear(); fur.push_back(e); }
            else if (t == bt) fur.push_back(e);
        }
        std::swap(sel, fur);
        fur.clear();
    }
    if (sel.size() > 1) {
        for (const auto& e : sel)
            if ((e.first == u0 && e.second == u1) || (e.first == u1 && e.second == u0)) {
                fur.emplace_back(u0, u1);
                std::swap(sel, fur);
                fur.clear();
                break;
            }
    }
    // The reference's last tie-break can empty the selection when every candidate's neighbour
    // degree sum is NaN; it then reads element 0 of the cleared vector, which libstdc++ leaves
    // holding the selection from before the cardinality tie-break.  That element is kept here.
    const std::pair<uint32_t, uint32_t> stale_first = sel.empty() ? std::make_pair(kNoVertex, kNoVertex) : sel[0];
    if (sel.size() > 1) {
        uint32_t bc = UINT32_MAX;
        for (const auto& e : sel) {
            const uint32_t c = card[q.rel(e.first, e.second)];
            if (c < bc) { bc = c; fur.clear(); fur.push_back(e); }
            else if (c == bc) fur.push_back(e);
        }
        std::swap(sel, fur);
        fur.clear();
    }
    if (sel.size() > 1) {
        float bs = std::numeric_limits<float>::max();
        for (const auto& e : sel) {
            float s = 0.0f;
            for (const auto& x : q.nbrs[e.first]) {
                if (x.first == e.second) continue;
                s += avg[q.rel(e.first, x.first)];
            }
            for (const auto& x : q.nbrs[e.second]) {
                if (x.first == e.first) continue;
                s += avg[q.rel(e.second, x.first)];
            }
            if (s < bs) { bs = s; fur.clear(); fur.push_back(e); }
            else if (s == bs) fur.push_back(e);
        }
        std::swap(sel, fur);
        fur.clear();
    }
    if (sel.empty()) {
        if (stale_first.first == kNoVertex) return tagged(Code::kState, "planner: empty starting-edge set");
        out = stale_first;
        return "";
    }
    out = sel[0];
    return "";
}

// plan.cpp:195-344: RI order (maximum-degree start, then the backward-neighbour heuristics).
std::vector<uint32_t> ri_order(const Query& q) {
    std::vector<uint32_t> order;
    uint32_t visited = 0;
    uint32_t best = 0, start = 0;
    bool any = false;
    for (uint32_t v = 0; v < q.n; v++) {
        const uint32_t d = q.degree(v);
        if (d > best) { best = d; start = v; any = true; }
        else if (d == best && !any) { start = v; any = true; }
    }
    visited |= 1u << start;
    order.push_back(start);
    auto vis = [&](uint32_t v) { return (visited >> v & 1u) != 0; };
    for (uint32_t step = 1; step < q.n; step++) {
        std::vector<uint32_t> sel, fur;
        {
            uint32_t bb = 0;
            for (uint32_t u = 0; u < q.n; u++) {
                if (vis(u)) continue;
                uint32_t c = 0;
                for (const auto& e : q.nbrs[u]) c += vis(e.first) ? 1u : 0u;
                if (c > bb) { bb = c; fur.clear(); fur.push_back(u); }
                else if (c == bb) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        if (sel.size() > 1) {
            uint32_t bb = 0;
            for (uint32_t u : sel) {
                uint32_t c = 0, seen = 0;
                for (const auto& e : q.nbrs[u]) {
                    if (vis(e.first)) continue;
                    for (const auto& f : q.nbrs[e.first])
                        if (vis(f.first) && !(seen >> f.first & 1u)) { seen |= 1u << f.first; c++; }
                }
                if (c > bb) { bb = c; fur.clear(); fur.push_back(u); }
                else if (c == bb) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        if (sel.size() > 1) {
            uint32_t bb = 0;
            for (uint32_t u : sel) {
                uint32_t c = 0;
                for (const auto& e : q.nbrs[u]) {
                    if (vis(e.first)) continue;
                    bool dangling = true;
                    for (const auto& f : q.nbrs[e.first])
                        if (vis(f.first)) { dangling = false; break; }
                    if (dangling) c++;
                }
                if (c > bb) { bb = c; fur.clear(); fur.push_back(u); }
                else if (c == bb) fur.push_back(u);
            }
            std::swap(fur, sel);
            fur.clear();
        }
        order.push_back(sel[0]);
        visited |= 1u << sel[0];
    }
    return order;
}

}  // namespace

uint32_t tail_leaf_depth(const Query& q, const Plan& p, uint32_t i) {
    uint32_t d = 0;
    while (d < q.n && p.cp_info[i][d] != kCpTreeProduct && p.cp_info[i][d] != kCpTreeSingle) d++;
    return d;
}

std::string make_plan(const Query& q, MatchingOrder kind, const uint32_t* card, const float* avg, Plan& p) {
    p = Plan{};
    p.order_kind = kind;
    for (uint32_t i = 0; i < q.m; i++) build_indexing_order(q, q.qe_list[i].first, q.qe_list[i].second, p.indexing[i]);

    switch (kind) {
        case MatchingOrder::kIndexing: {
            for (uint32_t i = 0; i < q.m; i++) {
                p.orders[i] = p.indexing[i];
                std::vector<uint32_t> order(p.orders[i].vs, p.orders[i].vs + q.n);
                cartesian_product_info(q, order, p.cp_info[i]);
            }
            non_tail_leaf_depths(q, p);
            break;
        }
        case MatchingOrder::kGamma:
        case MatchingOrder::kGcsm: {
            for (uint32_t i = 0; i < q.m; i++) {
                const auto [u0, u1] = q.qe_list[i];
                std::vector<uint32_t> order;
                uint32_t visited = 0;
                if (kind == MatchingOrder::kGamma) {
                    order = {u0, u1};
                } else {
                    std::pair<uint32_t, uint32_t> s;
                    std::string err = gcsm_start_edge(q, i, card, avg, s);
                    if (!err.empty()) return err;
                    bool reverse = false;
                    for (uint32_t j = 0; j < q.n; j++) {
                        if (p.indexing[i].vs[j] == s.first) break;
                        if (p.indexing[i].vs[j] == s.second) { reverse = true; break; }
                    }
                    order = reverse ? std::vector<uint32_t>{s.second, s.first} : std::vector<uint32_t>{s.first, s.second};
                }
                visited = (1u << order[0]) | (1u << order[1]);
                std::string err = extend_order(q, order, visited, avg);
                if (!err.empty()) return err;
                rebuild_flags_v2(q, i, order, p);
                cartesian_product_info(q, order, p.cp_info[i]);
                fill_order_meta(q, i, order, avg, p);
            }
            non_tail_leaf_depths(q, p);
            break;
        }
        case MatchingOrder::kRi: {
            for (uint32_t i = 0; i < q.m; i++) {
                const auto [u0, u1] = q.qe_list[i];
                std::vector<uint32_t> order = ri_order(q);
                std::vector<char> flags(q.n * q.n, 0);
                rebuild_edge_flags(q, u0, u1, {order[0], order[1]}, flags);
                std::string err = rebuild_vertex_flags_check(q, i, p, flags);
                if (!err.empty()) return err;
                cartesian_product_info(q, order, p.cp_info[i]);
                fill_order_meta(q, i, order, avg, p);
            }
            // The RI path never computes non-tail-leaf depths (plan.cpp:195-384).
            break;
        }
        default:
            return tagged(Code::kArgument, "unknown matching order");
    }
    return "";
}

std::string print_orders(const Query& q, const Plan& p) {
    std::ostringstream os;
    os << "Orders:\n";
    for (uint32_t i = 0; i < q.m; i++) {
        const auto [u0, u1] = q.qe_list[i];
        const OrderPerEdge& io = p.indexing[i];
        const OrderPerEdge& o = p.orders[i];
        os << '[' << u0 << ',' << u1 << "]:\n    indexing order:";
        for (uint32_t j = 0; j < q.n; j++) {
            const uint32_t v = io.vs[j];
            os << v;
            os << ((p.rebuild_R[i] >> v & 1u) ? "-buildR" : "");
            os << ((p.rebuild_B[i] >> v & 1u) ? "-buildB-(" : "-(");
            for (uint32_t k = io.bni_offs[j]; k < io.bni_offs[j + 1]; k++) {
                os << (uint32_t)io.vs[io.bni[k]] << '@' << (uint32_t)io.bni[k];
                if (k != static_cast<uint32_t>(io.bni_offs[j + 1] - 1)) os << ' ';
            }
            os << ") ";
        }
        os << "\n    order: ";
        for (uint32_t j = 0; j < q.n; j++) {
            os << (uint32_t)o.vs[j] << "-(";
            for (uint32_t k = o.bni_offs[j]; k < o.bni_offs[j + 1]; k++) {
                os << (uint32_t)o.vs[o.bni[k]] << '@' << (uint32_t)o.bni[k];
                if (k != static_cast<uint32_t>(o.bni_offs[j + 1] - 1)) os << ' ';
            }
            os << ") ";
        }
        os << "\n    cartesian product info: ";
        for (uint32_t j = 0; j < q.n; j++) {
            os << (uint32_t)o.vs[j];
            if (j == 0) os << ' ';
            else os << "-(" << (uint32_t)p.cp_info[i][j] << ") ";
        }
        os << "\n";
    }
    return os.str();
}

std::string dump_plan(const Query& q, const Plan& p) {
    std::ostringstream os;
    const uint32_t n = q.n;
    for (uint32_t i = 0; i < q.m; i++) {
        const OrderPerEdge& io = p.indexing[i];
        const OrderPerEdge& o = p.orders[i];
        os << "edge " << i << "\n  indexing vs:";
        for (uint32_t j = 0; j < n; j++) os << ' ' << (uint32_t)io.vs[j];
        os << "\n  indexing bni_offs:";
        for (uint32_t j = 0; j <= n; j++) os << ' ' << (uint32_t)io.bni_offs[j];
        os << "\n  indexing bni:";
        for (uint32_t k = 0; k < io.bni_offs[n]; k++) os << ' ' << (uint32_t)io.bni[k];
        os << "\n  order vs:";
        for (uint32_t j = 0; j < n; j++) os << ' ' << (uint32_t)o.vs[j];
        os << "\n  order bni_offs:";
        for (uint32_t j = 0; j <= n; j++) os << ' ' << (uint32_t)o.bni_offs[j];
        os << "\n  order bni:";
        for (uint32_t k = 0; k < o.bni_offs[n]; k++) os << ' ' << (uint32_t)o.bni[k];
        os << "\n  rebuild_R:";
        for (uint32_t u = 0; u < n; u++) os << ((p.rebuild_R[i] >> u & 1u) ? '1' : '0');
        os << "\n  rebuild_B:";
        for (uint32_t u = 0; u < n; u++) os << ((p.rebuild_B[i] >> u & 1u) ? '1' : '0');
        os << "\n  cp:";
        for (uint32_t j = 0; j < n; j++) os << ' ' << (uint32_t)p.cp_info[i][j];
        os << "\n";
    }
    os << "non_tail_leaf:";
    for (uint32_t i = 0; i < q.m; i++) os << ' ' << (p.order_kind == MatchingOrder::kRi ? 0u : (uint32_t)p.non_tail_leaf[i]);
    os << "\n";
    return os.str();
}

}  // namespace csm
