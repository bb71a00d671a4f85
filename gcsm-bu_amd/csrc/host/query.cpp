#include "query.h"

#include "../common/errors.h"

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <unordered_map>

namespace csm {

std::pair<uint32_t, uint32_t> Query::rel_endpoints(uint32_t r) const {
    for (uint32_t u = 0; u < n; u++)
        if (r >= qv_offs[u] && r < qv_offs[u + 1]) return {u, qv_nbrs[r]};
    return {kNoVertex, kNoVertex};
}

uint32_t Query::rel_step(uint32_t r) const {
    for (uint32_t i = 0; i < m; i++)
        if (qe_eidx[i].first == r || qe_eidx[i].second == r) return i;
    return kNoVertex;
}

namespace {

// graph.cpp:271-357: CSR, eidx, NLF groups, step order.
std::string set_query_meta(Query& q) {
    if (q.n > kMaxQV || q.m > kMaxQE) {
        return tagged(Code::kLimit, "The query graph should have at most 16 vertices and 42 edges.");
    }
    q.eidx.fill(0xFF);
    uint32_t pos = 0;
    for (uint32_t u = 0; u < q.n; u++) {
        q.qv_offs[u] = static_cast<uint8_t>(pos);
        for (const auto& e : q.nbrs[u]) {
            q.eidx[u * q.n + e.first] = static_cast<uint8_t>(pos);
            q.qv_nbrs[pos] = static_cast<uint8_t>(e.first);
            q.rel_out[pos] = static_cast<uint8_t>((q.out_mask[u] >> e.first) & 1u);
            pos++;
        }
    }
    q.qv_offs[q.n] = static_cast<uint8_t>(pos);

    // A group is the set of relations of u with one (edge label, neighbour label); the first
    // relation carries the group's size as its NLF bound (graph.cpp:295-330).  Groups of more than
    // two relations record first/last relation ids; smaller ones point at themselves.
    q.nlf.fill(0);
    for (uint32_t u = 0; u < q.n; u++) {
        struct Group { uint32_t first, last, count; };
        std::map<std::pair<uint32_t, uint32_t>, Group> groups;
        for (const auto& e : q.nbrs[u]) {
            const auto key = std::make_pair(e.second, q.vlabel[e.first]);
            const uint32_t r = q.eidx[u * q.n + e.first];
            auto it = groups.find(key);
            if (it == groups.end()) groups[key] = Group{r, r, 1};
            else { it->second.last = r; it->second.count++; }
        }
        for (const auto& e : q.nbrs[u]) {
            const Group& g = groups.at(std::make_pair(e.second, q.vlabel[e.first]));
            const uint32_t r = q.eidx[u * q.n + e.first];
            if (g.count > 2) { q.first_nl[r] = static_cast<uint8_t>(g.first); q.last_nl[r] = static_cast<uint8_t>(g.last); }
            else { q.first_nl[r] = static_cast<uint8_t>(r); q.last_nl[r] = static_cast<uint8_t>(r); }
        }
        for (const auto& kv : groups) q.nlf[kv.second.first] = static_cast<uint8_t>(kv.second.count);
    }

    pos = 0;
    for (uint32_t u = 0; u < q.n; u++) {
        for (const auto& e : q.nbrs[u]) {
            const uint32_t w = e.first;
            if (u > w) continue;
            q.qe_list[pos] = {u, w};
            q.qe_labels[pos] = std::make_tuple(q.vlabel[u], e.second, q.vlabel[w]);
            q.qe_rlabels[pos] = std::make_tuple(q.vlabel[w], e.second, q.vlabel[u]);
            q.qe_eidx[pos] = {q.eidx[u * q.n + w], q.eidx[w * q.n + u]};
            pos++;
        }
    }
    return "";
}

bool connected(const Query& q) {
    if (q.n == 0) return false;
    std::vector<char> seen(q.n, 0);
    std::vector<uint32_t> st{0};
    seen[0] = 1;
    uint32_t cnt = 1;
    while (!st.empty()) {
        uint32_t u = st.back();
        st.pop_back();
        for (const auto& e : q.nbrs[u])
            if (!seen[e.first]) { seen[e.first] = 1; cnt++; st.push_back(e.first); }
    }
    return cnt == q.n;
}

std::string parse_stream(std::istream& is, Query& q) {
    q = Query{};
    std::unordered_map<uint32_t, uint32_t> vmap, emap;
    std::vector<char> declared;
    char type;
    // Token-level parse identical to graph.cpp:207-262: 't' skips two tokens, 'v' reads two
    // numbers, any other leading character reads an edge.
    while (is >> type) {
        if (type == 't') {
            char c;
            uint32_t x;
            is >> c >> x;
        } else if (type == 'v') {
            uint32_t id, label;
            if (!(is >> id >> label)) return tagged(Code::kFormat, "malformed vertex line in query");
            auto it = vmap.find(label);
            if (it == vmap.end()) it = vmap.emplace(label, static_cast<uint32_t>(vmap.size())).first;
            if (id >= kMaxQV) return tagged(Code::kLimit, "query vertex id out of range (ids must be 0..15)");
            if (id >= declared.size()) {
                declared.resize(id + 1, 0);
                q.vlabel.resize(id + 1, kNoVertex);
                q.vlabel_raw.resize(id + 1, kNoVertex);
                q.nbrs.resize(id + 1);
            }
            if (declared[id]) return tagged(Code::kFormat, "query vertex declared twice");
            declared[id] = 1;
            q.vlabel[id] = it->second;
            q.vlabel_raw[id] = label;
            q.n += 1;
        } else {
            uint32_t a, b, label;
            if (!(is >> a >> b >> label)) return tagged(Code::kFormat, "malformed edge line in query");
            if (a >= declared.size() || b >= declared.size() || !declared[a] || !declared[b])
                return tagged(Code::kFormat, "query edge references an undeclared vertex");
            if (a == b) return tagged(Code::kFormat, "query self-loops are not supported");
            auto it = emap.find(label);
            if (it == emap.end()) {
                it = emap.emplace(label, static_cast<uint32_t>(emap.size())).first;
                q.elabel_raw.push_back(label);
            }
            const uint32_t l = it->second;
            for (const auto& e : q.nbrs[a])
                if (e.first == b) return tagged(Code::kFormat, "duplicate query edge");
            q.out_mask[a] |= static_cast<uint16_t>(1u << b);
            auto p = std::make_pair(b, l);
            q.nbrs[a].insert(std::lower_bound(q.nbrs[a].begin(), q.nbrs[a].end(), p), p);
            p = std::make_pair(a, l);
            q.nbrs[b].insert(std::lower_bound(q.nbrs[b].begin(), q.nbrs[b].end(), p), p);
            q.m += 1;
        }
    }
    if (q.n < 2) return tagged(Code::kFormat, "the number of query verticex should not be less than 2!");
    if (declared.size() != q.n) return tagged(Code::kFormat, "query vertex ids must be contiguous from 0");
    std::string err = set_query_meta(q);
    if (!err.empty()) return err;
    if (!connected(q)) return tagged(Code::kFormat, "query graph must be connected");
    return "";
}

}  // namespace

std::string parse_query(const std::string& path, Query& q) {
    std::ifstream ifs(path);
    if (!ifs) return tagged(Code::kIo, "Failed to open: " + path);
    std::string err = parse_stream(ifs, q);
    q.path = path;
    return err;
}

std::string parse_query_text(const std::string& text, Query& q) {
    std::istringstream is(text);
    return parse_stream(is, q);
}

std::string query_meta_lines(const Query& q) {
    std::ostringstream os;
    for (uint32_t u = 0; u < q.n; u++) {
        for (const auto& e : q.nbrs[u]) os << "[" << u << "," << e.first << "]: " << (uint32_t)q.eidx[u * q.n + e.first] << ' ';
        os << '\n';
    }
    return os.str();
}

std::string dump_query_meta(const Query& q) {
    std::ostringstream os;
    os << "qv " << q.n << " qe " << q.m << "\n";
    os << "vlabels:";
    for (uint32_t u = 0; u < q.n; u++) os << ' ' << q.vlabel[u];
    os << "\nnbrs:";
    for (uint32_t u = 0; u < q.n; u++) {
        os << " [";
        for (size_t k = 0; k < q.nbrs[u].size(); k++) {
            if (k) os << ' ';
            os << q.nbrs[u][k].first << ':' << q.nbrs[u][k].second;
        }
        os << ']';
    }
    os << "\nqv_offs:";
    for (uint32_t u = 0; u <= q.n; u++) os << ' ' << (uint32_t)q.qv_offs[u];
    os << "\nqv_nbrs:";
    for (uint32_t r = 0; r < 2 * q.m; r++) os << ' ' << (uint32_t)q.qv_nbrs[r];
    os << "\neidx:";
    for (uint32_t u = 0; u < q.n; u++)
        for (uint32_t w = 0; w < q.n; w++) os << ' ' << (uint32_t)q.eidx[u * q.n + w];
    os << "\nnlf:";
    for (uint32_t r = 0; r < 2 * q.m; r++) os << ' ' << (uint32_t)q.nlf[r];
    os << "\nfirst_nl:";
    for (uint32_t r = 0; r < 2 * q.m; r++) os << ' ' << (uint32_t)q.first_nl[r];
    os << "\nlast_nl:";
    for (uint32_t r = 0; r < 2 * q.m; r++) os << ' ' << (uint32_t)q.last_nl[r];
    os << "\nqe_list:";
    for (uint32_t i = 0; i < q.m; i++)
        os << ' ' << q.qe_list[i].first << ',' << q.qe_list[i].second << ',' << q.qe_eidx[i].first << ','
           << q.qe_eidx[i].second << ',' << std::get<0>(q.qe_labels[i]) << '/' << std::get<1>(q.qe_labels[i])
           << '/' << std::get<2>(q.qe_labels[i]);
    os << "\n";
    return os.str();
}

}  // namespace csm
