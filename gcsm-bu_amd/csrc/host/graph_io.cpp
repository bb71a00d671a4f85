#include "graph_io.h"

#include "../common/errors.h"

#include <cstdio>
#include <sstream>
#include <unordered_map>

namespace csm {
namespace {

// Reads the whole file; the tokenizer below keeps operator>> semantics for well-formed files
// (a one-character type token, then unsigned decimal numbers separated by whitespace).
bool read_file(const std::string& path, std::string& buf) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    buf.resize(sz > 0 ? static_cast<size_t>(sz) : 0);
    size_t got = sz > 0 ? std::fread(&buf[0], 1, buf.size(), f) : 0;
    std::fclose(f);
    buf.resize(got);
    return true;
}

struct Tok {
    const char* p;
    const char* e;
    void ws() { while (p < e && (*p == ' ' || *p == '\n' || *p == '\t' || *p == '\r' || *p == '\f' || *p == '\v')) p++; }
    bool chr(char& c) { ws(); if (p >= e) return false; c = *p++; return true; }
    bool u32(uint32_t& v) {
        ws();
        if (p >= e || *p < '0' || *p > '9') return false;
        uint64_t x = 0;
        while (p < e && *p >= '0' && *p <= '9') { x = x * 10 + static_cast<uint64_t>(*p - '0'); p++; if (x > 0xFFFFFFFFull) return false; }
        v = static_cast<uint32_t>(x);
        return true;
    }
};

}  // namespace

std::string load_data_graph(const std::string& path, DataGraph& g) {
    std::string buf;
    if (!read_file(path, buf)) return tagged(Code::kIo, "Failed to open: " + path);
    g = DataGraph{};
    Tok t{buf.data(), buf.data() + buf.size()};
    char type;
    while (t.chr(type)) {
        if (type == 't') {
            char c; uint32_t x;
            t.chr(c); t.u32(x);
        } else if (type == 'v') {
            uint32_t id, label;
            if (!t.u32(id) || !t.u32(label)) return tagged(Code::kFormat, "malformed vertex line in data graph");
            if (id == kNoVertex) return tagged(Code::kLimit, "vertex id out of range");
            if (id >= g.vlabel.size()) g.vlabel.resize(static_cast<size_t>(id) + 1, kNoVertex);
            if (g.vlabel[id] != kNoVertex) return tagged(Code::kFormat, "data vertex declared twice");
            g.vlabel[id] = label;
        } else {
            uint32_t a, b, label;
            if (!t.u32(a) || !t.u32(b) || !t.u32(label)) return tagged(Code::kFormat, "malformed edge line in data graph");
            if (a >= g.vlabel.size() || b >= g.vlabel.size() || g.vlabel[a] == kNoVertex || g.vlabel[b] == kNoVertex)
                return tagged(Code::kFormat, "data edge references an undeclared vertex");
            if (a == b) return tagged(Code::kFormat, "data graph self-loops are not supported");
            g.src.push_back(a);
            g.dst.push_back(b);
            g.elabel.push_back(label);
        }
    }
    g.dv = static_cast<uint32_t>(g.vlabel.size());
    return "";
}

std::string load_update_stream(const std::string& path, const DataGraph& g, UpdateStream& s) {
    std::string buf;
    if (!read_file(path, buf)) return tagged(Code::kIo, "Failed to open: " + path);
    s = UpdateStream{};
    Tok t{buf.data(), buf.data() + buf.size()};
    char type;
    while (t.chr(type)) {
        if (type == 't') {
            char c; uint32_t x;
            t.chr(c); t.u32(x);
        } else if (type == 'v') {
            return tagged(Code::kFormat, "vertex update.");
        } else {
            uint8_t op = 0;
            if (type == '-') {
                char e;
                if (!t.chr(e) || e != 'e') return tagged(Code::kFormat, "malformed deletion line in update stream");
                op = 1;
            }
            uint32_t a, b, label;
            if (!t.u32(a) || !t.u32(b) || !t.u32(label)) return tagged(Code::kFormat, "malformed edge line in update stream");
            if (a >= g.dv || b >= g.dv || g.vlabel[a] == kNoVertex || g.vlabel[b] == kNoVertex)
                return tagged(Code::kFormat, "update edge references an undeclared vertex");
            if (a == b) return tagged(Code::kFormat, "self-loops are not supported");
            s.src.push_back(a);
            s.dst.push_back(b);
            s.elabel.push_back(label);
            s.op.push_back(op);
        }
    }
    return "";
}

namespace {

struct Grouper {
    const Query& q;
    std::unordered_map<uint32_t, uint32_t> vmap, emap;
    explicit Grouper(const Query& qq) : q(qq) {
        for (uint32_t u = 0; u < q.n; u++) vmap.emplace(q.vlabel_raw[u], q.vlabel[u]);
        for (uint32_t l = 0; l < q.elabel_raw.size(); l++) emap.emplace(q.elabel_raw[l], l);
    }
    // graph.cpp:421-459: calls push(relation, from, to) in the reference's push order.
    template <class F>
    void edge(const DataGraph& g, uint32_t a, uint32_t b, uint32_t el, F push) const {
        auto ia = vmap.find(g.vlabel[a]);
        auto ib = vmap.find(g.vlabel[b]);
        auto ie = emap.find(el);
        if (ia == vmap.end() || ib == vmap.end() || ie == emap.end()) return;
        const auto tup = std::make_tuple(ia->second, ie->second, ib->second);
        for (uint32_t i = 0; i < q.m; i++) {
            if (q.qe_labels[i] == tup) {
                push(q.qe_eidx[i].first, a, b);
                push(q.qe_eidx[i].second, b, a);
            }
            if (q.qe_rlabels[i] == tup) {
                push(q.qe_eidx[i].first, b, a);
                push(q.qe_eidx[i].second, a, b);
            }
        }
    }
};

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    uint64_t n = 0;
    void add(uint32_t a, uint32_t b) {
        const uint32_t pair[2] = {a, b};
        const unsigned char* p = reinterpret_cast<const unsigned char*>(pair);
        for (int k = 0; k < 8; k++) { h ^= p[k]; h *= 1099511628211ull; }
        n++;
    }
};

}  // namespace

void relation_edge_counts(const Query& q, const DataGraph& g, const uint32_t* src, const uint32_t* dst,
                          const uint32_t* elabel, size_t n, std::vector<uint64_t>& counts) {
    Grouper gr(q);
    counts.assign(2 * q.m, 0);
    for (size_t k = 0; k < n; k++)
        gr.edge(g, src[k], dst[k], elabel[k], [&](uint32_t r, uint32_t, uint32_t) { counts[r]++; });
}

std::string dump_reference_grouping(const Query& q, const DataGraph& g, const UpdateStream& s, uint32_t batch_size) {
    Grouper gr(q);
    std::ostringstream os;
    os << "dv " << g.dv << "\n";
    auto dump = [&](const std::vector<Fnv>& f) {
        for (uint32_t r = 0; r < 2 * q.m; r++) os << ' ' << r << ':' << f[r].n << ':' << std::hex << f[r].h << std::dec;
        os << "\n";
    };
    {
        std::vector<Fnv> f(2 * q.m);
        for (size_t k = 0; k < g.src.size(); k++)
            gr.edge(g, g.src[k], g.dst[k], g.elabel[k], [&](uint32_t r, uint32_t a, uint32_t b) { f[r].add(a, b); });
        os << "initial:";
        dump(f);
    }
    const size_t nb = batch_size == 0 ? 0 : (s.size() + batch_size - 1) / batch_size;
    os << "batches " << nb << "\n";
    for (size_t b = 0; b < nb; b++) {
        std::vector<Fnv> f(2 * q.m);
        const size_t lo = b * batch_size, hi = std::min(s.size(), lo + batch_size);
        for (size_t k = lo; k < hi; k++)
            gr.edge(g, s.src[k], s.dst[k], s.elabel[k], [&](uint32_t r, uint32_t a, uint32_t bb) { f[r].add(a, bb); });
        os << "batch " << b << ":";
        dump(f);
    }
    return os.str();
}

}  // namespace csm
