// CPU oracle for continuous subgraph matching on batch updates.
//
// TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
// bench.py as the checker and the reported CPU baseline.  The product never links or runs it.
// It shares no code with the product (own parser, own enumerator, own plan-independent order).
//
// Commands (results as text on stdout):
//   oracle count   <data> <query> [--directed]
//       |Emb(Q, G)|: injective, vertex-label and edge-label preserving, non-induced maps; every
//       automorphic image counts (reference: no automorphism handling, match_gpu.cu:279-286).
//   oracle stream  <data> <update> <batch> <query>... [--threads T] [--directed] [--max-batches K]
//                  [--batch-sizes FILE]  (variable batches: one size per line, replaces <batch>)
//                  [--anchor-stride S] [--time-limit-ms T]
//       Per batch t and query: N_t (deletion-anchored) and P_t (insertion-anchored) with the
//       embedding-set hashes, following the exactly-once interleaving of
//       technique_permutation.h:127-215: for the step of query edge e_i (qe_list order,
//       graph.cpp:332-347), f(e_i) is a batch edge, edges e_j (j < i) see the batch applied,
//       edges e_j (j > i) do not.  Deletions mirror it on G_t: f(e_j) in G_t \ D_t for j < i and in
//       G_t for j > i; they are processed before insertions.  --anchor-stride S processes only
//       every S-th anchor (a bounded CPU-baseline sample; counts are then partial); --time-limit-ms
//       stops taking anchors after T ms per batch (partial counts, reported with the anchors done).
//   oracle summary <data> <query> [--literal]
//       Offline global-index summary (match_gpu.cu:589-609): per relation r the cardinality and the
//       float average degree card/#sources.  Default: closed form (candidates = vertices passing
//       the NLF test on G_0, relation = arcs with both endpoints candidates).  --literal replays the
//       reference's step-by-step UpdateGlobalIndex (match_gpu.cu:611-834, indexing.cu:303-439) on
//       sets; the two must agree.
//   oracle nlf     <data> <query>
//       Per query vertex: number of NLF candidates and an FNV hash of the sorted candidate ids.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace {

// ---- embedding hash (MurmurHash3 fmix64, public domain) --------------------------------------
uint64_t fmix64(uint64_t k) {
    k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
    return k;
}
uint64_t pair_hash(uint32_t q, uint32_t v) { return fmix64(((uint64_t)v << 4 | q) + 0x9E3779B97F4A7C15ull); }

// ---- files -------------------------------------------------------------------------------------
bool slurp(const char* path, std::string& s) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return true;
}
struct Lex {
    const char* p; const char* e;
    void ws() { while (p < e && (unsigned char)*p <= ' ') p++; }
    bool ch(char& c) { ws(); if (p >= e) return false; c = *p++; return true; }
    bool num(uint32_t& v) { ws(); if (p >= e || *p < '0' || *p > '9') return false; uint64_t x = 0; while (p < e && *p >= '0' && *p <= '9') x = x * 10 + (*p++ - '0'); v = (uint32_t)x; return true; }
};

struct Query {
    uint32_t n = 0;
    std::vector<uint32_t> lab;                    // raw labels
    std::vector<std::array<uint32_t, 3>> edges;   // (a, b, elabel) as in the file (a -> b if directed)
    // step order (qe_list): u ascending, neighbour ascending, u < w
    std::vector<std::array<uint32_t, 3>> steps;   // (u, w, elabel), u < w
    std::vector<bool> step_forward;               // directed: is the file edge u -> w ?
};

bool load_query(const char* path, Query& q) {
    std::string s;
    if (!slurp(path, s)) return false;
    Lex L{s.data(), s.data() + s.size()};
    char c;
    while (L.ch(c)) {
        if (c == 't') { char x; uint32_t y; L.ch(x); L.num(y); }
        else if (c == 'v') { uint32_t id = 0, l = 0; L.num(id); L.num(l); if (id >= q.lab.size()) q.lab.resize(id + 1); q.lab[id] = l; q.n++; }
        else { uint32_t a = 0, b = 0, l = 0; L.num(a); L.num(b); L.num(l); q.edges.push_back({a, b, l}); }
    }
    // adjacency sorted by neighbour: the undirected edges in (min, max) order sorted
    std::vector<std::array<uint32_t, 4>> st;
    for (auto& e : q.edges) st.push_back({std::min(e[0], e[1]), std::max(e[0], e[1]), e[2], e[0] < e[1] ? 1u : 0u});
    std::sort(st.begin(), st.end());
    for (auto& e : st) { q.steps.push_back({e[0], e[1], e[2]}); q.step_forward.push_back(e[3] != 0); }
    return true;
}

struct Graph {
    bool directed = false;
    uint32_t dv = 0;
    std::vector<uint32_t> lab;
    // out[v] / in[v]: sorted (nbr << 32 | elabel)... stored as pairs
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> out, in;
    static bool has(const std::vector<std::pair<uint32_t, uint32_t>>& l, uint32_t w, uint32_t el) {
        auto it = std::lower_bound(l.begin(), l.end(), std::make_pair(w, el));
        return it != l.end() && it->first == w && it->second == el;
    }
    void add_arc(uint32_t a, uint32_t b, uint32_t el) {
        auto& o = out[a];
        o.insert(std::lower_bound(o.begin(), o.end(), std::make_pair(b, el)), std::make_pair(b, el));
        auto& i = in[b];
        i.insert(std::lower_bound(i.begin(), i.end(), std::make_pair(a, el)), std::make_pair(a, el));
    }
    void del_arc(uint32_t a, uint32_t b, uint32_t el) {
        auto& o = out[a];
        auto it = std::lower_bound(o.begin(), o.end(), std::make_pair(b, el));
        if (it != o.end() && *it == std::make_pair(b, el)) o.erase(it);
        auto& i = in[b];
        auto jt = std::lower_bound(i.begin(), i.end(), std::make_pair(a, el));
        if (jt != i.end() && *jt == std::make_pair(a, el)) i.erase(jt);
    }
    void add_edge(uint32_t a, uint32_t b, uint32_t el) { add_arc(a, b, el); if (!directed) add_arc(b, a, el); }
    void del_edge(uint32_t a, uint32_t b, uint32_t el) { del_arc(a, b, el); if (!directed) del_arc(b, a, el); }
};

bool load_graph(const char* path, bool directed, Graph& g) {
    std::string s;
    if (!slurp(path, s)) return false;
    Lex L{s.data(), s.data() + s.size()};
    char c;
    g.directed = directed;
    std::vector<std::array<uint32_t, 3>> es;
    while (L.ch(c)) {
        if (c == 't') { char x; uint32_t y; L.ch(x); L.num(y); }
        else if (c == 'v') { uint32_t id = 0, l = 0; L.num(id); L.num(l); if (id >= g.lab.size()) g.lab.resize(id + 1, UINT32_MAX); g.lab[id] = l; }
        else { uint32_t a = 0, b = 0, l = 0; L.num(a); L.num(b); L.num(l); es.push_back({a, b, l}); }
    }
    g.dv = (uint32_t)g.lab.size();
    g.out.assign(g.dv, {});
    g.in.assign(g.dv, {});
    for (auto& e : es) { g.out[e[0]].push_back({e[1], e[2]}); g.in[e[1]].push_back({e[0], e[2]}); if (!directed) { g.out[e[1]].push_back({e[0], e[2]}); g.in[e[0]].push_back({e[1], e[2]}); } }
    for (uint32_t v = 0; v < g.dv; v++) { std::sort(g.out[v].begin(), g.out[v].end()); std::sort(g.in[v].begin(), g.in[v].end()); }
    return true;
}

struct Update { uint32_t a, b, l; bool del; };
bool load_updates(const char* path, std::vector<Update>& u) {
    std::string s;
    if (!slurp(path, s)) return false;
    Lex L{s.data(), s.data() + s.size()};
    char c;
    while (L.ch(c)) {
        if (c == 't') { char x; uint32_t y; L.ch(x); L.num(y); continue; }
        bool del = false;
        if (c == '-') { char e; L.ch(e); del = true; }
        uint32_t a = 0, b = 0, l = 0;
        L.num(a); L.num(b); L.num(l);
        u.push_back({a, b, l, del});
    }
    return true;
}

// ---- brute force -----------------------------------------------------------------------------
// Query-vertex order: BFS from vertex 0; each later vertex is extended from its first earlier
// neighbour and checked against all other earlier neighbours.
struct Brute {
    const Graph& g; const Query& q;
    std::vector<uint32_t> order, pos;
    std::vector<std::vector<std::array<uint32_t, 3>>> back;  // for order[k]: (earlier vertex, elabel, dir: 1 if earlier->this)
    uint64_t count = 0, hash = 0;
    std::vector<uint32_t> f;
    Brute(const Graph& gg, const Query& qq) : g(gg), q(qq) {
        pos.assign(q.n, UINT32_MAX);
        order.push_back(0); pos[0] = 0;
        for (size_t k = 0; k < order.size(); k++)
            for (auto& e : q.edges) {
                uint32_t x = UINT32_MAX;
                if (e[0] == order[k]) x = e[1]; else if (e[1] == order[k]) x = e[0];
                if (x != UINT32_MAX && pos[x] == UINT32_MAX) { pos[x] = (uint32_t)order.size(); order.push_back(x); }
            }
        back.resize(q.n);
        for (uint32_t k = 0; k < q.n; k++)
            for (auto& e : q.edges) {
                const uint32_t u = order[k];
                if (e[1] == u && pos[e[0]] < k) back[k].push_back({e[0], e[2], 1u});
                else if (e[0] == u && pos[e[1]] < k) back[k].push_back({e[1], e[2], 0u});
            }
        f.assign(q.n, 0);
    }
    void rec(uint32_t k) {
        if (k == q.n) {
            uint64_t s = 0;
            for (uint32_t u = 0; u < q.n; u++) s += pair_hash(u, f[u]);
            count++; hash += fmix64(s);
            return;
        }
        const uint32_t u = order[k];
        auto ok = [&](uint32_t c) {
            if (g.lab[c] != q.lab[u]) return false;
            for (uint32_t j = 0; j < k; j++) if (f[order[j]] == c) return false;
            for (auto& b : back[k]) {
                const uint32_t fb = f[b[0]];
                bool present = b[2] ? Graph::has(g.out[fb], c, b[1]) : Graph::has(g.out[c], fb, b[1]);
                if (!present) return false;
            }
            return true;
        };
        if (k == 0) {
            for (uint32_t c = 0; c < g.dv; c++) if (g.lab[c] != UINT32_MAX && ok(c)) { f[u] = c; rec(k + 1); }
            return;
        }
        const auto& b0 = back[k][0];
        const auto& lst = b0[2] ? g.out[f[b0[0]]] : g.in[f[b0[0]]];
        for (auto& p : lst) if (p.second == b0[1] && ok(p.first)) { f[u] = p.first; rec(k + 1); }
    }
};

// ---- anchored incremental enumeration -----------------------------------------------------------
enum ViewKind { kBase, kBasePlus, kDelta, kBaseMinus };

struct DeltaSet {  // batch arcs of one kind (insertions or deletions) with adjacency for iteration
    // Membership is exact: an arc is the pair (a << 32 | b, label) and set equality compares both
    // members; the hash only picks the bucket.
    struct Arc {
        uint64_t ab;
        uint32_t l;
        bool operator==(const Arc& o) const { return ab == o.ab && l == o.l; }
    };
    struct ArcHash {
        size_t operator()(const Arc& x) const { return static_cast<size_t>(fmix64(x.ab ^ ((uint64_t)x.l << 1))); }
    };
    std::unordered_set<Arc, ArcHash> arcs;
    std::unordered_map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>> out, in;
    static Arc key(uint32_t a, uint32_t b, uint32_t l) { return Arc{(uint64_t)a << 32 | b, l}; }
    void add_arc(uint32_t a, uint32_t b, uint32_t l) { arcs.insert(key(a, b, l)); out[a].push_back({b, l}); in[b].push_back({a, l}); }
    bool has(uint32_t a, uint32_t b, uint32_t l) const { return arcs.count(key(a, b, l)) != 0; }
};

struct Anchored {
    const Graph& g; const Query& q; const DeltaSet& d; bool deletion;
    uint32_t step;
    std::vector<uint32_t> order, pos;
    struct Back { uint32_t w; uint32_t el; uint32_t dir; ViewKind view; };  // dir 1: f(w) -> f(u)
    std::vector<std::vector<Back>> back;
    std::vector<uint32_t> f;
    uint64_t count = 0, hash = 0;

    Anchored(const Graph& gg, const Query& qq, const DeltaSet& dd, bool del, uint32_t i)
        : g(gg), q(qq), d(dd), deletion(del), step(i) {
        const uint32_t u0 = q.steps[i][0], u1 = q.steps[i][1];
        pos.assign(q.n, UINT32_MAX);
        order = {u0, u1}; pos[u0] = 0; pos[u1] = 1;
        for (size_t k = 0; k < order.size(); k++)
            for (auto& e : q.edges) {
                uint32_t x = UINT32_MAX;
                if (e[0] == order[k]) x = e[1]; else if (e[1] == order[k]) x = e[0];
                if (x != UINT32_MAX && pos[x] == UINT32_MAX) { pos[x] = (uint32_t)order.size(); order.push_back(x); }
            }
        back.resize(q.n);
        for (uint32_t k = 2; k < q.n; k++) {
            const uint32_t u = order[k];
            for (uint32_t s = 0; s < q.steps.size(); s++) {
                const auto& e = q.steps[s];
                uint32_t w;
                if (e[0] == u) w = e[1]; else if (e[1] == u) w = e[0]; else continue;
                if (pos[w] >= k) continue;
                // direction of the file edge relative to (w -> u)
                const bool fwd_file = q.step_forward[s];  // file edge is steps[s][0] -> steps[s][1]
                const uint32_t from = fwd_file ? e[0] : e[1];
                ViewKind v = s < step ? (deletion ? kBaseMinus : kBasePlus) : (s == step ? kDelta : kBase);
                back[k].push_back({w, e[2], from == w ? 1u : 0u, v});
            }
        }
        f.assign(q.n, 0);
    }
    bool in_view(uint32_t a, uint32_t b, uint32_t l, ViewKind v) const {
        switch (v) {
            case kBase: return Graph::has(g.out[a], b, l);
            case kBasePlus: return Graph::has(g.out[a], b, l) || d.has(a, b, l);
            case kDelta: return d.has(a, b, l);
            case kBaseMinus: return Graph::has(g.out[a], b, l) && !d.has(a, b, l);
        }
        return false;
    }
    void emit() {
        uint64_t s = 0;
        for (uint32_t u = 0; u < q.n; u++) s += pair_hash(u, f[u]);
        count++; hash += fmix64(s);
    }
    void rec(uint32_t k) {
        if (k == q.n) { emit(); return; }
        const uint32_t u = order[k];
        const Back& p = back[k][0];
        auto ok = [&](uint32_t c) {
            if (g.lab[c] != q.lab[u]) return false;
            for (uint32_t j = 0; j < k; j++) if (f[order[j]] == c) return false;
            for (const Back& b : back[k]) {
                const uint32_t fb = f[b.w];
                if (!(b.dir ? in_view(fb, c, b.el, b.view) : in_view(c, fb, b.el, b.view))) return false;
            }
            return true;
        };
        const uint32_t fp = f[p.w];
        auto scan = [&](const std::vector<std::pair<uint32_t, uint32_t>>& lst) {
            for (auto& x : lst) if (x.second == p.el && ok(x.first)) { f[u] = x.first; rec(k + 1); }
        };
        const auto& base = p.dir ? g.out[fp] : g.in[fp];
        const auto& dm = p.dir ? d.out : d.in;
        auto it = dm.find(fp);
        if (p.view != kDelta) scan(base);
        if ((p.view == kBasePlus || p.view == kDelta) && it != dm.end()) scan(it->second);
    }
    // anchor: f(u0) = a, f(u1) = b for the batch arc a -> b (oriented to the query edge)
    void anchor(uint32_t a, uint32_t b) {
        const uint32_t u0 = order[0], u1 = order[1];
        if (g.lab[a] != q.lab[u0] || g.lab[b] != q.lab[u1]) return;
        f[u0] = a; f[u1] = b;
        rec(2);
    }
};

struct StepResult { uint64_t count = 0, hash = 0; };

// Anchored counts for one kind of batch arcs (deletions or insertions) over all steps.
std::chrono::steady_clock::time_point g_deadline = std::chrono::steady_clock::time_point::max();
std::atomic<uint64_t> g_anchors_done{0}, g_anchors_total{0};

StepResult run_phase(const Graph& g, const Query& q, const std::vector<Update>& batch, bool deletion,
                     int threads, uint32_t stride) {
    DeltaSet d;
    for (auto& e : batch) if (e.del == deletion) { d.add_arc(e.a, e.b, e.l); if (!g.directed) d.add_arc(e.b, e.a, e.l); }
    // anchors per step: arcs (a, b) with f(u0)=a, f(u1)=b respecting direction and edge label
    std::vector<std::pair<uint32_t, std::pair<uint32_t, uint32_t>>> work;  // (step, (a, b))
    for (uint32_t i = 0; i < q.steps.size(); i++) {
        const uint32_t el = q.steps[i][2];
        for (auto& e : batch) {
            if (e.del != deletion || e.l != el) continue;
            if (g.directed) {
                if (q.step_forward[i]) work.push_back({i, {e.a, e.b}}); else work.push_back({i, {e.b, e.a}});
            } else {
                work.push_back({i, {e.a, e.b}});
                work.push_back({i, {e.b, e.a}});
            }
        }
    }
    std::atomic<size_t> next{0};
    std::vector<StepResult> part(threads);
    auto body = [&](int t) {
        std::vector<Anchored*> ctx(q.steps.size(), nullptr);
        for (uint32_t i = 0; i < q.steps.size(); i++) ctx[i] = new Anchored(g, q, d, deletion, i);
        while (true) {
            const size_t k0 = next.fetch_add(64);
            if (k0 >= work.size()) break;
            if (std::chrono::steady_clock::now() > g_deadline) break;
            for (size_t k = k0; k < std::min(work.size(), k0 + 64); k++) {
                if (stride > 1 && (k % stride) != 0) continue;
                if (std::chrono::steady_clock::now() > g_deadline) break;
                ctx[work[k].first]->anchor(work[k].second.first, work[k].second.second);
                g_anchors_done++;
            }
        }
        for (auto* c : ctx) { part[t].count += c->count; part[t].hash += c->hash; delete c; }
    };
    for (size_t k = 0; k < work.size(); k++) if (stride <= 1 || k % stride == 0) g_anchors_total++;
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; t++) ts.emplace_back(body, t);
    for (auto& t : ts) t.join();
    StepResult r;
    for (auto& p : part) { r.count += p.count; r.hash += p.hash; }
    return r;
}

// ---- global-index summary --------------------------------------------------------------------
// Query metadata restated from graph.cpp:271-357 (only what the summary needs).
struct QMeta {
    uint32_t n, m;
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> nbrs;  // (nbr, elabel raw), sorted by (nbr, remapped elabel)
    std::vector<uint32_t> off, rnbr, rel_el;   // relation r = position in the CSR
    std::vector<uint32_t> nlf, first_nl, last_nl;
    std::vector<std::array<uint32_t, 4>> qe;   // (u, w, r_uw, r_wu)
    std::vector<uint32_t> src_of;
};

QMeta qmeta(const Query& q) {
    QMeta M;
    M.n = q.n; M.m = (uint32_t)q.edges.size();
    // the reference sorts adjacency by (neighbour, remapped edge label); remapping is first-seen
    std::map<uint32_t, uint32_t> emap;
    for (auto& e : q.edges) if (!emap.count(e[2])) { uint32_t k = (uint32_t)emap.size(); emap[e[2]] = k; }
    std::vector<std::vector<std::array<uint32_t, 3>>> adj(q.n);
    for (auto& e : q.edges) { adj[e[0]].push_back({e[1], emap[e[2]], e[2]}); adj[e[1]].push_back({e[0], emap[e[2]], e[2]}); }
    M.nbrs.resize(q.n);
    M.off.push_back(0);
    for (uint32_t u = 0; u < q.n; u++) {
        std::sort(adj[u].begin(), adj[u].end());
        for (auto& x : adj[u]) { M.nbrs[u].push_back({x[0], x[2]}); M.rnbr.push_back(x[0]); M.rel_el.push_back(x[2]); M.src_of.push_back(u); }
        M.off.push_back((uint32_t)M.rnbr.size());
    }
    const uint32_t R = M.off[q.n];
    M.nlf.assign(R, 0); M.first_nl.assign(R, 0); M.last_nl.assign(R, 0);
    for (uint32_t u = 0; u < q.n; u++) {
        std::map<std::pair<uint32_t, uint32_t>, std::array<uint32_t, 3>> grp;
        for (uint32_t r = M.off[u]; r < M.off[u + 1]; r++) {
            auto key = std::make_pair(M.rel_el[r], q.lab[M.rnbr[r]]);
            auto it = grp.find(key);
            if (it == grp.end()) grp[key] = {r, r, 1}; else { it->second[1] = r; it->second[2]++; }
        }
        for (uint32_t r = M.off[u]; r < M.off[u + 1]; r++) {
            auto& t = grp[std::make_pair(M.rel_el[r], q.lab[M.rnbr[r]])];
            if (t[2] > 2) { M.first_nl[r] = t[0]; M.last_nl[r] = t[1]; } else { M.first_nl[r] = r; M.last_nl[r] = r; }
        }
        for (auto& kv : grp) M.nlf[kv.second[0]] = kv.second[2];
    }
    auto rel = [&](uint32_t u, uint32_t w) { for (uint32_t r = M.off[u]; r < M.off[u + 1]; r++) if (M.rnbr[r] == w) return r; return UINT32_MAX; };
    for (uint32_t u = 0; u < q.n; u++)
        for (uint32_t r = M.off[u]; r < M.off[u + 1]; r++)
            if (u < M.rnbr[r]) M.qe.push_back({u, M.rnbr[r], r, rel(M.rnbr[r], u)});
    return M;
}

// Data arcs of relation r (u -> w): v with label(u), x with label(w), edge label of r
// (undirected interpretation, graph.cpp:435-459).
std::vector<std::pair<uint32_t, uint32_t>> relation_arcs(const Graph& g, const Query& q, const QMeta& M, uint32_t r) {
    std::vector<std::pair<uint32_t, uint32_t>> out;
    const uint32_t u = M.src_of[r], w = M.rnbr[r], el = M.rel_el[r];
    for (uint32_t v = 0; v < g.dv; v++) {
        if (g.lab[v] != q.lab[u]) continue;
        for (auto& p : g.out[v]) if (p.second == el && g.lab[p.first] == q.lab[w]) out.push_back({v, p.first});
    }
    return out;
}

void summary_closed(const Graph& g, const Query& q, const QMeta& M, std::vector<uint32_t>& card, std::vector<float>& avg,
                    std::vector<std::vector<char>>* cand_out = nullptr) {
    const uint32_t R = M.off[q.n];
    std::vector<std::vector<uint32_t>> deg(R, std::vector<uint32_t>(g.dv, 0));
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> arcs(R);
    for (uint32_t r = 0; r < R; r++) { arcs[r] = relation_arcs(g, q, M, r); for (auto& a : arcs[r]) deg[r][a.first]++; }
    std::vector<std::vector<char>> cand(q.n, std::vector<char>(g.dv, 0));
    for (uint32_t u = 0; u < q.n; u++)
        for (uint32_t v = 0; v < g.dv; v++) {
            if (g.lab[v] != q.lab[u]) continue;
            bool pass = true, any = false;
            for (uint32_t r = M.off[u]; r < M.off[u + 1]; r++) { if (deg[r][v] < M.nlf[r]) pass = false; if (deg[r][v]) any = true; }
            cand[u][v] = pass && any;
        }
    card.assign(R, 0); avg.assign(R, 0.f);
    for (uint32_t r = 0; r < R; r++) {
        const uint32_t u = M.src_of[r], w = M.rnbr[r];
        std::vector<uint32_t> per(g.dv, 0);
        uint32_t c = 0, ns = 0;
        for (auto& a : arcs[r]) if (cand[u][a.first] && cand[w][a.second]) { per[a.first]++; c++; }
        for (uint32_t v = 0; v < g.dv; v++) if (per[v]) ns++;
        card[r] = c; avg[r] = (float)c / (float)ns;
    }
    if (cand_out) *cand_out = cand;
}

// Literal replay of the offline UpdateGlobalIndex over steps i = 0..m-1 (match_gpu.cu:611-834).
void summary_literal(const Graph& g, const Query& q, const QMeta& M, std::vector<uint32_t>& card, std::vector<float>& avg) {
    const uint32_t R = M.off[q.n];
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> csr(R), gi(R);
    std::vector<std::vector<std::vector<uint32_t>>> data(R, std::vector<std::vector<uint32_t>>(g.dv));
    for (uint32_t r = 0; r < R; r++) csr[r] = relation_arcs(g, q, M, r);
    std::vector<std::vector<uint32_t>> ddeg(R, std::vector<uint32_t>(g.dv, 0));
    std::vector<std::vector<char>> cand(q.n, std::vector<char>(g.dv, 0));
    auto rel = [&](uint32_t u, uint32_t w) { for (uint32_t r = M.off[u]; r < M.off[u + 1]; r++) if (M.rnbr[r] == w) return r; return UINT32_MAX; };
    for (uint32_t i = 0; i < M.qe.size(); i++) {
        for (int j = 0; j < 2; j++) {
            const uint32_t idx = j == 0 ? M.qe[i][2] : M.qe[i][3];
            const uint32_t u = j == 0 ? M.qe[i][0] : M.qe[i][1];
            if (M.nlf[idx] == 0) continue;
            std::map<uint32_t, uint32_t> bdeg;  // batch sources of relation idx with their batch degree
            for (auto& a : csr[idx]) bdeg[a.first]++;
            std::vector<uint32_t> fresh;
            for (auto& kv : bdeg) {
                const uint32_t v = kv.first;
                bool pass = true;
                for (uint32_t k = M.off[u]; k < M.off[u + 1]; k++) {
                    const uint32_t d = ddeg[k][v] + (k == idx ? kv.second : 0);
                    if (d < M.nlf[k]) { pass = false; break; }
                }
                if (pass && !cand[u][v]) { cand[u][v] = 1; fresh.push_back(v); }
            }
            for (uint32_t k = M.off[u]; k < M.off[u + 1]; k++) {
                const uint32_t uo = M.rnbr[k];
                const uint32_t rk = rel(uo, u);
                for (uint32_t v : fresh)
                    for (uint32_t x : data[k][v]) if (cand[uo][x]) { gi[k].push_back({v, x}); gi[rk].push_back({x, v}); }
            }
        }
        for (int j = 0; j < 2; j++) {
            const uint32_t idx = j == 0 ? M.qe[i][2] : M.qe[i][3];
            const uint32_t u = j == 0 ? M.qe[i][0] : M.qe[i][1];
            const uint32_t uu = j == 0 ? M.qe[i][1] : M.qe[i][0];
            for (auto& a : csr[idx]) { data[idx][a.first].push_back(a.second); ddeg[idx][a.first]++; }
            for (auto& a : csr[idx]) if (cand[u][a.first] && cand[uu][a.second]) gi[idx].push_back(a);
        }
    }
    card.assign(R, 0); avg.assign(R, 0.f);
    for (uint32_t r = 0; r < R; r++) {
        std::set<uint32_t> srcs;
        for (auto& a : gi[r]) srcs.insert(a.first);
        card[r] = (uint32_t)gi[r].size();
        avg[r] = (float)card[r] / (float)(uint32_t)srcs.size();
    }
}

int usage() {
    std::fprintf(stderr, "usage: oracle count|stream|summary|nlf ...\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) return usage();
    const std::string cmd = argv[1];
    bool directed = false;
    int threads = 1;
    uint32_t max_batches = UINT32_MAX, stride = 1;
    bool literal = false;
    double time_limit_ms = 0;
    const char* batch_sizes = nullptr;
    std::vector<const char*> pos;
    for (int i = 2; i < argc; i++) {
        if (!std::strcmp(argv[i], "--directed")) directed = true;
        else if (!std::strcmp(argv[i], "--literal")) literal = true;
        else if (!std::strcmp(argv[i], "--threads") && i + 1 < argc) threads = std::max(1, std::atoi(argv[++i]));
        else if (!std::strcmp(argv[i], "--max-batches") && i + 1 < argc) max_batches = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        else if (!std::strcmp(argv[i], "--anchor-stride") && i + 1 < argc) stride = std::max(1u, (uint32_t)std::strtoul(argv[++i], nullptr, 10));
        else if (!std::strcmp(argv[i], "--time-limit-ms") && i + 1 < argc) time_limit_ms = std::atof(argv[++i]);
        else if (!std::strcmp(argv[i], "--batch-sizes") && i + 1 < argc) batch_sizes = argv[++i];
        else pos.push_back(argv[i]);
    }
    if (cmd == "count" && pos.size() == 2) {
        Graph g; Query q;
        if (!load_graph(pos[0], directed, g) || !load_query(pos[1], q)) return 1;
        Brute b(g, q);
        b.rec(0);
        std::printf("count %llu hash %016llx\n", (unsigned long long)b.count, (unsigned long long)b.hash);
        return 0;
    }
    if (cmd == "summary" && pos.size() == 2) {
        Graph g; Query q;
        if (!load_graph(pos[0], false, g) || !load_query(pos[1], q)) return 1;
        QMeta M = qmeta(q);
        std::vector<uint32_t> card; std::vector<float> avg;
        if (literal) summary_literal(g, q, M, card, avg); else summary_closed(g, q, M, card, avg);
        for (uint32_t r = 0; r < card.size(); r++) { uint32_t bits; std::memcpy(&bits, &avg[r], 4); std::printf("%u %u %x\n", r, card[r], bits); }
        return 0;
    }
    if (cmd == "nlf" && pos.size() == 2) {
        Graph g; Query q;
        if (!load_graph(pos[0], false, g) || !load_query(pos[1], q)) return 1;
        QMeta M = qmeta(q);
        std::vector<uint32_t> card; std::vector<float> avg; std::vector<std::vector<char>> cand;
        summary_closed(g, q, M, card, avg, &cand);
        for (uint32_t u = 0; u < q.n; u++) {
            uint64_t h = 1469598103934665603ull, c = 0;
            for (uint32_t v = 0; v < g.dv; v++) if (cand[u][v]) { c++; h ^= v; h *= 1099511628211ull; }
            std::printf("u %u count %llu hash %016llx\n", u, (unsigned long long)c, (unsigned long long)h);
        }
        return 0;
    }
    if (cmd == "stream" && pos.size() >= 4) {
        Graph g;
        if (!load_graph(pos[0], directed, g)) return 1;
        std::vector<Update> ups;
        if (!load_updates(pos[1], ups)) return 1;
        const uint32_t bs = (uint32_t)std::strtoul(pos[2], nullptr, 10);
        std::vector<Query> qs(pos.size() - 3);
        for (size_t k = 3; k < pos.size(); k++) if (!load_query(pos[k], qs[k - 3])) return 1;
        // batch b covers updates [bounds[b], bounds[b + 1]): fixed size bs, or the sizes listed one
        // per line in --batch-sizes (variable batches, as a serving loop applies them)
        std::vector<size_t> bounds{0};
        if (batch_sizes) {
            FILE* f = std::fopen(batch_sizes, "r");
            if (!f) { std::fprintf(stderr, "cannot read %s\n", batch_sizes); return 1; }
            unsigned long long x;
            while (std::fscanf(f, "%llu", &x) == 1 && bounds.back() < ups.size()) bounds.push_back(std::min<size_t>(ups.size(), bounds.back() + x));
            std::fclose(f);
        } else if (bs) {
            while (bounds.back() < ups.size()) bounds.push_back(std::min<size_t>(ups.size(), bounds.back() + bs));
        }
        const size_t nb = bounds.size() - 1;
        std::printf("threads %d batches %zu\n", threads, nb);
        for (size_t b = 0; b < nb && b < max_batches; b++) {
            std::vector<Update> batch(ups.begin() + bounds[b], ups.begin() + bounds[b + 1]);
            auto t0 = std::chrono::steady_clock::now();
            if (time_limit_ms > 0) g_deadline = t0 + std::chrono::microseconds((long long)(time_limit_ms * 1000));
            g_anchors_done = 0;
            g_anchors_total = 0;
            std::vector<StepResult> neg(qs.size()), posr(qs.size());
            for (size_t k = 0; k < qs.size(); k++) neg[k] = run_phase(g, qs[k], batch, true, threads, stride);
            for (auto& e : batch) if (e.del) g.del_edge(e.a, e.b, e.l);
            for (size_t k = 0; k < qs.size(); k++) posr[k] = run_phase(g, qs[k], batch, false, threads, stride);
            for (auto& e : batch) if (!e.del) g.add_edge(e.a, e.b, e.l);
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            for (size_t k = 0; k < qs.size(); k++)
                std::printf("batch %zu query %zu neg %llu nhash %016llx pos %llu phash %016llx\n", b, k,
                            (unsigned long long)neg[k].count, (unsigned long long)neg[k].hash,
                            (unsigned long long)posr[k].count, (unsigned long long)posr[k].hash);
            std::printf("batch %zu edges %zu ms %.3f\n", b, batch.size(), ms);
            std::printf("batch %zu anchors %llu of %llu\n", b, (unsigned long long)g_anchors_done.load(), (unsigned long long)g_anchors_total.load());
            std::fflush(stdout);
        }
        return 0;
    }
    return usage();
}
