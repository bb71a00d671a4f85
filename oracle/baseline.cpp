// CPU baseline for continuous subgraph matching on batch updates: a multithreaded backtracking
// enumerator over the planner's matching orders.
//
// TEST INFRASTRUCTURE ONLY: run by the cpu_baseline leg of bench.py as the reported CPU baseline.
// The product never links or runs it.  It shares no code with the product: the matching orders
// come in as the PrintOrders text of each query's plan (plan.cpp:1200-1291 format), which the
// caller obtains from the planner.
//
//   baseline <data> <update> <batch> <query> <orders> [<query> <orders>]...
//            [--threads T] [--max-batches K] [--time-limit-ms T] [--seed S]
//
// Semantics per batch (technique_permutation.h:127-215, with deletions mirrored before the
// insertions): for the step of query edge e_i (the i-th "[u0,u1]:" entry of the orders text,
// which is qe_list order), f(e_i) is a batch arc of the phase, query edges e_j with j < i see the
// phase's batch applied and edges with j > i do not.  The step's search starts at its anchor
// edge (u0, u1) and then places the remaining query vertices in the order the plan gives for that
// step; each vertex's candidates come from the run of its first planned backward neighbour that is
// already placed, and every other placed neighbour is checked by binary search.  The last vertex
// is counted, not placed.
//
// Work items are (query, step, anchor arc) triples, processed in a seeded random order by the
// threads; with --time-limit-ms the threads stop taking items after that time, so the items done
// form a uniform random sample of the batch's anchors.  Output per batch:
//   batch <b> items <done> of <total> matches <count> ms <wall time of the items>
// (counts only: positive plus negative matches of the items done; the deletion and insertion
// phases' items are mixed, both reading G_t through their views).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace {

bool slurp(const char* path, std::string& s) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    char buf[1 << 16];
    size_t n;
    s.clear();
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, n);
    std::fclose(f);
    return true;
}

struct Reader {
    const char* p;
    const char* e;
    bool word(std::string& w) {
        while (p < e && std::isspace(static_cast<unsigned char>(*p))) p++;
        if (p >= e) return false;
        const char* s = p;
        while (p < e && !std::isspace(static_cast<unsigned char>(*p))) p++;
        w.assign(s, p);
        return true;
    }
    uint32_t num() {
        while (p < e && std::isspace(static_cast<unsigned char>(*p))) p++;
        uint32_t x = 0;
        while (p < e && *p >= '0' && *p <= '9') x = x * 10 + static_cast<uint32_t>(*p++ - '0');
        return x;
    }
};

[[noreturn]] void die(const std::string& m) {
    std::fprintf(stderr, "baseline: %s\n", m.c_str());
    std::exit(1);
}

struct Query {
    uint32_t n = 0;
    std::vector<uint32_t> vlab;                          // raw labels
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> adj;  // (neighbour, raw edge label)
    uint32_t elabel(uint32_t a, uint32_t b) const {
        for (auto& x : adj[a]) if (x.first == b) return x.second;
        return UINT32_MAX;
    }
};

Query read_query(const char* path) {
    std::string s, w;
    if (!slurp(path, s)) die(std::string("cannot read ") + path);
    Reader r{s.data(), s.data() + s.size()};
    Query q;
    while (r.word(w)) {
        if (w == "t") { r.word(w); r.word(w); }
        else if (w == "v") { const uint32_t id = r.num(), l = r.num(); if (id >= q.vlab.size()) q.vlab.resize(id + 1); q.vlab[id] = l; }
        else if (w == "e") {
            const uint32_t a = r.num(), b = r.num(), l = r.num();
            if (std::max(a, b) >= q.adj.size()) q.adj.resize(std::max(a, b) + 1);
            q.adj[a].emplace_back(b, l);
            q.adj[b].emplace_back(a, l);
        }
    }
    q.n = static_cast<uint32_t>(q.vlab.size());
    q.adj.resize(q.n);
    return q;
}

// One step of a plan: the anchor query edge and the planned order with backward neighbours.
struct PlanStep {
    uint32_t u0, u1;
    std::vector<uint32_t> order;                  // query vertices
    std::vector<std::vector<uint32_t>> back;      // per position: backward neighbours (query vertices), planned order
};

// Parses the "[u0,u1]:" entries and their "order:" lines of a PrintOrders text.
std::vector<PlanStep> read_orders(const char* path) {
    std::string s;
    if (!slurp(path, s)) die(std::string("cannot read ") + path);
    std::vector<PlanStep> steps;
    size_t pos = 0;
    while ((pos = s.find('[', pos)) != std::string::npos) {
        PlanStep st;
        if (std::sscanf(s.c_str() + pos, "[%u,%u]:", &st.u0, &st.u1) != 2) { pos++; continue; }
        const size_t o = s.find("order: ", s.find("indexing order:", pos) + 15);
        if (o == std::string::npos) die("orders text without an order line");
        const size_t end = s.find('\n', o);
        std::string line = s.substr(o + 7, end - o - 7);
        // tokens "v-(b@p b@p) " ; backward entries list query vertices before '@'
        size_t k = 0;
        while (k < line.size()) {
            while (k < line.size() && line[k] == ' ') k++;
            if (k >= line.size()) break;
            const uint32_t v = static_cast<uint32_t>(std::strtoul(line.c_str() + k, nullptr, 10));
            const size_t open = line.find('(', k), close = line.find(')', k);
            std::vector<uint32_t> bk;
            for (size_t t = open + 1; t < close;) {
                while (t < close && line[t] == ' ') t++;
                if (t >= close) break;
                bk.push_back(static_cast<uint32_t>(std::strtoul(line.c_str() + t, nullptr, 10)));
                t = line.find_first_of(" )", t);
            }
            st.order.push_back(v);
            st.back.push_back(bk);
            k = close + 1;
        }
        steps.push_back(st);
        pos = end;
    }
    return steps;
}

// Base adjacency: per vertex, keys (slot << 32 | neighbour) sorted; slot = vertex label * nel + edge
// label rank.  The batch's arcs live in hash maps and per-vertex sorted lists beside it.
struct Graph {
    uint32_t nv = 0, nel = 0;
    std::vector<uint32_t> vlab;
    std::unordered_map<uint32_t, uint32_t> erank;
    std::vector<std::vector<uint64_t>> adj;
    uint64_t key(uint32_t nbr, uint32_t el) const { return (static_cast<uint64_t>(vlab[nbr] * nel + el) << 32) | nbr; }
    bool has(uint32_t a, uint64_t k) const { return std::binary_search(adj[a].begin(), adj[a].end(), k); }
};

struct Batch {
    std::unordered_map<uint64_t, uint32_t> arcs[2];       // [phase]: (a << 32 | b) -> edge label rank
    std::unordered_map<uint32_t, std::vector<uint64_t>> out[2];  // [phase]: vertex -> sorted keys of its batch arcs
    bool has(int ph, uint32_t a, uint32_t b, uint32_t el) const {
        auto it = arcs[ph].find((static_cast<uint64_t>(a) << 32) | b);
        return it != arcs[ph].end() && it->second == el;
    }
};

struct Item { uint32_t ph, q, step, a, b; };

struct Step {  // a plan step compiled against one query: positions and views
    uint32_t qi, n;
    std::vector<uint32_t> order;  // anchor endpoints first
    struct Back { uint32_t pos, slot, el, edge; };
    std::vector<std::vector<Back>> back;  // per position >= 2; back[k][0] generates candidates
    uint32_t step;                // index of the anchor edge in qe_list order
    uint32_t el0;                 // label rank of the anchor edge
    uint32_t l0, l1;              // raw vertex labels of u0, u1
};

struct Ctx {
    const Graph& g;  // G_t: the batch is applied only after all its items
    const Batch& bt;
    int ph;  // 0 insertion phase (on G_t \ D_t), 1 deletion phase (on G_t)
};

// Whether the view of query edge index j during step i drops the batch's deleted arcs from G_t:
// the insertion phase runs on G_t \ D_t, and in the deletion phase edges before the step's see
// the deletions applied.
inline bool drops_deleted(const Ctx& c, uint32_t j, uint32_t i) { return c.ph == 0 || j < i; }

// Is arc a -> b with label rank el in the view of query edge index j during step i?
inline bool in_view(const Ctx& c, uint32_t j, uint32_t i, uint32_t a, uint32_t b, uint32_t el) {
    bool in = c.g.has(a, c.g.key(b, el));
    if (in && drops_deleted(c, j, i) && c.bt.has(1, a, b, el)) in = false;
    if (!in && c.ph == 0 && j < i) in = c.bt.has(0, a, b, el);  // insertions seen by earlier edges
    return in;
}

uint64_t extend(const Ctx& c, const Step& st, std::vector<uint32_t>& f, uint32_t k) {
    const auto& bk = st.back[k];
    const Step::Back& gb = bk[0];
    const uint32_t p = f[gb.pos];
    const uint64_t lo_key = static_cast<uint64_t>(gb.slot) << 32, hi_key = lo_key + (1ull << 32);
    const auto& base = c.g.adj[p];
    auto s = std::lower_bound(base.begin(), base.end(), lo_key);
    auto e = std::lower_bound(s, base.end(), hi_key);
    const std::vector<uint64_t>* extra = nullptr;  // batch arcs of p that the view adds
    if (gb.edge < st.step && c.ph == 0) {
        auto it = c.bt.out[0].find(p);
        if (it != c.bt.out[0].end()) extra = &it->second;
    }
    const bool last = k + 1 == st.n;
    uint64_t total = 0;
    const bool drop = drops_deleted(c, gb.edge, st.step);
    auto try_x = [&](uint32_t x, bool from_base) {
        if (from_base && drop && c.bt.has(1, p, x, gb.el)) return;  // deleted in this view
        for (uint32_t t = 0; t < k; t++) if (f[t] == x) return;
        for (size_t t = 1; t < bk.size(); t++)
            if (!in_view(c, bk[t].edge, st.step, f[bk[t].pos], x, bk[t].el)) return;
        if (last) { total++; return; }
        f[k] = x;
        total += extend(c, st, f, k + 1);
    };
    for (auto it = s; it != e; ++it) try_x(static_cast<uint32_t>(*it), true);
    if (extra) {
        auto xs = std::lower_bound(extra->begin(), extra->end(), lo_key);
        for (; xs != extra->end() && *xs < hi_key; ++xs) try_x(static_cast<uint32_t>(*xs), false);
    }
    return total;
}

}  // namespace

int main(int argc, char** argv) {
    std::vector<std::string> pos;
    unsigned threads = std::max(1u, std::thread::hardware_concurrency());
    uint64_t max_batches = UINT64_MAX, time_limit_ms = 0, seed = 1;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--threads" && i + 1 < argc) threads = static_cast<unsigned>(std::strtoul(argv[++i], nullptr, 10));
        else if (a == "--max-batches" && i + 1 < argc) max_batches = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--time-limit-ms" && i + 1 < argc) time_limit_ms = std::strtoull(argv[++i], nullptr, 10);
        else if (a == "--seed" && i + 1 < argc) seed = std::strtoull(argv[++i], nullptr, 10);
        else pos.push_back(a);
    }
    if (pos.size() < 5 || (pos.size() - 3) % 2) die("usage: baseline <data> <update> <batch> (<query> <orders>)... [options]");
    const uint64_t batch = std::strtoull(pos[2].c_str(), nullptr, 10);
    if (!batch) die("batch size must be positive");
    threads = std::max(1u, threads);

    // ---- data graph ----
    Graph g;
    std::string s, w;
    std::vector<std::pair<uint64_t, uint32_t>> edges;  // (a << 32 | b, raw label)
    if (!slurp(pos[0].c_str(), s)) die("cannot read " + pos[0]);
    {
        Reader r{s.data(), s.data() + s.size()};
        while (r.p < r.e) {  // one type character per line, then numbers
            while (r.p < r.e && std::isspace(static_cast<unsigned char>(*r.p))) r.p++;
            if (r.p >= r.e) break;
            const char t = *r.p++;
            if (t == 't') { r.word(w); r.word(w); }
            else if (t == 'v') { const uint32_t id = r.num(), l = r.num(); if (id >= g.vlab.size()) g.vlab.resize(id + 1, UINT32_MAX); g.vlab[id] = l; }
            else if (t == 'e') { const uint32_t a = r.num(), b = r.num(), l = r.num(); edges.emplace_back((static_cast<uint64_t>(a) << 32) | b, l); }
            else die(std::string("unexpected line type in data graph: ") + t);
        }
    }
    g.nv = static_cast<uint32_t>(g.vlab.size());
    std::vector<std::tuple<uint32_t, uint32_t, uint32_t, uint8_t>> stream;  // a, b, raw label, op
    if (!slurp(pos[1].c_str(), s)) die("cannot read " + pos[1]);
    {
        Reader r{s.data(), s.data() + s.size()};
        while (r.word(w)) {
            if (w == "e" || w == "-e") { const uint32_t a = r.num(), b = r.num(), l = r.num(); stream.emplace_back(a, b, l, w == "-e" ? 1 : 0); }
            else die("unexpected token in update stream: " + w);
        }
    }
    auto rank_of = [&](uint32_t l) {
        auto it = g.erank.find(l);
        if (it != g.erank.end()) return it->second;
        const uint32_t k = static_cast<uint32_t>(g.erank.size());
        g.erank.emplace(l, k);
        return k;
    };
    for (auto& e : edges) rank_of(e.second);
    for (auto& e : stream) rank_of(std::get<2>(e));
    std::vector<Query> qs;
    std::vector<std::vector<PlanStep>> plans;
    for (size_t k = 3; k < pos.size(); k += 2) {
        qs.push_back(read_query(pos[k].c_str()));
        plans.push_back(read_orders(pos[k + 1].c_str()));
        for (auto& v : qs.back().adj) for (auto& x : v) rank_of(x.second);
    }
    g.nel = static_cast<uint32_t>(g.erank.size());
    g.adj.assign(g.nv, {});
    {
        std::vector<uint32_t> deg(g.nv, 0);
        for (auto& e : edges) { deg[e.first >> 32]++; deg[static_cast<uint32_t>(e.first)]++; }
        for (uint32_t v = 0; v < g.nv; v++) g.adj[v].reserve(deg[v]);
    }
    for (auto& e : edges) {
        const uint32_t a = static_cast<uint32_t>(e.first >> 32), b = static_cast<uint32_t>(e.first), el = g.erank[e.second];
        g.adj[a].push_back(g.key(b, el));
        g.adj[b].push_back(g.key(a, el));
    }
    {
        std::atomic<uint32_t> nextv{0};
        std::vector<std::thread> th;
        for (unsigned t = 0; t < threads; t++)
            th.emplace_back([&]() {
                for (uint32_t lo = nextv.fetch_add(4096); lo < g.nv; lo = nextv.fetch_add(4096))
                    for (uint32_t v = lo; v < std::min(g.nv, lo + 4096); v++) std::sort(g.adj[v].begin(), g.adj[v].end());
            });
        for (auto& x : th) x.join();
    }
    edges.clear();
    edges.shrink_to_fit();

    // ---- compiled steps ----
    std::vector<std::vector<Step>> steps(qs.size());
    for (size_t qi = 0; qi < qs.size(); qi++) {
        const Query& q = qs[qi];
        const auto& pl = plans[qi];
        std::vector<std::pair<uint32_t, uint32_t>> qe;  // qe_list order
        for (auto& p : pl) qe.emplace_back(p.u0, p.u1);
        auto edge_index = [&](uint32_t a, uint32_t b) {
            for (uint32_t j = 0; j < qe.size(); j++)
                if ((qe[j].first == a && qe[j].second == b) || (qe[j].first == b && qe[j].second == a)) return j;
            die("plan does not cover a query edge");
        };
        for (uint32_t i = 0; i < pl.size(); i++) {
            Step st;
            st.qi = static_cast<uint32_t>(qi);
            st.n = q.n;
            st.step = i;
            // anchor first, then repeatedly the earliest vertex of the planned order that has a
            // placed neighbour (the planned order itself when it starts at the anchor edge)
            st.order = {pl[i].u0, pl[i].u1};
            while (st.order.size() < q.n) {
                bool added = false;
                for (uint32_t v : pl[i].order) {
                    if (std::find(st.order.begin(), st.order.end(), v) != st.order.end()) continue;
                    bool linked = false;
                    for (auto& x : q.adj[v]) linked |= std::find(st.order.begin(), st.order.end(), x.first) != st.order.end();
                    if (!linked) continue;
                    st.order.push_back(v);
                    added = true;
                    break;
                }
                if (!added) die("plan order does not cover the query");
            }
            std::vector<uint32_t> at(q.n);
            for (uint32_t k = 0; k < q.n; k++) at[st.order[k]] = k;
            std::vector<std::vector<uint32_t>> planned(q.n);  // planned backward neighbours of each query vertex
            for (size_t k = 0; k < pl[i].order.size(); k++) planned[pl[i].order[k]] = pl[i].back[k];
            st.back.assign(q.n, {});
            for (uint32_t k = 2; k < q.n; k++) {
                const uint32_t u = st.order[k];
                std::vector<uint32_t> nb;
                for (uint32_t b : planned[u]) if (at[b] < k) nb.push_back(b);       // planned pivot first
                for (auto& x : q.adj[u]) if (at[x.first] < k && std::find(nb.begin(), nb.end(), x.first) == nb.end()) nb.push_back(x.first);
                if (nb.empty()) die("query vertex without a placed neighbour");
                for (uint32_t b : nb) {
                    const uint32_t el = g.erank[q.elabel(b, u)];
                    st.back[k].push_back(Step::Back{at[b], q.vlab[u] * g.nel + el, el, edge_index(b, u)});
                }
            }
            st.el0 = g.erank[q.elabel(pl[i].u0, pl[i].u1)];
            st.l0 = q.vlab[pl[i].u0];
            st.l1 = q.vlab[pl[i].u1];
            steps[qi].push_back(st);
        }
    }

    // ---- batches ----
    const uint64_t nb = (stream.size() + batch - 1) / batch;
    for (uint64_t b = 0; b < nb && b < max_batches; b++) {
        Batch bt;
        const size_t lo = b * batch, hi = std::min<size_t>(stream.size(), lo + batch);
        for (size_t k = lo; k < hi; k++) {
            const auto& [x, y, l, op] = stream[k];
            const uint32_t el = g.erank[l];
            bt.arcs[op][(static_cast<uint64_t>(x) << 32) | y] = el;
            bt.arcs[op][(static_cast<uint64_t>(y) << 32) | x] = el;
            bt.out[op][x].push_back(g.key(y, el));
            bt.out[op][y].push_back(g.key(x, el));
        }
        for (auto& m : bt.out) for (auto& kv : m) std::sort(kv.second.begin(), kv.second.end());
        uint64_t matches = 0, done = 0, total = 0;
        // every (phase, query, step, anchor) item of the batch in one seeded random order; all
        // items read G_t (the phases' views differ only in which batch arcs they add or drop)
        std::vector<Item> items;
        for (uint32_t ph : {1u, 0u}) {
            std::vector<std::pair<uint64_t, uint32_t>> arcs(bt.arcs[ph].begin(), bt.arcs[ph].end());
            std::sort(arcs.begin(), arcs.end());
            for (size_t qi = 0; qi < qs.size(); qi++)
                for (const Step& st : steps[qi])
                    for (const auto& kv : arcs) {
                        const uint32_t a = static_cast<uint32_t>(kv.first >> 32), c = static_cast<uint32_t>(kv.first);
                        if (kv.second == st.el0 && g.vlab[a] == st.l0 && g.vlab[c] == st.l1)
                            items.push_back(Item{ph, static_cast<uint32_t>(qi), st.step, a, c});
                    }
        }
        uint64_t r = seed * 0x9E3779B97F4A7C15ull + b;  // seeded shuffle (splitmix64)
        for (size_t k = items.size(); k > 1; k--) {
            r += 0x9E3779B97F4A7C15ull;
            uint64_t z = r;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            std::swap(items[k - 1], items[(z ^ (z >> 31)) % k]);
        }
        total = items.size();
        const auto t1 = std::chrono::steady_clock::now();
        std::atomic<size_t> next{0};
        std::atomic<uint64_t> found{0}, finished{0};
        std::vector<std::thread> th;
        for (unsigned t = 0; t < threads; t++)
            th.emplace_back([&]() {
                std::vector<uint32_t> f(32);
                uint64_t mine = 0, cnt = 0;
                for (size_t k = next++; k < items.size(); k = next++) {
                    if (time_limit_ms && std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count() > time_limit_ms) break;
                    const Item& it = items[k];
                    const Step& st = steps[it.q][it.step];
                    const Ctx c{g, bt, static_cast<int>(it.ph)};
                    f[0] = it.a;
                    f[1] = it.b;
                    mine += st.n == 2 ? 1 : extend(c, st, f, 2);
                    cnt++;
                }
                found += mine;
                finished += cnt;
            });
        for (auto& x : th) x.join();
        matches = found;
        done = finished;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        for (const auto& kv : bt.arcs[1]) {  // G_{t+1} = (G_t \ D_t) + I_t
            const uint32_t a = static_cast<uint32_t>(kv.first >> 32), y = static_cast<uint32_t>(kv.first);
            auto& l = g.adj[a];
            auto it = std::lower_bound(l.begin(), l.end(), g.key(y, kv.second));
            if (it != l.end() && *it == g.key(y, kv.second)) l.erase(it);
        }
        for (const auto& kv : bt.arcs[0]) {
            const uint32_t a = static_cast<uint32_t>(kv.first >> 32), y = static_cast<uint32_t>(kv.first);
            auto& l = g.adj[a];
            l.insert(std::lower_bound(l.begin(), l.end(), g.key(y, kv.second)), g.key(y, kv.second));
        }
        std::printf("batch %llu items %llu of %llu matches %llu ms %.3f\n", (unsigned long long)b, (unsigned long long)done,
                    (unsigned long long)total, (unsigned long long)matches, ms);
        std::fflush(stdout);
    }
    return 0;
}
