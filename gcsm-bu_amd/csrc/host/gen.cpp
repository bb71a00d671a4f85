#include "gen.h"
#include "gen_count.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <atomic>
#include <chrono>
#include <dirent.h>
#include <functional>
#include <thread>
#include <unordered_set>
#include <tuple>
#include <vector>

namespace csm {
namespace {

// xoshiro256** seeded through splitmix64 (public-domain generators by Blackman and Vigna).
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (int i = 0; i < 4; i++) {
            seed += 0x9E3779B97F4A7C15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            s[i] = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    uint64_t below(uint64_t n) { return n ? next() % n : 0; }  // modulo bias is irrelevant here
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

struct Alias {  // Walker's alias method over non-negative weights
    std::vector<double> prob;
    std::vector<uint32_t> alias;
    explicit Alias(const std::vector<double>& w) {
        const size_t n = w.size();
        prob.resize(n);
        alias.resize(n);
        double sum = 0;
        for (double x : w) sum += x;
        std::vector<double> p(n);
        std::vector<uint32_t> small, large;
        for (size_t i = 0; i < n; i++) {
            p[i] = w[i] * n / sum;
            (p[i] < 1.0 ? small : large).push_back(static_cast<uint32_t>(i));
        }
        while (!small.empty() && !large.empty()) {
            uint32_t s = small.back(), l = large.back();
            small.pop_back();
            prob[s] = p[s];
            alias[s] = l;
            p[l] = (p[l] + p[s]) - 1.0;
            if (p[l] < 1.0) { large.pop_back(); small.push_back(l); }
        }
        for (uint32_t i : large) { prob[i] = 1.0; alias[i] = i; }
        for (uint32_t i : small) { prob[i] = 1.0; alias[i] = i; }
    }
    uint32_t sample(Rng& r) const {
        const uint32_t i = static_cast<uint32_t>(r.below(prob.size()));
        return r.uniform() < prob[i] ? i : alias[i];
    }
};

struct Edge { uint32_t a, b, l; };

// Chung-Lu: expected degree of vertex i proportional to (i + 1)^(-1/(gamma-1)), capped at dmax,
// rescaled so that expected degrees sum to 2m; endpoints are drawn independently by weight and
// self-loops / duplicates are discarded.  With hubs > 0, the first `hubs` vertices instead get the
// expected degree hub_deg and the power law covers the remaining 2m - hubs * hub_deg.  Vertex ids are
// shuffled so hubs are spread out.
std::vector<uint64_t> chung_lu(uint32_t n, uint64_t m, double gamma, double dmax, Rng& rng, uint32_t hubs = 0, double hub_deg = 0) {
    std::vector<double> w(n);
    const double ex = 1.0 / (gamma - 1.0);
    for (uint32_t i = hubs; i < n; i++) w[i] = std::pow(static_cast<double>(i - hubs) + 1.0, -ex);
    const double bulk = 2.0 * m - hubs * hub_deg;
    for (int it = 0; it < 60; it++) {
        double sum = 0;
        for (uint32_t i = hubs; i < n; i++) sum += w[i];
        const double scale = bulk / sum;
        bool capped = false;
        for (uint32_t i = hubs; i < n; i++) {
            double& x = w[i];
            x *= scale;
            if (x > dmax) { x = dmax; capped = true; }
        }
        if (!capped) break;
    }
    for (uint32_t i = 0; i < hubs; i++) w[i] = hub_deg;
    std::vector<uint32_t> perm(n);
    for (uint32_t i = 0; i < n; i++) perm[i] = i;
    for (uint32_t i = n; i > 1; i--) std::swap(perm[i - 1], perm[rng.below(i)]);
    Alias al(w);
    std::vector<uint64_t> keys;
    keys.reserve(m + m / 8);
    while (keys.size() < m) {
        const uint64_t need = (m - keys.size()) + (m - keys.size()) / 16 + 16;
        for (uint64_t k = 0; k < need; k++) {
            uint32_t a = perm[al.sample(rng)], b = perm[al.sample(rng)];
            if (a == b) continue;
            if (a > b) std::swap(a, b);
            keys.push_back((static_cast<uint64_t>(a) << 32) | b);
        }
        std::sort(keys.begin(), keys.end());
        keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    }
    // Random subset of exactly m edges, in random order.
    for (size_t i = keys.size(); i > 1; i--) std::swap(keys[i - 1], keys[rng.below(i)]);
    keys.resize(m);
    return keys;
}

// Directed arcs a -> b: out-degree of every vertex lognormal(mu, sigma) (rounded, capped), targets
// uniform, no self-loops or duplicate arcs; returned in random order.
std::vector<uint64_t> lognormal_arcs(uint32_t n, double mu, double sigma, Rng& rng) {
    std::vector<uint64_t> keys;
    for (uint32_t a = 0; a < n; a++) {
        const double x = std::exp(mu + sigma * rng.normal());
        const uint64_t d = std::min<uint64_t>(std::min<uint64_t>(static_cast<uint64_t>(x + 0.5), n - 1), 100000);
        for (uint64_t k = 0; k < d; k++) {
            const uint32_t b = static_cast<uint32_t>(rng.below(n));
            if (b != a) keys.push_back((static_cast<uint64_t>(a) << 32) | b);
        }
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (size_t i = keys.size(); i > 1; i--) std::swap(keys[i - 1], keys[rng.below(i)]);
    return keys;
}

std::vector<uint32_t> labels(uint32_t n, uint32_t k, bool zipf, Rng& rng) {
    std::vector<uint32_t> lab(n);
    if (!zipf) {
        for (auto& x : lab) x = static_cast<uint32_t>(rng.below(k));
        return lab;
    }
    std::vector<double> cdf(k);
    double s = 0;
    for (uint32_t i = 0; i < k; i++) { s += 1.0 / (i + 1.0); cdf[i] = s; }
    for (auto& x : lab) {
        const double u = rng.uniform() * s;
        x = static_cast<uint32_t>(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        if (x >= k) x = k - 1;
    }
    return lab;
}

struct Csr {  // undirected adjacency of the final graph, sorted lists (used for query extraction)
    std::vector<uint64_t> off;
    std::vector<uint32_t> nbr;
    std::vector<uint32_t> lab;  // edge label aligned with nbr
    std::vector<uint8_t> out;   // 1 when the edge is the arc v -> nbr (directed graphs), aligned with nbr
    bool directed = false;
    void build(uint32_t n, const std::vector<Edge>& es) {
        off.assign(static_cast<size_t>(n) + 1, 0);
        for (const auto& e : es) { off[e.a + 1]++; off[e.b + 1]++; }
        for (uint32_t i = 0; i < n; i++) off[i + 1] += off[i];
        nbr.resize(off[n]);
        lab.resize(off[n]);
        out.resize(off[n]);
        std::vector<uint64_t> pos(off.begin(), off.end() - 1);
        for (const auto& e : es) {
            nbr[pos[e.a]] = e.b; out[pos[e.a]] = 1; lab[pos[e.a]++] = e.l;
            nbr[pos[e.b]] = e.a; out[pos[e.b]] = 0; lab[pos[e.b]++] = e.l;
        }
        for (uint32_t v = 0; v < n; v++) {
            std::vector<std::tuple<uint32_t, uint32_t, uint8_t>> t;
            for (uint64_t k = off[v]; k < off[v + 1]; k++) t.emplace_back(nbr[k], lab[k], out[k]);
            std::sort(t.begin(), t.end());
            for (uint64_t k = off[v]; k < off[v + 1]; k++) {
                nbr[k] = std::get<0>(t[k - off[v]]);
                lab[k] = std::get<1>(t[k - off[v]]);
                out[k] = std::get<2>(t[k - off[v]]);
            }
        }
    }
    uint64_t deg(uint32_t v) const { return off[v + 1] - off[v]; }
    int64_t find(uint32_t a, uint32_t b) const {
        auto lo = nbr.begin() + off[a], hi = nbr.begin() + off[a + 1];
        auto it = std::lower_bound(lo, hi, b);
        return (it != hi && *it == b) ? (it - nbr.begin()) : -1;
    }
};

enum Shape { kTree, kSparse, kDense, kCycles, kClosure, kC1 };

struct QueryOut {
    std::vector<uint32_t> verts;  // data vertices, query vertex id = position
    std::vector<Edge> edges;      // query vertex ids
    std::vector<Edge> spare;      // kCycles: the walk's other chords, in the order they may be added
};

// Random walk with restarts from visited vertices (so trees branch); returns the vertices and the
// first-visit (tree) edges, or false.
bool walk(const Csr& g, uint32_t n_target, const std::vector<uint64_t>& edge_keys, Rng& rng, bool uniform_start,
          std::vector<uint32_t>& vs, std::vector<std::pair<uint32_t, uint32_t>>& tree) {
    const uint64_t key = edge_keys[rng.below(edge_keys.size())];
    uint32_t cur = static_cast<uint32_t>(key >> 32);
    if (uniform_start) {  // a uniformly chosen vertex keeps walks out of the hub core
        for (int t = 0; t < 64; t++) {
            const uint32_t v = static_cast<uint32_t>(rng.below(g.off.size() - 1));
            if (g.deg(v) >= 1) { cur = v; break; }
        }
    }
    vs.assign(1, cur);
    tree.clear();
    for (uint32_t step = 0; step < 64 * n_target && vs.size() < n_target; step++) {
        if (g.deg(cur) == 0) return false;
        const uint32_t nx = g.nbr[g.off[cur] + rng.below(g.deg(cur))];
        auto it = std::find(vs.begin(), vs.end(), nx);
        if (it == vs.end()) {
            const uint32_t pi = static_cast<uint32_t>(std::find(vs.begin(), vs.end(), cur) - vs.begin());
            vs.push_back(nx);
            tree.emplace_back(pi, static_cast<uint32_t>(vs.size() - 1));
            cur = nx;
        } else {
            cur = nx;
        }
        if (rng.uniform() < 0.35) cur = vs[rng.below(vs.size())];
    }
    return vs.size() == n_target;
}

bool extract(const Csr& g, const std::vector<uint32_t>& vlab, uint32_t nq, Shape shape,
             const std::vector<uint64_t>& edge_keys, Rng& rng, QueryOut& q) {
    for (int attempt = 0; attempt < 20000; attempt++) {
        std::vector<uint32_t> vs;
        std::vector<std::pair<uint32_t, uint32_t>> tree;
        if (shape == kC1) {
            // 4-cycle a-b-c-d with chord a-c, plus a fifth vertex hanging off the cycle.
            const uint64_t key = edge_keys[rng.below(edge_keys.size())];
            const uint32_t a = static_cast<uint32_t>(key >> 32), c = static_cast<uint32_t>(key);
            std::vector<uint32_t> common;
            for (uint64_t k = g.off[a]; k < g.off[a + 1]; k++)
                if (g.nbr[k] != c && g.find(c, g.nbr[k]) >= 0) common.push_back(g.nbr[k]);
            if (common.size() < 2) continue;
            const uint32_t b = common[rng.below(common.size())];
            uint32_t d = b;
            for (int t = 0; t < 16 && d == b; t++) d = common[rng.below(common.size())];
            if (d == b) continue;
            vs = {a, b, c, d};
            const uint32_t hang = static_cast<uint32_t>(rng.below(4));
            const uint32_t h = vs[hang];
            if (g.deg(h) < 2) continue;
            uint32_t e = vs[0];
            for (int t = 0; t < 32; t++) {
                const uint32_t x = g.nbr[g.off[h] + rng.below(g.deg(h))];
                if (std::find(vs.begin(), vs.end(), x) == vs.end()) { e = x; break; }
            }
            if (std::find(vs.begin(), vs.end(), e) != vs.end()) continue;
            vs.push_back(e);
            q.verts = vs;
            q.edges.clear();
            const uint32_t pairs[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 2}, {hang, 4}};
            for (auto& p : pairs) {
                const int64_t pos = g.find(vs[p[0]], vs[p[1]]);
                q.edges.push_back(Edge{p[0], p[1], g.lab[pos]});
            }
            return true;
        }
        if (shape == kClosure) {
            // From a random edge, each further vertex but the last 2-3 is a common neighbour of two
            // query vertices and joins with both edges (a tree edge and a cycle edge); the last
            // ones join as tree leaves of random query vertices.
            const uint64_t key = edge_keys[rng.below(edge_keys.size())];
            vs = {static_cast<uint32_t>(key >> 32), static_cast<uint32_t>(key)};
            std::vector<std::pair<uint32_t, uint32_t>> es{{0, 1}};
            const uint32_t leaves = 2 + static_cast<uint32_t>(rng.below(2));
            bool ok = nq >= leaves + 3;
            while (vs.size() < nq - leaves && ok) {
                std::vector<std::pair<uint32_t, uint32_t>> hits;  // (data vertex, query position)
                for (uint32_t i = 0; i < vs.size(); i++)
                    for (uint64_t k = g.off[vs[i]]; k < g.off[vs[i] + 1]; k++)
                        if (std::find(vs.begin(), vs.end(), g.nbr[k]) == vs.end()) hits.emplace_back(g.nbr[k], i);
                std::sort(hits.begin(), hits.end());
                std::vector<size_t> starts;  // runs of one data vertex adjacent to >= 2 query vertices
                for (size_t k = 0; k < hits.size();) {
                    size_t e = k;
                    while (e < hits.size() && hits[e].first == hits[k].first) e++;
                    if (e - k >= 2) starts.push_back(k);
                    k = e;
                }
                if (starts.empty()) { ok = false; break; }
                const size_t k = starts[rng.below(starts.size())];
                size_t e = k;
                while (e < hits.size() && hits[e].first == hits[k].first) e++;
                const size_t a = k + rng.below(e - k);
                size_t b = k + rng.below(e - k - 1);
                if (b >= a) b++;
                const uint32_t p = static_cast<uint32_t>(vs.size());
                vs.push_back(hits[k].first);
                es.emplace_back(hits[a].second, p);
                es.emplace_back(hits[b].second, p);
            }
            for (int tries = 0; vs.size() < nq && ok && tries < 256; tries++) {
                const uint32_t i = static_cast<uint32_t>(rng.below(vs.size()));
                const uint64_t d = g.deg(vs[i]);
                uint32_t x = vs[0];
                for (int t = 0; t < 16 && d; t++) {
                    x = g.nbr[g.off[vs[i]] + rng.below(d)];
                    if (std::find(vs.begin(), vs.end(), x) == vs.end()) break;
                }
                if (std::find(vs.begin(), vs.end(), x) != vs.end()) continue;
                es.emplace_back(i, static_cast<uint32_t>(vs.size()));
                vs.push_back(x);
            }
            if (!ok || vs.size() != nq) continue;
            q.verts = vs;
            q.edges.clear();
            q.spare.clear();
            for (const auto& t : es) {
                const int64_t pos = g.find(vs[t.first], vs[t.second]);
                q.edges.push_back(Edge{t.first, t.second, g.lab[pos]});
            }
            return true;
        }
        if (!walk(g, nq, edge_keys, rng, shape != kDense && shape != kCycles, vs, tree)) continue;
        std::vector<std::pair<uint32_t, uint32_t>> extra;
        for (uint32_t i = 0; i < nq; i++)
            for (uint32_t j = i + 1; j < nq; j++) {
                if (g.find(vs[i], vs[j]) < 0) continue;
                bool in_tree = false;
                for (auto& t : tree)
                    if ((t.first == i && t.second == j) || (t.first == j && t.second == i)) in_tree = true;
                if (!in_tree) extra.emplace_back(i, j);
            }
        for (size_t i = extra.size(); i > 1; i--) std::swap(extra[i - 1], extra[rng.below(i)]);
        size_t take = 0;
        const size_t half3 = (3 * nq + 1) / 2;  // |E| >= 1.5 n  <=>  average degree >= 3
        if (shape == kTree) take = 0;
        else if (shape == kSparse) {
            const size_t max_extra = half3 - 1 - (nq - 1);  // keep |E| < 1.5 n
            if (extra.empty() || max_extra == 0) continue;
            take = 1 + rng.below(std::min(extra.size(), max_extra));
        } else if (shape == kDense) {
            if (nq - 1 + extra.size() < half3) continue;
            take = extra.size();
        } else if (shape == kCycles) {
            const size_t want = 1 + rng.below(2);
            if (extra.size() < want) continue;
            take = want;
        }
        q.verts = vs;
        q.edges.clear();
        // a query edge takes the label and (for directed graphs) the direction of its data edge
        auto add = [&](uint32_t i, uint32_t j) {
            const int64_t p = g.find(vs[i], vs[j]);
            if (!g.directed || g.out[p]) q.edges.push_back(Edge{i, j, g.lab[p]});
            else q.edges.push_back(Edge{j, i, g.lab[p]});
        };
        for (auto& t : tree) add(t.first, t.second);
        for (size_t k = 0; k < take; k++) add(extra[k].first, extra[k].second);
        q.spare.clear();
        if (shape == kCycles) {
            const size_t base = q.edges.size();
            for (size_t k = take; k < extra.size(); k++) add(extra[k].first, extra[k].second);
            q.spare.assign(q.edges.begin() + base, q.edges.end());
            q.edges.resize(base);
        }
        (void)vlab;
        return true;
    }
    return false;
}

// Estimated number of embeddings of q in g (injective, label and edge-label preserving): exact
// completion counts from `samples` uniformly drawn data arcs standing for query edge 0, scaled by the
// number of arcs.  Returns false when one sample needs more than node_budget search nodes.
bool estimate_embeddings(const Csr& g, const std::vector<uint32_t>& vlab, const QueryOut& q, Rng& rng, uint32_t samples,
                         uint64_t node_budget, double& est) {
    const uint32_t n = static_cast<uint32_t>(q.verts.size());
    std::vector<uint32_t> qlab(n);
    for (uint32_t i = 0; i < n; i++) qlab[i] = vlab[q.verts[i]];
    // BFS order from query edge 0; back[k] = earlier neighbours (position, edge label)
    std::vector<uint32_t> order{q.edges[0].a, q.edges[0].b}, pos(n, UINT32_MAX);
    pos[order[0]] = 0;
    pos[order[1]] = 1;
    for (size_t k = 0; k < order.size(); k++)
        for (const auto& e : q.edges) {
            uint32_t x = UINT32_MAX;
            if (e.a == order[k]) x = e.b; else if (e.b == order[k]) x = e.a;
            if (x != UINT32_MAX && pos[x] == UINT32_MAX) { pos[x] = static_cast<uint32_t>(order.size()); order.push_back(x); }
        }
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> back(n);
    for (uint32_t k = 1; k < n; k++)
        for (const auto& e : q.edges) {
            const uint32_t u = order[k];
            if (e.a == u && pos[e.b] < k) back[k].emplace_back(pos[e.b], e.l);
            else if (e.b == u && pos[e.a] < k) back[k].emplace_back(pos[e.a], e.l);
        }
    const uint64_t arcs = g.nbr.size();
    std::vector<uint32_t> f(n);
    uint64_t nodes = 0;
    bool over = false;
    double total = 0;
    std::function<double(uint32_t)> rec = [&](uint32_t k) -> double {
        if (k == n) return 1.0;
        if (++nodes > node_budget) { over = true; return 0.0; }
        const auto& b0 = back[k][0];
        const uint32_t p = f[b0.first];
        double c = 0;
        for (uint64_t i = g.off[p]; i < g.off[p + 1] && !over; i++) {
            const uint32_t x = g.nbr[i];
            if (g.lab[i] != b0.second || vlab[x] != qlab[order[k]]) continue;
            bool ok = true;
            for (uint32_t j = 0; j < k && ok; j++) ok = f[j] != x;
            for (size_t t = 1; t < back[k].size() && ok; t++) {
                const int64_t e = g.find(f[back[k][t].first], x);
                ok = e >= 0 && g.lab[e] == back[k][t].second;
            }
            if (!ok) continue;
            f[k] = x;
            c += rec(k + 1);
        }
        return c;
    };
    uint64_t spent = 0;
    const uint64_t total_budget = 25 * node_budget;
    auto completions = [&](uint64_t a, double& c) {
        const uint32_t x = g.nbr[a];
        const uint32_t v = static_cast<uint32_t>(std::upper_bound(g.off.begin(), g.off.end(), a) - g.off.begin() - 1);
        c = 0;
        if (vlab[v] != qlab[order[0]] || vlab[x] != qlab[order[1]] || g.lab[a] != q.edges[0].l) return true;
        f[0] = v;
        f[1] = x;
        nodes = 0;
        c = rec(2);
        spent += nodes;
        return !over && spent <= total_budget;
    };
    // Stratified by log2 of the source degree: every stratum gets the same number of samples, so
    // the hub-anchored arcs that dominate heavy queries are well represented; each stratum's mean is
    // weighted by its number of arcs.
    const uint32_t nv = static_cast<uint32_t>(g.off.size() - 1);
    std::vector<std::vector<uint32_t>> strata(40);
    for (uint32_t v = 0; v < nv; v++) {
        const uint64_t d = g.deg(v);
        if (!d || vlab[v] != qlab[order[0]]) continue;
        uint32_t b = 0;
        while ((2ull << b) <= d) b++;
        strata[b].push_back(v);
    }
    for (const auto& st : strata) {
        if (st.empty()) continue;
        uint64_t st_arcs = 0;
        for (uint32_t v : st) st_arcs += g.deg(v);
        const uint32_t per = std::max<uint32_t>(8, samples / 12);
        double sum = 0;
        for (uint32_t k = 0; k < per; k++) {
            // uniform arc of the stratum: vertex by degree via rejection against the stratum's max
            uint32_t v = st[rng.below(st.size())];
            const uint64_t dmax = 2ull * g.deg(st[0]) + 2;
            for (int t = 0; t < 64 && rng.below(dmax) >= g.deg(v); t++) v = st[rng.below(st.size())];
            double c;
            if (!completions(g.off[v] + rng.below(g.deg(v)), c)) return false;
            sum += c;
        }
        total += sum / per * static_cast<double>(st_arcs);
    }
    (void)arcs;
    est = total;
    return true;
}

bool write_query(const std::string& path, const QueryOut& q, const std::vector<uint32_t>& vlab) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fprintf(f, "t # 0\n");
    for (size_t i = 0; i < q.verts.size(); i++) std::fprintf(f, "v %zu %u\n", i, vlab[q.verts[i]]);
    for (const auto& e : q.edges) std::fprintf(f, "e %u %u %u\n", e.a, e.b, e.l);
    std::fclose(f);
    return true;
}

struct Spec {
    const char* name;
    uint32_t n;
    uint64_t m;
    double gamma;
    double dmax;
    uint32_t nlabels;
    bool zipf;
    double holdout;     // fraction of edges held out as insertions
    double delete_frac; // fraction of the stream that are deletions
    uint32_t batch;
    std::vector<std::pair<Shape, uint32_t>> queries;  // (shape, vertices)
    double emb_cap = 0;  // reject queries with more estimated embeddings in the final graph (0: none)
    uint32_t hubs = 0;   // planted hub vertices of expected degree hub_deg (dmax then caps the rest)
    double hub_deg = 0;
    // exact selection: keep a query only if its exact number of embeddings in the union of all
    // graphs the stream passes through is at most emb_cap, found within work_budget search steps
    bool exact = false;
    double work_budget = 0;
    // directed configurations: arcs a -> b, out-degree lognormal(mu, sigma), uniform edge labels, and
    // a batch-size file of lognormal sizes clipped to [batch_min, batch_max]
    bool directed = false;
    uint32_t nelabels = 1;
    double mu = 0, sigma = 0;
    uint32_t batch_min = 0, batch_max = 0;
};

Spec spec_of(const std::string& c) {
    Spec s{};
    if (c == "tiny") {
        s = Spec{"tiny", 300, 1500, 2.3, 60, 3, false, 0.2, 0.1, 100, {}};
        s.queries = {{kTree, 4}, {kSparse, 5}, {kDense, 5}};
    } else if (c == "tinyins") {
        s = Spec{"tinyins", 200, 900, 2.3, 40, 3, false, 0.2, 0.0, 100, {}};
        s.queries = {{kTree, 4}, {kSparse, 5}, {kDense, 5}};
    } else if (c == "c1") {
        s = Spec{"c1", 10000, 50000, 2.3, 1000, 8, false, 0.10, 0.0, 0xFFFFFFFFu, {}};
        s.queries = {{kC1, 5}};
    } else if (c == "c2" || c == "c4") {
        s = Spec{c == "c2" ? "c2" : "c4", 400000, 3400000, 2.1, 3000, 32, true, 0.10, 0.10, 65536, {}};
        s.emb_cap = 2e8;
        if (c == "c2") {
            s.queries = {{kSparse, 6}, {kSparse, 7}, {kSparse, 8}, {kDense, 6}, {kDense, 7}, {kDense, 8}};
        } else {
            s.exact = true;
            s.emb_cap = 2e9;
            s.work_budget = 4e8;
            // Trees of 9 or more vertices and sparse cyclic queries of 10 or more exceed the cap on
            // this graph (none of 4,096 candidates each), so the sizes stop at 8 and 9.
            for (int i = 0; i < 16; i++) s.queries.push_back({kTree, 4 + (i % 5)});
            for (int i = 0; i < 24; i++) s.queries.push_back({kSparse, 4 + (i % 6)});
            for (int i = 0; i < 24; i++) s.queries.push_back({kDense, 4 + (i % 6)});
        }
    } else if (c == "c5") {
        s = Spec{"c5", 3000000, 0, 0, 0, 4, false, 0.10, 0.15, 16384, {}};
        s.directed = true;
        s.nelabels = 8;
        s.mu = 1.5;
        s.sigma = 1.2;
        s.batch_min = 1024;
        s.batch_max = 262144;
        s.emb_cap = 2e8;
        // uniform random targets leave few cycles: sparse shapes fall back to trees when none is found
        for (int i = 0; i < 32; i++) s.queries.push_back({i % 2 == 0 ? kTree : kSparse, 4 + (i % 5)});
    } else if (c == "c3") {
        s = Spec{"c3", 4000000, 35000000, 2.0, 2000, 16, false, 0.05, 0.05, 1048576, {}};
        s.hubs = 8;
        s.hub_deg = 20000;
        // Trees with one or two added cycle edges have 10^12 or more embeddings on this graph (exact
        // counts), beyond any enumerator; the queries keep 8-10 vertices and 2-3 tree leaves, and
        // every other vertex closes a cycle.
        s.queries = {{kClosure, 8}, {kClosure, 9}, {kClosure, 10}, {kClosure, 8}, {kClosure, 9}, {kClosure, 10}};
        s.exact = true;
        s.emb_cap = 1e9;
        s.work_budget = 5e8;
    }
    return s;
}

}  // namespace

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

unsigned worker_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

// Appends the decimal digits of x.
inline void put_u32(std::string& out, uint32_t x) {
    char buf[10];
    int n = 0;
    do { buf[n++] = static_cast<char>('0' + x % 10); x /= 10; } while (x);
    while (n) out.push_back(buf[--n]);
}

// Writes `count` text lines produced by line(k, out) in order; chunks of lines are formatted in
// parallel and written one round of chunks at a time.
bool write_lines(const std::string& path, const char* head, size_t count, const std::function<void(size_t, std::string&)>& line) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    if (head) std::fputs(head, f);
    const unsigned nth = worker_threads();
    constexpr size_t kChunk = 1 << 18;
    std::vector<std::string> buf(nth);
    for (size_t base = 0; base < count; base += nth * kChunk) {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nth; t++)
            th.emplace_back([&, t]() {
                std::string& b = buf[t];
                b.clear();
                const size_t lo = base + t * kChunk, hi = std::min(count, lo + kChunk);
                for (size_t k = lo; k < hi; k++) line(k, b);
            });
        for (auto& x : th) x.join();
        for (auto& b : buf) std::fwrite(b.data(), 1, b.size(), f);
    }
    return std::fclose(f) == 0;
}

}  // namespace

int gen_main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: gen <tiny|tinyins|c1|c2|c3|c4|c5> <out_dir> [--seed S] [--scale F] [--queries-from DIR]\n");
        return 2;
    }
    const std::string cfg = argv[1];
    const std::string out = argv[2];
    uint64_t seed = 1;
    double scale = 1.0;
    std::string queries_from;
    for (int i = 3; i + 1 < argc; i += 2) {
        if (!std::strcmp(argv[i], "--seed")) seed = std::strtoull(argv[i + 1], nullptr, 10);
        else if (!std::strcmp(argv[i], "--scale")) scale = std::atof(argv[i + 1]);
        else if (!std::strcmp(argv[i], "--queries-from")) queries_from = argv[i + 1];
    }
    Spec sp = spec_of(cfg);
    if (!sp.name) { std::fprintf(stderr, "unknown config %s\n", cfg.c_str()); return 2; }
    if (scale != 1.0) {
        sp.n = std::max<uint32_t>(64, static_cast<uint32_t>(sp.n * scale));
        sp.m = std::max<uint64_t>(256, static_cast<uint64_t>(sp.m * scale));
        sp.dmax = std::max(16.0, sp.dmax * std::sqrt(scale));
        sp.hub_deg *= std::sqrt(scale);
        if (sp.batch != 0xFFFFFFFFu) sp.batch = std::max<uint32_t>(64, static_cast<uint32_t>(sp.batch * scale));
        if (sp.batch_max) {
            sp.batch_min = std::max<uint32_t>(16, static_cast<uint32_t>(sp.batch_min * scale));
            sp.batch_max = std::max<uint32_t>(sp.batch_min, static_cast<uint32_t>(sp.batch_max * scale));
        }
        sp.emb_cap *= sp.exact ? scale : std::sqrt(scale);  // scaled-down graphs get lighter queries
        sp.work_budget *= std::max(scale, 0.1);
    }
    mkdir(out.c_str(), 0755);
    const double t_start = now_s();
    auto progress = [&](const char* what) { std::fprintf(stderr, "gen %s: %s (%.1fs)\n", sp.name, what, now_s() - t_start); };
    Rng rng(seed * 0x100000001B3ull + 7);

    std::vector<uint64_t> keys = sp.directed ? lognormal_arcs(sp.n, sp.mu, sp.sigma, rng) : chung_lu(sp.n, sp.m, sp.gamma, sp.dmax, rng, sp.hubs, sp.hub_deg);
    std::vector<uint32_t> vlab = labels(sp.n, sp.nlabels, sp.zipf, rng);
    std::vector<uint32_t> elab(keys.size(), 0);  // edge label of keys[k]
    if (sp.nelabels > 1)
        for (auto& x : elab) x = static_cast<uint32_t>(rng.below(sp.nelabels));
    progress("edges sampled");
    const size_t n_hold = static_cast<size_t>(keys.size() * sp.holdout);
    const size_t n_init = keys.size() - n_hold;
    // Stream: all held-out edges are inserted; deletions are distinct initial edges, so every
    // deletion removes an existing edge and nothing is reinserted.
    const size_t n_del = sp.delete_frac > 0 ? static_cast<size_t>(n_hold * sp.delete_frac / (1.0 - sp.delete_frac)) : 0;
    std::vector<std::pair<size_t, uint8_t>> stream;  // (index into keys, 1 = deletion)
    stream.reserve(n_hold + n_del);
    for (size_t k = n_init; k < keys.size(); k++) stream.emplace_back(k, 0);
    {
        std::vector<size_t> pick(n_init);
        for (size_t k = 0; k < n_init; k++) pick[k] = k;
        for (size_t k = 0; k < n_del && k < n_init; k++) {
            const size_t j = k + rng.below(n_init - k);
            std::swap(pick[k], pick[j]);
            stream.emplace_back(pick[k], 1);
        }
    }
    for (size_t i = stream.size(); i > 1; i--) std::swap(stream[i - 1], stream[rng.below(i)]);

    // data graph file
    {
        const bool ok = write_lines(out + "/data.graph", "t # 0\n", sp.n + n_init, [&](size_t k, std::string& o) {
            if (k < sp.n) {
                o += "v "; put_u32(o, static_cast<uint32_t>(k)); o += ' '; put_u32(o, vlab[k]); o += '\n';
            } else {
                const uint64_t e = keys[k - sp.n];
                o += "e "; put_u32(o, static_cast<uint32_t>(e >> 32)); o += ' '; put_u32(o, static_cast<uint32_t>(e)); o += ' ';
                put_u32(o, elab[k - sp.n]); o += '\n';
            }
        });
        if (!ok) { std::fprintf(stderr, "cannot write %s\n", out.c_str()); return 1; }
    }
    {
        const bool ok = write_lines(out + "/update.stream", nullptr, stream.size(), [&](size_t k, std::string& o) {
            const auto& e = stream[k];
            if (e.second) o += '-';
            o += "e "; put_u32(o, static_cast<uint32_t>(keys[e.first] >> 32)); o += ' '; put_u32(o, static_cast<uint32_t>(keys[e.first]));
            o += ' '; put_u32(o, elab[e.first]); o += '\n';
        });
        if (!ok) { std::fprintf(stderr, "cannot write %s\n", out.c_str()); return 1; }
    }
    progress("graph and stream written");
    std::vector<uint64_t> qcounts, qtried;
    uint32_t qk = 0;
    if (!queries_from.empty()) {
        // queries selected earlier for this configuration, seed and scale
        for (; qk < sp.queries.size(); qk++) {
            const std::string src = queries_from + "/q" + std::to_string(qk) + ".graph";
            FILE* in = std::fopen(src.c_str(), "r");
            if (!in) { std::fprintf(stderr, "missing %s\n", src.c_str()); return 1; }
            FILE* o = std::fopen((out + "/q" + std::to_string(qk) + ".graph").c_str(), "w");
            char b[4096];
            for (size_t r; (r = std::fread(b, 1, sizeof b, in)) > 0;) std::fwrite(b, 1, r, o);
            std::fclose(in);
            std::fclose(o);
        }
    } else {
        // queries are extracted from the final graph (initial minus deletions plus insertions)
        std::vector<Edge> fin;
        {
            std::unordered_set<uint64_t> dels;
            for (const auto& e : stream) if (e.second) dels.insert(keys[e.first]);
            fin.reserve(keys.size());
            for (size_t k = 0; k < keys.size(); k++)
                if (k >= n_init || !dels.count(keys[k])) fin.push_back(Edge{static_cast<uint32_t>(keys[k] >> 32), static_cast<uint32_t>(keys[k]), elab[k]});
        }
        Csr g;
        g.build(sp.n, fin);
        g.directed = sp.directed;
        std::vector<uint64_t> fin_keys;
        fin_keys.reserve(fin.size());
        for (const auto& e : fin) fin_keys.push_back((static_cast<uint64_t>(e.a) << 32) | e.b);
        CountGraph cg;
        if (sp.exact) {  // every graph the stream passes through is a subgraph of initial + all insertions
            std::vector<CountEdge> all(keys.size());
            for (size_t k = 0; k < keys.size(); k++) all[k] = CountEdge{static_cast<uint32_t>(keys[k] >> 32), static_cast<uint32_t>(keys[k]), elab[k]};
            cg.build(sp.n, vlab, std::max(1u, sp.nelabels), sp.directed, all, worker_threads());
        }
        progress("query search structures built");
        double last_note = now_s();
        const bool verbose = std::getenv("CSM_GEN_VERBOSE") != nullptr;
        if (const char* x = std::getenv("CSM_GEN_CAP")) sp.emb_cap = std::atof(x);
        if (const char* x = std::getenv("CSM_GEN_BUDGET")) sp.work_budget = std::atof(x);
        for (const auto& qs : sp.queries) {
            QueryOut q;
            bool found = false;
            uint64_t tried = 0, qcount = 0;
            if (sp.exact) {
                // Candidates in generator order; the first whose exact count stays within the cap and
                // the work budget is kept (the outcome does not depend on the number of threads).
                CountQuery cq;
                for (int attempt = 0; attempt < 4096 && !found; attempt++) {
                    QueryOut c;
                    if (!extract(g, vlab, qs.second, qs.first, fin_keys, rng, c)) break;
                    tried++;
                    cq.n = static_cast<uint32_t>(c.verts.size());
                    cq.vlab.assign(cq.n, 0);
                    for (uint32_t i = 0; i < cq.n; i++) cq.vlab[i] = vlab[c.verts[i]];
                    // the shape's edges first; then, while the count is over the cap, the walk's
                    // further chords one at a time (each closes another cycle)
                    for (size_t extra = 0;; extra++) {
                        cq.edges.clear();
                        for (const auto& e : c.edges) cq.edges.push_back(CountEdge{e.a, e.b, e.l});
                        const CountResult r = count_embeddings(cg, cq, static_cast<uint64_t>(sp.emb_cap),
                                                               static_cast<uint64_t>(sp.work_budget), worker_threads());
                        if (verbose)
                            std::fprintf(stderr, "  candidate %llu: %zu edges, within %d, count %llu, work %llu\n", (unsigned long long)tried,
                                         c.edges.size(), r.within ? 1 : 0, (unsigned long long)r.count, (unsigned long long)r.work);
                        if (r.within) { q = c; found = true; qcount = r.count; break; }
                        if (extra >= c.spare.size()) break;
                        c.edges.push_back(c.spare[extra]);
                    }
                    if (now_s() - last_note > 30) {
                        std::fprintf(stderr, "gen %s: query %u, %llu candidates tried (%.1fs)\n", sp.name, qk, (unsigned long long)tried,
                                     now_s() - t_start);
                        last_note = now_s();
                    }
                }
            } else {
                // Candidates are drawn in fixed groups of 16 from the main generator; their estimates run
                // in parallel with per-candidate seeds, and the first accepted candidate in group order is
                // kept, so the outcome does not depend on the number of threads.
                constexpr int kGroup = 16;
                for (int group = 0; group < 256 && !found; group++) {
                    std::vector<QueryOut> cands;
                    std::vector<uint64_t> seeds;
                    for (int k = 0; k < kGroup; k++) {
                        QueryOut c;
                        if (!extract(g, vlab, qs.second, qs.first, fin_keys, rng, c)) break;
                        cands.push_back(c);
                        seeds.push_back(rng.next());
                    }
                    if (cands.empty()) break;
                    tried += cands.size();
                    if (sp.emb_cap <= 0) { q = cands[0]; found = true; tried -= cands.size() - 1; break; }
                    std::vector<char> ok(cands.size(), 0);
                    std::vector<std::thread> th;
                    const unsigned nth = std::max(1u, std::min<unsigned>(kGroup, std::thread::hardware_concurrency()));
                    std::atomic<size_t> next{0};
                    for (unsigned t = 0; t < nth; t++)
                        th.emplace_back([&]() {
                            for (size_t i = next++; i < cands.size(); i = next++) {
                                Rng r(seeds[i]);
                                double est = 0;
                                ok[i] = estimate_embeddings(g, vlab, cands[i], r, 480, 200000, est) && est <= sp.emb_cap;
                            }
                        });
                    for (auto& x : th) x.join();
                    for (size_t i = 0; i < cands.size(); i++)
                        if (ok[i]) {  // the group's later candidates were never needed
                            q = cands[i];
                            found = true;
                            tried -= cands.size() - 1 - i;
                            break;
                        }
                }
                if (!found && sp.directed && qs.first != kTree) {  // no cycle of this size: a tree instead
                    for (int t = 0; t < 64 && !found; t++) found = extract(g, vlab, qs.second, kTree, fin_keys, rng, q);
                }
            }
            if (!found) {
                std::fprintf(stderr, "could not extract query %u (%llu candidates tried)\n", qk, (unsigned long long)tried);
                return 1;
            }
            write_query(out + "/q" + std::to_string(qk) + ".graph", q, vlab);
            qcounts.push_back(qcount);
            qtried.push_back(tried);
            if (sp.exact)
                std::fprintf(stderr, "gen %s: query %u: %zu vertices, %zu edges, %llu embeddings in the union graph, %llu candidates tried (%.1fs)\n",
                             sp.name, qk, q.verts.size(), q.edges.size(), (unsigned long long)qcount, (unsigned long long)tried, now_s() - t_start);
            qk++;
        }
    }
    if (sp.batch_max) {  // variable batches: lognormal sizes (median batch, sigma 1) clipped to [min, max]
        FILE* bf = std::fopen((out + "/batches.txt").c_str(), "w");
        for (size_t left = stream.size(); left > 0;) {
            const double x = sp.batch * std::exp(rng.normal());
            const size_t b = std::min<size_t>(left, std::min<double>(sp.batch_max, std::max<double>(sp.batch_min, x)));
            std::fprintf(bf, "%zu\n", b);
            left -= b;
        }
        std::fclose(bf);
    }
    FILE* f = std::fopen((out + "/meta.txt").c_str(), "w");
    std::fprintf(f, "config %s\nseed %llu\nn %u\nm_initial %zu\nstream %zu\ninserts %zu\ndeletes %zu\nbatch_size %u\nqueries %u\n",
                 sp.name, (unsigned long long)seed, sp.n, n_init, stream.size(), n_hold, n_del, sp.batch, qk);
    std::fprintf(f, "directed %d\nbatch_max %u\n", sp.directed ? 1 : 0, sp.batch_max ? sp.batch_max : sp.batch);
    if (sp.exact && queries_from.empty()) {
        std::fprintf(f, "embedding_cap %.0f\n", sp.emb_cap);
        for (size_t k = 0; k < qcounts.size(); k++) std::fprintf(f, "union_embeddings_q%zu %llu\n", k, (unsigned long long)qcounts[k]);
    } else if (!sp.exact && sp.emb_cap > 0) {
        std::fprintf(f, "estimated_embedding_cap %.0f\n", sp.emb_cap);
    }
    // the selection's bias: candidates drawn per kept query (every one before the kept one was rejected)
    for (size_t k = 0; k < qtried.size(); k++) std::fprintf(f, "candidates_tried_q%zu %llu\n", k, (unsigned long long)qtried[k]);
    std::fclose(f);
    progress("done");
    return 0;
}

}  // namespace csm
