#include "gen_count.h"

#include <algorithm>
#include <atomic>
#include <thread>

namespace csm {

void CountGraph::build(uint32_t nn, const std::vector<uint32_t>& vl, uint32_t ne, bool dir, const std::vector<CountEdge>& es,
                       unsigned threads) {
    n = nn;
    vlab = vl;
    nel = ne;
    directed = dir;
    const int sides = dir ? 2 : 1;
    for (int s = 0; s < 2; s++) { off[s].clear(); key[s].clear(); }
    for (int s = 0; s < sides; s++) off[s].assign(static_cast<size_t>(n) + 1, 0);
    for (const auto& e : es) {
        off[0][e.a + 1]++;
        off[dir ? 1 : 0][e.b + 1]++;
    }
    std::vector<uint64_t> pos[2];
    for (int s = 0; s < sides; s++) {
        for (uint32_t v = 0; v < n; v++) off[s][v + 1] += off[s][v];
        key[s].resize(off[s][n]);
        pos[s].assign(off[s].begin(), off[s].end() - 1);
    }
    auto slot_key = [&](uint32_t x, uint32_t l) { return (static_cast<uint64_t>(vlab[x] * nel + l) << 32) | x; };
    for (const auto& e : es) {
        key[0][pos[0][e.a]++] = slot_key(e.b, e.l);
        const int s = dir ? 1 : 0;
        key[s][pos[s][e.b]++] = slot_key(e.a, e.l);
    }
    threads = std::max(1u, threads);
    for (int s = 0; s < sides; s++) {
        std::atomic<uint32_t> next{0};
        std::vector<std::thread> th;
        for (unsigned t = 0; t < threads; t++)
            th.emplace_back([&, s]() {
                for (uint32_t lo = next.fetch_add(4096); lo < n; lo = next.fetch_add(4096))
                    for (uint32_t v = lo; v < std::min(n, lo + 4096); v++)
                        std::sort(key[s].begin() + off[s][v], key[s].begin() + off[s][v + 1]);
            });
        for (auto& x : th) x.join();
    }
    uint32_t maxl = 0;
    for (uint32_t x : vlab) maxl = std::max(maxl, x);
    by_label.assign(static_cast<size_t>(maxl) + 1, {});
    for (uint32_t v = 0; v < n; v++) by_label[vlab[v]].push_back(v);
}

namespace {

constexpr uint32_t kMaxTail = 8;   // trailing leaves counted by inclusion-exclusion
constexpr uint32_t kMaxGroup = 6;  // same-label leaves per group (Bell(6) = 203 partitions)

struct Back { uint32_t pos, slot, side; };

struct Group {
    std::vector<uint32_t> leaves;                // tail positions
    std::vector<std::vector<uint32_t>> parts;    // set partitions as lists of block masks over `leaves`
    std::vector<int64_t> mu;                     // Moebius coefficient of each partition
};

struct Plan {
    uint32_t n = 0, tail = 0;
    std::vector<uint32_t> order;
    std::vector<std::vector<Back>> back;
    std::vector<uint32_t> slot_label;  // data label of the vertex at each position
    std::vector<Group> groups;
};

// Order: the query vertex of highest degree first, then repeatedly the unvisited vertex with the
// most visited neighbours (ties: higher degree, then lower id), so cycles close early and tree
// leaves come last.  Returns false for a disconnected query.
bool make_plan(const CountGraph& g, const CountQuery& q, Plan& p) {
    const uint32_t n = q.n;
    p.n = n;
    std::vector<uint32_t> deg(n, 0), pos(n, UINT32_MAX);
    for (const auto& e : q.edges) { deg[e.a]++; deg[e.b]++; }
    uint32_t start = 0;
    for (uint32_t u = 1; u < n; u++)
        if (deg[u] > deg[start]) start = u;
    p.order.assign(1, start);
    pos[start] = 0;
    while (p.order.size() < n) {
        uint32_t best = UINT32_MAX, best_back = 0;
        for (uint32_t u = 0; u < n; u++) {
            if (pos[u] != UINT32_MAX) continue;
            uint32_t nb = 0;
            for (const auto& e : q.edges)
                if ((e.a == u && pos[e.b] != UINT32_MAX) || (e.b == u && pos[e.a] != UINT32_MAX)) nb++;
            if (nb == 0) continue;
            if (best == UINT32_MAX || nb > best_back || (nb == best_back && deg[u] > deg[best])) { best = u; best_back = nb; }
        }
        if (best == UINT32_MAX) return false;
        pos[best] = static_cast<uint32_t>(p.order.size());
        p.order.push_back(best);
    }
    p.back.assign(n, {});
    p.slot_label.assign(n, 0);
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t u = p.order[k];
        p.slot_label[k] = q.vlab[u];
        for (const auto& e : q.edges) {
            // undirected: x in N(f(other)); directed arc other -> u: x in out(f(other)); u -> other: x in in(f(other))
            if (e.b == u && pos[e.a] < k) p.back[k].push_back(Back{pos[e.a], q.vlab[u] * g.nel + e.l, 0});
            else if (e.a == u && pos[e.b] < k) p.back[k].push_back(Back{pos[e.b], q.vlab[u] * g.nel + e.l, g.directed ? 1u : 0u});
        }
    }
    // trailing leaves: degree-1 vertices at the end whose groups stay small
    uint32_t t = n;
    std::vector<uint32_t> per_label;
    while (t > 1 && deg[p.order[t - 1]] == 1 && n - (t - 1) <= kMaxTail) {
        const uint32_t l = q.vlab[p.order[t - 1]];
        uint32_t same = 1;
        for (uint32_t k = t; k < n; k++) same += q.vlab[p.order[k]] == l;
        if (same > kMaxGroup) break;
        t--;
    }
    p.tail = t;
    for (uint32_t k = t; k < n; k++) {
        Group* gr = nullptr;
        for (auto& x : p.groups)
            if (p.slot_label[x.leaves[0]] == p.slot_label[k]) gr = &x;
        if (!gr) { p.groups.emplace_back(); gr = &p.groups.back(); }
        gr->leaves.push_back(k);
    }
    for (auto& gr : p.groups) {  // restricted growth strings enumerate the set partitions
        const uint32_t m = static_cast<uint32_t>(gr.leaves.size());
        std::vector<uint32_t> a(m, 0);
        while (true) {
            uint32_t blocks = 0;
            for (uint32_t i = 0; i < m; i++) blocks = std::max(blocks, a[i] + 1);
            std::vector<uint32_t> masks(blocks, 0);
            for (uint32_t i = 0; i < m; i++) masks[a[i]] |= 1u << i;
            int64_t mu = 1;
            for (uint32_t b : masks) {
                const int s = __builtin_popcount(b);
                for (int x = 2; x < s; x++) mu *= x;
                if (s % 2 == 0) mu = -mu;
            }
            gr.parts.push_back(masks);
            gr.mu.push_back(mu);
            int i = static_cast<int>(m) - 1;  // next string: increment the last position that may grow
            while (i > 0) {
                uint32_t mx = 0;
                for (int j = 0; j < i; j++) mx = std::max(mx, a[j]);
                if (a[i] <= mx) break;
                a[i] = 0;
                i--;
            }
            if (i <= 0) break;
            a[i]++;
        }
    }
    return true;
}

struct Shared {
    std::atomic<uint64_t> count{0}, work{0};
    std::atomic<bool> stop{false};
    uint64_t cap, budget;
};

struct Worker {
    const CountGraph& g;
    const Plan& p;
    Shared& sh;
    uint32_t f[32];
    uint64_t count = 0, work = 0, flushed_work = 0;

    Worker(const CountGraph& gg, const Plan& pp, Shared& s) : g(gg), p(pp), sh(s) {}

    void flush() {
        if (count) {
            if (sh.count.fetch_add(count) + count > sh.cap) sh.stop = true;
            count = 0;
        }
        if (work != flushed_work) {
            if (sh.work.fetch_add(work - flushed_work) + (work - flushed_work) > sh.budget) sh.stop = true;
            flushed_work = work;
        }
    }
    bool add(unsigned __int128 c) {
        if (c > sh.cap || count + c > sh.cap) { sh.count.fetch_add(sh.cap + 1); sh.stop = true; return false; }
        count += static_cast<uint64_t>(c);
        if (work - flushed_work > (1u << 16)) flush();
        return !sh.stop.load(std::memory_order_relaxed);
    }
    void run(uint32_t v, const Back& b, const uint64_t*& lo, const uint64_t*& hi) {
        const uint64_t* base = g.key[b.side].data();
        const uint64_t* s = base + g.off[b.side][v];
        const uint64_t* e = base + g.off[b.side][v + 1];
        const uint64_t k = static_cast<uint64_t>(b.slot) << 32;
        lo = std::lower_bound(s, e, k);
        hi = std::lower_bound(lo, e, k + (1ull << 32));
        work += 2;
    }
    bool used(uint32_t x, uint32_t k) const {
        for (uint32_t j = 0; j < k; j++)
            if (f[j] == x) return true;
        return false;
    }
    unsigned __int128 tail() {
        const uint32_t t = p.tail;
        const uint64_t* lo[kMaxTail];
        const uint64_t* hi[kMaxTail];
        uint32_t pref[kMaxTail];
        for (uint32_t k = t; k < p.n; k++) {
            const uint32_t i = k - t;
            run(f[p.back[k][0].pos], p.back[k][0], lo[i], hi[i]);
            if (lo[i] == hi[i]) return 0;
            pref[i] = 0;
            for (uint32_t j = 0; j < t; j++) {
                if (g.vlab[f[j]] != p.slot_label[k]) continue;
                const uint64_t key = (static_cast<uint64_t>(p.back[k][0].slot) << 32) | f[j];
                work++;
                if (std::binary_search(lo[i], hi[i], key)) pref[i] |= 1u << j;
            }
        }
        unsigned __int128 total = 1;
        for (const auto& gr : p.groups) {
            const uint32_t m = static_cast<uint32_t>(gr.leaves.size());
            int64_t size[1u << kMaxGroup];
            for (uint32_t b = 0; b < (1u << m); b++) size[b] = -1;
            auto block_size = [&](uint32_t b) -> int64_t {
                if (size[b] >= 0) return size[b];
                uint32_t shortest = UINT32_MAX, pm = ~0u;
                bool same = true;
                for (uint32_t i = 0; i < m; i++) {
                    if (!(b >> i & 1)) continue;
                    const uint32_t li = gr.leaves[i] - t;
                    pm &= pref[li];
                    if (shortest == UINT32_MAX || hi[li] - lo[li] < hi[shortest] - lo[shortest]) shortest = li;
                    if (lo[li] != lo[gr.leaves[__builtin_ctz(b)] - t]) same = false;
                }
                int64_t c = 0;
                if (same) {
                    c = hi[shortest] - lo[shortest];
                } else {  // intersection of the runs: scan the shortest, search the others
                    for (const uint64_t* x = lo[shortest]; x < hi[shortest]; x++) {
                        const uint32_t id = static_cast<uint32_t>(*x);
                        bool in = true;
                        for (uint32_t i = 0; i < m && in; i++) {
                            if (!(b >> i & 1)) continue;
                            const uint32_t li = gr.leaves[i] - t;
                            if (li == shortest) continue;
                            const uint64_t key = (static_cast<uint64_t>(p.back[gr.leaves[i]][0].slot) << 32) | id;
                            in = std::binary_search(lo[li], hi[li], key);
                        }
                        c += in;
                    }
                    work += static_cast<uint64_t>(hi[shortest] - lo[shortest]) * (__builtin_popcount(b) - 1);
                }
                c -= __builtin_popcount(pm);
                return size[b] = c;
            };
            __int128 sum = 0;
            for (size_t k = 0; k < gr.parts.size(); k++) {
                __int128 prod = gr.mu[k];
                for (uint32_t b : gr.parts[k]) {
                    prod *= block_size(b);
                    if (prod == 0) break;
                }
                sum += prod;
            }
            if (sum <= 0) return 0;
            total *= static_cast<unsigned __int128>(sum);
            if (total > sh.cap) return total;
        }
        return total;
    }
    // extends positions k.. of the current prefix; false stops the search
    bool level(uint32_t k) {
        const auto& bk = p.back[k];
        const uint64_t* lo[16];
        const uint64_t* hi[16];
        uint32_t piv = 0;
        for (uint32_t i = 0; i < bk.size(); i++) {
            run(f[bk[i].pos], bk[i], lo[i], hi[i]);
            if (lo[i] == hi[i]) return true;
            if (hi[i] - lo[i] < hi[piv] - lo[piv]) piv = i;
        }
        work += hi[piv] - lo[piv];
        const bool last = k + 1 == p.n;
        uint64_t leaf = 0;
        for (const uint64_t* x = lo[piv]; x < hi[piv]; x++) {
            if (work - flushed_work > (1u << 16)) {  // searches that find nothing still spend the budget
                flush();
                if (sh.stop.load(std::memory_order_relaxed)) return false;
            }
            const uint32_t id = static_cast<uint32_t>(*x);
            if (used(id, k)) continue;
            bool ok = true;
            for (uint32_t i = 0; i < bk.size() && ok; i++) {
                if (i == piv) continue;
                const uint64_t key = (static_cast<uint64_t>(bk[i].slot) << 32) | id;
                lo[i] = std::lower_bound(lo[i], hi[i], key);  // candidates ascend: the window only moves up
                ok = lo[i] < hi[i] && *lo[i] == key;
                work++;
            }
            if (!ok) continue;
            if (last) { leaf++; continue; }
            f[k] = id;
            if (k + 1 == p.tail) {
                if (!add(tail())) return false;
            } else if (!level(k + 1)) {
                return false;
            }
        }
        if (leaf && !add(leaf)) return false;
        return true;
    }
    bool root(uint32_t v) {
        f[0] = v;
        work++;
        if (p.n == 1) return add(1);
        if (p.tail == 1) return add(tail());
        return level(1);
    }
};

}  // namespace

CountResult count_embeddings(const CountGraph& g, const CountQuery& q, uint64_t cap, uint64_t work_budget, unsigned threads) {
    CountResult r;
    Plan p;
    if (q.n == 0 || q.n > 32 || !make_plan(g, q, p)) return r;
    const uint32_t l0 = p.slot_label[0];
    if (l0 >= g.by_label.size()) { r.within = true; return r; }
    const auto& roots = g.by_label[l0];
    Shared sh;
    sh.cap = cap;
    sh.budget = work_budget;
    std::atomic<size_t> next{0};
    threads = std::max(1u, threads);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < threads; t++)
        th.emplace_back([&]() {
            Worker w(g, p, sh);
            for (size_t lo = next.fetch_add(64); lo < roots.size() && !sh.stop; lo = next.fetch_add(64)) {
                for (size_t i = lo; i < std::min(roots.size(), lo + 64); i++)
                    if (!w.root(roots[i])) break;
                w.flush();
            }
            w.flush();
        });
    for (auto& x : th) x.join();
    r.count = sh.count.load();
    r.work = sh.work.load();
    r.within = !sh.stop && r.count <= cap && r.work <= work_budget;
    return r;
}

}  // namespace csm
