// Fill- and depth-reducing symmetric orderings on the host, behind the
// reference's CsrAmdOrder / CsrNdOrder (include/csr_matrix.h:45-49 there;
// src/csr_matrix.cpp:173-225 delegates to SuiteSparse AMD and CHOLMOD
// nested dissection, which this build does not ship).  Both work on the
// pattern of A + A^T without the diagonal and return p (p[k] = old index of
// the k-th row of the permuted matrix) and ip (ip[p[k]] = k), as the
// reference does; the benchmark applies them symmetrically
// (test/benchmark.cpp:44-80).
//
//   * amd_order: approximate minimum degree on a quotient graph (Amestoy,
//     Davis and Duff, SIAM J. Matrix Anal. Appl. 17(4), 1996): eliminated
//     nodes become elements whose variable lists replace the fill edges;
//     elements reached through the pivot are absorbed; a variable's degree
//     is the approximate external degree |A_i| + |L_p \ i| + sum over its
//     other elements of |L_e \ L_p|; variables with identical adjacency
//     (indistinguishable nodes, found by hashing) merge into supervariables
//     that are eliminated together.
//   * nd_order: nested dissection by recursive bisection: breadth-first
//     level structure from a pseudo-peripheral vertex, the level holding the
//     median vertex as separator (thinned to the vertices that touch the
//     next level), parts ordered before their separator; parts of at most
//     kNdLeaf vertices keep their natural order.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "csr_matrix.h"

namespace {

// Symmetrised adjacency (no diagonal, no duplicates), CSR.
struct Graph {
    int n = 0;
    std::vector<int> ptr, adj;
};

Graph symmetric_graph(const CsrMatrix* m) {
    Graph g;
    const int n = m->size;
    g.n = n;
    std::vector<int> deg(static_cast<size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i)
        for (int q = m->row_ptr[i]; q < m->row_ptr[i + 1]; ++q) {
            const int j = m->col_idx[q];
            if (j == i || j < 0 || j >= n) continue;
            ++deg[i];
            ++deg[j];
        }
    g.ptr.assign(static_cast<size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) g.ptr[i + 1] = g.ptr[i] + deg[i];
    g.adj.resize(static_cast<size_t>(g.ptr[n]));
    std::vector<int> fill(g.ptr.begin(), g.ptr.end() - 1);
    for (int i = 0; i < n; ++i)
        for (int q = m->row_ptr[i]; q < m->row_ptr[i + 1]; ++q) {
            const int j = m->col_idx[q];
            if (j == i || j < 0 || j >= n) continue;
            g.adj[fill[i]++] = j;
            g.adj[fill[j]++] = i;
        }
    // sort and deduplicate each list in place, then compact
    int out = 0;
    std::vector<int> nptr(static_cast<size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) {
        auto b = g.adj.begin() + g.ptr[i], e = g.adj.begin() + g.ptr[i + 1];
        std::sort(b, e);
        const auto u = std::unique(b, e);
        nptr[i] = out;
        for (auto it = b; it != u; ++it) g.adj[out++] = *it;
    }
    nptr[n] = out;
    g.ptr.swap(nptr);
    g.adj.resize(static_cast<size_t>(out));
    return g;
}

// ---------------------------------------------------------------------
// Approximate minimum degree.
// ---------------------------------------------------------------------
class Amd {
public:
    explicit Amd(const Graph& g) : n_(g.n) {
        adjv_.resize(n_);
        adje_.resize(n_);
        elem_.resize(n_);
        nv_.assign(n_, 1);
        deg_.resize(n_);
        alive_.assign(n_, 1);
        is_elem_.assign(n_, 0);
        elem_w_.assign(n_, 0);
        wstamp_.assign(n_, 0);
        w_.assign(n_, 0);
        mark_.assign(n_, 0);
        members_.resize(n_);
        for (int i = 0; i < n_; ++i) {
            adjv_[i].assign(g.adj.begin() + g.ptr[i], g.adj.begin() + g.ptr[i + 1]);
            deg_[i] = static_cast<int>(adjv_[i].size());
            members_[i] = {i};
        }
        head_.assign(static_cast<size_t>(n_) + 1, -1);
        next_.assign(n_, -1);
        prev_.assign(n_, -1);
        for (int i = 0; i < n_; ++i) push(i);
    }

    std::vector<int> order() {
        std::vector<int> p;
        p.reserve(n_);
        int mindeg = 0, left = n_;
        while (left > 0) {
            while (mindeg <= n_ && head_[mindeg] < 0) ++mindeg;
            if (mindeg > n_) throw std::runtime_error("amd: degree lists empty");
            const int piv = head_[mindeg];
            pop(piv);
            min_degree_hint_ = n_ + 1;
            eliminate(piv, p, left);
            mindeg = std::min(mindeg, min_degree_hint_);
        }
        return p;
    }

private:
    int n_;
    std::vector<std::vector<int>> adjv_, adje_, elem_, members_;
    std::vector<int> nv_, deg_, head_, next_, prev_, elem_w_, wstamp_, w_, mark_;
    std::vector<char> alive_, is_elem_;
    int stamp_ = 0;
    int min_degree_hint_ = 0;  // smallest degree pushed since the last pivot choice

    void push(int i) {
        const int d = std::min(std::max(deg_[i], 0), n_);
        deg_[i] = d;
        prev_[i] = -1;
        next_[i] = head_[d];
        if (head_[d] >= 0) prev_[head_[d]] = i;
        head_[d] = i;
        min_degree_hint_ = std::min(min_degree_hint_, d);
    }
    void pop(int i) {
        if (prev_[i] >= 0) next_[prev_[i]] = next_[i];
        else head_[deg_[i]] = next_[i];
        if (next_[i] >= 0) prev_[next_[i]] = prev_[i];
        prev_[i] = next_[i] = -1;
    }

    void eliminate(int piv, std::vector<int>& p, int& left) {
        // order the pivot's supervariable
        for (int v : members_[piv]) p.push_back(v);
        left -= nv_[piv];
        alive_[piv] = 0;
        is_elem_[piv] = 1;
        // L_p = variables of A_p and of the elements reached through p
        ++stamp_;
        mark_[piv] = stamp_;
        std::vector<int> lp;
        auto take = [&](int v) {
            if (alive_[v] && nv_[v] > 0 && mark_[v] != stamp_) {
                mark_[v] = stamp_;
                lp.push_back(v);
            }
        };
        for (int v : adjv_[piv]) take(v);
        for (int e : adje_[piv]) {
            if (!is_elem_[e] || elem_w_[e] < 0) continue;
            for (int v : elem_[e]) take(v);
            elem_w_[e] = -1;  // absorbed into the new element
            std::vector<int>().swap(elem_[e]);
        }
        std::vector<int>().swap(adjv_[piv]);
        std::vector<int>().swap(adje_[piv]);
        int lp_w = 0;
        for (int v : lp) {
            lp_w += nv_[v];
            pop(v);
        }
        // |L_e \ L_p| (weighted) of the other elements of L_p's variables
        const int ws = ++stamp_;
        for (int v : lp)
            for (int e : adje_[v]) {
                if (elem_w_[e] < 0 || e == piv) continue;
                if (wstamp_[e] != ws) {
                    wstamp_[e] = ws;
                    w_[e] = elem_w_[e];
                }
                w_[e] -= nv_[v];
            }
        // update the variables of L_p
        for (int v : lp) {
            auto& ae = adje_[v];
            size_t k = 0;
            int ext = 0;
            for (int e : ae) {
                if (elem_w_[e] < 0 || e == piv) continue;  // absorbed (or the pivot, re-added below)
                if (wstamp_[e] == ws && w_[e] <= 0) {
                    elem_w_[e] = -1;  // L_e inside L_p: absorbed into p
                    std::vector<int>().swap(elem_[e]);
                    continue;
                }
                ae[k++] = e;
                ext += wstamp_[e] == ws ? w_[e] : elem_w_[e];
            }
            ae.resize(k);
            ae.push_back(piv);
            auto& av = adjv_[v];
            size_t m = 0;
            int vw = 0;
            for (int u : av) {
                if (!alive_[u] || nv_[u] == 0 || mark_[u] == stamp_ - 1 || u == piv) continue;  // in L_p or gone
                av[m++] = u;
                vw += nv_[u];
            }
            av.resize(m);
            const int d = vw + (lp_w - nv_[v]) + ext;
            deg_[v] = std::min({d, deg_[v] + lp_w - nv_[v], left - nv_[v]});
            if (deg_[v] < 0) deg_[v] = 0;
        }
        // new element
        elem_[piv] = lp;
        elem_w_[piv] = lp_w;
        // supervariables: variables of L_p with identical element and variable lists
        merge_indistinguishable(lp);
        for (int v : lp)
            if (alive_[v] && nv_[v] > 0) push(v);
        // keep the element's list to live supervariables only
        auto& le = elem_[piv];
        size_t k = 0;
        int lw = 0;
        for (int v : le)
            if (alive_[v] && nv_[v] > 0) {
                le[k++] = v;
                lw += nv_[v];
            }
        le.resize(k);
        elem_w_[piv] = lw;
    }

    void merge_indistinguishable(const std::vector<int>& lp) {
        if (lp.size() < 2) return;
        std::vector<std::pair<uint64_t, int>> h;
        h.reserve(lp.size());
        for (int v : lp) {
            auto& ae = adje_[v];
            auto& av = adjv_[v];
            std::sort(ae.begin(), ae.end());
            std::sort(av.begin(), av.end());
            uint64_t x = 1469598103934665603ull ^ (ae.size() * 1315423911ull) ^ (av.size() << 32);
            for (int e : ae) x = (x ^ static_cast<uint64_t>(e)) * 1099511628211ull;
            x ^= 0x9e3779b97f4a7c15ull;
            for (int u : av) x = (x ^ static_cast<uint64_t>(u)) * 1099511628211ull;
            h.emplace_back(x, v);
        }
        std::sort(h.begin(), h.end());
        for (size_t a = 0; a < h.size();) {
            size_t b = a + 1;
            while (b < h.size() && h[b].first == h[a].first) ++b;
            for (size_t i = a; i < b; ++i) {
                const int vi = h[i].second;
                if (nv_[vi] == 0) continue;
                for (size_t j = i + 1; j < b; ++j) {
                    const int vj = h[j].second;
                    if (nv_[vj] == 0) continue;
                    // adjacency lists exclude each other?  i and j adjacent as
                    // variables: indistinguishable if the lists agree apart from
                    // each other
                    if (adje_[vi] != adje_[vj]) continue;
                    if (!same_without(adjv_[vi], adjv_[vj], vi, vj)) continue;
                    const int wj = nv_[vj];
                    nv_[vi] += wj;
                    nv_[vj] = 0;
                    alive_[vj] = 0;
                    members_[vi].insert(members_[vi].end(), members_[vj].begin(), members_[vj].end());
                    std::vector<int>().swap(members_[vj]);
                    std::vector<int>().swap(adjv_[vj]);
                    std::vector<int>().swap(adje_[vj]);
                    deg_[vi] = std::max(0, deg_[vi] - wj);
                }
            }
            a = b;
        }
        // drop merged variables from the surviving lists
        for (int v : lp) {
            if (nv_[v] == 0) continue;
            auto& av = adjv_[v];
            av.erase(std::remove_if(av.begin(), av.end(), [&](int u) { return nv_[u] == 0 || !alive_[u]; }),
                     av.end());
        }
    }

    static bool same_without(const std::vector<int>& a, const std::vector<int>& b, int x, int y) {
        size_t i = 0, j = 0;
        while (true) {
            while (i < a.size() && (a[i] == x || a[i] == y)) ++i;
            while (j < b.size() && (b[j] == x || b[j] == y)) ++j;
            if (i == a.size() || j == b.size()) return i == a.size() && j == b.size();
            if (a[i] != b[j]) return false;
            ++i;
            ++j;
        }
    }
};

// ---------------------------------------------------------------------
// Nested dissection.
// ---------------------------------------------------------------------
constexpr int kNdLeaf = 64;

class Nd {
public:
    explicit Nd(const Graph& g) : g_(g), part_(g.n, 0), level_(g.n, -1), seen_(g.n, 0) {}

    std::vector<int> order() {
        std::vector<int> p;
        p.reserve(g_.n);
        // every connected component, then recursive bisection
        std::vector<int> all(g_.n);
        std::iota(all.begin(), all.end(), 0);
        int tag = 1;
        for (auto& c : components(all, 0, tag)) dissect(std::move(c), p, tag);
        return p;
    }

private:
    const Graph& g_;
    std::vector<int> part_, level_, seen_;
    int stamp_ = 0;

    // Connected components of `verts` (all tagged `tag` in part_) as
    // separate lists, each tagged with a fresh id.
    std::vector<std::vector<int>> components(const std::vector<int>& verts, int tag, int& next_tag) {
        std::vector<std::vector<int>> out;
        ++stamp_;
        for (int s : verts) {
            if (seen_[s] == stamp_) continue;
            std::vector<int> comp{s};
            seen_[s] = stamp_;
            for (size_t h = 0; h < comp.size(); ++h) {
                const int v = comp[h];
                for (int q = g_.ptr[v]; q < g_.ptr[v + 1]; ++q) {
                    const int u = g_.adj[q];
                    if (part_[u] == tag && seen_[u] != stamp_) {
                        seen_[u] = stamp_;
                        comp.push_back(u);
                    }
                }
            }
            const int t = next_tag++;
            for (int v : comp) part_[v] = t;
            out.push_back(std::move(comp));
        }
        return out;
    }

    // Breadth-first levels of the vertices tagged `tag`, from `root`;
    // returns the vertices in visit order, level_ set, the last level in h.
    std::vector<int> bfs(int root, int tag, int& h) {
        std::vector<int> q{root};
        ++stamp_;
        seen_[root] = stamp_;
        level_[root] = 0;
        for (size_t k = 0; k < q.size(); ++k) {
            const int v = q[k];
            for (int e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) {
                const int u = g_.adj[e];
                if (part_[u] == tag && seen_[u] != stamp_) {
                    seen_[u] = stamp_;
                    level_[u] = level_[v] + 1;
                    q.push_back(u);
                }
            }
        }
        h = level_[q.back()];
        return q;
    }

    int degree_in(int v, int tag) const {
        int d = 0;
        for (int e = g_.ptr[v]; e < g_.ptr[v + 1]; ++e) d += part_[g_.adj[e]] == tag;
        return d;
    }

    void dissect(std::vector<int> verts, std::vector<int>& p, int& next_tag) {
        // explicit stack: (vertices, separator to append after both parts)
        struct Task {
            std::vector<int> verts;
            std::vector<int> sep;  // appended when `verts` is empty (a finished pair)
        };
        std::vector<Task> st;
        st.push_back({std::move(verts), {}});
        while (!st.empty()) {
            Task t = std::move(st.back());
            st.pop_back();
            if (t.verts.empty()) {
                p.insert(p.end(), t.sep.begin(), t.sep.end());
                continue;
            }
            std::vector<int>& v = t.verts;
            if (static_cast<int>(v.size()) <= kNdLeaf) {
                std::sort(v.begin(), v.end());
                p.insert(p.end(), v.begin(), v.end());
                continue;
            }
            const int tag = part_[v[0]];
            // pseudo-peripheral root: a few sweeps from a minimum-degree vertex
            int root = v[0], best = degree_in(root, tag);
            for (int x : v) {
                const int d = degree_in(x, tag);
                if (d < best) { best = d; root = x; }
            }
            int h = 0;
            std::vector<int> q = bfs(root, tag, h);
            for (int it = 0; it < 4; ++it) {
                int cand = q.back(), cd = degree_in(cand, tag);
                for (auto r = q.rbegin(); r != q.rend() && level_[*r] == h; ++r) {
                    const int d = degree_in(*r, tag);
                    if (d < cd) { cd = d; cand = *r; }
                }
                int h2 = 0;
                std::vector<int> q2 = bfs(cand, tag, h2);
                if (h2 <= h) break;
                h = h2;
                q.swap(q2);
                root = cand;
            }
            (void)bfs(root, tag, h);
            if (h < 2) {  // too dense to split by levels
                std::sort(v.begin(), v.end());
                p.insert(p.end(), v.begin(), v.end());
                continue;
            }
            // separator level: the one holding the median vertex
            std::vector<int> cnt(static_cast<size_t>(h) + 1, 0);
            for (int x : v) ++cnt[level_[x]];
            int m = 0;
            for (long acc = 0; m <= h; ++m) {
                acc += cnt[m];
                if (2 * acc >= static_cast<long>(v.size())) break;
            }
            m = std::min(std::max(m, 1), h - 1);
            std::vector<int> left, right, sep;
            for (int x : v) {
                const int l = level_[x];
                if (l < m) {
                    left.push_back(x);
                } else if (l > m) {
                    right.push_back(x);
                } else {
                    bool touches_next = false;
                    for (int e = g_.ptr[x]; e < g_.ptr[x + 1] && !touches_next; ++e) {
                        const int u = g_.adj[e];
                        touches_next = part_[u] == tag && level_[u] == m + 1;
                    }
                    (touches_next ? sep : left).push_back(x);
                }
            }
            std::sort(sep.begin(), sep.end());
            for (int x : sep) part_[x] = -1;
            // children: left and right components (left first in the order)
            std::vector<std::vector<int>> parts;
            for (auto* side : {&left, &right}) {
                const int stag = next_tag++;
                for (int x : *side) part_[x] = stag;
                for (auto& c : components(*side, stag, next_tag)) parts.push_back(std::move(c));
            }
            st.push_back({{}, std::move(sep)});
            for (auto it = parts.rbegin(); it != parts.rend(); ++it) st.push_back({std::move(*it), {}});
        }
    }
};

void fill_inverse(int n, const std::vector<int>& p, int* out_p, int* ip) {
    for (int k = 0; k < n; ++k) out_p[k] = p[k];
    if (ip)
        for (int k = 0; k < n; ++k) ip[p[k]] = k;
}

bool valid_input(const CsrMatrix* m, const int* p) {
    return m && p && m->size >= 0 && (m->size == 0 || (m->row_ptr && m->col_idx));
}

}  // namespace

extern "C" {

int CsrAmdOrder(const CsrMatrix* mat, int* p, int* ip) {
    if (!valid_input(mat, p)) return -1;
    try {
        const Graph g = symmetric_graph(mat);
        Amd amd(g);
        const std::vector<int> order = amd.order();
        if (static_cast<int>(order.size()) != g.n) return -1;
        fill_inverse(g.n, order, p, ip);
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int CsrNdOrder(const CsrMatrix* mat, int* p, int* ip) {
    if (!valid_input(mat, p)) return 0;
    try {
        const Graph g = symmetric_graph(mat);
        Nd nd(g);
        const std::vector<int> order = nd.order();
        if (static_cast<int>(order.size()) != g.n) return 0;
        fill_inverse(g.n, order, p, ip);
        return 1;
    } catch (const std::exception&) {
        return 0;
    }
}

}  // extern "C"
