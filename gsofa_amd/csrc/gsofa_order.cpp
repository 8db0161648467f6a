// gsofa_order.cpp -- fill-reducing ordering for inputs that arrive without
// one, and the symmetric permutation that applies it (host code).
//
// No counterpart in the reference, whose inputs come already ordered by
// SuperLU_DIST (README.md, "Dataset generation").  The method is automatic
// nested dissection on level structures (A. George and J. W. H. Liu, "An
// automatic nested dissection algorithm for irregular finite element
// problems", SIAM J. Numer. Anal. 15 (1978)), on the symmetrized graph of A
// (an edge i-j for every entry A(i,j) or A(j,i), i != j):
//   - a region (a connected set of unnumbered vertices) of at most `leaf`
//     vertices is numbered by minimum degree on its own elimination graph
//     (counting its neighbours outside the region, which are separator
//     vertices numbered later);
//   - otherwise a breadth-first level structure is rooted at a
//     pseudo-peripheral vertex (repeated BFS from a minimum-degree vertex of
//     the last level while the depth grows), the separator is the level where
//     the cumulative count first reaches half the region, trimmed to the
//     vertices with a neighbour in the next level, and the vertices before it
//     and after it are dissected in turn (each connected component on its
//     own) and numbered before the separator.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gsofa.h"

namespace {

struct Graph {
    int32_t n = 0;
    std::vector<int64_t> xadj;  // n+1
    std::vector<int32_t> adj;   // sorted, no duplicates, no self loops
};

int build_graph(const gsofa_matrix* A, Graph& G) {
    const int32_t n = A->n;
    G.n = n;
    std::vector<int64_t> deg((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; i++)
        for (int64_t k = A->colbeg[i]; k < A->colend[i]; k++) {
            const int32_t j = A->csr[k];
            if (j < 0 || j >= n) return GSOFA_ERR_INVALID;
            if (j != i) {
                deg[(size_t)i]++;
                deg[(size_t)j]++;
            }
        }
    std::vector<int64_t> pos((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; i++) pos[(size_t)i + 1] = pos[(size_t)i] + deg[(size_t)i];
    std::vector<int32_t> tmp((size_t)pos[(size_t)n]);
    std::vector<int64_t> cur(pos.begin(), pos.end() - 1);
    for (int32_t i = 0; i < n; i++)
        for (int64_t k = A->colbeg[i]; k < A->colend[i]; k++) {
            const int32_t j = A->csr[k];
            if (j != i) {
                tmp[(size_t)cur[(size_t)i]++] = j;
                tmp[(size_t)cur[(size_t)j]++] = i;
            }
        }
    G.xadj.assign((size_t)n + 1, 0);
    G.adj.clear();
    G.adj.reserve(tmp.size());
    for (int32_t i = 0; i < n; i++) {
        auto b = tmp.begin() + pos[(size_t)i], e = tmp.begin() + pos[(size_t)i + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        G.adj.insert(G.adj.end(), b, e);
        G.xadj[(size_t)i + 1] = (int64_t)G.adj.size();
    }
    return GSOFA_OK;
}

class Dissector {
public:
    Dissector(const Graph& G, int32_t leaf, int32_t* perm)
        : G_(G), leaf_(std::max(leaf, 1)), perm_(perm), region_((size_t)G.n, -1), level_((size_t)G.n, -1) {}

    void run() {
        std::vector<int32_t> all((size_t)G_.n);
        for (int32_t v = 0; v < G_.n; v++) all[(size_t)v] = v;
        split_components(all);
        while (!stack_.empty()) {
            Task t = std::move(stack_.back());
            stack_.pop_back();
            if (t.number_only) number(t.v);
            else dissect(t.v);
        }
    }

private:
    struct Task {
        std::vector<int32_t> v;
        bool number_only;
    };
    const Graph& G_;
    int32_t leaf_;
    int32_t* perm_;
    int32_t next_ = 0;
    int32_t rid_ = 0;
    std::vector<int32_t> region_;  // region id of each unnumbered vertex
    std::vector<int32_t> level_;   // BFS depth of the vertices the last search reached
    std::vector<Task> stack_;

    void number(std::vector<int32_t>& v) {
        std::sort(v.begin(), v.end());
        for (int32_t x : v) {
            perm_[next_++] = x;
            region_[(size_t)x] = -2;  // numbered
        }
    }
    int32_t degree(int32_t v) const { return (int32_t)(G_.xadj[(size_t)v + 1] - G_.xadj[(size_t)v]); }

    // Level structure of region `rid` from root: the levels' vertex lists.
    // Afterwards tagged_[w] == tag_ marks the reached vertices and level_[w]
    // holds their depth, until the next search.
    std::vector<std::vector<int32_t>> bfs(int32_t root, int32_t rid) {
        std::vector<std::vector<int32_t>> lv;
        std::vector<int32_t> frontier{root};
        tag_++;
        tagged_[(size_t)root] = tag_;
        level_[(size_t)root] = 0;
        int32_t depth = 0;
        while (!frontier.empty()) {
            lv.push_back(frontier);
            std::vector<int32_t> nextf;
            for (int32_t v : frontier)
                for (int64_t k = G_.xadj[(size_t)v]; k < G_.xadj[(size_t)v + 1]; k++) {
                    const int32_t w = G_.adj[(size_t)k];
                    if (region_[(size_t)w] != rid || tagged_[(size_t)w] == tag_) continue;
                    tagged_[(size_t)w] = tag_;
                    level_[(size_t)w] = depth + 1;
                    nextf.push_back(w);
                }
            frontier.swap(nextf);
            depth++;
        }
        return lv;
    }
    std::vector<int64_t> tagged_ = std::vector<int64_t>((size_t)G_.n, 0);
    int64_t tag_ = 0;

    // Push each connected component of v as its own region (in reverse, so
    // the first component is dissected first).
    void split_components(const std::vector<int32_t>& v) {
        const int32_t rid = ++rid_;
        for (int32_t x : v) region_[(size_t)x] = rid;
        std::vector<std::vector<int32_t>> comps;
        tag_++;
        for (int32_t s : v) {
            if (tagged_[(size_t)s] == tag_) continue;
            std::vector<int32_t> comp{s};
            tagged_[(size_t)s] = tag_;
            for (size_t h = 0; h < comp.size(); h++) {
                const int32_t u = comp[h];
                for (int64_t k = G_.xadj[(size_t)u]; k < G_.xadj[(size_t)u + 1]; k++) {
                    const int32_t w = G_.adj[(size_t)k];
                    if (region_[(size_t)w] == rid && tagged_[(size_t)w] != tag_) {
                        tagged_[(size_t)w] = tag_;
                        comp.push_back(w);
                    }
                }
            }
            comps.push_back(std::move(comp));
        }
        for (size_t c = comps.size(); c-- > 0;) stack_.push_back(Task{std::move(comps[c]), false});
    }

    // A leaf by minimum degree on its own elimination graph: each step
    // numbers an unnumbered leaf vertex with the fewest neighbours (its leaf
    // neighbours in the elimination graph plus its original neighbours
    // outside the leaf, all numbered later), ties to the smaller index, and
    // joins its leaf neighbours into a clique.
    void number_leaf(std::vector<int32_t>& v) {
        const size_t L = v.size();
        if (L <= 2) {
            number(v);
            return;
        }
        std::sort(v.begin(), v.end());
        const int32_t rid = ++rid_;
        for (int32_t x : v) region_[(size_t)x] = rid;
        std::vector<std::vector<char>> adj(L, std::vector<char>(L, 0));
        std::vector<int32_t> ext(L, 0);
        for (size_t a = 0; a < L; a++)
            for (int64_t k = G_.xadj[(size_t)v[a]]; k < G_.xadj[(size_t)v[a] + 1]; k++) {
                const int32_t w = G_.adj[(size_t)k];
                if (region_[(size_t)w] == rid) {
                    const size_t b = (size_t)(std::lower_bound(v.begin(), v.end(), w) - v.begin());
                    adj[a][b] = 1;
                } else if (region_[(size_t)w] != -2) {
                    ext[a]++;
                }
            }
        std::vector<char> done(L, 0);
        for (size_t step = 0; step < L; step++) {
            size_t best = L;
            int32_t bestd = 0;
            for (size_t a = 0; a < L; a++) {
                if (done[a]) continue;
                int32_t d = ext[a];
                for (size_t b = 0; b < L; b++) d += (!done[b] && adj[a][b]) ? 1 : 0;
                if (best == L || d < bestd) {
                    best = a;
                    bestd = d;
                }
            }
            done[best] = 1;
            perm_[next_++] = v[best];
            region_[(size_t)v[best]] = -2;
            for (size_t a = 0; a < L; a++) {
                if (done[a] || !adj[best][a]) continue;
                for (size_t b = a + 1; b < L; b++)
                    if (!done[b] && adj[best][b]) adj[a][b] = adj[b][a] = 1;
            }
        }
    }

    void dissect(std::vector<int32_t>& v) {
        if ((int64_t)v.size() <= leaf_) {
            number_leaf(v);
            return;
        }
        const int32_t rid = ++rid_;
        for (int32_t x : v) region_[(size_t)x] = rid;
        // pseudo-peripheral root
        int32_t root = *std::min_element(v.begin(), v.end());
        std::vector<std::vector<int32_t>> lv = bfs(root, rid);
        bool marks_current = true;  // tagged_/level_ describe lv
        for (int it = 0; it < 8; it++) {
            const std::vector<int32_t>& last = lv.back();
            int32_t cand = last[0];
            for (int32_t x : last)
                if (degree(x) < degree(cand) || (degree(x) == degree(cand) && x < cand)) cand = x;
            std::vector<std::vector<int32_t>> l2 = bfs(cand, rid);
            if (l2.size() <= lv.size()) {
                marks_current = false;
                break;
            }
            root = cand;
            lv.swap(l2);
        }
        if (!marks_current) lv = bfs(root, rid);
        if (lv.size() < 3) {  // too shallow to separate: number as one piece
            number(v);
            return;
        }
        // separator level: where the cumulative count first reaches half
        const int64_t half = (int64_t)v.size() / 2;
        int64_t acc = 0;
        size_t m = 1;
        for (size_t l = 0; l < lv.size(); l++) {
            acc += (int64_t)lv[l].size();
            if (acc >= half) {
                m = l;
                break;
            }
        }
        m = std::min(std::max<size_t>(m, 1), lv.size() - 2);
        // the separator: vertices of level m with a neighbour in level m+1
        std::vector<int32_t> sep, lo, hi;
        for (int32_t x : lv[m]) {
            bool touches = false;
            for (int64_t k = G_.xadj[(size_t)x]; k < G_.xadj[(size_t)x + 1] && !touches; k++) {
                const int32_t w = G_.adj[(size_t)k];
                touches = region_[(size_t)w] == rid && tagged_[(size_t)w] == tag_ &&
                          level_[(size_t)w] == (int32_t)m + 1;
            }
            (touches ? sep : lo).push_back(x);
        }
        for (size_t l = 0; l < m; l++) lo.insert(lo.end(), lv[l].begin(), lv[l].end());
        for (size_t l = m + 1; l < lv.size(); l++) hi.insert(hi.end(), lv[l].begin(), lv[l].end());
        // numbered: lower side, upper side, then the separator
        stack_.push_back(Task{std::move(sep), true});
        split_components_later(hi);
        split_components_later(lo);
    }

    // split_components, deferred to keep the lower side on top of the stack
    void split_components_later(std::vector<int32_t>& v) {
        if (v.empty()) return;
        split_components(v);
    }
};

}  // namespace

extern "C" {

int gsofa_order_nd(const gsofa_matrix* A, int32_t leaf, int32_t* perm) {
    if (!A || A->n < 0 || (A->n > 0 && (!A->colbeg || !A->colend || (A->nnz > 0 && !A->csr) || !perm)))
        return GSOFA_ERR_INVALID;
    for (int32_t i = 0; i < A->n; i++)
        if (A->colbeg[i] < 0 || A->colend[i] < A->colbeg[i] || A->colend[i] > A->nnz) return GSOFA_ERR_INVALID;
    Graph G;
    int rc = build_graph(A, G);
    if (rc) return rc;
    Dissector d(G, leaf, perm);
    d.run();
    return GSOFA_OK;
}

int gsofa_permute(const gsofa_matrix* A, const int32_t* perm, gsofa_csr* out) {
    if (!A || !out || A->n < 0 || (A->n > 0 && (!A->colbeg || !A->colend || !perm || (A->nnz > 0 && !A->csr))))
        return GSOFA_ERR_INVALID;
    memset(out, 0, sizeof(*out));
    const int32_t n = A->n;
    std::vector<int32_t> inv((size_t)n, -1);
    for (int32_t k = 0; k < n; k++) {
        const int32_t p = perm[k];
        if (p < 0 || p >= n || inv[(size_t)p] != -1) return GSOFA_ERR_INVALID;
        inv[(size_t)p] = k;
    }
    int64_t nnz = 0;
    for (int32_t i = 0; i < n; i++) {
        if (A->colbeg[i] < 0 || A->colend[i] < A->colbeg[i] || A->colend[i] > A->nnz) return GSOFA_ERR_INVALID;
        nnz += A->colend[i] - A->colbeg[i];
    }
    out->n = n;
    out->nnz = nnz;
    out->colbeg = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max(n, 1));
    out->colend = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max(n, 1));
    out->csr = (int32_t*)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
    if (!out->colbeg || !out->colend || !out->csr) {
        gsofa_csr_free(out);
        return GSOFA_ERR_ALLOC;
    }
    int64_t p = 0;
    for (int32_t k = 0; k < n; k++) {
        const int32_t i = perm[k];
        out->colbeg[k] = (int32_t)p;
        for (int64_t e = A->colbeg[i]; e < A->colend[i]; e++) {
            const int32_t j = A->csr[e];
            if (j < 0 || j >= n) {
                gsofa_csr_free(out);
                return GSOFA_ERR_INVALID;
            }
            out->csr[p++] = inv[(size_t)j];
        }
        std::sort(out->csr + out->colbeg[k], out->csr + p);
        out->colend[k] = (int32_t)p;
    }
    return GSOFA_OK;
}

}  // extern "C"
