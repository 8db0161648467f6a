// gsofa_bound.cpp -- per-row upper bounds of the L/U structure, on the host,
// before any traversal.
//
// No counterpart in the reference, which sizes its frontier pool from a
// memory budget and recovers from overflow by spilling to the host
// (with_space_optimization/symbfact.cuh:2079, 2170-2380).  Here the fill
// stream and the output columns are sized once, from bounds that hold for
// every input:
//
// A fill path of A (a directed path i -> ... -> j whose intermediates are all
// below min(i, j); Rose and Tarjan, SIAM J. Appl. Math. 34, 1978) is also a
// path of the undirected graph of S = A + A^T with the same property, so
// struct(L + U) lies inside struct(C + C^T) for the Cholesky factor C of S.
// Row i of L therefore has at most rc(i) entries and row i of U at most
// cc(i), where rc / cc are the strict row / column counts of C.  Both come
// from the elimination tree of S in O(nnz(S) alpha(n)) time with the
// skeleton-leaf method of Gilbert, Ng and Peyton (SIAM J. Matrix Anal.
// Appl. 15, 1994):
//   * row i's pattern in C is its row subtree T_i: the union of the tree
//     paths from each j < i adjacent to i up to (not including) i.  Its
//     leaves are the adjacent j that have no adjacent descendant; in
//     postorder, j is one iff first(j) (the smallest postorder number in j's
//     subtree) exceeds the postorder number of the previous adjacent vertex
//     of row i;
//   * |T_i| = sum over its leaves l_1 < l_2 < ... (postorder) of
//     depth(l_t) - depth(q_t), with q_1 = i and q_t = lca(l_{t-1}, l_t);
//   * cc(k) = #{i : k in T_i} = the subtree sum at k of weights +1 at every
//     leaf, -1 at every lca(l_{t-1}, l_t) and -1 at i (once per nonempty T_i);
//   * the lca of an earlier vertex with the current one (in postorder) is
//     the representative of its set in a union-find whose vertices are
//     linked to their parent when finished (Tarjan's offline lca).
#include <algorithm>
#include <cstdint>
#include <vector>

#include "gsofa.h"

namespace gsofa {

namespace {

int32_t find_root(std::vector<int32_t>& anc, int32_t x) {
    int32_t r = x;
    while (anc[r] != r) r = anc[r];
    while (anc[x] != r) {  // compress the path just walked
        const int32_t nx = anc[x];
        anc[x] = r;
        x = nx;
    }
    return r;
}

}  // namespace

// rp/ci: packed rows (sorted, duplicate-free), n rows.  On return
// Lb[i] >= |L_i|, Ub[i] >= |U_i| (both strict, A's own entries included).
void lu_row_bounds(int32_t n, const int32_t* rp, const int32_t* ci, std::vector<int64_t>& Lb,
                   std::vector<int64_t>& Ub) {
    Lb.assign((size_t)n, 0);
    Ub.assign((size_t)n, 0);
    if (n <= 0) return;
    // S's strictly lower entries as (hi, lo) pairs, in two layouts:
    // by hi (row hi's lower neighbours) and by lo (column lo's rows below)
    std::vector<int64_t> cnt_hi((size_t)n + 1, 0), cnt_lo((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; i++)
        for (int32_t p = rp[i]; p < rp[i + 1]; p++) {
            const int32_t j = ci[p];
            if (j == i) continue;
            cnt_hi[(size_t)std::max(i, j) + 1]++;
            cnt_lo[(size_t)std::min(i, j) + 1]++;
        }
    for (int32_t i = 0; i < n; i++) {
        cnt_hi[(size_t)i + 1] += cnt_hi[(size_t)i];
        cnt_lo[(size_t)i + 1] += cnt_lo[(size_t)i];
    }
    std::vector<int32_t> by_hi((size_t)cnt_hi[(size_t)n]), by_lo((size_t)cnt_lo[(size_t)n]);
    {
        std::vector<int64_t> ph(cnt_hi.begin(), cnt_hi.end() - 1), pl(cnt_lo.begin(), cnt_lo.end() - 1);
        for (int32_t i = 0; i < n; i++)
            for (int32_t p = rp[i]; p < rp[i + 1]; p++) {
                const int32_t j = ci[p];
                if (j == i) continue;
                const int32_t hi = std::max(i, j), lo = std::min(i, j);
                by_hi[(size_t)ph[(size_t)hi]++] = lo;
                by_lo[(size_t)pl[(size_t)lo]++] = hi;
            }
    }
    // elimination tree of S (Liu's ancestor algorithm with path compression)
    std::vector<int32_t> parent((size_t)n, -1), anc((size_t)n, -1);
    for (int32_t i = 0; i < n; i++)
        for (int64_t p = cnt_hi[(size_t)i]; p < cnt_hi[(size_t)i + 1]; p++) {
            int32_t r = by_hi[(size_t)p];
            while (anc[(size_t)r] != -1 && anc[(size_t)r] != i) {
                const int32_t t = anc[(size_t)r];
                anc[(size_t)r] = i;
                r = t;
            }
            if (anc[(size_t)r] == -1) {
                anc[(size_t)r] = i;
                parent[(size_t)r] = i;
            }
        }
    // depth (parents are numbered above their children) and a postorder
    std::vector<int32_t> depth((size_t)n, 0), head((size_t)n, -1), nxt((size_t)n, -1);
    for (int32_t j = n - 1; j >= 0; j--) {
        const int32_t pj = parent[(size_t)j];
        depth[(size_t)j] = pj < 0 ? 0 : depth[(size_t)pj] + 1;
        if (pj >= 0) {  // children listed in increasing order
            nxt[(size_t)j] = head[(size_t)pj];
            head[(size_t)pj] = j;
        }
    }
    std::vector<int32_t> post((size_t)n), stack;
    int32_t k = 0;
    for (int32_t root = 0; root < n; root++) {
        if (parent[(size_t)root] != -1) continue;
        stack.push_back(root);
        while (!stack.empty()) {
            const int32_t v = stack.back();
            const int32_t c = head[(size_t)v];
            if (c >= 0) {  // descend into the next unvisited child
                head[(size_t)v] = nxt[(size_t)c];
                stack.push_back(c);
            } else {
                stack.pop_back();
                post[(size_t)k++] = v;
            }
        }
    }
    std::vector<int32_t> pnum((size_t)n), first((size_t)n, -1);
    for (int32_t q = 0; q < n; q++) {
        pnum[(size_t)post[(size_t)q]] = q;
        for (int32_t v = post[(size_t)q]; v >= 0 && first[(size_t)v] < 0; v = parent[(size_t)v])
            first[(size_t)v] = q;
    }
    // skeleton leaves, row counts and column-count weights, in postorder
    std::vector<int32_t> last_adj((size_t)n, -1), prev_leaf((size_t)n, -1), set((size_t)n);
    std::vector<int64_t> w((size_t)n, 0);
    for (int32_t v = 0; v < n; v++) set[(size_t)v] = v;
    for (int32_t q = 0; q < n; q++) {
        const int32_t j = post[(size_t)q];
        for (int64_t p = cnt_lo[(size_t)j]; p < cnt_lo[(size_t)j + 1]; p++) {
            const int32_t i = by_lo[(size_t)p];  // S(i, j) != 0 with i > j
            const bool leaf = first[(size_t)j] > last_adj[(size_t)i];
            last_adj[(size_t)i] = q;
            if (!leaf) continue;
            w[(size_t)j] += 1;
            const int32_t pl = prev_leaf[(size_t)i];
            int32_t top;  // where the path from j joins what row i already holds
            if (pl < 0) {
                top = i;
                w[(size_t)i] -= 1;
            } else {
                top = find_root(set, pl);
                w[(size_t)top] -= 1;
            }
            Lb[(size_t)i] += depth[(size_t)j] - depth[(size_t)top];
            prev_leaf[(size_t)i] = j;
        }
        if (parent[(size_t)j] >= 0) set[(size_t)j] = parent[(size_t)j];
    }
    for (int32_t j = 0; j < n; j++)  // subtree sums: children precede parents
        if (parent[(size_t)j] >= 0) w[(size_t)parent[(size_t)j]] += w[(size_t)j];
    for (int32_t j = 0; j < n; j++) Ub[(size_t)j] = w[(size_t)j];
}

}  // namespace gsofa
