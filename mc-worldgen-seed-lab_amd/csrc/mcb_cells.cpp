// mcb_cells.cpp - the climate cell table that replaces the per-sample R-tree walk.
//
// The reference resolves a climate target by searching the R-tree for the leaf of least squared
// distance (engine/mc_biomes.h:31-62); the engine's result is the leaf minimising (distance, pre-order
// index).  Every leaf box is a product of intervals whose ends come from a small set of breakpoints
// per dimension (6-16), so the target space splits into cells (elementary intervals between
// consecutive breakpoints, per dimension).  For each cell the table lists, in ascending pre-order
// index, every leaf that can be the minimum at some target inside the cell; a sample then evaluates
// only its cell's list (typically 2-4 leaves) instead of a few hundred tree nodes.
//
// A leaf L is left out of a cell B only when another leaf S dominates it over all of B:
//     min over p in B of  d_L(p) - d_S(p)  >  0       (or = 0 when S has the smaller index).
// Both distances are sums over dimensions of squared interval distances, so the minimum of the
// difference is the sum of per-dimension minima, each found exactly among the interval ends and the
// cell ends (the difference is piecewise quadratic, convex pieces have their vertex at an interval
// end, concave and linear pieces take their minimum at a piece end).  A dominated leaf is never the
// (distance, index) minimum of any target in B, so the list keeps the exact result; the same test on
// a subtree (its node box bounds every leaf below it from below) prunes whole subtrees while the list
// is collected.  Cells are bounded by +-kCellTMax; a target outside that range walks the tree.
// The strict variant drops a leaf only when another is nearer everywhere in the cell (tot > 0), so its
// lists also hold every leaf that can tie at the least distance: ties are decided from the list alone.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "mcb_internal.h"

namespace mcb {
namespace {

struct Iv {
    int64_t lo, hi;
};

inline int64_t dsq(int64_t x, int64_t lo, int64_t hi) {
    const int64_t d = x < lo ? lo - x : (x > hi ? x - hi : 0);
    return d * d;
}

// min over integer x in [a.lo, a.hi] of dist(x, A)^2 - dist(x, S)^2
inline int64_t gmin(Iv c, int64_t al, int64_t ah, int64_t sl, int64_t sh) {
    if (sl <= c.lo && c.hi <= sh) {               // S covers the cell: difference = dist(x, A)^2 >= 0
        const int64_t gap = al > c.hi ? al - c.hi : (ah < c.lo ? c.lo - ah : 0);
        return gap * gap;
    }
    const int64_t xs[6] = {c.lo, c.hi, std::clamp(al, c.lo, c.hi), std::clamp(ah, c.lo, c.hi),
                           std::clamp(sl, c.lo, c.hi), std::clamp(sh, c.lo, c.hi)};
    int64_t m = INT64_MAX;
    for (int64_t x : xs) m = std::min(m, dsq(x, al, ah) - dsq(x, sl, sh));
    return m;
}

// Does leaf S dominate node `ni` (and so every leaf of its subtree [ni, esc)) over the cell?  Strict:
// only when S is nearer everywhere (an equally near leaf may tie, so it stays listed).
inline bool dominates(const McbNode &S, int si, const McbNode &N, int ni, int esc, const Iv *B, bool strict) {
    if (si >= ni && si < esc) return false;
    int64_t tot = (int64_t)N.d6 - (int64_t)S.d6;
    for (int k = 0; k < 6; k++) tot += gmin(B[k], N.lo[k], N.hi[k], S.lo[k], S.hi[k]);
    return tot > 0 || (!strict && tot == 0 && si < ni);
}

// exact (distance, pre-order index) minimum for one target, by walking the pre-order array
int best_leaf(const std::vector<McbNode> &T, const int64_t *t) {
    const int n = (int)T.size();
    int bi = n;
    int64_t bd = INT64_MAX;
    int cur = 1;
    while (cur < n) {
        const McbNode &nd = T[cur];
        int64_t d = nd.d6;
        for (int k = 0; k < 6; k++) d += dsq(t[k], nd.lo[k], nd.hi[k]);
        const bool better = d < bd || (d == bd && cur < bi);
        const bool leaf = nd.leaf != 0;
        if (better && leaf) {
            bd = d;
            bi = cur;
        }
        cur = (better && !leaf) ? cur + 1 : (int)nd.escape;
    }
    return bi;
}

struct Builder {
    const std::vector<McbNode> &T;
    const CellTable &ct;
    bool strict = false;
    int ncell[6];
    std::vector<std::vector<uint16_t>> lists;     // per cell
    std::atomic<int> next{0};

    Builder(const std::vector<McbNode> &t, const CellTable &c) : T(t), ct(c) {
        for (int k = 0; k < 6; k++) ncell[k] = c.nb[k] + 1;
    }
    Iv cell_iv(int k, int i) const {
        const int32_t *b = ct.bp.data() + ct.off[k];
        return {i == 0 ? -(int64_t)kCellTMax : (int64_t)b[i - 1], i == ct.nb[k] ? (int64_t)kCellTMax : (int64_t)b[i] - 1};
    }

    // one 5-dimension cell (all depth cells)
    void build5(int c5) {
        int ci[6];
        int r = c5;
        const int order5[5] = {5, 3, 2, 1, 0};     // index = ((((cT*nH + cH)*nC + cC)*nE + cE)*nW + cW)
        for (int q = 0; q < 5; q++) {
            const int k = order5[q];
            ci[k] = r % ncell[k];
            r /= ncell[k];
        }
        Iv B[6];
        for (int k = 0; k < 6; k++) B[k] = k == 4 ? Iv{-(int64_t)kCellTMax, (int64_t)kCellTMax} : cell_iv(k, ci[k]);
        // dominators: exact minima at the cell centre for probe depths across every depth cell
        std::vector<int> D;
        int64_t p[6];
        for (int k = 0; k < 6; k++) p[k] = (B[k].lo + B[k].hi) / 2;
        for (int dc = 0; dc < ncell[4]; dc++) {
            const Iv d = cell_iv(4, dc);
            for (int64_t v : {d.lo, (d.lo + d.hi) / 2, d.hi}) {
                p[4] = v;
                const int b = best_leaf(T, p);
                if (std::find(D.begin(), D.end(), b) == D.end()) D.push_back(b);
            }
        }
        // superset over the whole depth range: subtrees dominated by some probe minimum are skipped
        std::vector<int> C;
        const int n = (int)T.size();
        int cur = 1;
        while (cur < n) {
            const McbNode &nd = T[cur];
            const int esc = nd.escape;
            bool dom = false;
            for (size_t q = 0; q < D.size() && !dom; q++)
                if (dominates(T[D[q]], D[q], nd, cur, esc, B, strict)) {
                    dom = true;
                    if (q) std::swap(D[q], D[0]);      // the last successful dominator is tried first
                }
            if (dom) {
                cur = esc;
            } else if (nd.leaf) {
                C.push_back(cur);
                cur = esc;
            } else {
                cur++;
            }
        }
        // per depth cell: drop every leaf dominated by another leaf of the superset (first against the
        // probe minima, then pairwise among what is left)
        std::vector<int> C1;
        for (int dc = 0; dc < ncell[4]; dc++) {
            Iv Bd[6];
            for (int k = 0; k < 6; k++) Bd[k] = B[k];
            Bd[4] = cell_iv(4, dc);
            C1.clear();
            for (int l : C) {
                bool dom = false;
                for (size_t q = 0; q < D.size() && !dom; q++)
                    if (D[q] != l) dom = dominates(T[D[q]], D[q], T[l], l, l + 1, Bd, strict);
                if (!dom) C1.push_back(l);
            }
            std::vector<uint16_t> &out = lists[(size_t)c5 * ncell[4] + dc];
            for (size_t a = 0; a < C1.size(); a++) {
                const McbNode &L = T[C1[a]];
                bool dom = false;
                for (size_t b = 0; b < C1.size() && !dom; b++)
                    if (b != a) dom = dominates(T[C1[b]], C1[b], L, C1[a], C1[a] + 1, Bd, strict);
                if (!dom) out.push_back((uint16_t)C1[a]);
            }
        }
    }
};

}  // namespace

int build_cells(const std::vector<McbNode> &tree, CellTable &ct, std::string &err, bool strict) {
    const auto t0 = std::chrono::steady_clock::now();
    ct = CellTable();
    if (tree.size() >= 65535) {
        err = "tree too large for the cell table";
        return -1;
    }
    // breakpoints: every distinct interval end of the leaves, per dimension (a leaf ending at b touches
    // the cell starting at b in one point; the dominance test handles that point like any other)
    std::vector<int32_t> bps[6];
    for (const McbNode &nd : tree) {
        if (!nd.leaf) continue;
        for (int k = 0; k < 6; k++) {
            bps[k].push_back(nd.lo[k]);
            bps[k].push_back(nd.hi[k]);
        }
    }
    for (int k = 0; k < 6; k++) {
        std::sort(bps[k].begin(), bps[k].end());
        bps[k].erase(std::unique(bps[k].begin(), bps[k].end()), bps[k].end());
        if ((int)bps[k].size() > kCellMaxBp) {
            err = "too many breakpoints for the cell table";
            return -1;
        }
        ct.off[k] = (int)ct.bp.size();
        ct.nb[k] = (int)bps[k].size();
        ct.bp.insert(ct.bp.end(), bps[k].begin(), bps[k].end());
    }
    Builder b(tree, ct);
    b.strict = strict;
    int64_t n5 = 1;
    for (int k : {0, 1, 2, 3, 5}) n5 *= b.ncell[k];
    b.lists.resize((size_t)n5 * b.ncell[4]);
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nth; i++)
        th.emplace_back([&b, n5] {
            for (int c; (c = b.next.fetch_add(1)) < n5;) b.build5(c);
        });
    for (auto &x : th) x.join();
    ct.entry.resize(b.lists.size());
    for (size_t c = 0; c < b.lists.size(); c++) {
        const auto &l = b.lists[c];
        const size_t first = ct.list.size();
        if (l.empty() || l.size() > 255 || first >= (1u << 24)) {
            ct.entry[c] = 0;                    // count 0: the kernel walks the tree for this cell
            continue;
        }
        ct.entry[c] = (uint32_t)(first << 8) | (uint32_t)l.size();
        ct.list.insert(ct.list.end(), l.begin(), l.end());
    }
    ct.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // namespace mcb
