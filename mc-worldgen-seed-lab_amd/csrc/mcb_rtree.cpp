// mcb_rtree.cpp - the biome R-tree built from a version's parameter list (Climate.RTree.create; the
// reference's port is engine/mc_biomes.h:64-212, loaded by engine/mc_data.h:153-183), laid out for the
// device in DFS pre-order with escape links (McbNode, mcb_common.h).
//
// The steps are the game's and must stay so: the tree's shape decides which leaf a tie picks.  A group
// of at most `per_node` entries becomes one subtree ordered by centre magnitude; a larger group is sorted
// along each of the seven dimensions in turn (stable, every sort continuing from the previous order, with
// the remaining dimensions as tie-breaks), cut into buckets of per_node^floor(log(n - 0.01)/log(per_node))
// entries, the dimension whose buckets have the least summed extent wins (the first one on equal cost),
// and its buckets, ordered by the absolute centres of their boxes from that dimension on, are built
// recursively.  Centres are (lo + hi) / 2 truncated toward zero (Java long division).  per_node = 6 before
// 26.3, 19 from 26.3.  26.4-snapshot-2 measures with an exclusive upper bound; the tree is built from the
// listed values and every upper bound lowered by one afterwards (engine/mc_data.h:178-180).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "mcb_internal.h"

namespace mcb {
namespace {

constexpr int kDims = 7;                // temperature, humidity, continentalness, erosion, depth, weirdness, offset

struct Box {
    int64_t lo[kDims], hi[kDims];
};

// nodes under construction: leaves first (in list order), then subtrees as they are made
struct Forest {
    std::vector<Box> box;
    std::vector<int> biome;             // leaf biome id, -1 for a subtree
    std::vector<std::vector<int>> kids;
};

int64_t centre(const Box &b, int d) { return (b.lo[d] + b.hi[d]) / 2; }

Box enclose(const Forest &F, const int *ids, int n) {
    Box r = F.box[ids[0]];
    for (int i = 1; i < n; i++)
        for (int d = 0; d < kDims; d++) {
            r.lo[d] = std::min(r.lo[d], F.box[ids[i]].lo[d]);
            r.hi[d] = std::max(r.hi[d], F.box[ids[i]].hi[d]);
        }
    return r;
}

int64_t extent(const Box &b) {
    int64_t s = 0;
    for (int d = 0; d < kDims; d++) s += std::llabs(b.hi[d] - b.lo[d]);
    return s;
}

// lexicographic over the centres of dimensions first, first+1, ... (mod 7), optionally by magnitude
bool centre_less(const Box &a, const Box &b, int first, bool absolute) {
    for (int k = 0; k < kDims; k++) {
        const int d = (first + k) % kDims;
        int64_t ca = centre(a, d), cb = centre(b, d);
        if (absolute) { ca = std::llabs(ca); cb = std::llabs(cb); }
        if (ca != cb) return ca < cb;
    }
    return false;
}

int64_t magnitude(const Box &b) {
    int64_t s = 0;
    for (int d = 0; d < kDims; d++) s += std::llabs(centre(b, d));
    return s;
}

int subtree(Forest &F, const std::vector<int> &kids) {
    const int id = (int)F.box.size();
    F.box.push_back(enclose(F, kids.data(), (int)kids.size()));
    F.biome.push_back(-1);
    F.kids.push_back(kids);
    return id;
}

int grow(Forest &F, std::vector<int> ids, int per_node) {
    const int n = (int)ids.size();
    if (n == 1) return ids[0];
    if (n <= per_node) {
        std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return magnitude(F.box[a]) < magnitude(F.box[b]); });
        return subtree(F, ids);
    }
    const int bucket = (int)std::pow((double)per_node, std::floor(std::log((double)n - 0.01) / std::log((double)per_node)));
    std::vector<int> order = ids, best;
    int64_t best_cost = INT64_MAX;
    int best_dim = -1;
    for (int d = 0; d < kDims; d++) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return centre_less(F.box[a], F.box[b], d, false); });
        int64_t cost = 0;
        for (int s = 0; s < n; s += bucket) cost += extent(enclose(F, order.data() + s, std::min(bucket, n - s)));
        if (cost < best_cost) {
            best_cost = cost;
            best_dim = d;
            best = order;
        }
    }
    struct Bucket { Box box; int start, len; };
    std::vector<Bucket> buckets;
    for (int s = 0; s < n; s += bucket) {
        const int len = std::min(bucket, n - s);
        buckets.push_back({enclose(F, best.data() + s, len), s, len});
    }
    std::stable_sort(buckets.begin(), buckets.end(),
                     [&](const Bucket &a, const Bucket &b) { return centre_less(a.box, b.box, best_dim, true); });
    std::vector<int> kids;
    for (const Bucket &b : buckets) kids.push_back(grow(F, std::vector<int>(best.begin() + b.start, best.begin() + b.start + b.len), per_node));
    return subtree(F, kids);
}

// distance contribution of one dimension as the game measures it: above = t - hi, below = lo - t
int64_t axis_distance(int64_t lo, int64_t hi, int64_t t) {
    const int64_t above = t - hi, below = lo - t;
    return above > 0 ? above : (below > 0 ? below : 0);
}

bool emit(const Forest &F, int id, int64_t hi_shift, std::vector<McbNode> &out, std::string &err) {
    const size_t my = out.size();
    out.push_back(McbNode());
    McbNode n;
    memset(&n, 0, sizeof n);
    const Box &b = F.box[id];
    for (int d = 0; d < 6; d++) {
        const int64_t lo = b.lo[d], hi = b.hi[d] - hi_shift;
        if (lo < -32768 || lo > 32767 || hi < -32768 || hi > 32767) { err = "parameter outside int16"; return false; }
        n.lo[d] = (int16_t)lo;
        n.hi[d] = (int16_t)hi;
    }
    const int64_t x = axis_distance(b.lo[6], b.hi[6] - hi_shift, 0);      // offset: target 0
    if (x > 65535) { err = "offset parameter out of range"; return false; }
    n.d6 = (uint32_t)(x * x);
    n.leaf = F.biome[id] >= 0 ? (uint16_t)(0x8000 | F.biome[id]) : 0;
    for (int k : F.kids[id])
        if (!emit(F, k, hi_shift, out, err)) return false;
    if (out.size() > 65535) { err = "tree too large"; return false; }
    n.escape = (uint16_t)out.size();
    out[my] = n;
    return true;
}

}  // namespace

int build_tree(const std::vector<ParamEntry> &list, int per_node, bool exclusive_hi, std::vector<McbNode> &out,
               std::string &err) {
    out.clear();
    if (list.empty()) { err = "empty parameter list"; return -1; }
    Forest F;
    std::vector<int> leaves;
    for (const ParamEntry &p : list) {
        Box b;
        for (int d = 0; d < kDims; d++) { b.lo[d] = p.lo[d]; b.hi[d] = p.hi[d]; }
        leaves.push_back((int)F.box.size());
        F.box.push_back(b);
        F.biome.push_back(p.biome);
        F.kids.emplace_back();
    }
    const int root = grow(F, leaves, per_node);
    return emit(F, root, exclusive_hi ? 1 : 0, out, err) ? 0 : -1;
}

}  // namespace mcb
