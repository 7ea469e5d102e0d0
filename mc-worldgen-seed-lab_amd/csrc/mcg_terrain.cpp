// mcg_terrain.cpp - host side of the terrain volume executor: the mcgpu_terrain_* exports of the libmcgen
// GPU plugin ABI (include/mcgpu.h) and the plan they run (mcg_terrain.h).
//
// Reference interface: libmcgen/gpu/mcgpu_terrain.cu:556-862 (mcgpu_terrain_new, _free, _max_chunks,
// _cell_floats, _cell_info, _cell_cand, _batch, mcgpu_host_alloc / _free), bound by
// libmcgen/src/gpu_bridge.c:39-47, 141-144 and used by gpu_bridge.c:503-551, 581-656.  What a batch returns
// is the reference's: final_density per chunk, the ore-vein patch per block, and the cache cells the CPU
// would hold after the density (which candidate volume of each cache, and its values).  How it is computed
// is this file's:
//  * the recursion of sampleVolume is walked once per world (relative to the chunk origin), not once per
//    batch, and becomes a short list of passes (mcg_terrain.h); each pass is one launch over all chunks of
//    the batch, with intermediate values in LDS instead of a kernel and an HBM buffer per node;
//  * the batch arena holds only values another pass or the caller reads;
//  * the cache-state trace follows the CPU's visiting order with the spline coordinates each chunk
//    actually read (masks collected by the pass kernel), as the reference's trace does
//    (mcgpu_terrain.cu:602-655).
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mcb_internal.h"
#include "mcg_plugin.h"
#include "mcg_terrain.h"

using namespace mcb::plug;

namespace {

int fl_div(int a, int b) { int q = a / b; if ((a % b != 0) && ((a < 0) != (b < 0))) q--; return q; }
int fl_mod(int a, int b) { int m = a % b; if (m != 0 && ((m < 0) != (b < 0))) m += b; return m; }

// Interpolation on volume v (df_new.c:1084-1100): 0 = v lies on the cell grid (the input's own values),
// 1 = unit-step volume filled cell by cell from the corner volume *cv, -1 = not covered (the reference
// executor covers the same geometry, mcgpu_terrain.cu:456-472)
int interp_plan(const McgNode &s, const TVol &v, TVol *cv) {
    const int cxz = s.i0, cy = s.i1;
    if (cxz <= 0 || cy <= 0 || 16 % cxz != 0) return -1;
    if ((v.dx == cxz || v.sx == 1) && (v.dy == cy || v.sy == 1) && (v.dz == cxz || v.sz == 1) && fl_mod(v.xoff, cxz) == 0 &&
        fl_mod(v.y0, cy) == 0 && fl_mod(v.zoff, cxz) == 0)
        return 0;
    if (!(v.dx == 1 && v.dy == 1 && v.dz == 1)) return -1;
    if (v.sx == 1 && v.xrel) return -1;
    const int maxx = v.xoff + v.sx - 1, maxy = v.y0 + v.sy - 1, maxz = v.zoff + v.sz - 1;
    const int mnx = fl_div(v.xoff, cxz), mny = fl_div(v.y0, cy), mnz = fl_div(v.zoff, cxz);
    const int ncx = fl_div(maxx, cxz) - mnx + 1, ncy = fl_div(maxy, cy) - mny + 1, ncz = fl_div(maxz, cxz) - mnz + 1;
    cv->sx = fl_mod(maxx, cxz) == 0 ? ncx : ncx + 1;
    cv->sy = fl_mod(maxy, cy) == 0 ? ncy : ncy + 1;
    cv->sz = fl_mod(maxz, cxz) == 0 ? ncz : ncz + 1;
    cv->xrel = v.xrel; cv->xoff = mnx * cxz; cv->y0 = mny * cy; cv->zrel = v.zrel; cv->zoff = mnz * cxz;
    cv->dx = cxz; cv->dy = cy; cv->dz = cxz;
    return 1;
}
// Slices (df_new.c:1102-1136): 0 = v already has the fixed coordinate, 1 = broadcast from *iv (axis 0 x,
// 1 y, 2 z, 3 x and z), -1 = not covered (a one-wide relative axis; mcgpu_terrain.cu:474-491)
int slice_plan(const McgNode &s, const TVol &v, TVol *iv, int *axis) {
    *iv = v;
    switch (s.k) {
    case MCGK_SLICE_X:
        if (v.sx == 1) return (!v.xrel && v.xoff == s.i0) ? 0 : -1;
        iv->sx = 1; iv->xrel = 0; iv->xoff = s.i0; *axis = 0;
        return 1;
    case MCGK_SLICE_Y:
        if (v.sy == 1 && v.y0 == s.i0) return 0;
        iv->sy = 1; iv->y0 = s.i0; *axis = 1;
        return 1;
    case MCGK_SLICE_Z:
        if (v.sz == 1) return (!v.zrel && v.zoff == s.i0) ? 0 : -1;
        iv->sz = 1; iv->zrel = 0; iv->zoff = s.i0; *axis = 2;
        return 1;
    default:
        if (v.sx == 1 && v.sz == 1 && !v.xrel && !v.zrel && v.xoff == s.i0 && v.zoff == s.i1) return 0;
        if (v.sx == 1 || v.sz == 1) return -1;
        iv->sx = 1; iv->sz = 1; iv->xrel = 0; iv->zrel = 0; iv->xoff = s.i0; iv->zoff = s.i1; *axis = 3;
        return 1;
    }
}

typedef std::array<int32_t, 12> VKey;       // node and volume
VKey vkey(int node, const TVol &v) { return {node, v.sx, v.sy, v.sz, v.dx, v.dy, v.dz, v.y0, v.xrel, v.xoff, v.zrel, v.zoff}; }

enum { VK_NODE, VK_INTERP, VK_SLICE, VK_BEARD, VK_ALIAS };
struct Val {
    int kind = VK_NODE;
    int node = -1;
    TVol v{};
    std::vector<int> ops;         // operands (resolved), in the node's field order: a, b, c, or its list
    int src = -1, axis = 0;       // fetched value (interp, slice)
    int alias = -1;
    int level = 0;
    bool store = false;
    int64_t off = -1;             // arena offset of the batch buffer
    int pass = -1;
    int flag = -1;                // spline instance
};
struct Pass {
    TVol v{};
    int level = 0;
    std::vector<int> vals;        // in creation order (operands first)
    // built for the device
    std::vector<TIns> ins;
    std::vector<int32_t> arr;
    int nslot = 0;
    int feat = 0, spd = 1;
    bool fused = false;           // run as k_tinterp
    TInterp fi{};
    size_t ins_off = 0, arr_off = 0;
};
struct CellInfo {
    int cid;
    std::vector<TVol> cand;
    size_t cap = 0, off = 0;
};

// The device resources of a terrain world: a stream and buffers that only grow.  A freed world hands them
// to a per-device pool and the next world made on that device takes them (a world per seed was dominated
// by stream creation, allocations and frees, 0.6 ms of its 0.9 ms).
struct Kit {
    hipStream_t st = nullptr;
    char *d_prog = nullptr, *d_gap = nullptr;   // program tables + pass plan; the veins' gap program
    float *d_arena = nullptr, *d_stage = nullptr;
    int32_t *d_cx0 = nullptr, *d_cz0 = nullptr;
    uint32_t *d_flags = nullptr;
    uint16_t *d_veins = nullptr;
    TGather *d_items = nullptr;
    size_t prog_cap = 0, gap_cap = 0, arena_cap = 0, stage_cap = 0, cx_cap = 0, cz_cap = 0, flags_cap = 0, veins_cap = 0,
           items_cap = 0;
    size_t bytes() const {
        return prog_cap + gap_cap + 4 * (arena_cap + stage_cap + cx_cap + cz_cap + flags_cap) + 2 * veins_cap + items_cap * sizeof(TGather);
    }
    void destroy() {
        if (st) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
        (void)hipFree(d_prog); (void)hipFree(d_gap); (void)hipFree(d_arena); (void)hipFree(d_stage); (void)hipFree(d_cx0);
        (void)hipFree(d_cz0); (void)hipFree(d_flags); (void)hipFree(d_veins); (void)hipFree(d_items);
        *this = Kit();
    }
};
// at most kKeepKits kits per device, none over kKeepBytes, held until the process ends
constexpr size_t kKeepKits = 2, kKeepBytes = (size_t)8 << 30;
std::mutex g_kit_mu;
std::map<int, std::vector<Kit>> g_kits;
Kit take_kit(int dev) {
    std::lock_guard<std::mutex> lk(g_kit_mu);
    std::vector<Kit> &v = g_kits[dev];
    if (v.empty()) return Kit();
    Kit k = v.back();
    v.pop_back();
    return k;
}
void give_kit(int dev, Kit k) {
    if (k.st && hipStreamSynchronize(k.st) != hipSuccess) { k.destroy(); return; }
    std::lock_guard<std::mutex> lk(g_kit_mu);
    std::vector<Kit> &v = g_kits[dev];
    if (v.size() < kKeepKits && k.bytes() <= kKeepBytes) v.push_back(k);
    else k.destroy();
}
bool hchk(hipError_t e, const char *what);
// *p holds at least n elements: kept when it does, else freed and allocated again (contents not kept)
template <typename T>
bool grow_buf(T **p, size_t *cap, size_t n, const char *what) {
    n = std::max<size_t>(n, 1);
    if (*p && *cap >= n) return true;
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (!hchk(hipMalloc((void **)p, n * sizeof(T)), what)) return false;
    *cap = n;
    return true;
}

struct Terrain {
    McgTerrainWorld T{};
    const McgProg *P = nullptr;                 // the caller's volume program, read during planning only
    std::vector<McgNode> nodes;
    std::vector<int32_t> parr;                  // the program's node lists
    int proots[MCG_MAX_ROOTS] = {};
    // plan
    std::vector<Val> vals;
    std::map<VKey, int> memo;
    std::vector<std::pair<VKey, int>> memo_order;   // (node, volume) -> value, in the order values were planned
    std::vector<Pass> passes;
    std::map<std::pair<VKey, int>, int> pass_of;    // (volume, level) -> pass
    std::vector<int> beards;
    std::vector<int> splops;                    // spline values, by instance
    std::string fail;
    int density = -1, vdens[MCG_MAX_VEINS] = {}, vrich[MCG_MAX_VEINS] = {};
    TVol vfull{};
    std::vector<CellInfo> cells;
    size_t cell_floats = 0;
    int Bmax = 0;
    size_t arena_floats = 0;
    // device: the kit's stream and buffers (d_plan points into d_prog)
    int dev = -1;
    Kit kit;
    hipStream_t &st = kit.st;
    char *&d_prog = kit.d_prog, *&d_gap = kit.d_gap;
    char *d_plan = nullptr;
    TProg prog{};
    PWorld gapw{};
    int gap_lds = 0;
    int gap_slot[MCG_MAX_VEINS] = {};
    float *&d_arena = kit.d_arena, *&d_stage = kit.d_stage;
    int32_t *&d_cx0 = kit.d_cx0, *&d_cz0 = kit.d_cz0;
    uint32_t *&d_flags = kit.d_flags;
    uint16_t *&d_veins = kit.d_veins;
    TGather *&d_items = kit.d_items;
    size_t &items_cap = kit.items_cap;
    std::vector<float> h_beard;

    Terrain() {
        if (hipGetDevice(&dev) == hipSuccess) kit = take_kit(dev);
        else dev = -1;
    }
    Terrain(const Terrain &) = delete;
    Terrain &operator=(const Terrain &) = delete;
    ~Terrain() {
        if (dev >= 0) give_kit(dev, kit);
        else kit.destroy();
    }

    int resolve(int id) const {
        while (vals[id].kind == VK_ALIAS) id = vals[id].alias;
        return id;
    }
    bool bad(const char *msg) {
        if (fail.empty()) fail = msg;
        return false;
    }
    int add(Val v) {
        vals.push_back(v);
        return (int)vals.size() - 1;
    }
    int plan(int ni, const TVol &v);
    bool trace(int chunk, bool lazy, const std::vector<uint32_t> &masks, std::vector<std::pair<bool, TVol>> &cs, int ni,
               const TVol &v);
};

// The value of node ni over v: sampleVolume's recursion (df_new.c:917-1150), children in the order it
// evaluates them.  -1 on a node kind or geometry the executor does not cover.
int Terrain::plan(int ni, const TVol &v) {
    if (!fail.empty()) return -1;
    if (ni < 0 || ni >= (int)nodes.size()) { bad("node index out of range"); return -1; }
    const VKey key = vkey(ni, v);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    const McgNode &s = nodes[ni];
    auto child = [&](int j, const TVol &w) -> int {
        if (j < 0 || j >= ni) { bad("node reads a later node"); return -1; }
        const int r = plan(j, w);
        return r < 0 ? -1 : resolve(r);
    };
    Val x;
    x.node = ni;
    x.v = v;
    int id = -1;
    switch (s.k) {
    case MCGK_CACHE: case MCGK_BLEND_DENSITY: {
        const int a = child(s.a, v);
        if (a < 0) return -1;
        x.kind = VK_ALIAS; x.alias = a;
        id = add(x);
        break;
    }
    case MCGK_INTERP: {
        TVol cv;
        const int mode = interp_plan(s, v, &cv);
        if (mode < 0) { bad("interpolation: volume geometry not covered"); return -1; }
        const int a = child(s.a, mode == 0 ? v : cv);
        if (a < 0) return -1;
        if (mode == 0) { x.kind = VK_ALIAS; x.alias = a; }
        else { x.kind = VK_INTERP; x.src = a; x.level = vals[a].level + 1; }
        id = add(x);
        break;
    }
    case MCGK_SLICE_X: case MCGK_SLICE_Y: case MCGK_SLICE_Z: case MCGK_SLICE_XZ: {
        TVol iv;
        int axis = 0;
        const int mode = slice_plan(s, v, &iv, &axis);
        if (mode < 0) { bad("slice: volume geometry not covered"); return -1; }
        const int a = child(s.a, mode == 0 ? v : iv);
        if (a < 0) return -1;
        if (mode == 0) { x.kind = VK_ALIAS; x.alias = a; }
        else { x.kind = VK_SLICE; x.src = a; x.axis = axis; x.level = vals[a].level + 1; }
        id = add(x);
        break;
    }
    case MCGK_CTX_BEARD:
        if (s.f0 != 0.0f) { bad("Beardifier with a non-zero default value"); return -1; }
        x.kind = VK_BEARD;
        x.store = true;
        id = add(x);
        beards.push_back(id);
        break;
    case MCGK_LOG: case MCGK_ROUND_INT: case MCGK_ROUND: case MCGK_POW_CB: case MCGK_POW_CE: case MCGK_POW: case MCGK_FTS: {
        char m[96];
        snprintf(m, sizeof m, "node kind %d is not covered by the volume executor", s.k);
        bad(m);
        return -1;
    }
    default: {
        if (s.k < 0 || s.k >= MCGK__COUNT) { bad("unknown node kind"); return -1; }
        auto op = [&](int j) {
            const int r = child(j, v);
            if (r >= 0) x.ops.push_back(r);
            return r >= 0;
        };
        bool ok = true;
        switch (s.k) {
        case MCGK_CONST: case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: case MCGK_NOISE: case MCGK_SHIFT_B: case MCGK_END:
        case MCGK_DIST: case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR:
            break;
        case MCGK_NOISE_XZ: ok = op(s.a) && op(s.c); break;
        case MCGK_NOISE_XYZ: ok = op(s.a) && op(s.b) && op(s.c); break;
        case MCGK_ADD: case MCGK_SUB: case MCGK_MUL: case MCGK_DIV: case MCGK_MIN: case MCGK_MAX: ok = op(s.a) && op(s.b); break;
        case MCGK_LERP: ok = op(s.a) && op(s.b) && op(s.c); break;
        case MCGK_LERP_CF: ok = op(s.a) && op(s.c); break;
        case MCGK_LERP_CS: ok = op(s.a) && op(s.b); break;
        case MCGK_RANGE: ok = op(s.b) && op(s.a) && op(s.c); break;        // when_in first, as the CPU does
        case MCGK_ISEL1: ok = op(s.a) && op(s.b) && op(s.c); break;
        case MCGK_ISEL: case MCGK_SPLINE:
            if (s.narr < 1 || s.arr0 < 0 || s.arr0 + s.narr > (int)parr.size()) { bad("node list out of range"); return -1; }
            if (s.k == MCGK_SPLINE && s.narr > 32) { bad("spline with more than 32 coordinates"); return -1; }
            if (s.k == MCGK_ISEL) ok = op(s.a);
            for (int q = 0; ok && q < s.narr; q++) ok = op(parr[s.arr0 + q]);
            break;
        default: ok = op(s.a); break;           // the unary kinds
        }
        if (!ok) return -1;
        for (int o : x.ops) x.level = std::max(x.level, vals[o].level);
        id = add(x);
        if (s.k == MCGK_SPLINE) { vals[id].flag = (int)splops.size(); splops.push_back(id); }
        break;
    }
    }
    memo[key] = id;
    memo_order.push_back({key, id});
    return id;
}

// Cache-cell state of one chunk as the CPU's sampleVolume leaves it (mcgpu_terrain.cu:602-655): a cache
// visited with the volume it holds is a hit and its input is not visited again; spline coordinates are
// visited when the chunk's mask says the spline read them (lazy) or always.
bool Terrain::trace(int chunk, bool lazy, const std::vector<uint32_t> &masks, std::vector<std::pair<bool, TVol>> &cs,
                    int ni, const TVol &v) {
    if (ni < 0) return true;
    const McgNode &s = nodes[ni];
    auto rec = [&](int j, const TVol &w) { return trace(chunk, lazy, masks, cs, j, w); };
    switch (s.k) {
    case MCGK_CONST: case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: case MCGK_CTX_BEARD: case MCGK_NOISE: case MCGK_SHIFT_B:
    case MCGK_END: case MCGK_DIST: case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR:
        return true;
    case MCGK_NOISE_XZ: return rec(s.a, v) && rec(s.c, v);
    case MCGK_NOISE_XYZ: return rec(s.a, v) && rec(s.b, v) && rec(s.c, v);
    case MCGK_ADD: case MCGK_SUB: case MCGK_MUL: case MCGK_DIV: case MCGK_MIN: case MCGK_MAX: return rec(s.a, v) && rec(s.b, v);
    case MCGK_LERP: case MCGK_ISEL1: return rec(s.a, v) && rec(s.b, v) && rec(s.c, v);
    case MCGK_LERP_CF: return rec(s.a, v) && rec(s.c, v);
    case MCGK_LERP_CS: return rec(s.a, v) && rec(s.b, v);
    case MCGK_RANGE: return rec(s.b, v) && rec(s.a, v) && rec(s.c, v);
    case MCGK_ISEL: {
        if (!rec(s.a, v)) return false;
        for (int k = 0; k < s.narr; k++)
            if (!rec(parr[s.arr0 + k], v)) return false;
        return true;
    }
    case MCGK_SPLINE: {
        uint32_t mask = 0xFFFFFFFFu;
        if (lazy) {
            auto it = memo.find(vkey(ni, v));
            if (it == memo.end() || vals[it->second].flag < 0) return false;
            mask = masks[(size_t)vals[it->second].flag * Bmax + chunk];
        }
        for (int k = 0; k < s.narr; k++)
            if ((mask >> k & 1) && !rec(parr[s.arr0 + k], v)) return false;
        return true;
    }
    case MCGK_CACHE: {
        if (s.cid < 0) return rec(s.a, v);
        if (s.cid >= (int)cs.size()) cs.resize((size_t)s.cid + 1, {false, TVol{}});
        if (cs[s.cid].first && cs[s.cid].second == v) return true;
        cs[s.cid] = {true, v};
        return rec(s.a, v);
    }
    case MCGK_BLEND_DENSITY: return rec(s.a, v);
    case MCGK_INTERP: {
        TVol cv;
        const int mode = interp_plan(s, v, &cv);
        return mode >= 0 && rec(s.a, mode == 0 ? v : cv);
    }
    case MCGK_SLICE_X: case MCGK_SLICE_Y: case MCGK_SLICE_Z: case MCGK_SLICE_XZ: {
        TVol iv;
        int axis = 0;
        const int mode = slice_plan(s, v, &iv, &axis);
        return mode >= 0 && rec(s.a, mode == 0 ? v : iv);
    }
    case MCGK_FTS: return rec(s.b, v);
    case MCGK_ABS: case MCGK_SQUARE: case MCGK_CUBE: case MCGK_SQRT: case MCGK_LEAKY: case MCGK_RECIP: case MCGK_NEG:
    case MCGK_SQUEEZE: case MCGK_SIGN: case MCGK_CADD: case MCGK_CSUB: case MCGK_CMUL: case MCGK_CDIV: case MCGK_CMIN:
    case MCGK_CMAX: case MCGK_CLAMP: case MCGK_RANGE_C:
        return rec(s.a, v);
    default: return false;
    }
}

// ---------------------------------------------------------------------------------------------------
// passes: values grouped by (volume, level), instructions with LDS slots, arena offsets
// ---------------------------------------------------------------------------------------------------
bool build_passes(Terrain &g) {
    for (size_t i = 0; i < g.vals.size(); i++) {
        Val &x = g.vals[i];
        if (x.kind == VK_ALIAS || x.kind == VK_BEARD) continue;
        const auto key = std::make_pair(vkey(0, x.v), x.level);
        auto it = g.pass_of.find(key);
        int p;
        if (it == g.pass_of.end()) {
            p = (int)g.passes.size();
            g.passes.push_back(Pass());
            g.passes[p].v = x.v;
            g.passes[p].level = x.level;
            g.pass_of[key] = p;
        } else {
            p = it->second;
        }
        x.pass = p;
        g.passes[p].vals.push_back((int)i);
    }
    // reads across passes: the operand is written to the arena
    for (const Val &x : g.vals) {
        if (x.kind == VK_ALIAS || x.kind == VK_BEARD) continue;
        for (int o : x.ops)
            if (g.vals[o].pass != x.pass) g.vals[o].store = true;
        if (x.src >= 0) g.vals[x.src].store = true;
    }
    std::stable_sort(g.passes.begin(), g.passes.end(), [](const Pass &a, const Pass &b) { return a.level < b.level; });
    for (size_t p = 0; p < g.passes.size(); p++)
        for (int i : g.passes[p].vals) g.vals[i].pass = (int)p;
    return true;
}

bool is_unary(int k) {
    switch (k) {
    case MCGK_ABS: case MCGK_SQUARE: case MCGK_CUBE: case MCGK_SQRT: case MCGK_LEAKY: case MCGK_RECIP: case MCGK_NEG:
    case MCGK_SQUEEZE: case MCGK_SIGN: case MCGK_CADD: case MCGK_CSUB: case MCGK_CMUL: case MCGK_CDIV: case MCGK_CMIN:
    case MCGK_CMAX: case MCGK_CLAMP: case MCGK_RANGE_C:
        return true;
    default:
        return false;
    }
}

// A pass of one interpolation over a cell-aligned unit-step volume and a chain of unary nodes, with only
// the chain's end written out, runs as k_tinterp (mcg_terrain.h).
void fuse_interp(Pass &P) {
    const std::vector<TIns> &I = P.ins;
    if (I.empty() || I[0].op != TOP_INTERP || (int)I.size() - 1 > kMaxTail || I[0].g4 < 0) return;
    const TVol &v = P.v;
    const int cxz = I[0].i0, cy = I[0].i1;
    if (cy % 4 || v.dx != 1 || v.dy != 1 || v.dz != 1 || fl_mod(v.xoff, cxz) || fl_mod(v.zoff, cxz) || fl_mod(v.y0, cy) ||
        v.sx % cxz || v.sz % cxz || v.sy % cy || (!v.xrel && fl_mod(v.xoff, cxz)) || (!v.zrel && fl_mod(v.zoff, cxz)))
        return;
    for (size_t k = 1; k < I.size(); k++)
        if (!is_unary(I[k].op) || I[k].a != I[k - 1].dst || I[k - 1].store >= 0) return;
    if (I.back().store < 0) return;
    TInterp &f = P.fi;
    f.src = I[0].src;
    f.store = I.back().store;
    f.ncv = I[0].g0;
    f.cvx = I[0].g1; f.cvy = I[0].g2; f.cvz = I[0].g3;
    f.lx = I[0].g4 & 255; f.ly = I[0].g4 >> 8;
    f.f0 = I[0].f0; f.f1 = I[0].f1;
    f.ntail = (int)I.size() - 1;
    for (int k = 0; k < f.ntail; k++) {
        const TIns &q = I[k + 1];
        f.op[k] = q.op; f.k0[k] = q.f0; f.k1[k] = q.f1; f.k2[k] = q.f2; f.k3[k] = q.f3;
    }
    P.fused = true;
}

// levels of spline calls from spline spi down to its deepest leaf (constants count as a level)
int spline_nesting(const McgProg *p, int spi) {
    int best = 0;
    std::vector<std::pair<int, int>> st{{spi, 1}};
    while (!st.empty()) {
        const auto [i, d] = st.back();
        st.pop_back();
        best = std::max(best, d);
        const McgSpline &q = p->sp[i];
        if (q.is_const) continue;
        for (int k = 0; k < q.n; k++) st.push_back({p->spc[q.val0 + k], d + 1});
    }
    return best;
}

// arena offsets for a batch of Bmax chunks, then each pass's instruction list
bool emit_passes(Terrain &g) {
    size_t total = 0;
    for (Val &x : g.vals) {
        if (!x.store || x.kind == VK_ALIAS) continue;
        x.off = (int64_t)total;
        total += ((size_t)g.Bmax * (size_t)x.v.size() + 63) & ~(size_t)63;
    }
    g.arena_floats = std::max<size_t>(total, 64);
    for (Pass &P : g.passes) {
        // instruction order: the pass's values, each preceded by loads of operands from other passes
        struct Item { int val; int load; };           // load >= 0: a load of that value
        std::vector<Item> items;
        std::map<int, int> loaded;                    // value -> item index of its load
        for (int i : P.vals) {
            const Val &x = g.vals[i];
            for (int o : x.ops)
                if (g.vals[o].pass != x.pass && !loaded.count(o)) {
                    loaded[o] = (int)items.size();
                    items.push_back({-1, o});
                }
            items.push_back({i, -1});
        }
        // the item that defines an operand's value inside this pass
        std::map<int, int> def;
        for (size_t k = 0; k < items.size(); k++) def[items[k].val >= 0 ? items[k].val : items[k].load] = (int)k;
        const int ni = (int)items.size();
        std::vector<int> last(ni, -1);
        for (int k = 0; k < ni; k++)
            if (items[k].val >= 0)
                for (int o : g.vals[items[k].val].ops) last[def[o]] = std::max(last[def[o]], k);
        // linear-scan slots
        std::vector<int> slot(ni, -1), freed;
        std::vector<std::vector<int>> dies(ni);
        for (int k = 0; k < ni; k++)
            if (last[k] >= 0) dies[last[k]].push_back(k);
        int ns = 0;
        for (int k = 0; k < ni; k++) {
            for (int d : dies[k]) freed.push_back(slot[d]);         // the kernel reads operands before it writes
            std::sort(freed.begin(), freed.end(), std::greater<int>());
            int s;
            if (!freed.empty()) { s = freed.back(); freed.pop_back(); }
            else s = ns++;
            slot[k] = s;
            if (last[k] < 0) freed.push_back(s);
        }
        P.nslot = std::max(ns, 1);
        if (256 * 8 + P.nslot * kTWG * 4 > kMaxLdsBytes) return g.bad("a pass keeps too many values alive at once");
        for (int k = 0; k < ni; k++) {
            TIns q;
            memset(&q, 0, sizeof q);
            q.dst = slot[k];
            q.store = -1;
            q.src = -1;
            q.flag = -1;
            if (items[k].load >= 0) {
                const Val &u = g.vals[items[k].load];
                q.op = TOP_LOAD;
                q.src = u.off;
                q.g0 = u.v.size();
                P.ins.push_back(q);
                continue;
            }
            const Val &x = g.vals[items[k].val];
            if (x.store) q.store = x.off;
            if (x.kind == VK_INTERP || x.kind == VK_SLICE) {
                const Val &u = g.vals[x.src];
                q.src = u.off;
                q.g0 = u.v.size();
                if (x.kind == VK_INTERP) {
                    const McgNode &s = g.nodes[x.node];
                    q.op = TOP_INTERP;
                    q.i0 = s.i0; q.i1 = s.i1; q.f0 = s.f0; q.f1 = s.f1;
                    q.g1 = u.v.sx; q.g2 = u.v.sy; q.g3 = u.v.sz;
                    auto log2_of = [](int v) { return (v > 0 && (v & (v - 1)) == 0) ? __builtin_ctz((unsigned)v) : -1; };
                    const int lx = log2_of(s.i0), ly = log2_of(s.i1);
                    q.g4 = (lx >= 0 && ly >= 0) ? (lx | ly << 8) : -1;
                } else {
                    q.op = TOP_SLICE;
                    q.g1 = x.axis;
                }
                P.ins.push_back(q);
                continue;
            }
            const McgNode &s = g.nodes[x.node];
            q.op = s.k;
            q.ns = s.ns; q.sp = s.sp; q.thr0 = s.thr0; q.nthr = s.nthr;
            q.i0 = s.i0; q.i1 = s.i1; q.i2 = s.i2; q.i3 = s.i3; q.oct = s.oct;
            q.f0 = s.f0; q.f1 = s.f1; q.f2 = s.f2; q.f3 = s.f3; q.d0 = s.d0; q.d1 = s.d1;
            q.flag = x.flag;
            if (s.k == MCGK_NOISE || s.k == MCGK_NOISE_XZ || s.k == MCGK_NOISE_XYZ || s.k == MCGK_SHIFT_B || s.k == MCGK_END)
                P.feat |= TF_NOISE;
            if (s.k == MCGK_SPLINE) {
                P.feat |= TF_NOISE | TF_SPLINE;
                P.spd = std::max(P.spd, spline_nesting(g.P, s.sp));
            }
            auto sl = [&](int o) { return slot[def[o]]; };
            size_t o = 0;
            switch (s.k) {
            case MCGK_NOISE_XZ: q.a = sl(x.ops[0]); q.c = sl(x.ops[1]); break;
            case MCGK_NOISE_XYZ: case MCGK_LERP: case MCGK_ISEL1: q.a = sl(x.ops[0]); q.b = sl(x.ops[1]); q.c = sl(x.ops[2]); break;
            case MCGK_LERP_CF: q.a = sl(x.ops[0]); q.c = sl(x.ops[1]); break;
            case MCGK_LERP_CS: case MCGK_ADD: case MCGK_SUB: case MCGK_MUL: case MCGK_DIV: case MCGK_MIN: case MCGK_MAX:
                q.a = sl(x.ops[0]); q.b = sl(x.ops[1]);
                break;
            case MCGK_RANGE: q.b = sl(x.ops[0]); q.a = sl(x.ops[1]); q.c = sl(x.ops[2]); break;
            case MCGK_ISEL: case MCGK_SPLINE:
                if (s.k == MCGK_ISEL) q.a = sl(x.ops[o++]);
                q.arr0 = (int)P.arr.size();
                q.narr = s.narr;
                for (; o < x.ops.size(); o++) P.arr.push_back(sl(x.ops[o]));
                break;
            default:
                if (!x.ops.empty()) q.a = sl(x.ops[0]);
                break;
            }
            P.ins.push_back(q);
        }
        fuse_interp(P);
    }
    return true;
}

bool hchk(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    (void)hipGetLastError();
    return plugin_fail("%s: %s", what, hipGetErrorString(e));
}

// validation of the parts of a program the kernels index
bool check_prog(const McgProg *p) {
    if (!p || p->old) return plugin_fail("the volume program must be a 26.3+ (float) program");
    if (p->nnodes < 1 || !p->nodes || p->nnoise < 0 || p->nlayers < 0 || p->noct < 0 || p->nsp < 0 || p->nsplf < 0 ||
        p->nspc < 0 || p->narr < 0 || p->nthr < 0 || (p->nnoise && !p->noise) || (p->nlayers && !p->layers) ||
        (p->noct && !p->octs) || (p->nsp && !p->sp) || (p->nsplf && !p->splf) || (p->nspc && !p->spc) ||
        (p->narr && !p->arr) || (p->nthr && !p->thr) || p->nroots < 0 || p->nroots > MCG_MAX_ROOTS)
        return plugin_fail("malformed program (counts or arrays)");
    for (int i = 0; i < p->nnoise; i++) {
        const McgNoise &m = p->noise[i];
        if (m.n0 < 0 || m.l0 < 0 || m.l0 + m.n0 > p->nlayers) return plugin_fail("noise %d: layers out of range", i);
        for (int l = 0; l < m.n0; l++)
            if (p->layers[m.l0 + l].oct < 0 || p->layers[m.l0 + l].oct >= p->noct) return plugin_fail("noise %d: octave out of range", i);
    }
    for (int i = 0; i < p->nsp; i++) {
        const McgSpline &s = p->sp[i];
        if (s.is_const) continue;
        if (s.n < 1 || s.loc0 < 0 || s.loc0 + 2 * s.n > p->nsplf || s.val0 < 0 || s.val0 + s.n > p->nspc || s.coord < 0 || s.coord >= 32)
            return plugin_fail("spline %d out of range", i);
        for (int k = 0; k < s.n; k++)
            if (p->spc[s.val0 + k] < 0 || p->spc[s.val0 + k] >= p->nsp) return plugin_fail("spline %d: child out of range", i);
    }
    for (int i = 0; i < p->nnodes; i++) {
        const McgNode &s = p->nodes[i];
        const bool noise = s.k == MCGK_NOISE || s.k == MCGK_NOISE_XZ || s.k == MCGK_NOISE_XYZ || s.k == MCGK_SHIFT_B;
        if (noise && (s.ns < 0 || s.ns >= p->nnoise)) return plugin_fail("node %d: noise out of range", i);
        if (s.k == MCGK_END && (s.oct < 0 || s.oct >= p->noct)) return plugin_fail("node %d: octave out of range", i);
        if ((s.k == MCGK_GRAD_REPEAT || s.k == MCGK_GRAD_MIRROR) && s.i2 == 0) return plugin_fail("node %d: gradient period 0", i);
        if (s.k == MCGK_ISEL && (s.nthr < 0 || s.thr0 < 0 || s.thr0 + s.nthr > p->nthr || s.nthr >= s.narr))
            return plugin_fail("node %d: thresholds out of range", i);
        if (s.k == MCGK_SPLINE) {
            // every coordinate a spline of this node reads exists, and the nesting fits the kernel
            std::vector<std::pair<int, int>> st{{s.sp, 1}};
            while (!st.empty()) {
                const auto [spi, depth] = st.back();
                st.pop_back();
                if (spi < 0 || spi >= p->nsp) return plugin_fail("node %d: spline out of range", i);
                if (depth > kSplineDepth) return plugin_fail("spline nesting deeper than %d", kSplineDepth);
                const McgSpline &q = p->sp[spi];
                if (q.is_const) continue;
                if (q.coord >= s.narr) return plugin_fail("node %d: spline coordinate out of range", i);
                for (int k = 0; k < q.n; k++) st.push_back({p->spc[q.val0 + k], depth + 1});
            }
        }
    }
    return true;
}

Terrain *terrain_new(const McgTerrainWorld *W, int max_chunks) {
    std::unique_ptr<Terrain> g(new Terrain());
    if (!W || !W->vol) { plugin_fail("null world"); return nullptr; }
    if (max_chunks < 1) { plugin_fail("max_chunks must be positive"); return nullptr; }
    if (W->nh <= 0 || W->nh > 4096) { plugin_fail("noise height %d out of range", W->nh); return nullptr; }
    if (W->veins_on && (W->nveins < 0 || W->nveins > MCG_MAX_VEINS || !W->gap)) { plugin_fail("malformed vein rules"); return nullptr; }
    if (W->ncache_ids < 0 || W->ncache_ids > 32) { plugin_fail("too many cache ids"); return nullptr; }
    if (!check_prog(W->vol)) return nullptr;
    g->T = *W;
    const McgProg *p = W->vol;
    g->P = p;
    g->nodes.assign(p->nodes, p->nodes + p->nnodes);
    g->parr.assign(p->arr, p->arr + p->narr);
    for (int i = 0; i < p->nroots; i++) g->proots[i] = p->root[i];
    auto root = [&](int k) -> int {
        if (k < 0 || k >= p->nroots || p->root[k] < 0) { g->bad("root out of range"); return -1; }
        return p->root[k];
    };
    g->vfull = TVol{16, W->nh, 16, 1, 1, 1, W->nmin, 1, 0, 1, 0};
    // the plan: density, then each vein's density and richness (the reference's dry run order,
    // mcgpu_terrain.cu:540-550)
    int r = root(W->root_density);
    g->density = r >= 0 ? g->plan(r, g->vfull) : -1;
    const int nveins = W->veins_on ? W->nveins : 0;
    for (int k = 0; k < nveins && g->fail.empty(); k++) {
        const int rd = root(W->vein[k].density_root), rr = root(W->vein[k].richness_root);
        g->vdens[k] = rd >= 0 ? g->plan(rd, g->vfull) : -1;
        g->vrich[k] = rr >= 0 ? g->plan(rr, g->vfull) : -1;
    }
    if (!g->fail.empty()) { plugin_fail("execution plan: %s", g->fail.c_str()); return nullptr; }
    g->density = g->resolve(g->density);
    g->vals[g->density].store = true;
    for (int k = 0; k < nveins; k++) {
        g->vdens[k] = g->resolve(g->vdens[k]);
        g->vrich[k] = g->resolve(g->vrich[k]);
        g->vals[g->vdens[k]].store = g->vals[g->vrich[k]].store = true;
    }
    // cache cells the CPU needs: every volume each listed cache was planned on (candidates, in plan order)
    size_t off = 0;
    for (int k = 0; k < W->ncache_ids; k++) {
        CellInfo ci;
        ci.cid = W->cache_ids[k];
        for (const auto &m : g->memo_order) {
            const McgNode &s = g->nodes[m.first[0]];
            if (s.k != MCGK_CACHE || s.cid != ci.cid) continue;
            const TVol v{m.first[1], m.first[2], m.first[3], m.first[4], m.first[5], m.first[6], m.first[7], m.first[8],
                         m.first[9], m.first[10], m.first[11]};
            if (std::find(ci.cand.begin(), ci.cand.end(), v) != ci.cand.end()) continue;
            ci.cand.push_back(v);
            ci.cap = std::max<size_t>(ci.cap, (size_t)v.size());
            g->vals[g->resolve(m.second)].store = true;
        }
        ci.off = off;
        off += ci.cap;
        g->cells.push_back(ci);
    }
    g->cell_floats = off;
    // the cache trace must cover the pre-density volume and the density's geometry
    {
        std::vector<std::pair<bool, TVol>> cs;
        std::vector<uint32_t> none;
        bool ok = true;
        if (W->pre_root >= 0) {
            const int pr = root(W->pre_root);
            ok = pr >= 0 && g->trace(0, false, none, cs,
                                     pr, TVol{W->pre_sx, W->pre_sy, W->pre_sz, W->pre_dx, W->pre_dy, W->pre_dz, W->pre_y0, 1,
                                              W->pre_xoff, 1, W->pre_zoff});
        }
        if (ok) ok = g->trace(0, false, none, cs, root(W->root_density), g->vfull);
        if (!ok) { plugin_fail("cache trace: volume geometry not covered"); return nullptr; }
    }
    if (!build_passes(*g)) { plugin_fail("execution plan: %s", g->fail.c_str()); return nullptr; }
    // batch size from free device memory (a quarter of it, at most 16 GiB), at most max_chunks
    size_t per_chunk = 0;
    for (const Val &x : g->vals)
        if (x.store && x.kind != VK_ALIAS) per_chunk += (size_t)x.v.size();
    const size_t nd = (size_t)256 * (size_t)W->nh;
    size_t chunk_bytes = per_chunk * 4 + nd * 2 + g->cell_floats * 4 + 64;
    size_t fr = 0, tot = 0;
    if (!hchk(hipMemGetInfo(&fr, &tot), "hipMemGetInfo")) return nullptr;
    const size_t budget = std::min<size_t>(fr / 4, (size_t)16 << 30);
    g->Bmax = (int)std::max<size_t>(1, std::min<size_t>((size_t)max_chunks, budget / chunk_bytes));
    // the kernels index a pass's (chunk, element) pairs in 32 bits
    size_t largest = nd;
    for (const Pass &P : g->passes) largest = std::max<size_t>(largest, (size_t)P.v.size());
    g->Bmax = (int)std::min<size_t>((size_t)g->Bmax, std::max<size_t>(1, ((size_t)1 << 31) / largest - 1));
    if (!emit_passes(*g)) { plugin_fail("execution plan: %s", g->fail.c_str()); return nullptr; }
    // device: program tables, pass instructions, arena, chunk origins, masks, vein patches, staging
    struct Part { const void *src; size_t bytes; size_t off; };
    Part parts[] = {{p->noise, (size_t)p->nnoise * sizeof(McgNoise), 0}, {p->layers, (size_t)p->nlayers * sizeof(McgLayer), 0},
                    {p->octs, (size_t)p->noct * sizeof(McgOct), 0},       {p->sp, (size_t)p->nsp * sizeof(McgSpline), 0},
                    {p->splf, (size_t)p->nsplf * sizeof(float), 0},       {p->spc, (size_t)p->nspc * sizeof(int32_t), 0},
                    {p->thr, (size_t)p->nthr * sizeof(double), 0}};
    size_t total = 0;
    for (auto &q : parts) { q.off = total; total += (q.bytes + 255) & ~(size_t)255; }
    // the pass plan after the tables, all in one host image and one copy
    size_t ptotal = total;
    for (Pass &P : g->passes) {
        P.ins_off = ptotal - total;
        ptotal += (P.ins.size() * sizeof(TIns) + 255) & ~(size_t)255;
        P.arr_off = ptotal - total;
        ptotal += (std::max<size_t>(P.arr.size(), 1) * 4 + 255) & ~(size_t)255;
    }
    if (!g->st && !hchk(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking), "hipStreamCreate")) return nullptr;
    if (!grow_buf(&g->d_prog, &g->kit.prog_cap, ptotal, "program tables")) return nullptr;
    {
        std::vector<char> image(ptotal);
        for (auto &q : parts)
            if (q.bytes) memcpy(image.data() + q.off, q.src, q.bytes);
        for (const Pass &P : g->passes) {
            memcpy(image.data() + total + P.ins_off, P.ins.data(), P.ins.size() * sizeof(TIns));
            if (!P.arr.empty()) memcpy(image.data() + total + P.arr_off, P.arr.data(), P.arr.size() * 4);
        }
        if (!hchk(hipMemcpy(g->d_prog, image.data(), ptotal, hipMemcpyHostToDevice), "program upload")) return nullptr;
    }
    g->d_plan = g->d_prog + total;
    g->prog = TProg{(const McgNoise *)(g->d_prog + parts[0].off), (const McgLayer *)(g->d_prog + parts[1].off),
                    (const McgOct *)(g->d_prog + parts[2].off),   (const McgSpline *)(g->d_prog + parts[3].off),
                    (const float *)(g->d_prog + parts[4].off),    (const int32_t *)(g->d_prog + parts[5].off),
                    (const double *)(g->d_prog + parts[6].off)};
    Kit &k = g->kit;
    if (!grow_buf(&g->d_arena, &k.arena_cap, g->arena_floats, "batch arena") ||
        !grow_buf(&g->d_cx0, &k.cx_cap, (size_t)g->Bmax, "chunk origins") ||
        !grow_buf(&g->d_cz0, &k.cz_cap, (size_t)g->Bmax, "chunk origins") ||
        !grow_buf(&g->d_flags, &k.flags_cap, g->splops.size() * (size_t)g->Bmax, "spline masks") ||
        !grow_buf(&g->d_stage, &k.stage_cap, (size_t)g->Bmax * std::max<size_t>(g->cell_floats, 1), "cell staging"))
        return nullptr;
    if (nveins > 0) {
        if (!grow_buf(&g->d_veins, &k.veins_cap, (size_t)g->Bmax * nd, "vein patches")) return nullptr;
        if (!point_program(W->gap, g->gapw, &g->d_gap, &k.gap_cap, &g->gap_lds)) return nullptr;
        for (int k = 0; k < nveins; k++) {
            const int gr = W->vein[k].gap_root;
            if (gr < 0 || gr >= W->gap->nroots || gr >= 6 || g->gapw.root_slot[gr] < 0) { plugin_fail("vein %d: gap root out of range", k); return nullptr; }
            g->gap_slot[k] = g->gapw.root_slot[gr];
        }
    }
    g->P = nullptr;                         // the caller's program is not kept
    return g.release();
}

bool terrain_batch(Terrain *g, int n, const int *cx, const int *cz, McgBeardFn beard_cb, void *beard_ud, float *out_dens,
                   uint16_t *out_veins, float *out_cells, signed char *out_which) {
    const McgTerrainWorld &W = g->T;
    std::vector<int32_t> hx(n), hz(n);
    for (int i = 0; i < n; i++) { hx[i] = cx[i] * 16; hz[i] = cz[i] * 16; }
    if (!hchk(hipMemcpyAsync(g->d_cx0, hx.data(), 4 * (size_t)n, hipMemcpyHostToDevice, g->st), "chunk upload") ||
        !hchk(hipMemcpyAsync(g->d_cz0, hz.data(), 4 * (size_t)n, hipMemcpyHostToDevice, g->st), "chunk upload"))
        return false;
    // Beardifier volumes: the callback's values where it has them, zero (the default) elsewhere
    for (int b : g->beards) {
        const Val &x = g->vals[b];
        const int nv = x.v.size();
        g->h_beard.assign((size_t)nv * n, 0.0f);
        for (int c = 0; c < n && beard_cb; c++) {
            const int vol[9] = {x.v.sx, x.v.sy, x.v.sz, x.v.xrel ? hx[c] + x.v.xoff : x.v.xoff, x.v.y0,
                                x.v.zrel ? hz[c] + x.v.zoff : x.v.zoff, x.v.dx, x.v.dy, x.v.dz};
            float *dst = g->h_beard.data() + (size_t)c * nv;
            if (!beard_cb(beard_ud, c, vol, dst)) std::fill(dst, dst + nv, 0.0f);
        }
        if (!hchk(hipMemcpyAsync(g->d_arena + x.off, g->h_beard.data(), 4 * g->h_beard.size(), hipMemcpyHostToDevice, g->st),
                  "Beardifier upload") ||
            !hchk(hipStreamSynchronize(g->st), "Beardifier upload"))
            return false;
    }
    const bool want_cells = out_cells && out_which && !g->cells.empty();
    if (want_cells && !g->splops.empty() &&
        !hchk(hipMemsetAsync(g->d_flags, 0, 4 * g->splops.size() * (size_t)g->Bmax, g->st), "mask clear"))
        return false;
    const TBatch B{g->d_cx0, g->d_cz0, n, g->d_arena, want_cells ? g->d_flags : nullptr, g->Bmax};
    for (const Pass &P : g->passes) {
        const TPass A{(const TIns *)(g->d_plan + P.ins_off), (const int32_t *)(g->d_plan + P.arr_off), (int32_t)P.ins.size(),
                      P.nslot, P.v, P.v.size(), P.feat, P.spd};
        const hipError_t e = P.fused ? launch_tinterp(P.fi, P.v, B, g->st) : launch_tpass(g->prog, A, B, g->st);
        if (!hchk(e, "pass launch")) return false;
    }
    const size_t nd = (size_t)256 * (size_t)W.nh;
    const int nveins = W.veins_on ? W.nveins : 0;
    if (nveins > 0 && out_veins) {
        TVeins V{};
        V.n = nveins;
        for (int k = 0; k < nveins; k++) {
            V.dens[k] = g->vals[g->vdens[k]].off;
            V.rich[k] = g->vals[g->vrich[k]].off;
            V.ore[k] = W.vein[k].ore; V.raw[k] = W.vein[k].raw; V.filler[k] = W.vein[k].filler;
            V.raw_chance[k] = W.vein[k].raw_chance;
            V.gap_slot[k] = g->gap_slot[k];
        }
        V.lo = W.ore_lo; V.hi = W.ore_hi; V.nmin = W.nmin; V.nh = W.nh;
        if (!hchk(launch_tveins(g->gapw, g->gap_lds, V, B, g->d_veins, g->st), "vein launch")) return false;
    }
    if (out_dens && !hchk(mcb::copy_to_host(out_dens, g->d_arena + g->vals[g->density].off, 4 * nd * (size_t)n, g->st),
                          "density copy"))
        return false;
    if (nveins > 0 && out_veins &&
        !hchk(mcb::copy_to_host(out_veins, g->d_veins, 2 * nd * (size_t)n, g->st), "vein copy"))
        return false;
    if (want_cells) {
        std::vector<uint32_t> masks(g->splops.size() * (size_t)g->Bmax, 0u);
        if (!masks.empty() &&
            !hchk(hipMemcpyAsync(masks.data(), g->d_flags, 4 * masks.size(), hipMemcpyDeviceToHost, g->st), "mask copy"))
            return false;
        if (!hchk(hipStreamSynchronize(g->st), "batch")) return false;
        std::vector<TGather> items;
        const int nc = (int)g->cells.size();
        for (int c = 0; c < n; c++) {
            std::vector<std::pair<bool, TVol>> cs;
            bool ok = true;
            if (W.pre_root >= 0)
                ok = g->trace(c, false, masks, cs, g->proots[W.pre_root],
                              TVol{W.pre_sx, W.pre_sy, W.pre_sz, W.pre_dx, W.pre_dy, W.pre_dz, W.pre_y0, 1, W.pre_xoff, 1, W.pre_zoff});
            const std::vector<std::pair<bool, TVol>> before = cs;
            if (ok) ok = g->trace(c, true, masks, cs, g->proots[W.root_density], g->vfull);
            for (int k = 0; k < nc; k++) {
                const CellInfo &ci = g->cells[k];
                signed char which = -1;
                if (!ok) {
                    which = -2;
                } else if (ci.cid >= 0 && ci.cid < (int)cs.size() && cs[ci.cid].first) {
                    const bool same = ci.cid < (int)before.size() && before[ci.cid].first && before[ci.cid].second == cs[ci.cid].second;
                    if (!same) {
                        which = -2;
                        for (size_t j = 0; j < ci.cand.size(); j++)
                            if (ci.cand[j] == cs[ci.cid].second) which = (signed char)j;
                    }
                }
                if (which >= 0) {
                    auto it = g->memo.end();
                    for (const auto &m : g->memo_order)
                        if (g->nodes[m.first[0]].k == MCGK_CACHE && g->nodes[m.first[0]].cid == ci.cid &&
                            m.first == vkey(m.first[0], ci.cand[which])) {
                            it = g->memo.find(m.first);
                            break;
                        }
                    if (it == g->memo.end()) which = -2;
                    else {
                        const Val &u = g->vals[g->resolve(it->second)];
                        const int cnt = u.v.size();
                        items.push_back(TGather{u.off + (int64_t)c * cnt, (int64_t)((size_t)c * g->cell_floats + ci.off), cnt, 0});
                    }
                }
                out_which[(size_t)c * nc + k] = which;
            }
        }
        if (!items.empty()) {
            if (items.size() > g->items_cap && !grow_buf(&g->d_items, &g->items_cap, items.size() * 2, "gather list")) return false;
            if (!hchk(hipMemsetAsync(g->d_stage, 0, 4 * (size_t)n * g->cell_floats, g->st), "staging clear") ||
                !hchk(hipMemcpyAsync(g->d_items, items.data(), items.size() * sizeof(TGather), hipMemcpyHostToDevice, g->st), "gather upload") ||
                !hchk(launch_tgather(g->d_items, (int)items.size(), g->d_arena, g->d_stage, g->st), "gather launch"))
                return false;
        } else if (!hchk(hipMemsetAsync(g->d_stage, 0, 4 * (size_t)n * g->cell_floats, g->st), "staging clear")) {
            return false;
        }
        if (!hchk(mcb::copy_to_host(out_cells, g->d_stage, 4 * (size_t)n * g->cell_floats, g->st), "cell copy"))
            return false;
    }
    return hchk(hipStreamSynchronize(g->st), "batch");
}

}  // namespace

extern "C" {

void *mcgpu_terrain_new(const McgTerrainWorld *T, int max_chunks, char *err, size_t errlen) {
    PluginGate gate(true);
    Terrain *g = gate.ready ? terrain_new(T, max_chunks) : nullptr;
    if (!g && err && errlen) snprintf(err, errlen, "%s", mcb_last_error());
    return g;
}

void mcgpu_terrain_free(void *h) {
    if (!h) return;
    PluginGate gate(true);
    delete (Terrain *)h;
}

int mcgpu_terrain_max_chunks(void *h) { return h ? ((Terrain *)h)->Bmax : 0; }

size_t mcgpu_terrain_cell_floats(void *h) { return h ? ((Terrain *)h)->cell_floats : 0; }

int mcgpu_terrain_cell_info(void *h, int k, int *cid, int *ncand, size_t *cap, size_t *off) {
    const Terrain *g = (const Terrain *)h;
    if (!g || k < 0 || k >= (int)g->cells.size()) return -1;
    const CellInfo &c = g->cells[k];
    *cid = c.cid;
    *ncand = (int)c.cand.size();
    *cap = c.cap;
    *off = c.off;
    return 0;
}

int mcgpu_terrain_cell_cand(void *h, int k, int j, int vol[9], int *xrel, int *zrel, int *n) {
    const Terrain *g = (const Terrain *)h;
    if (!g || k < 0 || k >= (int)g->cells.size() || j < 0 || j >= (int)g->cells[k].cand.size()) return -1;
    const TVol &v = g->cells[k].cand[j];
    const int o[9] = {v.sx, v.sy, v.sz, v.xoff, v.y0, v.zoff, v.dx, v.dy, v.dz};
    memcpy(vol, o, sizeof o);
    *xrel = v.xrel;
    *zrel = v.zrel;
    *n = v.size();
    return 0;
}

void *mcgpu_host_alloc(size_t bytes) {
    PluginGate gate(true);
    void *p = nullptr;
    if (!gate.ready) return nullptr;
    if (hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocPortable) != hipSuccess) {
        (void)hipGetLastError();
        plugin_fail("hipHostMalloc of %zu bytes failed", bytes);
        return nullptr;
    }
    return p;
}

void mcgpu_host_free(void *p) {
    if (!p) return;
    PluginGate gate(true);
    (void)hipHostFree(p);
}

int mcgpu_terrain_batch(void *h, int n, const int *cx, const int *cz, McgBeardFn beard_cb, void *beard_ud, float *out_dens,
                        unsigned short *out_veins, float *out_cells, signed char *out_which) {
    PluginGate gate(true);
    Terrain *g = (Terrain *)h;
    if (!g || n < 1 || n > g->Bmax || !cx || !cz) return plugin_fail("bad batch (n must be 1..%d)", g ? g->Bmax : 0) ? 0 : 1;
    if (!gate.ready) return 1;
    return terrain_batch(g, n, cx, cz, beard_cb, beard_ud, out_dens, out_veins, out_cells, out_which) ? 0 : 3;
}

// The density of n chunks left in device memory (no host copies, no veins or cells): *dens = chunk i's
// 16 * nh * 16 floats at *dens + i * 16*nh*16, valid until the next batch on the handle.
int mcb_terrain_batch_device(void *h, int n, const int *cx, const int *cz, const float **dens) {
    PluginGate gate(true);
    Terrain *g = (Terrain *)h;
    if (!g || n < 1 || n > g->Bmax || !cx || !cz || !dens) return plugin_fail("bad batch (n must be 1..%d)", g ? g->Bmax : 0) ? 0 : 1;
    if (!gate.ready) return 1;
    if (!terrain_batch(g, n, cx, cz, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr)) return 3;
    *dens = g->d_arena + g->vals[g->density].off;
    return 0;
}

}  // extern "C"
