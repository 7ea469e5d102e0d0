// mcg_plugin.cpp - host side of the libmcgen GPU plugin (include/mcgpu.h): device selection, errors,
// compilation of a McgBiomeWorld into device form, and the grid / points entry points.
//
// Reference interface: libmcgen/gpu/mcgpu_core.cu:1-70 (device, errors), mcgpu_biome.cu:167-280 (world,
// grid, points), bound by libmcgen/src/gpu_bridge.c:123-148.  Differences in how, not what:
//  * the program is validated completely (operand order, index ranges, spline depth, tree shape) so a
//    malformed description is an error instead of an out-of-range access on the device;
//  * node values live in slots assigned by liveness (mcg_plugin.h), so there is no fixed cap on the
//    number of nodes, only on the number of values alive at once (what fits in LDS);
//  * a grid evaluates each quart cell (End: each chunk) it needs once and then applies the zoom,
//    instead of evaluating a cell per output point; per point when the grid step makes that cheaper;
//  * device buffers of a world persist between calls and grow on demand.
#include <algorithm>
#include <map>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mcb_internal.h"
#include "mcg_plugin.h"
#include "mcg_terrain.h"

using namespace mcb::plug;

namespace {

// ---------------------------------------------------------------------------------------------------
// errors and the device the plugin works on (contract of include/mcgpu.h)
// ---------------------------------------------------------------------------------------------------
// The plugin reports through the engine's per-thread error slot: mcgpu_last_error() and
// mcb_last_error() read the same text.  Always returns false, so a failing step can `return
// plugin_error(...)`.
bool plugin_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    mcb::failv(MCB_E_DEVICE, fmt, ap);
    va_end(ap);
    return false;
}

// One device per process.  mcgpu_select binds it, once, before any work; the first entry point that
// needs a device without a selection binds device 0.  Binding is checked by a memory query on the
// device, which also makes the runtime create its context there, so a bad device fails at selection
// rather than inside the first grid.
class DeviceBinding {
public:
    int device() const { return dev_; }
    bool bind(int want) {
        const int n = mcb_device_count();
        if (n <= 0) return plugin_error("no HIP device");
        if (want < 0 || want >= n) return plugin_error("no device %d (this process sees %d)", want, n);
        size_t free_b = 0, total_b = 0;
        hipError_t e = hipSetDevice(want);
        if (e == hipSuccess) e = hipMemGetInfo(&free_b, &total_b);
        if (e != hipSuccess) return plugin_error("device %d unusable: %s", want, hipGetErrorString(e));
        dev_ = want;
        return true;
    }
    // make the bound device (binding device 0 if none is) current on the calling thread
    bool make_current() {
        if (dev_ < 0) return bind(0);
        const hipError_t e = hipSetDevice(dev_);
        return e == hipSuccess || plugin_error("device %d: %s", dev_, hipGetErrorString(e));
    }

private:
    int dev_ = -1;
};

DeviceBinding &binding() {
    static DeviceBinding b;
    return b;
}

// Every entry point runs inside a Gate: the process-wide lock (entries are serialised, as the
// reference's are) and, for entries that touch the device, the bound device made current.
std::mutex &gate_mutex() {
    static std::mutex m;
    return m;
}
struct Gate {
    std::unique_lock<std::mutex> lock;
    bool ready;
    explicit Gate(bool needs_device) : lock(gate_mutex()), ready(!needs_device || binding().make_current()) {}
};

#define HCHK(call)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            plugin_error("%s: %s", #call, hipGetErrorString(e_));                                     \
            return false;                                                                           \
        }                                                                                           \
    } while (0)

// ---------------------------------------------------------------------------------------------------
// program compilation
// ---------------------------------------------------------------------------------------------------
enum { OP_A = 1, OP_B = 2, OP_C = 4, OP_NS = 8, OP_OCT = 16, OP_SPL = 32, OP_ISEL = 64 };

// operands each kind reads (mcgpu_dev.cuh:307-373 and 390-435)
int uses_new(int k) {
    switch (k) {
    case MCGK_CONST: case MCGK_DIST: case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR:
    case MCGK_CTX_ALPHA: case MCGK_CTX_OFFSET: case MCGK_CTX_BEARD: return 0;
    case MCGK_NOISE: case MCGK_SHIFT_B: return OP_NS;
    case MCGK_NOISE_XZ: return OP_NS | OP_A | OP_C;
    case MCGK_NOISE_XYZ: return OP_NS | OP_A | OP_B | OP_C;
    case MCGK_END: return OP_OCT;
    case MCGK_ADD: case MCGK_SUB: case MCGK_MUL: case MCGK_DIV: case MCGK_MIN: case MCGK_MAX: return OP_A | OP_B;
    case MCGK_LERP: case MCGK_RANGE: case MCGK_ISEL1: return OP_A | OP_B | OP_C;
    case MCGK_LERP_CF: return OP_A | OP_C;
    case MCGK_LERP_CS: return OP_A | OP_B;
    case MCGK_SPLINE: return OP_SPL;
    case MCGK_ISEL: return OP_A | OP_ISEL;
    default: return OP_A;         // unary kinds, CACHE, BLEND_DENSITY, SLICE_*
    }
}
int uses_old(int k) {
    switch (k) {
    case MCGO_CONST: case MCGO_Y_GRADIENT: case MCGO_BLEND_ALPHA: case MCGO_ZERO: return 0;
    case MCGO_NOISE: case MCGO_SHIFT_A: case MCGO_SHIFT_B: case MCGO_SHIFT: return OP_NS;
    case MCGO_SHIFTED_NOISE: return OP_NS | OP_A | OP_B | OP_C;
    case MCGO_WEIRD_SCALED: return OP_NS | OP_A;
    case MCGO_END_ISLANDS: return OP_OCT;
    case MCGO_ADD: case MCGO_MUL: case MCGO_MIN: case MCGO_MAX: return OP_A | OP_B;
    case MCGO_RANGE: return OP_A | OP_B | OP_C;
    case MCGO_SPLINE: return OP_SPL;
    case MCGO_ISEL: return OP_A | OP_ISEL;
    default: return OP_A;         // unary kinds, CLAMP, CADD, CMUL, PASS
    }
}

struct Compiled {
    std::vector<PIns> ins;           // the program, then its y-dependent instructions again
    int nins = 0;
    std::vector<int32_t> arr;        // slot numbers
    int nslot = 0, spl_depth = 1;
    int root_slot[6] = {-1, -1, -1, -1, -1, -1};
};

bool spline_depth(const McgProg *p, int spi, int narr_node, int level, int *depth) {
    if (spi < 0 || spi >= p->nsp) { plugin_error("spline index %d out of range", spi); return false; }
    if (level > kSplineDepth) { plugin_error("spline nesting deeper than %d", kSplineDepth); return false; }
    *depth = std::max(*depth, level);
    const McgSpline &s = p->sp[spi];
    if (s.is_const) return true;
    if (s.n < 1 || s.loc0 < 0 || (int64_t)s.loc0 + 2 * (int64_t)s.n > p->nsplf || s.val0 < 0 ||
        (int64_t)s.val0 + s.n > p->nspc || s.coord < 0 || s.coord >= narr_node) {
        plugin_error("spline %d: points or coordinate out of range", spi);
        return false;
    }
    for (int i = 0; i < s.n; i++)
        if (!spline_depth(p, p->spc[s.val0 + i], narr_node, level + 1, depth)) return false;
    return true;
}

bool check_noise(const McgProg *p, int ns) {
    if (ns < 0 || ns >= p->nnoise) { plugin_error("noise index %d out of range", ns); return false; }
    const McgNoise &mn = p->noise[ns];
    auto range_ok = [&](int l0, int n, bool allow_empty) {
        if (n < 0 || l0 < 0 || (int64_t)l0 + n > p->nlayers) return false;
        for (int i = 0; i < n; i++) {
            const int o = p->layers[l0 + i].oct;
            if (o == -1 && allow_empty) continue;
            if (o < 0 || o >= p->noct) return false;
        }
        return true;
    };
    const bool ok = p->old ? (range_ok(mn.l0, mn.n0, true) && range_ok(mn.l1, mn.n1, true)) : range_ok(mn.l0, mn.n0, false);
    if (!ok) plugin_error("noise %d: layers or octaves out of range", ns);
    return ok;
}

// false when node s may read the block y (the y-reading kinds of eval_new / eval_old)
bool y_free(bool old, const McgNode &s) {
    if (!old) {
        switch (s.k) {
        case MCGK_NOISE: case MCGK_NOISE_XZ: case MCGK_NOISE_XYZ: return (s.cm & 2) != 0 || s.d1 == 0.0;
        case MCGK_DIST: return (s.cm & 2) != 0;
        case MCGK_GRAD_CLAMP: case MCGK_GRAD_REPEAT: case MCGK_GRAD_MIRROR: return s.i0 != 1 || (s.cm & 2) != 0;
        default: return true;
        }
    }
    switch (s.k) {
    case MCGO_NOISE: case MCGO_SHIFTED_NOISE: case MCGO_SHIFT: return s.d1 == 0.0;
    case MCGO_Y_GRADIENT: case MCGO_WEIRD_SCALED: return false;
    default: return true;
    }
}

bool compile_prog(const McgProg *p, Compiled &C) {
    const int n = p->nnodes;
    if (n < 0 || (n && !p->nodes) || p->nnoise < 0 || p->nlayers < 0 || p->noct < 0 || p->nsp < 0 || p->nsplf < 0 ||
        p->nspc < 0 || p->narr < 0 || p->nthr < 0 || (p->nnoise && !p->noise) || (p->nlayers && !p->layers) ||
        (p->noct && !p->octs) || (p->nsp && !p->sp) || (p->nsplf && !p->splf) || (p->nspc && !p->spc) ||
        (p->narr && !p->arr) || (p->nthr && !p->thr)) {
        plugin_error("malformed program (counts or arrays)");
        return false;
    }
    std::vector<int> last(n, -1);
    auto use = [&](int user, int j) -> bool {
        if (j < 0 || j >= user) { plugin_error("node %d reads node %d (must be an earlier node)", user, j); return false; }
        last[j] = std::max(last[j], user);
        return true;
    };
    for (int i = 0; i < n; i++) {
        const McgNode &s = p->nodes[i];
        if (p->old) {
            if (s.k < 0 || s.k >= MCGO__COUNT) { plugin_error("unknown node kind %d", s.k); return false; }
        } else {
            switch (s.k) {
            case MCGK_INTERP: case MCGK_FTS: case MCGK_LOG: case MCGK_POW_CB: case MCGK_POW_CE: case MCGK_POW:
            case MCGK_ROUND_INT: case MCGK_ROUND:
                plugin_error("node kind %d is not supported by the point evaluator", s.k);
                return false;
            default:
                if (s.k < 0 || s.k >= MCGK__COUNT) { plugin_error("unknown node kind %d", s.k); return false; }
            }
            if ((s.k == MCGK_GRAD_REPEAT || s.k == MCGK_GRAD_MIRROR) && s.i2 == 0) {
                plugin_error("node %d: gradient period 0", i);
                return false;
            }
        }
        const int u = p->old ? uses_old(s.k) : uses_new(s.k);
        if ((u & OP_A) && !use(i, s.a)) return false;
        if ((u & OP_B) && !use(i, s.b)) return false;
        if ((u & OP_C) && !use(i, s.c)) return false;
        if ((u & OP_NS) && !check_noise(p, s.ns)) return false;
        if ((u & OP_OCT) && (s.oct < 0 || s.oct >= p->noct)) { plugin_error("node %d: octave out of range", i); return false; }
        if (u & (OP_SPL | OP_ISEL)) {
            if (s.narr < 1 || s.arr0 < 0 || (int64_t)s.arr0 + s.narr > p->narr) { plugin_error("node %d: list out of range", i); return false; }
            for (int q = 0; q < s.narr; q++)
                if (!use(i, p->arr[s.arr0 + q])) return false;
        }
        if ((u & OP_ISEL) && (s.nthr < 0 || s.thr0 < 0 || (int64_t)s.thr0 + s.nthr > p->nthr)) {
            plugin_error("node %d: thresholds out of range", i);
            return false;
        }
        if ((u & OP_ISEL) && s.nthr >= s.narr) { plugin_error("node %d: more thresholds than choices", i); return false; }
        if ((u & OP_SPL) && !spline_depth(p, s.sp, s.narr, 1, &C.spl_depth)) return false;
    }
    for (int k = 0; k < 6; k++) {
        const int r = k < p->nroots ? p->root[k] : -1;
        if (r < -1 || r >= n) { plugin_error("root %d out of range", k); return false; }
        if (r >= 0) last[r] = n;                              // live to the end
    }
    // y-independent nodes: no read of the block y (or y fixed by the coordinate mask, or scaled by 0, which
    // gives +-0 and the same noise value), and every operand y-independent.  A lane that evaluates the
    // same (x, z) again at another y skips them; the ones a y-dependent node or a root reads keep their
    // values between evaluations in slots of their own.
    std::vector<char> yind(n, 0), keep(n, 0);
    auto operands = [&](int i, auto &&f) {
        const McgNode &s = p->nodes[i];
        const int u = p->old ? uses_old(s.k) : uses_new(s.k);
        if (u & OP_A) f(s.a);
        if (u & OP_B) f(s.b);
        if (u & OP_C) f(s.c);
        if (u & (OP_SPL | OP_ISEL))
            for (int q = 0; q < s.narr; q++) f(p->arr[s.arr0 + q]);
    };
    for (int i = 0; i < n; i++) {
        bool ind = y_free(p->old != 0, p->nodes[i]);
        operands(i, [&](int j) { ind = ind && yind[j]; });
        yind[i] = ind;
    }
    for (int i = 0; i < n; i++)
        if (!yind[i]) operands(i, [&](int j) { if (yind[j]) keep[j] = 1; });
    for (int k = 0; k < 6; k++) {
        const int r = k < p->nroots ? p->root[k] : -1;
        if (r >= 0 && yind[r]) keep[r] = 1;
    }
    // slots: the kept values first, one each; the rest by linear scan (a value's slot is free again
    // after its last reader)
    auto assign = [&](const std::vector<char> &kept, std::vector<int> &slot) {
        std::vector<int> free_slots;
        int ns = 0;
        slot.assign(n, -1);
        for (int j = 0; j < n; j++)
            if (kept[j]) slot[j] = ns++;
        std::vector<std::vector<int>> dies(n + 1);
        for (int j = 0; j < n; j++)
            if (!kept[j] && last[j] >= 0 && last[j] < n) dies[last[j]].push_back(j);
        for (int i = 0; i < n; i++) {
            for (int j : dies[i]) free_slots.push_back(slot[j]);  // read by node i: the kernel reads before it writes
            if (kept[i]) continue;
            std::sort(free_slots.begin(), free_slots.end(), std::greater<int>());
            free_slots.erase(std::unique(free_slots.begin(), free_slots.end()), free_slots.end());
            int d;
            if (!free_slots.empty()) { d = free_slots.back(); free_slots.pop_back(); }
            else d = ns++;
            slot[i] = d;
            if (last[i] < 0) free_slots.push_back(d);             // never read: overwritten by a later node
        }
        return ns;
    };
    // The kept slots cost LDS: they are used when they add at most two slots and stay within 16 (which fits
    // beside any table); otherwise the program is evaluated in full every time, as without them.
    std::vector<int> slot, slot_plain;
    int nslot = assign(keep, slot);
    const int nplain = assign(std::vector<char>(n, 0), slot_plain);
    if (nslot > nplain + 2 || nslot > 16) {
        slot.swap(slot_plain);
        nslot = nplain;
        std::fill(yind.begin(), yind.end(), 0);
    }
    C.nslot = std::max(nslot, 1);
    C.arr.assign(std::max(p->narr, 1), 0);
    C.ins.resize(n);
    for (int i = 0; i < n; i++) {
        const McgNode &s = p->nodes[i];
        const int u = p->old ? uses_old(s.k) : uses_new(s.k);
        PIns &q = C.ins[i];
        memset(&q, 0, sizeof q);
        q.op = s.k;
        q.dst = slot[i];
        q.a = (u & OP_A) ? slot[s.a] : 0;
        q.b = (u & OP_B) ? slot[s.b] : 0;
        q.c = (u & OP_C) ? slot[s.c] : 0;
        q.ns = s.ns; q.sp = s.sp; q.arr0 = s.arr0; q.narr = s.narr; q.thr0 = s.thr0; q.nthr = s.nthr;
        q.i0 = s.i0; q.i1 = s.i1; q.i2 = s.i2; q.i3 = s.i3;
        q.cm = (p->old ? 0 : (s.cm & 7)) | (yind[i] ? kCmYFree : 0); q.cx = s.cx; q.cy = s.cy; q.cz = s.cz; q.oct = s.oct;
        q.f0 = s.f0; q.f1 = s.f1; q.f2 = s.f2; q.f3 = s.f3;
        q.d0 = s.d0; q.d1 = s.d1;
        if (u & (OP_SPL | OP_ISEL))
            for (int k = 0; k < s.narr; k++) C.arr[s.arr0 + k] = slot[p->arr[s.arr0 + k]];
    }
    for (int k = 0; k < 6; k++) {
        const int r = k < p->nroots ? p->root[k] : -1;
        C.root_slot[k] = r >= 0 ? slot[r] : -1;
    }
    // after the program, copies of its y-dependent instructions: what a wave whose lanes all repeat their
    // x and z runs
    C.nins = n;
    for (int i = 0; i < n; i++)
        if (!yind[i]) C.ins.push_back(C.ins[i]);
    return true;
}

// BFS tree (children consecutive) -> pre-order with escape links, children in their stored order
bool convert_tree(const McgTreeNode *T, int n, int root, std::vector<PNode> &out) {
    if (n <= 0 || !T || root < 0 || root >= n) { plugin_error("malformed R-tree"); return false; }
    std::vector<char> seen(n, 0);
    std::vector<std::pair<int, int>> st;          // (node, next child) ; out index kept in order stack
    std::vector<int> outidx;
    out.clear();
    auto enter = [&](int v) -> bool {
        if (seen[v]) { plugin_error("R-tree node %d reached twice", v); return false; }
        seen[v] = 1;
        const McgTreeNode &m = T[v];
        if (m.count < 0 || (m.count > 0 && (m.first < 0 || (int64_t)m.first + m.count > n))) {
            plugin_error("R-tree node %d: children out of range", v);
            return false;
        }
        PNode p;
        memset(&p, 0, sizeof p);
        for (int k = 0; k < 6; k++) { p.lo[k] = m.lo[k]; p.hi[k] = m.hi[k]; }
        const int64_t above = -(int64_t)m.hi[6], below = (int64_t)m.lo[6];
        const int64_t x = above > 0 ? above : (below > 0 ? below : 0);
        p.d6 = (uint64_t)x * (uint64_t)x;
        p.leaf = m.count == 0;
        p.biome = m.biome;
        outidx.push_back((int)out.size());
        out.push_back(p);
        st.push_back({v, 0});
        return true;
    };
    if (!enter(root)) return false;
    while (!st.empty()) {
        auto &top = st.back();
        const McgTreeNode &m = T[top.first];
        if (top.second < m.count) {
            const int c = m.first + top.second++;
            if (!enter(c)) return false;
        } else {
            out[outidx.back()].escape = (int32_t)out.size();
            outidx.pop_back();
            st.pop_back();
        }
    }
    return true;
}

// Device resources a world can borrow instead of owning: a stream and grow-only buffers, kept per device
// for the process.  Worlds made for one call (chunk_biomes) borrow them; creating a stream costs about a
// millisecond, several times the chunk kernels themselves.
struct Loan {
    hipStream_t st = nullptr;
    char *d_data = nullptr;
    uint8_t *d_out = nullptr;
    int32_t *d_xyz = nullptr, *d_bad = nullptr;
    size_t data_cap = 0, out_cap = 0, xyz_cap = 0;
};

struct GBiome {
    Loan *loan = nullptr;                   // borrowed stream and buffers (returned on release), or owned
    PWorld W;
    bool newf = false, perm_lds = false;
    bool router = false;                    // the vanilla router runs through the engine's column code
    int end_oct = -1;                       // End: the islands octave of the End table (-1: none)
    int lds = 0;
    int source = 0;
    hipStream_t st = nullptr;
    char *d_data = nullptr;                 // program and tree, one allocation
    uint8_t *d_out = nullptr, *d_lat = nullptr;
    int32_t *d_xyz = nullptr, *d_bad = nullptr;
    float *d_etab = nullptr;                // End islands table (launch_end_table)
    size_t out_cap = 0, lat_cap = 0, xyz_cap = 0, etab_cap = 0;
    void release() {
        if (loan) {                         // hand the (possibly grown) buffers back
            if (st) (void)hipStreamSynchronize(st);
            loan->d_out = d_out; loan->out_cap = out_cap;
            loan->d_xyz = d_xyz; loan->xyz_cap = xyz_cap;
            (void)hipFree(d_lat);
            (void)hipFree(d_etab);
            st = nullptr; d_data = nullptr; d_out = d_lat = nullptr; d_xyz = d_bad = nullptr; d_etab = nullptr;
            return;
        }
        if (st) { (void)hipStreamSynchronize(st); (void)hipStreamDestroy(st); }
        (void)hipFree(d_data); (void)hipFree(d_out); (void)hipFree(d_lat); (void)hipFree(d_xyz); (void)hipFree(d_bad);
        (void)hipFree(d_etab);
        st = nullptr; d_data = nullptr; d_out = d_lat = nullptr; d_xyz = d_bad = nullptr; d_etab = nullptr;
    }
};

template <typename T>
bool grow(T **p, size_t *cap, size_t need) {
    if (need <= *cap) return true;
    size_t c = std::max<size_t>(need, *cap * 2);
    (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    HCHK(hipMalloc((void **)p, c * sizeof(T)));
    *cap = c;
    return true;
}

// host half of mcgpu_biome_new: validation, compilation, tree layout (no device call)
struct HostWorld {
    Compiled C;
    std::vector<PNode> tree;
    std::shared_ptr<const mcb::CellTable> cells;   // strict cell table, when the tree admits one
    const McgProg *p = nullptr;
    // the vanilla Overworld router: the engine's constants and seed layout (null: interpret)
    std::unique_ptr<McbWorld> vw;
    std::unique_ptr<McbSlot> vslot;
    int end_oct = -1;                               // End: the islands nodes' octave (-1: no End table)
};

// The cell table is built over the engine's compact node form, so it is used only when that form holds
// the tree exactly (16-bit bounds, 32-bit offset distance, < 65535 nodes) and the table stays small;
// otherwise every search walks the tree.  Built once per distinct tree in the process (~1 s for a
// vanilla Overworld tree).
std::shared_ptr<const mcb::CellTable> plugin_cells(const std::vector<PNode> &T) {
    if (T.size() >= 65535) return nullptr;
    std::vector<McbNode> t16(T.size());
    for (size_t i = 0; i < T.size(); i++) {
        const PNode &p = T[i];
        McbNode &m = t16[i];
        memset(&m, 0, sizeof m);
        for (int k = 0; k < 6; k++) {
            if (p.lo[k] < -32768 || p.lo[k] > 32767 || p.hi[k] < -32768 || p.hi[k] > 32767) return nullptr;
            m.lo[k] = (int16_t)p.lo[k];
            m.hi[k] = (int16_t)p.hi[k];
        }
        if (p.d6 > 0xFFFFFFFFull) return nullptr;
        m.d6 = (uint32_t)p.d6;
        m.escape = (uint16_t)p.escape;
        m.leaf = p.leaf ? (uint16_t)(0x8000 | (p.biome & 0xFF)) : 0;
    }
    std::string err;
    return mcb::cells_for_tree(t16, err, true, (double)(1 << 22));
}
bool prepare_world(const McgBiomeWorld *W, HostWorld &H) {
    if (W->source != MCG_BS_MULTI && W->source != MCG_BS_END && W->source != MCG_BS_FIXED) {
        plugin_error("unknown biome source %d", W->source);
        return false;
    }
    if (W->qh <= 0) { plugin_error("world height must be positive"); return false; }
    if (W->source == MCG_BS_FIXED) return true;
    const McgProg *p = W->prog;
    if (!p) { plugin_error("no program"); return false; }
    if ((p->old != 0) == (W->newf != 0)) { plugin_error("program branch (old=%d) does not match newf=%d", p->old, W->newf); return false; }
    if (!compile_prog(p, H.C)) return false;
    if (W->source == MCG_BS_END && H.C.root_slot[3] < 0) { plugin_error("no erosion field"); return false; }
    if (W->source == MCG_BS_MULTI)
        for (int k = 0; k < 6; k++)
            if (H.C.root_slot[k] < 0) { plugin_error("no climate field %d", k); return false; }
    // MCB_PLUGIN_INTERPRET=1: interpret every program with no End table (comparisons of the paths)
    const char *force = getenv("MCB_PLUGIN_INTERPRET");
    const bool interpret = force && force[0] == '1';
    if (W->source == MCG_BS_MULTI) {
        if (!convert_tree(W->tree, W->ntree, W->tree_root, H.tree)) return false;
        H.cells = plugin_cells(H.tree);
        if (!interpret) {
            H.vw.reset(new McbWorld());
            H.vslot.reset(new McbSlot());
            if (!vanilla_router(W, *H.vw, *H.vslot)) { H.vw.reset(); H.vslot.reset(); }
        }
    } else if (!interpret) {
        // the End table serves the islands nodes of the first one's octave (others compute their windows)
        for (int i = 0; i < p->nnodes && H.end_oct < 0; i++)
            if (p->nodes[i].k == (p->old ? MCGO_END_ISLANDS : MCGK_END) && p->nodes[i].oct >= 0 && p->nodes[i].oct < p->noct)
                H.end_oct = p->nodes[i].oct;
    }
    const int vbytes = H.C.nslot * kWG * (W->newf ? 4 : 8);
    if (grad_lds_bytes(W->newf != 0) + vbytes > kMaxLdsBytes) {
        plugin_error("program keeps %d values alive at once; at most %d fit", H.C.nslot,
                  (kMaxLdsBytes - grad_lds_bytes(W->newf != 0)) / (kWG * (W->newf ? 4 : 8)));
        return false;
    }
    H.p = p;
    return true;
}

// The tree and its cell table depend only on the tree, not on the seed: one device copy per (device,
// table) serves every world built over it for the rest of the process (a world per call, as the chunk
// path makes, would otherwise upload several megabytes each time).  Guarded by the entry gate.
struct SharedTables {
    char *d = nullptr;
    size_t ntree = 0;
    const PNode *tree = nullptr;
    const int32_t *cbp = nullptr;
    const uint32_t *centry = nullptr;
    const uint16_t *clist = nullptr;
    const uint32_t *clut = nullptr;      // null: the table has no locator
    int32_t cstride[6] = {0, 0, 0, 0, 0, 0};
    const PRec *crec = nullptr;
    const uint32_t *cent1 = nullptr;
};
// The locator of targets within 16 bits: per dimension 256 buckets of 256 values, entry = (breakpoints at or
// below the bucket's first value) * stride | (offset of the one breakpoint inside the bucket, 256: none) << 20,
// so a dimension's cell offset is one load and a compare.  Usable only when no bucket holds two breakpoints
// and every entry index fits 20 bits.  recs: the listed leaves copied in list order.
bool cell_locator(const mcb::CellTable &ct, const std::vector<PNode> &T, std::vector<uint32_t> &lut, int32_t stride[6],
                  std::vector<PRec> &recs, std::vector<uint32_t> &ent1) {
    if (ct.entry.size() > (1u << 20) || ct.list.size() >= (1u << 24)) return false;
    int n[6];
    for (int k = 0; k < 6; k++) n[k] = ct.nb[k] + 1;
    // index = (((((cT*nH + cH)*nC + cC)*nE + cE)*nW + cW)*nD + cD)
    stride[4] = 1;
    stride[5] = n[4];
    stride[3] = n[5] * stride[5];
    stride[2] = n[3] * stride[3];
    stride[1] = n[2] * stride[2];
    stride[0] = n[1] * stride[1];
    if ((size_t)n[0] * (size_t)stride[0] != ct.entry.size()) return false;
    lut.assign(6 * 256, 0);
    for (int k = 0; k < 6; k++) {
        const int32_t *bp = ct.bp.data() + ct.off[k];
        for (int b = 0; b < 256; b++) {
            const int t0 = b * 256 - 32768;
            int below = 0, inside = 0, off = 256;
            for (int i = 0; i < ct.nb[k]; i++) {
                if (bp[i] <= t0) below++;
                else if (bp[i] <= t0 + 255) {
                    inside++;
                    off = bp[i] - t0;
                }
            }
            if (inside > 1) return false;
            lut[k * 256 + b] = (uint32_t)below * (uint32_t)stride[k] | (uint32_t)off << 20;
        }
    }
    recs.resize(std::max<size_t>(ct.list.size(), 1));
    for (size_t i = 0; i < ct.list.size(); i++) {
        const PNode &p = T[ct.list[i]];
        PRec &r = recs[i];
        for (int k = 0; k < 6; k++) {
            r.lo[k] = (int16_t)p.lo[k];
            r.hi[k] = (int16_t)p.hi[k];
        }
        r.d6 = (uint32_t)p.d6;
        r.leaf = (int32_t)ct.list[i] | (p.biome & 0xFF) << 16;
    }
    ent1.resize(ct.entry.size());
    for (size_t i = 0; i < ct.entry.size(); i++) {
        const uint32_t e = ct.entry[i], cnt = e & 255u;
        ent1[i] = cnt == 1 ? 1u | (uint32_t)ct.list[e >> 8] << 8 | (uint32_t)(T[ct.list[e >> 8]].biome & 0xFF) << 24 : e;
    }
    return true;
}
// null with *failed false: not shared (the world uploads its own copy); *failed: a device error
const SharedTables *shared_tables(const HostWorld &H, bool *failed) {
    *failed = false;
    static std::map<std::pair<int, const mcb::CellTable *>, SharedTables> cache;
    const auto key = std::make_pair(binding().device(), H.cells.get());
    auto it = cache.find(key);
    if (it != cache.end()) return it->second.ntree == H.tree.size() ? &it->second : nullptr;
    const mcb::CellTable &ct = *H.cells;
    SharedTables s;
    std::vector<uint32_t> lut, ent1;
    std::vector<PRec> recs;
    const bool located = cell_locator(ct, H.tree, lut, s.cstride, recs, ent1);
    const size_t b[7] = {H.tree.size() * sizeof(PNode), ct.bp.size() * sizeof(int32_t), ct.entry.size() * sizeof(uint32_t),
                         ct.list.size() * sizeof(uint16_t), located ? lut.size() * sizeof(uint32_t) : 0,
                         located ? recs.size() * sizeof(PRec) : 0, located ? ent1.size() * sizeof(uint32_t) : 0};
    const void *src[7] = {H.tree.data(), ct.bp.data(), ct.entry.data(), ct.list.data(), lut.data(), recs.data(), ent1.data()};
    size_t off[7], total = 0;
    for (int i = 0; i < 7; i++) { off[i] = total; total += (b[i] + 255) & ~(size_t)255; }
    if (hipMalloc((void **)&s.d, std::max<size_t>(total, 256)) != hipSuccess) {
        plugin_error("shared table allocation failed");
        *failed = true;
        return nullptr;
    }
    for (int i = 0; i < 7; i++)
        if (b[i] && hipMemcpy(s.d + off[i], src[i], b[i], hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(s.d);
            plugin_error("shared table upload failed");
            *failed = true;
            return nullptr;
        }
    s.ntree = H.tree.size();
    s.tree = (const PNode *)(s.d + off[0]);
    s.cbp = (const int32_t *)(s.d + off[1]);
    s.centry = (const uint32_t *)(s.d + off[2]);
    s.clist = (const uint16_t *)(s.d + off[3]);
    if (located) {
        s.clut = (const uint32_t *)(s.d + off[4]);
        s.crec = (const PRec *)(s.d + off[5]);
        s.cent1 = (const uint32_t *)(s.d + off[6]);
    }
    return &(cache[key] = s);
}

// device half: buffers, stream, uploads
bool build_world(GBiome *g, const McgBiomeWorld *W, const HostWorld &H) {
    memset(&g->W, 0, sizeof g->W);
    PWorld &P = g->W;
    P.source = W->source;
    P.fixed_biome = W->fixed_biome;
    memcpy(P.end_ids, W->end_ids, sizeof P.end_ids);
    P.min_qy = W->min_qy;
    P.qh = W->qh;
    P.zoom_seed = W->zoom_seed;
    P.chain_cap = kChainCap;
    for (int k = 0; k < 6; k++) P.root_slot[k] = -1;
    g->source = W->source;
    g->newf = W->newf != 0;
    if (Loan *L = g->loan) {
        if (!L->st) HCHK(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
        if (!L->d_bad) HCHK(hipMalloc((void **)&L->d_bad, 4 * sizeof(int32_t)));
        g->st = L->st;
        g->d_bad = L->d_bad;
        g->d_out = L->d_out; g->out_cap = L->out_cap;
        g->d_xyz = L->d_xyz; g->xyz_cap = L->xyz_cap;
    } else {
        HCHK(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
        HCHK(hipMalloc((void **)&g->d_bad, 4 * sizeof(int32_t)));  // [0] unresolved count, [1..2] the router's list lengths
    }
    if (W->source == MCG_BS_FIXED) return true;
    const McgProg *p = H.p;
    const Compiled &C = H.C;
    // LDS: gradient table and value slots always; the permutation tables with copies of the noise
    // records, layer records and octave offsets when they fit beside them
    const int gbytes = grad_lds_bytes(g->newf), vbytes = C.nslot * kWG * (g->newf ? 4 : 8);
    auto al16 = [](int64_t v) { return (v + 15) & ~(int64_t)15; };
    const int64_t pbytes = (int64_t)p->noct * (kPermDoubled ? 512 : 256) + al16((int64_t)p->nnoise * (int64_t)sizeof(McgNoise)) +
                           al16((int64_t)p->nlayers * (int64_t)sizeof(McgLayer)) + al16((int64_t)p->noct * 24);
    g->perm_lds = gbytes + vbytes + pbytes <= kMaxLdsBytes;
    g->lds = gbytes + vbytes + (g->perm_lds ? (int)pbytes : 0);
    // one device allocation: ins, noise, layers, octs, sp, splf, spc, arr, thr, tree (the tree and cell
    // table come from the shared copy when the tree has a cell table)
    bool failed = false;
    const SharedTables *sh = H.cells ? shared_tables(H, &failed) : nullptr;
    if (failed) return false;
    struct Part { const void *src; size_t bytes; size_t off; };
    std::vector<Part> parts = {
        {C.ins.data(), C.ins.size() * sizeof(PIns), 0},          {p->noise, (size_t)p->nnoise * sizeof(McgNoise), 0},
        {p->layers, (size_t)p->nlayers * sizeof(McgLayer), 0},   {p->octs, (size_t)p->noct * sizeof(McgOct), 0},
        {p->sp, (size_t)p->nsp * sizeof(McgSpline), 0},          {p->splf, (size_t)p->nsplf * sizeof(float), 0},
        {p->spc, (size_t)p->nspc * sizeof(int32_t), 0},          {C.arr.data(), C.arr.size() * sizeof(int32_t), 0},
        {p->thr, (size_t)p->nthr * sizeof(double), 0},           {H.tree.data(), H.tree.size() * sizeof(PNode), 0},
        {H.cells ? H.cells->bp.data() : nullptr, H.cells ? H.cells->bp.size() * sizeof(int32_t) : 0, 0},
        {H.cells ? H.cells->entry.data() : nullptr, H.cells ? H.cells->entry.size() * sizeof(uint32_t) : 0, 0},
        {H.cells ? H.cells->list.data() : nullptr, H.cells ? H.cells->list.size() * sizeof(uint16_t) : 0, 0},
        {H.vw.get(), H.vw ? sizeof(McbWorld) : 0, 0},             {H.vslot.get(), H.vslot ? sizeof(McbSlot) : 0, 0}};
    size_t total = 0;
    if (sh)
        for (int i = 9; i < 13; i++) parts[i].bytes = 0;
    for (auto &q : parts) { q.off = total; total += (q.bytes + 255) & ~(size_t)255; }
    if (Loan *L = g->loan) {
        if (!grow(&L->d_data, &L->data_cap, std::max<size_t>(total, 256))) return false;
        g->d_data = L->d_data;
    } else {
        HCHK(hipMalloc((void **)&g->d_data, std::max<size_t>(total, 256)));
    }
    // the parts packed into one host image and uploaded with one copy (one synchronous copy per part cost
    // about 20 us each)
    std::vector<char> image(total);
    for (auto &q : parts)
        if (q.bytes) memcpy(image.data() + q.off, q.src, q.bytes);
    if (total) HCHK(hipMemcpy(g->d_data, image.data(), total, hipMemcpyHostToDevice));
    auto at = [&](int i) { return (const void *)(g->d_data + parts[i].off); };
    P.ins = (const PIns *)at(0);
    P.noise = (const McgNoise *)at(1);
    P.layers = (const McgLayer *)at(2);
    P.octs = (const McgOct *)at(3);
    P.sp = (const McgSpline *)at(4);
    P.splf = (const float *)at(5);
    P.spc = (const int32_t *)at(6);
    P.arr = (const int32_t *)at(7);
    P.thr = (const double *)at(8);
    P.tree = sh ? sh->tree : (const PNode *)at(9);
    if (H.cells) {
        P.cbp = sh ? sh->cbp : (const int32_t *)at(10);
        P.centry = sh ? sh->centry : (const uint32_t *)at(11);
        P.clist = sh ? sh->clist : (const uint16_t *)at(12);
        for (int k = 0; k < 6; k++) { P.cnb[k] = H.cells->nb[k]; P.coff[k] = H.cells->off[k]; }
        if (sh && sh->clut) {
            P.clut = sh->clut;
            P.crec = sh->crec;
            P.cent1 = sh->cent1;
            for (int k = 0; k < 6; k++) P.cstride[k] = sh->cstride[k];
        }
    }
    g->end_oct = H.end_oct;
    if (H.vw) {
        P.vw = (const McbWorld *)at(13);
        P.vslot = (const McbSlot *)at(14);
        HCHK(finish_router_slot((McbSlot *)at(14), g->st));
        HCHK(hipStreamSynchronize(g->st));
        g->router = true;
    }
    P.nins = (int32_t)C.nins;
    P.nydep = (int32_t)C.ins.size() - C.nins;
    P.nnoise = p->nnoise;
    P.nlayers = p->nlayers;
    P.noct = p->noct;
    P.ntree = (int32_t)H.tree.size();
    P.nslot = C.nslot;
    P.spl_depth = C.spl_depth;
    P.chain_cap = kChainCap;
    for (int k = 0; k < 6; k++) P.root_slot[k] = C.root_slot[k];
    return true;
}

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// The router launches' working buffers (column targets, cells to follow along their chains, their chain
// windows): one grow-only set per device for the process.  Entries are serialised by the entry lock and
// each call waits for its stream before returning, so calls never overlap on them.
struct VanScratch {
    int32_t *tgt = nullptr, *fix = nullptr;
    WinCell *win = nullptr;
    size_t tgt_cap = 0, fix_cap = 0, win_cap = 0;
};
VanScratch &van_scratch() {
    static std::map<int, VanScratch> per_device;
    return per_device[binding().device()];
}

// A lattice or points job of a world with the vanilla router through its launches (launch_van_cells);
// false when the world has no router or its cell table no locator, or the job's cells exceed 2^27 (the
// launches keep up to 44 bytes per cell) or its buffers cannot be had: the caller runs the per-lane kernel
// instead.
bool van_run(GBiome *g, const PJob &J) {
    const int64_t nc = van_cells(J);
    if (!g->router || !g->W.clut || !g->W.cent1 || nc > ((int64_t)1 << 27)) return false;
    VanScratch &S = van_scratch();
    const bool grown = grow(&S.tgt, &S.tgt_cap, (size_t)van_planes(J) * (size_t)nc) && grow(&S.fix, &S.fix_cap, 2 * (size_t)nc) &&
                       grow(&S.win, &S.win_cap, (size_t)kWinCells * kWin);
    hipError_t e = grown ? hipMemsetAsync(g->d_bad + 1, 0, 2 * sizeof(int32_t), g->st) : hipErrorOutOfMemory;
    if (e == hipSuccess) e = launch_van_cells(g->W, J, g->newf, S.tgt, S.fix, S.win, g->d_bad + 1, g->st);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;                          // the per-lane kernel runs the job instead
    }
    return true;
}

// The device world of an End job: with the islands table over chunks [cx0, cx0 + ncx) x [cz0, cz0 + ncz)
// (every chunk the job's items can evaluate) when the table costs less than the windows the items would
// compute (about 6 simplex cells per table chunk against up to 625 per item); else g->W as it is.
bool end_world(GBiome *g, int64_t cx0, int64_t cz0, int64_t ncx, int64_t ncz, int64_t items, PWorld &W) {
    W = g->W;
    if (g->source != MCG_BS_END || g->end_oct < 0 || ncx <= 0 || ncz <= 0) return true;
    const int64_t area = ncx * ncz;
    if (area > 64 * items + 65536 || area >= (1ll << 31) || cx0 < -(1ll << 28) || cz0 < -(1ll << 28) ||
        cx0 + ncx > (1ll << 28) || cz0 + ncz > (1ll << 28))
        return true;
    if (!grow(&g->d_etab, &g->etab_cap, (size_t)area)) return false;
    HCHK(launch_end_table(g->W, g->end_oct, (int)cx0, (int)cz0, (int)ncx, (int)ncz, !g->newf, g->d_etab, g->st));
    W.etab = g->d_etab;
    W.etab_oct = g->end_oct;
    W.etab_x0 = (int32_t)cx0; W.etab_z0 = (int32_t)cz0; W.etab_nx = (int32_t)ncx; W.etab_nz = (int32_t)ncz;
    return true;
}

bool run_grid(GBiome *g, int x0, int z0, int nx, int nz, int step, int y, unsigned char *out, int *nunres) {
    const int64_t n = (int64_t)nx * nz;
    if (!grow(&g->d_out, &g->out_cap, (size_t)n)) return false;
    HCHK(hipMemsetAsync(g->d_bad, 0, sizeof(int32_t), g->st));
    PJob J;
    memset(&J, 0, sizeof J);
    J.nbad = g->d_bad;
    // the zoomed quart cells lie in [(x0-2)>>2, ((xmax-2)>>2)+1] (mcgpu_biome.cu:215-218)
    const int64_t xmax = (int64_t)x0 + (int64_t)(nx - 1) * step, zmax = (int64_t)z0 + (int64_t)(nz - 1) * step;
    const int64_t qxa = (int64_t)(x0 - 2) >> 2, qxb = ((xmax - 2) >> 2) + 1, qza = (int64_t)(z0 - 2) >> 2, qzb = ((zmax - 2) >> 2) + 1;
    const int py = (y - 2) >> 2, lo = g->W.min_qy, hi = g->W.min_qy + g->W.qh - 1;
    int64_t ax0 = qxa, az0 = qza, nax = qxb - qxa + 1, naz = qzb - qza + 1;
    if (g->source == MCG_BS_END) { ax0 = qxa >> 2; az0 = qza >> 2; nax = (qxb >> 2) - ax0 + 1; naz = (qzb >> 2) - az0 + 1; }
    const int64_t nlat = 2 * nax * naz;
    // the lattice path indexes by cell, so every block coordinate must be formed without int overflow
    const bool fits = (int64_t)x0 - 2 >= INT32_MIN && (int64_t)z0 - 2 >= INT32_MIN && xmax <= INT32_MAX && zmax <= INT32_MAX;
    if (g->source != MCG_BS_FIXED && fits && nlat <= n && nlat < (1ll << 31)) {
        if (!grow(&g->d_lat, &g->lat_cap, (size_t)nlat)) return false;
        J.gen = GEN_LATTICE;
        J.n = nlat;
        J.q0x = (int32_t)ax0; J.q0z = (int32_t)az0; J.nqx = (int32_t)nax; J.nqz = (int32_t)naz;
        J.qys[0] = clampi(py, lo, hi);
        J.qys[1] = clampi(py + 1, lo, hi);
        J.out = g->d_lat;
        J.nbad = nullptr;                    // counted per output point by the zoom pass
        if (g->source == MCG_BS_END) {              // the lattice is the table's rectangle of chunks
            PWorld W;
            if (!end_world(g, ax0, az0, nax, naz, nlat, W)) return false;
            HCHK(launch_cells(W, J, g->newf, g->perm_lds, g->lds, g->st));
        } else if (!van_run(g, J)) {
            HCHK(launch_cells(g->W, J, g->newf, g->perm_lds, g->lds, g->st));
        }
        HCHK(launch_zoom_lattice(g->W, J, x0, z0, nx, nz, step, y, g->d_out, g->d_bad, g->st));
    } else {
        J.gen = GEN_GRIDPTS;
        J.n = n;
        J.x0 = x0; J.z0 = z0; J.nx = nx; J.step = step; J.y = y;
        J.out = g->d_out;
        if (g->source == MCG_BS_END) {              // the chunks of the zoomed quarts, as for the lattice
            PWorld W;
            if (!end_world(g, ax0, az0, nax, naz, n, W)) return false;
            HCHK(launch_cells(W, J, g->newf, g->perm_lds, g->lds, g->st));
        } else if (!van_run(g, J)) {
            HCHK(launch_cells(g->W, J, g->newf, g->perm_lds, g->lds, g->st));
        }
    }
    int32_t bad = 0;
    HCHK(hipMemcpyAsync(&bad, g->d_bad, sizeof bad, hipMemcpyDeviceToHost, g->st));
    HCHK(mcb::copy_to_host(out, g->d_out, (size_t)n, g->st));
    HCHK(hipStreamSynchronize(g->st));
    if (nunres) *nunres = bad;
    return true;
}

bool run_points(GBiome *g, int n, const int *xyz, unsigned char *out, int *nunres) {
    if (n == 0) {
        if (nunres) *nunres = 0;
        return true;
    }
    if (!grow(&g->d_out, &g->out_cap, (size_t)n) || !grow(&g->d_xyz, &g->xyz_cap, (size_t)n * 3)) return false;
    HCHK(hipMemsetAsync(g->d_bad, 0, sizeof(int32_t), g->st));
    HCHK(hipMemcpyAsync(g->d_xyz, xyz, (size_t)n * 3 * sizeof(int32_t), hipMemcpyHostToDevice, g->st));
    PJob J;
    memset(&J, 0, sizeof J);
    J.gen = GEN_POINTS;
    J.n = n;
    J.xyz = g->d_xyz;
    J.out = g->d_out;
    J.nbad = g->d_bad;
    if (g->source == MCG_BS_END) {
        // the chunks of the quarts the zoom can pick: [(x - 2) >> 2, ((x - 2) >> 2) + 1] >> 2
        int64_t xa = INT64_MAX, xb = INT64_MIN, za = INT64_MAX, zb = INT64_MIN;
        for (int i = 0; i < n; i++) {
            const int64_t qx = ((int64_t)xyz[3 * i] - 2) >> 2, qz = ((int64_t)xyz[3 * i + 2] - 2) >> 2;
            xa = std::min(xa, qx >> 2); xb = std::max(xb, (qx + 1) >> 2);
            za = std::min(za, qz >> 2); zb = std::max(zb, (qz + 1) >> 2);
        }
        PWorld W;
        if (!end_world(g, xa, za, xb - xa + 1, zb - za + 1, n, W)) return false;
        HCHK(launch_cells(W, J, g->newf, g->perm_lds, g->lds, g->st));
    } else if (!van_run(g, J)) {
        HCHK(launch_cells(g->W, J, g->newf, g->perm_lds, g->lds, g->st));
    }
    int32_t bad = 0;
    HCHK(hipMemcpyAsync(&bad, g->d_bad, sizeof bad, hipMemcpyDeviceToHost, g->st));
    HCHK(mcb::copy_to_host(out, g->d_out, (size_t)n, g->st));
    HCHK(hipStreamSynchronize(g->st));
    if (nunres) *nunres = bad;
    return true;
}

bool run_chunks(GBiome *g, int n, const int32_t *cxz, uint8_t *out, int *nunres) {
    const int cells = g->W.qh * 16;
    const int64_t items = (int64_t)n * cells;
    if (!grow(&g->d_out, &g->out_cap, (size_t)items) || !grow(&g->d_xyz, &g->xyz_cap, (size_t)n * 2)) return false;
    HCHK(hipMemsetAsync(g->d_bad, 0, sizeof(int32_t), g->st));
    HCHK(hipMemcpyAsync(g->d_xyz, cxz, (size_t)n * 2 * sizeof(int32_t), hipMemcpyHostToDevice, g->st));
    PJob J;
    memset(&J, 0, sizeof J);
    J.gen = GEN_CHUNKS;
    J.n = items;
    J.xyz = g->d_xyz;
    J.cells = cells;
    J.out = g->d_out;
    J.nbad = g->d_bad;
    if (g->source == MCG_BS_END && n > 0) {    // every cell of a chunk evaluates at its chunk's section
        int64_t xa = INT64_MAX, xb = INT64_MIN, za = INT64_MAX, zb = INT64_MIN;
        for (int i = 0; i < n; i++) {
            xa = std::min<int64_t>(xa, cxz[2 * i]); xb = std::max<int64_t>(xb, cxz[2 * i]);
            za = std::min<int64_t>(za, cxz[2 * i + 1]); zb = std::max<int64_t>(zb, cxz[2 * i + 1]);
        }
        PWorld W;
        if (!end_world(g, xa, za, xb - xa + 1, zb - za + 1, items, W)) return false;
        HCHK(launch_cells(W, J, g->newf, g->perm_lds, g->lds, g->st));
    } else if (!van_run(g, J)) {
        HCHK(launch_cells(g->W, J, g->newf, g->perm_lds, g->lds, g->st));
    }
    int32_t bad = 0;
    HCHK(hipMemcpyAsync(&bad, g->d_bad, sizeof bad, hipMemcpyDeviceToHost, g->st));
    HCHK(mcb::copy_to_host(out, g->d_out, (size_t)items, g->st));
    HCHK(hipStreamSynchronize(g->st));
    if (nunres) *nunres = bad;
    return true;
}

}  // namespace

void *mcb::plug::world_new(const McgBiomeWorld *W, int chain_cap, char *err, size_t errlen) {
    Gate gate(true);
    HostWorld H;
    std::unique_ptr<GBiome> g(new GBiome());
    if (!W) plugin_error("null world");
    else if (gate.ready && prepare_world(W, H) && build_world(g.get(), W, H)) {
        g->W.chain_cap = chain_cap;
        return g.release();
    }
    g->release();
    if (err && errlen) snprintf(err, errlen, "%s", mcb_last_error());
    return nullptr;
}

int mcb::plug::world_chunks(void *h, int n, const int32_t *cxz, uint8_t *out, int *nunres) {
    Gate gate(true);
    if (!h || n < 0 || (n && (!cxz || !out))) return plugin_error("bad arguments") ? 0 : 1;
    if (n == 0) {
        if (nunres) *nunres = 0;
        return 0;
    }
    return gate.ready && run_chunks((GBiome *)h, n, cxz, out, nunres) ? 0 : 1;
}

int mcb::plug::chunk_biomes(const McgBiomeWorld *W, int n, const int32_t *cxz, uint8_t *out, int *nunres, char *err,
                            size_t errlen) {
    Gate gate(true);
    bool ok = false;
    HostWorld H;
    std::unique_ptr<GBiome> g(new GBiome());
    static std::map<int, Loan> loans;         // guarded by the entry gate
    if (gate.ready) g->loan = &loans[binding().device()];
    if (n < 0 || (n && (!cxz || !out))) plugin_error("bad arguments");
    else if (gate.ready && prepare_world(W, H) && build_world(g.get(), W, H)) {
        g->W.chain_cap = 1 << 30;            // a chunk's fill order is finite: the walk ends at its first cell
        ok = n == 0 || run_chunks(g.get(), n, cxz, out, nunres);
        if (n == 0 && nunres) *nunres = 0;
    }
    g->release();
    if (!ok && err && errlen) snprintf(err, errlen, "%s", mcb_last_error());
    return ok ? 0 : 1;
}

mcb::plug::PluginGate::PluginGate(bool needs_device) : lock(gate_mutex()), ready(!needs_device || binding().make_current()) {}

bool mcb::plug::plugin_fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    mcb::failv(MCB_E_DEVICE, fmt, ap);
    va_end(ap);
    return false;
}

bool mcb::plug::point_program(const McgProg *p, PWorld &P, char **d_data, size_t *cap, int *lds) {
    memset(&P, 0, sizeof P);
    if (!p || p->old) return plugin_error("the point program must be a 26.3+ (float) program");
    Compiled C;
    if (!compile_prog(p, C)) return false;
    *lds = grad_lds_bytes(true) + C.nslot * kWG * 4;
    if (*lds > kMaxLdsBytes) return plugin_error("point program keeps %d values alive at once", C.nslot);
    struct Part { const void *src; size_t bytes; size_t off; };
    Part parts[] = {{C.ins.data(), C.ins.size() * sizeof(PIns), 0},        {p->noise, (size_t)p->nnoise * sizeof(McgNoise), 0},
                    {p->layers, (size_t)p->nlayers * sizeof(McgLayer), 0}, {p->octs, (size_t)p->noct * sizeof(McgOct), 0},
                    {p->sp, (size_t)p->nsp * sizeof(McgSpline), 0},        {p->splf, (size_t)p->nsplf * sizeof(float), 0},
                    {p->spc, (size_t)p->nspc * sizeof(int32_t), 0},        {C.arr.data(), C.arr.size() * sizeof(int32_t), 0},
                    {p->thr, (size_t)p->nthr * sizeof(double), 0}};
    size_t total = 0;
    for (auto &q : parts) { q.off = total; total += (q.bytes + 255) & ~(size_t)255; }
    total = std::max<size_t>(total, 256);
    if (!*d_data || *cap < total) {
        (void)hipFree(*d_data);
        *d_data = nullptr;
        *cap = 0;
        HCHK(hipMalloc((void **)d_data, total));
        *cap = total;
    }
    std::vector<char> image(total);
    for (auto &q : parts)
        if (q.bytes) memcpy(image.data() + q.off, q.src, q.bytes);
    HCHK(hipMemcpy(*d_data, image.data(), total, hipMemcpyHostToDevice));
    auto at = [&](int i) { return (const void *)(*d_data + parts[i].off); };
    P.ins = (const PIns *)at(0);
    P.noise = (const McgNoise *)at(1);
    P.layers = (const McgLayer *)at(2);
    P.octs = (const McgOct *)at(3);
    P.sp = (const McgSpline *)at(4);
    P.splf = (const float *)at(5);
    P.spc = (const int32_t *)at(6);
    P.arr = (const int32_t *)at(7);
    P.thr = (const double *)at(8);
    P.nins = C.nins;
    P.nydep = (int32_t)C.ins.size() - C.nins;
    P.nnoise = p->nnoise;
    P.nlayers = p->nlayers;
    P.noct = p->noct;
    P.nslot = C.nslot;
    P.spl_depth = C.spl_depth;
    for (int k = 0; k < 6; k++) P.root_slot[k] = C.root_slot[k];
    return true;
}

namespace {
size_t to_mib(size_t bytes) { return bytes / ((size_t)1 << 20); }
}  // namespace

extern "C" {

int mcgpu_abi(void) { return MCGPU_ABI; }

int mcgpu_device_count(void) { return mcb_device_count(); }

int mcgpu_device_info(int index, char *name, size_t namelen, int *cc_major, int *cc_minor, size_t *mem_mb) {
    if (index < 0 || index >= mcb_device_count()) return plugin_error("no device %d", index) ? 0 : 1;
    int major = 0, minor = 0;
    size_t bytes = 0;
    char label[256] = {0};
    if (hipDeviceGetAttribute(&major, hipDeviceAttributeComputeCapabilityMajor, index) != hipSuccess ||
        hipDeviceGetAttribute(&minor, hipDeviceAttributeComputeCapabilityMinor, index) != hipSuccess ||
        hipDeviceTotalMem(&bytes, index) != hipSuccess || hipDeviceGetName(label, (int)sizeof label - 1, index) != hipSuccess)
        return plugin_error("device %d: properties unavailable", index) ? 0 : 1;
    if (name && namelen) snprintf(name, namelen, "%s", label);
    if (cc_major) *cc_major = major;
    if (cc_minor) *cc_minor = minor;
    if (mem_mb) *mem_mb = to_mib(bytes);
    return 0;
}

int mcgpu_select(int device) {
    Gate gate(false);
    const int bound = binding().device();
    if (bound >= 0)
        return device == bound ? 0 : (plugin_error("device %d is bound; the selection is made once, before any work", bound) ? 0 : 1);
    return binding().bind(device) ? 0 : 1;
}

int mcgpu_selected(void) {
    Gate gate(false);
    return binding().device();
}

const char *mcgpu_last_error(void) { return mcb_last_error(); }

int mcgpu_mem_info(size_t *free_mb, size_t *total_mb) {
    Gate gate(true);
    if (!gate.ready) return 1;
    size_t free_b = 0, total_b = 0;
    const hipError_t e = hipMemGetInfo(&free_b, &total_b);
    if (e != hipSuccess) return plugin_error("memory query: %s", hipGetErrorString(e)) ? 0 : 1;
    if (free_mb) *free_mb = to_mib(free_b);
    if (total_mb) *total_mb = to_mib(total_b);
    return 0;
}

void *mcgpu_biome_new(const McgBiomeWorld *W, char *err, size_t errlen) {
    Gate gate(false);
    HostWorld H;
    std::unique_ptr<GBiome> g(new GBiome());
    // the description is checked on the host before anything touches the device
    if (!W) plugin_error("null world");
    else if (prepare_world(W, H) && binding().make_current() && build_world(g.get(), W, H)) return g.release();
    g->release();
    if (err && errlen) snprintf(err, errlen, "%s", mcb_last_error());
    return nullptr;
}

int mcb_plugin_router(void *h) {
    Gate gate(false);
    return h && ((GBiome *)h)->router ? 1 : 0;
}

int mcb_plugin_router_world(const McgBiomeWorld *W) {
    std::unique_ptr<McbWorld> w(new McbWorld());
    std::unique_ptr<McbSlot> slot(new McbSlot());
    return W && vanilla_router(W, *w, *slot) ? 1 : 0;
}

void mcgpu_biome_free(void *h) {
    if (!h) return;
    Gate gate(true);
    GBiome *g = (GBiome *)h;
    if (gate.ready) g->release();
    delete g;
}

int mcgpu_biome_grid(void *h, int x0, int z0, int nx, int nz, int step, int y, unsigned char *out, int *nunres) {
    Gate gate(true);
    if (!h || nx <= 0 || nz <= 0 || step <= 0 || !out) return plugin_error("bad arguments") ? 0 : 1;
    return gate.ready && run_grid((GBiome *)h, x0, z0, nx, nz, step, y, out, nunres) ? 0 : 1;
}

int mcgpu_biome_points(void *h, int n, const int *xyz, unsigned char *out, int *nunres) {
    Gate gate(true);
    if (!h || n < 0 || (n > 0 && (!xyz || !out))) return plugin_error("bad arguments") ? 0 : 1;
    return gate.ready && run_points((GBiome *)h, n, xyz, out, nunres) ? 0 : 1;
}

}  // extern "C"
